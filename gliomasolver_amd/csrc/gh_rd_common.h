// gh_rd_common.h -- per-cell arithmetic of the RD operator shared by the HIP
// kernels and host-side utilities (__host__ __device__).
//
// Reference: Glioma/Operators/ReactionDiffusionOperator.h.  Every expression
// keeps the reference's rounding points for Real = float and Real = double;
// translation units including this header are compiled with
// -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#pragma clang fp contract(off)

namespace gh {

// Per-cell constants of one solve, one 4-Real record per cell so a lane
// fetches them with one vector load: harmonic diffusivity of the cell's +x,
// +y, +z faces and the outside threshold tau.
template <typename Real>
struct alignas(4 * sizeof(Real)) Stat4 {
    Real dx, dy, dz, tau;
};

// ordered integer keys: okey(a) < okey(b) iff a < b for non-NaN a, b
// (except -0 < +0, harmless: the predicates below agree on both zeros)
__host__ __device__ inline int64_t okey(float x) {
    const int32_t b = __builtin_bit_cast(int32_t, x);
    return b >= 0 ? (int64_t)b : (int64_t)(b ^ 0x7fffffff);
}
__host__ __device__ inline float okey_inv(int64_t k, float) {
    const int32_t kk = (int32_t)k;
    return __builtin_bit_cast(float, kk >= 0 ? kk : (kk ^ 0x7fffffff));
}
__host__ __device__ inline int64_t okey(double x) {
    const int64_t b = __builtin_bit_cast(int64_t, x);
    return b >= 0 ? b : (b ^ 0x7fffffffffffffffLL);
}
__host__ __device__ inline double okey_inv(int64_t k, double) {
    return __builtin_bit_cast(double, k >= 0 ? k : (k ^ 0x7fffffffffffffffLL));
}
template <typename Real> __host__ __device__ inline Real real_max();
template <> __host__ __device__ inline float real_max<float>() { return FLT_MAX; }
template <> __host__ __device__ inline double real_max<double>() { return DBL_MAX; }

// df = p_w*Dw + p_g*Dg in Real (ReactionDiffusionOperator.h:105-111)
template <typename Real>
__host__ __device__ inline Real diffusivity(Real pw, Real pg, Real Dw, Real Dg) {
    const Real a = pw * Dw;
    const Real b = pg * Dg;
    return a + b;
}

// _harmonic_mean (ReactionDiffusionOperator.h:148-163) seen from a cell with
// df_loc = dc and neighbour dn: (dn+dc < 1e-8) ? 0 : 2.*dn*dc/(dn+dc), the
// quotient in double, the sum in Real.  (2*dn)*dc and (2*dc)*dn are the same
// double and dn+dc commutes, so both cells of a face get the same value.
template <typename Real>
__host__ __device__ inline Real harmonic(Real dn, Real dc) {
    const Real s = dn + dc;
    if (s < (Real)1.0e-08) return (Real)0.;
    return (Real)(((2. * (double)dn) * (double)dc) / (double)s);
}

// chf = phi + p_w + p_g (Real, left to right) < 0.1 marks an outside neighbour
// (ReactionDiffusionOperator.h:115-120, 166-171).  The predicate is monotone
// in phi, so it equals "phi < tau" with tau the smallest Real for which it is
// false; tau is found by bisection over the ordered bit patterns.
template <typename Real>
__host__ __device__ inline bool chf_outside(Real phi, Real pw, Real pg) {
    Real t = phi + pw;
    t = t + pg;
    return t < (Real)0.1;
}

template <typename Real>
__host__ __device__ inline Real chf_threshold(Real pw, Real pg) {
    const Real M = real_max<Real>();
    if (!chf_outside(-M, pw, pg)) return -(Real)INFINITY;
    if (chf_outside(M, pw, pg)) return (Real)INFINITY;
    int64_t lo = okey(-M), hi = okey(M);  // outside at lo, not outside at hi
    while ((uint64_t)hi - (uint64_t)lo > 1) {
        const int64_t mid = lo + (int64_t)(((uint64_t)hi - (uint64_t)lo) >> 1);
        if (chf_outside(okey_inv(mid, (Real)0), pw, pg)) lo = mid;
        else hi = mid;
    }
    return okey_inv(hi, (Real)0);
}

// p_w + p_g + phi > 0 (ReactionDiffusionOperator.h:103) <=> phi > -(p_w+p_g):
// an exact sum of two Reals is > 0 iff its rounding is > 0.
template <typename Real>
__host__ __device__ inline Real domain_threshold(Real pw, Real pg) {
    const Real s = pw + pg;
    return -s;
}

}  // namespace gh
