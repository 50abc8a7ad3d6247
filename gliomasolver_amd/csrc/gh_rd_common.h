// gh_rd_common.h -- per-cell arithmetic of the RD operator shared by the HIP
// kernels and host-side utilities (__host__ __device__).
//
// Reference: Glioma/Operators/ReactionDiffusionOperator.h.  Every expression
// keeps the reference's rounding points for Real = float and Real = double;
// translation units including this header are compiled with
// -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "gh_internal.h"

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

// LDS index of cell (x, y, z) in a 10^3 block tile with a one-cell halo
__device__ inline int P(int x, int y, int z) { return (z + 1) * 100 + (y + 1) * 10 + (x + 1); }

// dphidt of one cell (ReactionDiffusionOperator.h:103-138) from its six face
// coefficients d0..d5 (-x,+x,-y,+y,-z,+z) after the no-flux rule, the
// neighbour values p0..p5, its own value pc and whether it lies in the brain
// domain.
template <typename Real>
__device__ inline Real rd_rhs_core(Real d0, Real d1, Real d2, Real d3, Real d4, Real d5, Real p0, Real p1, Real p2,
                                   Real p3, Real p4, Real p5, Real pc, bool dom, Real rho, double ih2) {
    // diffusionFluxIn = ih2 * (Real sum of df_i*phi_i, left to right)
    Real sin_ = d0 * p0;
    sin_ = sin_ + d1 * p1;
    sin_ = sin_ + d2 * p2;
    sin_ = sin_ + d3 * p3;
    sin_ = sin_ + d4 * p4;
    sin_ = sin_ + d5 * p5;
    const double fin = ih2 * (double)sin_;
    // diffusionFluxOut = -((Real sum of df_i) * phi * ih2)
    Real sdf = d0 + d1;
    sdf = sdf + d2;
    sdf = sdf + d3;
    sdf = sdf + d4;
    sdf = sdf + d5;
    const Real sdp = sdf * pc;
    const double fout = -((double)sdp * ih2);
    // reactionFlux = (Real)(rho*phi) * (1. - phi)
    const Real rp = rho * pc;
    const double react = (double)rp * (1. - (double)pc);
    const Real dphidt = (Real)((fout + fin) + react);
    return dom ? dphidt : (Real)0.;  // outside the brain domain (:103, :137-138)
}

// dphidt of one cell from its face coefficients d0..d5 before the no-flux
// rule, the neighbour values p0..p5 with their outside flags o0..o5, its own
// value pc and domain threshold sc.
template <typename Real>
__device__ inline Real rd_cell_rhs(Real d0, Real d1, Real d2, Real d3, Real d4, Real d5, Real p0, Real p1, Real p2,
                                   Real p3, Real p4, Real p5, bool o0, bool o1, bool o2, bool o3, bool o4, bool o5,
                                   Real pc, Real sc, Real rho, double ih2) {
    // _applyNoFluxBC: an outside neighbour doubles the opposite face (exact)
    if (o0) d1 = (Real)((double)d1 * 2.0);
    if (o1) d0 = (Real)((double)d0 * 2.0);
    if (o2) d3 = (Real)((double)d3 * 2.0);
    if (o3) d2 = (Real)((double)d2 * 2.0);
    if (o4) d5 = (Real)((double)d5 * 2.0);
    if (o5) d4 = (Real)((double)d4 * 2.0);
    return rd_rhs_core<Real>(d0, d1, d2, d3, d4, d5, p0, p1, p2, p3, p4, p5, pc, pc > sc, rho, ih2);
}

// UpdateTumor (ReactionDiffusionOperator.h:219-221): phi += dt*dphidt in
// double, then std::max((Real)0, .), std::min((Real)1, .)
template <typename Real>
__device__ inline Real rd_update(Real pc, Real dphidt, double dt) {
    Real p = (Real)((double)pc + dt * (double)dphidt);
    p = ((Real)0. < p) ? p : (Real)0.;
    return (p < (Real)1.) ? p : (Real)1.;
}

// Position of a block in a uniform periodic grid of bpd^3 blocks whose slots
// start at slot0 (slot = slot0 + bx + bpd*(by + bpd*bz)): the step kernels
// find neighbour blocks by arithmetic on the workgroup-uniform block index,
// so no halo load waits on a table load.
struct BlockPos {
    int slot0, bpd, bx, by, bz;
    __device__ inline int nb(int dx, int dy, int dz) const {
        int x = bx + dx, y = by + dy, z = bz + dz;
        x += x < 0 ? bpd : 0;
        x -= x >= bpd ? bpd : 0;
        y += y < 0 ? bpd : 0;
        y -= y >= bpd ? bpd : 0;
        z += z < 0 ? bpd : 0;
        z -= z >= bpd ? bpd : 0;
        return slot0 + x + bpd * (y + bpd * z);
    }
};
// magic multiplier of the division by bpd in block_pos: m = ceil(2^32 / bpd),
// m * bpd = 2^32 + r with 0 <= r < bpd.  floor(l * m / 2^32) = floor(l / bpd)
// whenever l * r < 2^32: always for a power of two (r = 0), and for l < bpd^3
// with bpd <= 64 otherwise (bpd = 1 has l = 0 only).
__host__ __device__ inline uint32_t bpd_magic(int bpd) {
    return bpd <= 1 ? 0xffffffffu : (uint32_t)((0x100000000ull + (uint64_t)bpd - 1) / (uint64_t)bpd);
}
__device__ inline BlockPos block_pos(int s, int slot0, int bpd, uint32_t magic) {
    const uint32_t l = (uint32_t)(s - slot0);
    const uint32_t q = __umulhi(l, magic);  // l / bpd
    const uint32_t z = __umulhi(q, magic);  // l / bpd^2
    return BlockPos{slot0, bpd, (int)(l - q * (uint32_t)bpd), (int)(q - z * (uint32_t)bpd), (int)z};
}
// unit offset of face f (FXM, FXP, FYM, FYP, FZM, FZP)
__device__ inline int face_dx(int f) { return f == FXM ? -1 : (f == FXP ? 1 : 0); }
__device__ inline int face_dy(int f) { return f == FYM ? -1 : (f == FYP ? 1 : 0); }
__device__ inline int face_dz(int f) { return f == FZM ? -1 : (f == FZP ? 1 : 0); }
__device__ inline int face_nb(const BlockPos& p, int f) { return p.nb(face_dx(f), face_dy(f), face_dz(f)); }

// Step parameters of a slot: one solve (the same rho, 1/h^2, dt for every
// slot) or many independent solves packed slot range after slot range.
template <typename Real>
struct UniformStepParams {
    Real rho;
    double ih2, dt;
    int bpd;
    uint32_t magic;  // bpd_magic(bpd)
    __device__ inline void get(int, Real& r, double& i, double& d) const {
        r = rho;
        i = ih2;
        d = dt;
    }
    __device__ inline BlockPos pos(int s) const { return block_pos(s, 0, bpd, magic); }
};

// BPW blocks per workgroup of 512 lanes; lane = cell (x fastest) of each
// block.  All loads of a lane (BPW cells, one vector load of constants each)
// are issued before the first wait.  Group mapping is XCD-aware: consecutive
// workgroups land on different XCDs, so workgroup w serves group
// (w % 8) * per_xcd + w / 8 and each XCD sweeps a contiguous range of blocks
// whose halos stay in its own L2.
template <typename Real, bool FUSED, int BPW, typename Params>
__global__ __launch_bounds__(512) void rd_step_kernel(int nslots, int per_xcd,
                                                      const Real* __restrict__ phi, Real* __restrict__ out,
                                                      const Stat4<Real>* __restrict__ st,
                                                      const Real* __restrict__ sig, Params prm) {
    const int grp = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (grp * BPW >= nslots) return;
    __shared__ Real s_phi[BPW][1000];
    __shared__ unsigned char s_out[BPW][1000];
    __shared__ Real s_Dx[BPW][576], s_Dy[BPW][576], s_Dz[BPW][576];

    const int t = threadIdx.x;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    Real pc[BPW], sc[BPW];
    Stat4<Real> sv[BPW];
#pragma unroll
    for (int j = 0; j < BPW; ++j) {
        const int s = grp * BPW + j;
        if (s >= nslots) break;
        const size_t g = (size_t)s * BCELLS + t;
        pc[j] = phi[g];
        sv[j] = st[g];
        sc[j] = sig[g];
    }
    if (t < 384) {  // six face halos of 64 cells, one lane each
        const int f = t >> 6, u = t & 7, v = (t >> 3) & 7;
        int src, dst;
        switch (f) {
            case FXM: src = 7 + 8 * u + 64 * v; dst = P(-1, u, v); break;
            case FXP: src = 0 + 8 * u + 64 * v; dst = P(8, u, v); break;
            case FYM: src = u + 8 * 7 + 64 * v; dst = P(u, -1, v); break;
            case FYP: src = u + 8 * 0 + 64 * v; dst = P(u, 8, v); break;
            case FZM: src = u + 8 * v + 64 * 7; dst = P(u, v, -1); break;
            default:  src = u + 8 * v + 64 * 0; dst = P(u, v, 8); break;
        }
#pragma unroll
        for (int j = 0; j < BPW; ++j) {
            const int s = grp * BPW + j;
            if (s >= nslots) break;
            const size_t nb = (size_t)face_nb(prm.pos(s), f) * BCELLS + src;
            const Real ph = phi[nb];
            const Stat4<Real> hs = st[nb];
            s_phi[j][dst] = ph;
            s_out[j][dst] = ph < hs.tau;
            if (f == FXM) s_Dx[j][v * 72 + u * 9] = hs.dx;
            else if (f == FYM) s_Dy[j][v * 72 + u] = hs.dy;
            else if (f == FZM) s_Dz[j][v * 8 + u] = hs.dz;
        }
    }
#pragma unroll
    for (int j = 0; j < BPW; ++j) {
        if (grp * BPW + j >= nslots) break;
        s_phi[j][P(x, y, z)] = pc[j];
        s_out[j][P(x, y, z)] = pc[j] < sv[j].tau;
        s_Dx[j][z * 72 + y * 9 + x + 1] = sv[j].dx;
        s_Dy[j][z * 72 + (y + 1) * 8 + x] = sv[j].dy;
        s_Dz[j][(z + 1) * 64 + y * 8 + x] = sv[j].dz;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < BPW; ++j) {
        const int s = grp * BPW + j;
        if (s >= nslots) break;
        Real rho;
        double ih2, dt;
        prm.get(s, rho, ih2, dt);
        const Real d = rd_cell_rhs<Real>(
            s_Dx[j][z * 72 + y * 9 + x], sv[j].dx, s_Dy[j][z * 72 + y * 8 + x], sv[j].dy,
            s_Dz[j][z * 64 + y * 8 + x], sv[j].dz, s_phi[j][P(x - 1, y, z)], s_phi[j][P(x + 1, y, z)],
            s_phi[j][P(x, y - 1, z)], s_phi[j][P(x, y + 1, z)], s_phi[j][P(x, y, z - 1)], s_phi[j][P(x, y, z + 1)],
            s_out[j][P(x - 1, y, z)], s_out[j][P(x + 1, y, z)], s_out[j][P(x, y - 1, z)],
            s_out[j][P(x, y + 1, z)], s_out[j][P(x, y, z - 1)], s_out[j][P(x, y, z + 1)], pc[j], sc[j], rho, ih2);
        out[(size_t)s * BCELLS + t] = FUSED ? rd_update(pc[j], d, dt) : d;
    }
}


// Two fused Euler steps per launch (temporal blocking): one workgroup per 8^3
// block.  Step 1 is evaluated on the block and its six depth-1 face halos
// (the cells step 2 reads), from the state on the block, the depth-1 and
// depth-2 face halos and the twelve edge rows next to the depth-1 halos; step
// 2 then runs on the block from the step-1 values in LDS.  A halo cell's
// step-1 value is computed with exactly the inputs and arithmetic its own
// block uses, so two launches of rd_step_kernel and one of this kernel give
// the same bits.  Per two steps a block reads 1,376 states and constant
// records (512 + 864 halo) instead of 2 x 896.
__device__ inline int T12(int x, int y, int z) { return (z + 2) * 144 + (y + 2) * 12 + (x + 2); }  // [-2,10)^3

template <typename Real, typename Params>
__global__ __launch_bounds__(512) void rd_step2_kernel(int nslots, int per_xcd,
                                                       const Real* __restrict__ phi, Real* __restrict__ out,
                                                       const Stat4<Real>* __restrict__ st,
                                                       const Real* __restrict__ sig, Params prm) {
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    __shared__ Real s_u0[1728], s_tau[1728], s_dx[1728], s_dy[1728], s_dz[1728];
    __shared__ Real s_u1[1000];

    const int t = threadIdx.x;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    const size_t g = (size_t)s * BCELLS + t;
    const Real pc0 = phi[g];
    const Stat4<Real> sv = st[g];
    const Real sc = sig[g];
    const BlockPos bp = prm.pos(s);
    // halo cells: [0,384) depth-1 faces, [384,768) depth-2 faces, [768,864) edge rows
    Real hsig = 0;         // sigma of this lane's depth-1 face cell (lanes < 384)
    int fx = 0, fy = 0, fz = 0;
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const int h = t + 512 * it;
        if (h >= 864) break;
        int src_slot, lx, ly, lz, tx, ty, tz;
        if (h < 768) {
            const int d = h < 384 ? 1 : 2;
            const int r = h % 384, f = r >> 6, a = r & 7, b = (r >> 3) & 7;
            src_slot = face_nb(bp, f);
            switch (f) {
                case FXM: tx = -d; ty = a; tz = b; lx = 8 - d; ly = a; lz = b; break;
                case FXP: tx = 7 + d; ty = a; tz = b; lx = d - 1; ly = a; lz = b; break;
                case FYM: tx = a; ty = -d; tz = b; lx = a; ly = 8 - d; lz = b; break;
                case FYP: tx = a; ty = 7 + d; tz = b; lx = a; ly = d - 1; lz = b; break;
                case FZM: tx = a; ty = b; tz = -d; lx = a; ly = b; lz = 8 - d; break;
                default:  tx = a; ty = b; tz = 7 + d; lx = a; ly = b; lz = d - 1; break;
            }
        } else {
            const int e = (h - 768) >> 3, i = (h - 768) & 7;
            const int pair = e >> 2, s1 = e & 1, s2 = (e >> 1) & 1;
            const int c1 = s1 ? 8 : -1, c2 = s2 ? 8 : -1;  // tile coordinate on the two crossed axes
            const int l1 = s1 ? 0 : 7, l2 = s2 ? 0 : 7;    // local coordinate in the diagonal block
            int f1, f2;
            if (pair == 0) {         // x and y crossed, z free
                f1 = s1 ? FXP : FXM; f2 = s2 ? FYP : FYM;
                tx = c1; ty = c2; tz = i; lx = l1; ly = l2; lz = i;
            } else if (pair == 1) {  // x and z crossed, y free
                f1 = s1 ? FXP : FXM; f2 = s2 ? FZP : FZM;
                tx = c1; ty = i; tz = c2; lx = l1; ly = i; lz = l2;
            } else {                 // y and z crossed, x free
                f1 = s1 ? FYP : FYM; f2 = s2 ? FZP : FZM;
                tx = i; ty = c1; tz = c2; lx = i; ly = l1; lz = l2;
            }
            src_slot = bp.nb(face_dx(f1) + face_dx(f2), face_dy(f1) + face_dy(f2), face_dz(f1) + face_dz(f2));
        }
        const size_t gs = (size_t)src_slot * BCELLS + lx + 8 * ly + 64 * lz;
        const int q = T12(tx, ty, tz);
        const Stat4<Real> hs = st[gs];
        s_u0[q] = phi[gs];
        s_tau[q] = hs.tau;
        s_dx[q] = hs.dx;
        s_dy[q] = hs.dy;
        s_dz[q] = hs.dz;
        if (h < 384) {
            hsig = sig[gs];
            fx = tx; fy = ty; fz = tz;
        }
    }
    {
        const int q = T12(x, y, z);
        s_u0[q] = pc0;
        s_tau[q] = sv.tau;
        s_dx[q] = sv.dx;
        s_dy[q] = sv.dy;
        s_dz[q] = sv.dz;
    }
    Real rho;
    double ih2, dt;
    prm.get(s, rho, ih2, dt);
    __syncthreads();

    // step 1 on the block cell of this lane and (lanes < 384) one depth-1 face cell
    auto step1 = [&](int cx, int cy, int cz, Real sc1) {
        const int q = T12(cx, cy, cz);
        const int n0 = T12(cx - 1, cy, cz), n1 = T12(cx + 1, cy, cz), n2 = T12(cx, cy - 1, cz),
                  n3 = T12(cx, cy + 1, cz), n4 = T12(cx, cy, cz - 1), n5 = T12(cx, cy, cz + 1);
        const Real p0 = s_u0[n0], p1 = s_u0[n1], p2 = s_u0[n2], p3 = s_u0[n3], p4 = s_u0[n4], p5 = s_u0[n5];
        const Real pc = s_u0[q];
        const Real d = rd_cell_rhs<Real>(s_dx[n0], s_dx[q], s_dy[n2], s_dy[q], s_dz[n4], s_dz[q], p0, p1, p2, p3, p4,
                                         p5, p0 < s_tau[n0], p1 < s_tau[n1], p2 < s_tau[n2], p3 < s_tau[n3],
                                         p4 < s_tau[n4], p5 < s_tau[n5], pc, sc1, rho, ih2);
        s_u1[P(cx, cy, cz)] = rd_update(pc, d, dt);
    };
    step1(x, y, z, sc);
    if (t < 384) step1(fx, fy, fz, hsig);
    __syncthreads();

    // step 2 on the block
    {
        const int q = T12(x, y, z);
        const int n0 = T12(x - 1, y, z), n1 = T12(x + 1, y, z), n2 = T12(x, y - 1, z), n3 = T12(x, y + 1, z),
                  n4 = T12(x, y, z - 1), n5 = T12(x, y, z + 1);
        const Real p0 = s_u1[P(x - 1, y, z)], p1 = s_u1[P(x + 1, y, z)], p2 = s_u1[P(x, y - 1, z)],
                   p3 = s_u1[P(x, y + 1, z)], p4 = s_u1[P(x, y, z - 1)], p5 = s_u1[P(x, y, z + 1)];
        const Real pc = s_u1[P(x, y, z)];
        const Real d = rd_cell_rhs<Real>(s_dx[n0], s_dx[q], s_dy[n2], s_dy[q], s_dz[n4], s_dz[q], p0, p1, p2, p3, p4,
                                         p5, p0 < s_tau[n0], p1 < s_tau[n1], p2 < s_tau[n2], p3 < s_tau[n3],
                                         p4 < s_tau[n4], p5 < s_tau[n5], pc, sc, rho, ih2);
        out[g] = rd_update(pc, d, dt);
    }
}

// ---------------------------------------------------------------------------
// Two fused steps per launch on clamped states (rd_step2f_kernel).
//
// Preconditions, checked by the callers: every state value has a clear sign
// bit (true of any state written by the fused update, whose clamp returns +0
// for p <= 0 and for NaN), and every cell has p_w + p_g >= 0.  Then
//   * "phi < tau" equals "phi < max(tau, +0)" and "phi > sigma" equals
//     "sigma < 0 or phi > 0", so one record {D+x, D+y, D+z, tau_enc} per cell
//     (tau_enc = max(tau, +0), negated for cells with sigma < 0) carries all
//     per-solve constants (StatFast);
//   * a neighbour's outside flag travels in the sign bit of its value in LDS
//     (-phi when phi < tau), so a stencil reads six values instead of six
//     values and six thresholds, and the flag doubles a face by ldexp.
// The arithmetic after the no-flux doubling is rd_rhs_core, and doubling is
// exact, so the results are those of rd_step2_kernel bit for bit.
//
// LDS: s_uu holds the encoded step-0 state on [-2,10)^3 and, in the second
// half of each 24-value row, the encoded step-1 state on [-1,9)^3; rows of 24
// and planes of 293 values make the 8x8 lane patterns of a wave conflict-free
// for the block interior.  s_dmx/y/z[c] = D+x(c-x), D+y(c-y), D+z(c-z): the
// minus face coefficients of every cell evaluated in step 1.
// ---------------------------------------------------------------------------
__device__ inline int UU(int x, int y, int z) { return (x + 2) + 24 * (y + 2) + 293 * (z + 2); }  // [-2,10)^3
constexpr int UU1 = 12;  // offset of the step-1 state in a row
constexpr int UU_SIZE = 3516;
__device__ inline int DMI(int x, int y, int z) { return (x + 1) + 10 * (y + 1) + 100 * (z + 1); }  // [-1,9)^3

__device__ inline uint32_t sign_of(float v) { return __builtin_bit_cast(uint32_t, v) >> 31; }
__device__ inline uint32_t sign_of(double v) { return (uint32_t)(__builtin_bit_cast(uint64_t, v) >> 63); }
__device__ inline float twice_if(float d, uint32_t o) { return __builtin_amdgcn_ldexpf(d, (int)o); }
__device__ inline double twice_if(double d, uint32_t o) { return __builtin_amdgcn_ldexp(d, (int)o); }
// value with the outside flag in its sign bit (value >= +0)
template <typename Real>
__device__ inline Real flag_in_sign(Real u, Real tau) { return u < tau ? -u : u; }

// dphidt from encoded neighbours e0..e5 (-x,+x,-y,+y,-z,+z) and the face
// coefficients d0..d5 before the no-flux rule
template <typename Real>
__device__ inline Real rd_rhs_enc(Real d0, Real d1, Real d2, Real d3, Real d4, Real d5, Real e0, Real e1, Real e2,
                                  Real e3, Real e4, Real e5, Real pc, bool inside, Real rho, double ih2) {
    d1 = twice_if(d1, sign_of(e0));
    d0 = twice_if(d0, sign_of(e1));
    d3 = twice_if(d3, sign_of(e2));
    d2 = twice_if(d2, sign_of(e3));
    d5 = twice_if(d5, sign_of(e4));
    d4 = twice_if(d4, sign_of(e5));
    return rd_rhs_core<Real>(d0, d1, d2, d3, d4, d5, std::fabs(e0), std::fabs(e1), std::fabs(e2), std::fabs(e3),
                             std::fabs(e4), std::fabs(e5), pc, inside || pc > (Real)0., rho, ih2);
}

// {dx, dy, dz, tau} + sigma -> the record of rd_step2f_kernel; counts the
// cells that break the precondition p_w + p_g >= 0 (sigma <= 0)
template <typename Real>
__global__ __launch_bounds__(256) void stat_fast_kernel(size_t n, const Stat4<Real>* __restrict__ st,
                                                        const Real* __restrict__ sig, Stat4<Real>* __restrict__ out,
                                                        unsigned int* __restrict__ bad) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Stat4<Real> r = st[i];
    const Real sg = sig[i];
    const Real t = r.tau > (Real)0. ? r.tau : (Real)0.;
    r.tau = sg < (Real)0. ? -t : t;
    out[i] = r;
    if (!(sg <= (Real)0.)) atomicAdd(bad, 1u);
}

// Blocks without diffusion.  When every face coefficient a block's cells use
// is +0 -- no tissue in the block, p_w*Dw + p_g*Dg == 0 in every cell, makes
// each harmonic mean with such a cell +0 -- the operator's flux sums are exact
// zeros: sum df_i*phi_i = +0, fin = ih2*(+0) = +0, fout = -(+0*ih2) = -0 and
// fout + fin = +0 (ReactionDiffusionOperator.h:125-135), so dphidt reduces to
// (Real)(+0 + reaction) in the domain and 0 outside it.  zero_flux_kernel
// marks such blocks from the clamped-pair records: the three plus-face
// coefficients of all 512 cells and the minus-face coefficients across the
// block's -x, -y, -z faces (the neighbours' plus-face values) must all be +0.
template <typename Real, typename Params>
__global__ __launch_bounds__(256) void zero_flux_kernel(int nslots, const Stat4<Real>* __restrict__ stf, Params prm,
                                                        unsigned char* __restrict__ out) {
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const Stat4<Real>* r = stf + (size_t)s * BCELLS;
    bool nz = false;
    for (int c = threadIdx.x; c < BCELLS; c += 256)
        nz |= (r[c].dx != (Real)0 || sign_of(r[c].dx)) || (r[c].dy != (Real)0 || sign_of(r[c].dy)) ||
              (r[c].dz != (Real)0 || sign_of(r[c].dz));
    if (threadIdx.x < 64) {
        const BlockPos bp = prm.pos(s);
        const int a = threadIdx.x & 7, b = threadIdx.x >> 3;
        const Real mx = stf[(size_t)bp.nb(-1, 0, 0) * BCELLS + 7 + 8 * a + 64 * b].dx;
        const Real my = stf[(size_t)bp.nb(0, -1, 0) * BCELLS + a + 8 * 7 + 64 * b].dy;
        const Real mz = stf[(size_t)bp.nb(0, 0, -1) * BCELLS + a + 8 * b + 64 * 7].dz;
        nz |= (mx != (Real)0 || sign_of(mx)) || (my != (Real)0 || sign_of(my)) || (mz != (Real)0 || sign_of(mz));
    }
    nz = __syncthreads_or(nz);
    if (threadIdx.x == 0) out[s] = nz ? 0 : 1;
}

// dphidt of a cell of a block without diffusion (see zero_flux_kernel)
template <typename Real>
__device__ inline Real rd_rhs_reaction_only(Real pc, bool dom, Real rho) {
    const Real rp = rho * pc;
    const double react = (double)rp * (1. - (double)pc);
    const Real dphidt = (Real)(0.0 + react);
    return dom ? dphidt : (Real)0.;
}

// Relative cost of a block without diffusion in rd_step2f_kernel (its own
// 24 B per cell, no halo, no LDS) against a full block (about 2.75 stencil
// evaluations per cell): the weight the XCD split gives it.
constexpr double kZeroFluxCost = 0.3;

// Work split of the two-step kernel over the eight XCDs: workgroup w runs on
// XCD w % 8 and serves block start[w % 8] + w / 8 while that is below
// start[w % 8 + 1]; contiguous slot ranges keep halos in one L2, and the split
// points balance the ranges' work when blocks differ in cost.
struct XcdSplit {
    int start[9];
};
// equal contiguous ranges of ceil(n / 8) blocks (grid: 8 * ceil(n / 8))
inline XcdSplit even_split(int n) {
    XcdSplit x{};
    const int per = (n + 7) / 8;
    for (int i = 0; i <= 8; ++i) x.start[i] = std::min(i * per, n);
    return x;
}

// Where cell (a, b) of face f at depth d (1 or 2) outside the block sits:
// LDS offsets uu (state tile) and dm (coefficient tile) and the cell index in
// the neighbour block across the face, each base + a * step_a + b * step_b
// with wave-uniform bases and steps.
struct FaceMap {
    int uu, uua, uub;
    int dm, dma, dmb;
    int lo, loa, lob;
};
__device__ inline FaceMap face_map(int f, int d) {
    const int u0 = UU(0, 0, 0), m0 = DMI(0, 0, 0);
    switch (f) {
        case FXM: return {u0 - d, 24, 293, m0 - d, 10, 100, 8 - d, 8, 64};
        case FXP: return {u0 + 7 + d, 24, 293, m0 + 7 + d, 10, 100, d - 1, 8, 64};
        case FYM: return {u0 - 24 * d, 1, 293, m0 - 10 * d, 1, 100, 8 * (8 - d), 1, 64};
        case FYP: return {u0 + 24 * (7 + d), 1, 293, m0 + 10 * (7 + d), 1, 100, 8 * (d - 1), 1, 64};
        case FZM: return {u0 - 293 * d, 1, 24, m0 - 100 * d, 1, 10, 64 * (8 - d), 1, 8};
        default:  return {u0 + 293 * (7 + d), 1, 24, m0 + 100 * (7 + d), 1, 10, 64 * (d - 1), 1, 8};
    }
}
__device__ inline int mad24(int a, int b, int c) { return (int)__umul24((unsigned)a, (unsigned)b) + c; }

// Work items: with items (may be null), workgroup item i is {slot, k}: k = 0
// a full block, k >= 1 the k consecutive blocks from slot on, all without
// diffusion (their two steps are each cell's reaction update alone: same
// bits, no halo).  Without items, item i is slot i, and zflux (may be null)
// marks the blocks without diffusion (k = 1).
constexpr int ZF_GROUP = 8;  // blocks without diffusion per workgroup
template <typename Real, typename Params>
__global__ __launch_bounds__(512) void rd_step2f_kernel(XcdSplit split, const int2* __restrict__ items,
                                                        const Real* __restrict__ phi, Real* __restrict__ out,
                                                        Real* __restrict__ mid, const Stat4<Real>* __restrict__ stf,
                                                        const unsigned char* __restrict__ zflux, Params prm) {
    const int x = (int)(blockIdx.x % 8);
    const int it = split.start[x] + (int)(blockIdx.x / 8);
    if (it >= split.start[x + 1]) return;
    int s = it, nzf = 0;
    if (items) {
        const int2 v = items[it];
        s = v.x;
        nzf = v.y;
    } else if (zflux && zflux[s]) {
        nzf = 1;
    }
    __shared__ Real s_uu[UU_SIZE];
    __shared__ Real s_dmx[1000], s_dmy[1000], s_dmz[1000];

    const int t = threadIdx.x;
    if (nzf > 0) {  // uniform per workgroup: nzf blocks, all loads issued before the updates
        Real u0[ZF_GROUP];
        bool in0[ZF_GROUP];
#pragma unroll
        for (int k = 0; k < ZF_GROUP; ++k)
            if (k < nzf) {
                const size_t bc = (size_t)(s + k) * BCELLS + (unsigned)t;
                u0[k] = phi[bc];
                in0[k] = sign_of(stf[bc].tau) != 0;
            }
#pragma unroll
        for (int k = 0; k < ZF_GROUP; ++k)
            if (k < nzf) {
                const size_t bc = (size_t)(s + k) * BCELLS + (unsigned)t;
                Real rho0;
                double ih20, dt0;
                prm.get(s + k, rho0, ih20, dt0);
                const Real u1 = rd_update(u0[k], rd_rhs_reaction_only(u0[k], in0[k] || u0[k] > (Real)0., rho0), dt0);
                if (mid) mid[bc] = u1;
                out[bc] = rd_update(u1, rd_rhs_reaction_only(u1, in0[k] || u1 > (Real)0., rho0), dt0);
            }
        return;
    }
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);  // wave = interior plane z; halo role
    const int a = t & 7, b = (t >> 3) & 7;
    const BlockPos bp = prm.pos(s);
    const size_t base = (size_t)s * BCELLS;
    const unsigned ut = (unsigned)t;

    // ---- loads: (waves 0-5) the depth-1 and depth-2 cells of face w, (waves
    // 6-7) the 96 edge cells next to the depth-1 faces, then the own cell; all
    // offsets are computed first so the loads issue back to back; block base
    // pointers are wave-uniform, lane offsets 32-bit
    Real u_1 = 0, u_2 = 0, u_e = 0;  // depth-1, depth-2 and edge cell
    Stat4<Real> st_1, st_2, st_e;
    int uu1 = 0, dm1 = 0, uu2 = 0, dm2 = 0;
    int ex = 0, ey = 0, ez = 0;  // edge cell: tile coordinates
    bool edge = false;
    if (w < 6) {
        const FaceMap m1 = face_map(w, 1), m2 = face_map(w, 2);
        const size_t nb = (size_t)face_nb(bp, w) * BCELLS;
        const unsigned l1 = (unsigned)mad24(b, m1.lob, mad24(a, m1.loa, m1.lo)) & (BCELLS - 1);
        const unsigned l2 = (l1 + (unsigned)(m2.lo - m1.lo)) & (BCELLS - 1);
        uu1 = mad24(b, m1.uub, mad24(a, m1.uua, m1.uu));
        dm1 = mad24(b, m1.dmb, mad24(a, m1.dma, m1.dm));
        uu2 = uu1 + (m2.uu - m1.uu);
        dm2 = dm1 + (m2.dm - m1.dm);
        const Real* pn = phi + nb;
        const Stat4<Real>* sn = stf + nb;
        u_1 = pn[l1];
        st_1 = sn[l1];
        u_2 = pn[l2];
        st_2 = sn[l2];
    } else {
        const int e = (w - 6) * 64 + (t & 63);
        edge = e < 96;
        if (edge) {
            const int k = e >> 3, i = e & 7;
            const int pair = k >> 2, s1 = k & 1, s2 = (k >> 1) & 1;
            const int c1 = s1 ? 8 : -1, c2 = s2 ? 8 : -1;  // tile coordinate on the two crossed axes
            const int m1 = s1 ? 0 : 7, m2 = s2 ? 0 : 7;    // coordinate in the diagonal block
            const int d1 = s1 ? 1 : -1, d2 = s2 ? 1 : -1;
            int sl, lc;
            if (pair == 0) {
                ex = c1; ey = c2; ez = i; lc = m1 + 8 * m2 + 64 * i; sl = bp.nb(d1, d2, 0);
            } else if (pair == 1) {
                ex = c1; ey = i; ez = c2; lc = m1 + 8 * i + 64 * m2; sl = bp.nb(d1, 0, d2);
            } else {
                ex = i; ey = c1; ez = c2; lc = i + 8 * m1 + 64 * m2; sl = bp.nb(0, d1, d2);
            }
            const size_t gs = (size_t)sl * BCELLS + lc;
            u_e = phi[gs];
            st_e = stf[gs];
        }
    }
    const Real u_c = (phi + base)[ut];
    const Stat4<Real> st_c = (stf + base)[ut];
    Real rho;
    double ih2, dt;
    prm.get(s, rho, ih2, dt);

    // ---- LDS: encoded step-0 state, minus-face coefficients of the step-1 cells
    const int uuc = mad24(w, 293, mad24(b, 24, a + UU(0, 0, 0)));
    const int dmc = mad24(w, 100, mad24(b, 10, a + DMI(0, 0, 0)));
    const Real tau_c = std::fabs(st_c.tau);
    s_uu[uuc] = flag_in_sign(u_c, tau_c);
    s_dmx[dmc + 1] = st_c.dx;
    s_dmy[dmc + 10] = st_c.dy;
    s_dmz[dmc + 100] = st_c.dz;
    Real tau_1 = 0;
    if (w < 6) {
        tau_1 = std::fabs(st_1.tau);
        s_uu[uu1] = flag_in_sign(u_1, tau_1);
        s_uu[uu2] = flag_in_sign(u_2, std::fabs(st_2.tau));
        // depth-1 cell: every component whose target stays in [-1,9)^3;
        // depth-2 cell: only the normal component of a minus face
        if (w != FXP) s_dmx[dm1 + 1] = st_1.dx;
        if (w != FYP) s_dmy[dm1 + 10] = st_1.dy;
        if (w != FZP) s_dmz[dm1 + 100] = st_1.dz;
        if (w == FXM) s_dmx[dm2 + 1] = st_2.dx;
        if (w == FYM) s_dmy[dm2 + 10] = st_2.dy;
        if (w == FZM) s_dmz[dm2 + 100] = st_2.dz;
    } else if (edge) {
        s_uu[UU(ex, ey, ez)] = flag_in_sign(u_e, std::fabs(st_e.tau));
        // an edge cell feeds the minus face of a depth-1 cell across each of
        // its crossed axes at -1
        if (ex == -1) s_dmx[DMI(0, ey, ez)] = st_e.dx;
        if (ey == -1) s_dmy[DMI(ex, 0, ez)] = st_e.dy;
        if (ez == -1) s_dmz[DMI(ex, ey, 0)] = st_e.dz;
    }
    __syncthreads();

    // ---- step 1 on the interior cell and (waves 0-5) the depth-1 face cell;
    // the step-1 half of a row (offsets 13..22) is disjoint from the step-0
    // half (offsets 0..11), so the writes race with no read
    const bool in_c = sign_of(st_c.tau) != 0;
    const Real dmx_c = s_dmx[dmc], dmy_c = s_dmy[dmc], dmz_c = s_dmz[dmc];
    Real u1c;
    {
        const int q = uuc;
        const Real d = rd_rhs_enc<Real>(dmx_c, st_c.dx, dmy_c, st_c.dy, dmz_c, st_c.dz, s_uu[q - 1], s_uu[q + 1],
                                        s_uu[q - 24], s_uu[q + 24], s_uu[q - 293], s_uu[q + 293], u_c, in_c, rho, ih2);
        u1c = rd_update(u_c, d, dt);
        s_uu[q + UU1] = flag_in_sign(u1c, tau_c);
        if (mid) (mid + base)[ut] = u1c;  // the state after the first step, when asked for
    }
    if (w < 6) {
        const int q = uu1;
        const Real d = rd_rhs_enc<Real>(s_dmx[dm1], st_1.dx, s_dmy[dm1], st_1.dy, s_dmz[dm1], st_1.dz, s_uu[q - 1], s_uu[q + 1],
                                        s_uu[q - 24], s_uu[q + 24], s_uu[q - 293], s_uu[q + 293], u_1,
                                        sign_of(st_1.tau) != 0, rho, ih2);
        s_uu[q + UU1] = flag_in_sign(rd_update(u_1, d, dt), tau_1);
    }
    __syncthreads();

    // ---- step 2 on the interior cell
    {
        const int q = uuc + UU1;
        const Real d = rd_rhs_enc<Real>(dmx_c, st_c.dx, dmy_c, st_c.dy, dmz_c, st_c.dz, s_uu[q - 1], s_uu[q + 1],
                                        s_uu[q - 24], s_uu[q + 24], s_uu[q - 293], s_uu[q + 293], u1c, in_c, rho, ih2);
        (out + base)[ut] = rd_update(u1c, d, dt);
    }
}

}  // namespace gh
