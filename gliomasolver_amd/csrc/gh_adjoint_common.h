// gh_adjoint_common.h -- reverse-step kernels of the discrete adjoint, shared
// by the single-solve adjoint (gh_adjoint.hip) and the batched one
// (gh_batch.hip).  The model and the per-cell formulas are documented at the
// top of gh_adjoint.hip.
#pragma once
#include "gh_rd_common.h"

#pragma clang fp contract(off)

namespace gh {

__device__ inline int Q(int x, int y, int z) { return (z + 2) * 144 + (y + 2) * 12 + (x + 2); }  // [-2,10)^3

// Doubles of parameter-gradient partials per slot: per wave of the 512-lane
// workgroup the two sums A = sum w_c Rdiff_c and B = sum w_c u_c (1 - u_c)
// (Rdiff including ih2); the parameter gradients are linear in them
// (adj_param_grads).
constexpr int ADJ_PART = 2 * 8;

// dJ/dDw = dt A / Dw, dJ/drho = dt B, dJ/dT = sum w_c R_c / N = (A + rho B) / N
inline void adj_param_grads(const double ab[2], double dt, double invDw, double invN, double rho, double out[3]) {
    out[0] = dt * ab[0] * invDw;
    out[1] = dt * ab[1];
    out[2] = (ab[0] + rho * ab[1]) * invN;
}

// v moved across lanes by the DPP pattern CTRL (both 32-bit halves)
template <int CTRL>
__device__ inline double dpp_move(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double read_lane(double v, int lane) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// Sum over the 64 lanes of a wave (all lanes active), valid in lanes 0-15.
// Each 16-lane row is reduced by butterflies (lane pairs, quad halves, row
// half mirror, row mirror) that leave the same row sum in every lane of the
// row (floating-point addition commutes); the rows are then added in order
// 0, 1, 2, 3.  Fixed order, no LDS, no barrier.
__device__ inline double wave_sum_dpp(double v) {
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_move<0x141>(v);  // row_half_mirror
    v += dpp_move<0x140>(v);  // row_mirror
    return ((v + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
}

__device__ inline double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x / 64); ++k) s += red[k];
    return s;  // valid in thread 0
}

// Per-slot parameters of the reverse step: one solve, or packed solves.
struct UniformAdjParams {
    double rho, ih2, dt, invN, invDw;
    int bpd;
    uint32_t magic;  // bpd_magic(bpd)
    __device__ inline void get(int, double& r, double& i, double& d, double& n, double& w) const {
        r = rho;
        i = ih2;
        d = dt;
        n = invN;
        w = invDw;
    }
    __device__ inline BlockPos pos(int s) const { return block_pos(s, 0, bpd, magic); }
};

// Domain test and outside flag of a cell with state u and record r.  Generic
// records (CLAMPED false) carry tau, with the domain threshold in sig.  Clamped
// records (stat_fast_kernel: every threshold <= 0, state u >= +0, i.e. any step
// k >= 1) carry +-max(tau, 0) with the sign set where the threshold is < 0:
// then u > sig holds there, and elsewhere it is u > 0; u < tau is
// u < max(tau, 0).
template <bool CLAMPED, typename Real>
__device__ inline bool adj_in_domain(Real u, const Stat4<Real>& r, const Real* sig, size_t g) {
    if (CLAMPED) return sign_of(r.tau) || u > (Real)0;
    return u > sig[g];
}
template <bool CLAMPED, typename Real>
__device__ inline bool adj_outside(Real u, const Stat4<Real>& r) {
    if (CLAMPED) return u < fabs(r.tau);
    return u < r.tau;
}

// Face halo cell (a, b) of face f at depth d (1 or 2): the cell index in the
// neighbour block across the face and the tile indices Q (state, flags) and,
// at depth 1, P (w), each base + a * step_a + b * step_b with wave-uniform
// bases and steps.
struct AdjHalo {
    int src, sa, sb, q, qa, qb, p, pa, pb;
};
__device__ inline AdjHalo adj_halo(int f, int d) {
    const int q0 = Q(0, 0, 0), p0 = P(0, 0, 0);
    switch (f) {
        case FXM: return {8 - d, 8, 64, q0 - d, 12, 144, p0 - d, 10, 100};
        case FXP: return {d - 1, 8, 64, q0 + 7 + d, 12, 144, p0 + 7 + d, 10, 100};
        case FYM: return {8 * (8 - d), 1, 64, q0 - 12 * d, 1, 144, p0 - 10 * d, 1, 100};
        case FYP: return {8 * (d - 1), 1, 64, q0 + 12 * (7 + d), 1, 144, p0 + 10 * (7 + d), 1, 100};
        case FZM: return {64 * (8 - d), 1, 8, q0 - 144 * d, 1, 12, p0 - 100 * d, 1, 10};
        default:  return {64 * (d - 1), 1, 8, q0 + 144 * (7 + d), 1, 12, p0 + 100 * (7 + d), 1, 10};
    }
}

// One cell of one reverse step: lambda_k (lk) and ih2 sum_i d'_i (u_i - u_c)
// (rdiff) from the face coefficients D (-x, +x, -y, +y, -z, +z), per face i
// the outside flags of the opposite neighbour (oo: face i ^ 1's neighbour) and
// of the cell two steps out (of), the neighbour state un and weight wn, and
// the cell's state u0, masked adjoint l1 and domain test.
// nb(i, oo, of, un, wn) reads face i's neighbour data where it is used.
template <typename Nb>
__device__ inline void adj_cell(const double D[6], Nb nb, double u0, double l1, bool dom, double rho, double ih2,
                                double dt, double& lk, double& rdiff) {
    double sdu = 0, sumd = 0, acc = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        int oo, of;
        double un, wn;
        nb(i, oo, of, un, wn);
        // face coefficients doubled (exactly: ldexp by the 0/1 flag) where the
        // opposite neighbour of c, or of the neighbour, is outside
        const double dp = ldexp(D[i], oo);
        sdu = fma(dp, un, sdu);
        sumd += dp;
        acc = fma(wn, ldexp(D[i], of), acc);
    }
    rdiff = fma(-sumd, u0, sdu) * ih2;  // ih2 sum_i d'_i (u_i - u_c)
    const double self = dom ? l1 * (1. + dt * (-ih2 * sumd + rho * (1. - 2. * u0))) : l1;
    lk = self + dt * ih2 * acc;
}

// lambda[c] = 0 where u[c] >= 1 (the upper clamp of the step that produced u
// is active): the masked adjoint the reverse step reads.
template <typename Real>
__device__ inline void adj_mask_cell(size_t c, double* __restrict__ lam, const Real* __restrict__ u) {
    if (!(u[c] < (Real)1)) lam[c] = 0.0;
}
template <typename Real>
__global__ __launch_bounds__(256) void adj_mask_kernel(size_t n, double* __restrict__ lam,
                                                       const Real* __restrict__ u) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) adj_mask_cell(c, lam, u);
}

// One reverse step; one workgroup per block (XCD-aware order).  lam1 holds
// m lambda_{k+1} with m = [u_{k+1} < 1] (the seed masked by adj_mask_kernel,
// later steps by the previous launch); lam0 receives lambda_k, masked by
// [u_k < 1] when mask_out (k >= 1: the next reverse step reads it), plain
// at k = 0 (the initial-condition gradient reads it).
// partial[ADJ_PART s + 2 w + q] accumulates wave w's sums A (q = 0) and B
// (q = 1) of block s, step after step.  CLAMPED: st holds the clamped-state records and
// sig is not read (steps k >= 1 when every threshold is <= 0).
// zflux (may be null) marks blocks without diffusion (zero_flux_kernel): every
// face coefficient their cells use is +0, so each of adj_cell's face terms is
// a signed zero whatever the neighbour's state, flag or weight (finite
// values), the sums come out +0, and the cell's update is adj_cell with zero
// neighbour data -- the same bits with no halo and no LDS.  (Grouping them
// several to a workgroup with the two-step kernel's cost-balanced split, as
// the forward does, made the reverse step slower: its short path is not
// cheap enough to batch.)
template <typename Real, typename Params, bool CLAMPED = false>
__global__ __launch_bounds__(512) void adj_step_kernel(int nslots, int per_xcd,
                                                       const Real* __restrict__ uk, int mask_out,
                                                       const double* __restrict__ lam1, double* __restrict__ lam0,
                                                       const Stat4<Real>* __restrict__ st,
                                                       const Real* __restrict__ sig, Params prm,
                                                       double* __restrict__ partial,
                                                       const unsigned char* __restrict__ zflux) {
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    double rho, ih2, dt, invN, invDw;
    prm.get(s, rho, ih2, dt, invN, invDw);
    if (zflux && zflux[s]) {  // uniform per workgroup
        const int t = threadIdx.x;
        const size_t g = (size_t)s * BCELLS + t;
        const Real uc = uk[g];
        const Stat4<Real> sv = st[g];
        const bool domc = adj_in_domain<CLAMPED>(uc, sv, sig, g);
        const double l1c = lam1[g];
        const double u0 = (double)uc;
        const double D[6] = {0., 0., 0., 0., 0., 0.};
        double lk, rdiff;
        adj_cell(
            D,
            [&](int, int& oo, int& of, double& un, double& wn) {
                oo = 0;
                of = 0;
                un = 0.;
                wn = 0.;
            },
            u0, l1c, domc, rho, ih2, dt, lk, rdiff);
        lam0[g] = (mask_out && !(uc < (Real)1)) ? 0.0 : lk;
        const double wc = domc ? l1c : 0.0;
        const double A = wave_sum_dpp(wc * rdiff), B = wave_sum_dpp(wc * (u0 * (1. - u0)));
        if ((t & 63) == 0) {
            double* pw = partial + ADJ_PART * (size_t)s + 2 * (t >> 6);
            pw[0] += A;
            pw[1] += B;
        }
        return;
    }
    __shared__ Real s_u[1728];
    __shared__ unsigned char s_o[1728];
    __shared__ double s_w[1000];
    __shared__ Real s_Dx[576], s_Dy[576], s_Dz[576];

    const int t = threadIdx.x;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    const size_t g = (size_t)s * BCELLS + t;
    const Real uc = uk[g];
    const Stat4<Real> sv = st[g];
    const bool domc = adj_in_domain<CLAMPED>(uc, sv, sig, g);
    const double l1c = lam1[g];
    s_u[Q(x, y, z)] = uc;
    s_o[Q(x, y, z)] = adj_outside<CLAMPED>(uc, sv);
    s_w[P(x, y, z)] = domc ? l1c : 0.0;
    s_Dx[z * 72 + y * 9 + x + 1] = sv.dx;
    s_Dy[z * 72 + (y + 1) * 8 + x] = sv.dy;
    s_Dz[(z + 1) * 64 + y * 8 + x] = sv.dz;
    // face halos: depth 1 (state, flags, w, minus-face D) and depth 2 (state,
    // flags).  Halo wave hw = wave (first pass) or wave + 8 (second pass,
    // waves 0-3) covers depth hw / 6, face hw % 6: wave-uniform (scalar) face
    // logic, and every load of both passes is issued before the first store.
    struct HaloCell {
        size_t src;  // global cell
        int q, p;    // tile indices (state/flags, w)
    };
    auto halo_cell = [&](int hw) {
        const int f = hw % 6, d = hw / 6 + 1, a = t & 7, b = (t >> 3) & 7;
        const AdjHalo m = adj_halo(f, d);
        return HaloCell{(size_t)face_nb(prm.pos(s), f) * BCELLS + (size_t)mad24(b, m.sb, mad24(a, m.sa, m.src)),
                        mad24(b, m.qb, mad24(a, m.qa, m.q)), mad24(b, m.pb, mad24(a, m.pa, m.p))};
    };
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const HaloCell h0 = halo_cell(wv);
    const Real u_h0 = uk[h0.src];
    const Stat4<Real> s_h0 = st[h0.src];
    double l_h0 = 0;
    bool dom_h0 = false;
    if (wv < 6) {  // depth 1
        dom_h0 = adj_in_domain<CLAMPED>(u_h0, s_h0, sig, h0.src);
        l_h0 = lam1[h0.src];
    }
    HaloCell h1{0, 0, 0};
    Real u_h1 = 0;
    Stat4<Real> s_h1{};
    if (wv < 4) {  // second pass: depth 2, faces 2-5
        h1 = halo_cell(wv + 8);
        u_h1 = uk[h1.src];
        s_h1 = st[h1.src];
    }
    s_u[h0.q] = u_h0;
    s_o[h0.q] = adj_outside<CLAMPED>(u_h0, s_h0);
    if (wv < 6) {
        const int a = t & 7, b = (t >> 3) & 7;
        s_w[h0.p] = dom_h0 ? l_h0 : 0.0;
        if (wv == FXM) s_Dx[b * 72 + a * 9] = s_h0.dx;
        else if (wv == FYM) s_Dy[b * 72 + a] = s_h0.dy;
        else if (wv == FZM) s_Dz[b * 8 + a] = s_h0.dz;
    }
    if (wv < 4) {
        s_u[h1.q] = u_h1;
        s_o[h1.q] = adj_outside<CLAMPED>(u_h1, s_h1);
    }
    __syncthreads();

    const double u0 = (double)uc;
    const double D[6] = {(double)s_Dx[z * 72 + y * 9 + x], (double)sv.dx, (double)s_Dy[z * 72 + y * 8 + x],
                         (double)sv.dy, (double)s_Dz[z * 64 + y * 8 + x], (double)sv.dz};
    const int nq[6] = {Q(x - 1, y, z), Q(x + 1, y, z), Q(x, y - 1, z), Q(x, y + 1, z), Q(x, y, z - 1), Q(x, y, z + 1)};
    const int np[6] = {P(x - 1, y, z), P(x + 1, y, z), P(x, y - 1, z), P(x, y + 1, z), P(x, y, z - 1), P(x, y, z + 1)};
    const int n2[6] = {Q(x - 2, y, z), Q(x + 2, y, z), Q(x, y - 2, z), Q(x, y + 2, z), Q(x, y, z - 2), Q(x, y, z + 2)};
    double lk, rdiff;
    adj_cell(
        D,
        [&](int i, int& oo, int& of, double& un, double& wn) {
            oo = s_o[nq[i ^ 1]];
            of = s_o[n2[i]];
            un = (double)s_u[nq[i]];
            wn = s_w[np[i]];
        },
        u0, l1c, domc, rho, ih2, dt, lk, rdiff);
    lam0[g] = (mask_out && !(uc < (Real)1)) ? 0.0 : lk;

    const double wc = s_w[P(x, y, z)];
    const double A = wave_sum_dpp(wc * rdiff), B = wave_sum_dpp(wc * (u0 * (1. - u0)));
    if ((t & 63) == 0) {
        double* pw = partial + ADJ_PART * (size_t)s + 2 * (t >> 6);
        pw[0] += A;
        pw[1] += B;
    }
}

// sum over cells of lam0 * u0 * (X - x0) / sigma^2 for X = ix, iy, iz, by
// workgroup wg of nwg (256 lanes): partial[3 wg + q]
template <typename Real>
__device__ inline void adj_ic_body(int wg, int nwg, size_t nc, int bpd, const double* __restrict__ lam,
                                   const Real* __restrict__ u0, double x0, double y0, double z0, double is2,
                                   double* __restrict__ partial) {
    double a[3] = {0, 0, 0};
    for (size_t c = (size_t)wg * 256 + threadIdx.x; c < nc; c += (size_t)nwg * 256) {
        const int s = (int)(c / BCELLS), l = (int)(c % BCELLS);
        const double ix = (s % bpd) * BS + (l & 7), iy = ((s / bpd) % bpd) * BS + ((l >> 3) & 7),
                     iz = (s / (bpd * bpd)) * BS + (l >> 6);
        const double v = lam[c] * (double)u0[c] * is2;
        a[0] += v * (ix - x0);
        a[1] += v * (iy - y0);
        a[2] += v * (iz - z0);
    }
    __shared__ double red[8];
    for (int q = 0; q < 3; ++q) {
        const double sq = block_sum(a[q], red);
        if (threadIdx.x == 0) partial[3 * wg + q] = sq;
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void adj_ic_partials(size_t nc, int bpd, const double* __restrict__ lam,
                                                       const Real* __restrict__ u0, double x0, double y0, double z0,
                                                       double is2, double* __restrict__ partial) {
    adj_ic_body<Real>((int)blockIdx.x, (int)gridDim.x, nc, bpd, lam, u0, x0, y0, z0, is2, partial);
}

// The host's summation of the per-wave partials: workgroup order g = 0..8*per_xcd-1
// of the step kernel's grid, slot = (g % 8) * per_xcd + g / 8, then wave order.
inline void sum_slot_partials(const double* part, int nslots, double ab[2]) {
    const int per_xcd = (nslots + 7) / 8;
    ab[0] = ab[1] = 0;
    for (int g = 0; g < 8 * per_xcd; ++g) {
        const int s = (g % 8) * per_xcd + g / 8;
        if (s >= nslots) continue;
        for (int w = 0; w < 8; ++w)
            for (int q = 0; q < 2; ++q) ab[q] += part[ADJ_PART * (size_t)s + 2 * w + q];
    }
}

}  // namespace gh
