// gh_adjoint_common.h -- reverse-step kernels of the discrete adjoint, shared
// by the single-solve adjoint (gh_adjoint.hip) and the batched one
// (gh_batch.hip).  The model and the per-cell formulas are documented at the
// top of gh_adjoint.hip.
#pragma once
#include "gh_rd_common.h"

#pragma clang fp contract(off)

namespace gh {

__device__ inline int Q(int x, int y, int z) { return (z + 2) * 144 + (y + 2) * 12 + (x + 2); }  // [-2,10)^3

// three workgroup sums at once (one LDS round trip, two barriers); red holds
// 3 * (blockDim / 64) doubles; the sums are valid in thread 0, each in the
// same fixed order as block_sum
__device__ inline void block_sum3(double& a, double& b, double& c, double* red) {
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        c += __shfl_xor(c, o, 64);
    }
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64, nw = (int)(blockDim.x / 64);
    __syncthreads();
    if (lane == 0) {
        red[w] = a;
        red[nw + w] = b;
        red[2 * nw + w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0;
        for (int k = 0; k < nw; ++k) {
            a += red[k];
            b += red[nw + k];
            c += red[2 * nw + k];
        }
    }
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

// One reverse step; one workgroup per block (XCD-aware order).  partial[3 s + q]
// accumulates the block's parameter-gradient terms, step after step.
template <typename Real, typename Params>
__global__ __launch_bounds__(512) void adj_step_kernel(int nslots, int per_xcd,
                                                       const Real* __restrict__ uk, const Real* __restrict__ uk1,
                                                       const double* __restrict__ lam1, double* __restrict__ lam0,
                                                       const Stat4<Real>* __restrict__ st,
                                                       const Real* __restrict__ sig, Params prm,
                                                       double* __restrict__ partial) {
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    double rho, ih2, dt, invN, invDw;
    prm.get(s, rho, ih2, dt, invN, invDw);
    __shared__ Real s_u[1728];
    __shared__ unsigned char s_o[1728];
    __shared__ double s_w[1000];
    __shared__ Real s_Dx[576], s_Dy[576], s_Dz[576];
    __shared__ double s_red[24];

    const int t = threadIdx.x;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    const size_t g = (size_t)s * BCELLS + t;
    const Real uc = uk[g];
    const Stat4<Real> sv = st[g];
    const Real sc = sig[g];
    const bool domc = uc > sc;
    const bool mc = uk1[g] < (Real)1;
    const double l1c = lam1[g];
    s_u[Q(x, y, z)] = uc;
    s_o[Q(x, y, z)] = uc < sv.tau;
    s_w[P(x, y, z)] = (mc && domc) ? l1c : 0.0;
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
        const int depth = hw / 6, f = hw % 6, a = t & 7, b = (t >> 3) & 7;
        const int in1 = depth == 0 ? 7 : 6, in0 = depth == 0 ? 0 : 1;  // source plane in the neighbour
        const int lo = depth == 0 ? -1 : -2, hi = depth == 0 ? 8 : 9;  // tile plane
        int src, qx, qy, qz;
        switch (f) {
            case FXM: src = in1 + 8 * a + 64 * b; qx = lo; qy = a; qz = b; break;
            case FXP: src = in0 + 8 * a + 64 * b; qx = hi; qy = a; qz = b; break;
            case FYM: src = a + 8 * in1 + 64 * b; qx = a; qy = lo; qz = b; break;
            case FYP: src = a + 8 * in0 + 64 * b; qx = a; qy = hi; qz = b; break;
            case FZM: src = a + 8 * b + 64 * in1; qx = a; qy = b; qz = lo; break;
            default:  src = a + 8 * b + 64 * in0; qx = a; qy = b; qz = hi; break;
        }
        return HaloCell{(size_t)face_nb(prm.pos(s), f) * BCELLS + src, Q(qx, qy, qz), depth == 0 ? P(qx, qy, qz) : 0};
    };
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const HaloCell h0 = halo_cell(wv);
    const Real u_h0 = uk[h0.src];
    const Stat4<Real> s_h0 = st[h0.src];
    Real sg_h0 = 0, u1_h0 = 0;
    double l_h0 = 0;
    if (wv < 6) {  // depth 1
        sg_h0 = sig[h0.src];
        u1_h0 = uk1[h0.src];
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
    s_o[h0.q] = u_h0 < s_h0.tau;
    if (wv < 6) {
        const int a = t & 7, b = (t >> 3) & 7;
        s_w[h0.p] = (u_h0 > sg_h0 && u1_h0 < (Real)1) ? l_h0 : 0.0;
        if (wv == FXM) s_Dx[b * 72 + a * 9] = s_h0.dx;
        else if (wv == FYM) s_Dy[b * 72 + a] = s_h0.dy;
        else if (wv == FZM) s_Dz[b * 8 + a] = s_h0.dz;
    }
    if (wv < 4) {
        s_u[h1.q] = u_h1;
        s_o[h1.q] = u_h1 < s_h1.tau;
    }
    __syncthreads();

    const double u0 = (double)uc;
    const double D[6] = {(double)s_Dx[z * 72 + y * 9 + x], (double)sv.dx, (double)s_Dy[z * 72 + y * 8 + x],
                         (double)sv.dy, (double)s_Dz[z * 64 + y * 8 + x], (double)sv.dz};
    const int nq[6] = {Q(x - 1, y, z), Q(x + 1, y, z), Q(x, y - 1, z), Q(x, y + 1, z), Q(x, y, z - 1), Q(x, y, z + 1)};
    const int np[6] = {P(x - 1, y, z), P(x + 1, y, z), P(x, y - 1, z), P(x, y + 1, z), P(x, y, z - 1), P(x, y, z + 1)};
    const int n2[6] = {Q(x - 2, y, z), Q(x + 2, y, z), Q(x, y - 2, z), Q(x, y + 2, z), Q(x, y, z - 2), Q(x, y, z + 2)};
    double rdiff = 0, sumd = 0, acc = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double dp = D[i] * (1.0 + (double)s_o[nq[i ^ 1]]);  // doubled if the opposite neighbour is outside
        rdiff += dp * ((double)s_u[nq[i]] - u0);
        sumd += dp;
        acc += s_w[np[i]] * D[i] * (1.0 + (double)s_o[n2[i]]);
    }
    rdiff *= ih2;
    const double R = rdiff + rho * u0 * (1. - u0);
    double self = 0.0;
    if (mc) self = domc ? l1c * (1. + dt * (-ih2 * sumd + rho * (1. - 2. * u0))) : l1c;
    lam0[g] = self + dt * ih2 * acc;

    const double wc = s_w[P(x, y, z)];
    double gd = wc * dt * rdiff * invDw, gr = wc * dt * u0 * (1. - u0), gt = wc * R * invN;
    block_sum3(gd, gr, gt, s_red);
    if (t == 0) {
        partial[3 * (size_t)s + 0] += gd;
        partial[3 * (size_t)s + 1] += gr;
        partial[3 * (size_t)s + 2] += gt;
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

// The host's summation of the per-slot partials: workgroup order w = 0..8*per_xcd-1
// of the step kernel's grid, slot = (w % 8) * per_xcd + w / 8.
inline void sum_slot_partials(const double* part, int nslots, double out[3]) {
    const int per_xcd = (nslots + 7) / 8;
    out[0] = out[1] = out[2] = 0;
    for (int w = 0; w < 8 * per_xcd; ++w) {
        const int s = (w % 8) * per_xcd + w / 8;
        if (s >= nslots) continue;
        for (int q = 0; q < 3; ++q) out[q] += part[3 * (size_t)s + q];
    }
}

}  // namespace gh
