// gh_adjoint.hip -- discrete adjoint of the reaction-diffusion solve and the
// gradient of J = -logL w.r.t. (Dw, rho, T, x0, y0, z0).
//
// No reference counterpart (the reference has no derivatives).  Differentiable
// model: N fused steps u_{k+1} = F(u_k) = clamp(u_k + dt R(u_k)) of the
// operator of Glioma/Operators/ReactionDiffusionOperator.h with dt = T/N, where
//   R_c = ih2 sum_i d'_i (u_{n_i} - u_c) + rho u_c (1 - u_c)   (in the domain)
//   d'_i = Dbar_i (1 + o(opposite neighbour of face i))        (no-flux rule)
//   o(x) = [u_x < tau_x]                                       (chf < 0.1)
// The flags o, the in-domain test and the upper clamp are frozen along the
// trajectory (a.e. derivative); the lower clamp is taken as inactive.
// Dbar is homogeneous of degree 1 in (Dw, Dg = Dw/10), so dDbar/dDw = Dbar/Dw.
//
// Reverse sweep, per step, per cell c (lambda in FP64):
//   lambda_k[c] = m_c l_c (dom_c ? 1 + dt(-ih2 sum_i d'_i + rho(1 - 2u_c)) : 1)
//               + dt ih2 sum_i w_{n_i} Dbar_i (1 + o(c + 2 dir_i))
//   with w_e = m_e dom_e lambda_{k+1}[e], m_e = [u_{k+1,e} < 1]
// (the face of neighbour e towards c is doubled when e's opposite neighbour,
// c + 2 dir_i, is outside).  Parameter gradients accumulate per cell
//   dJ/dDw += w_c dt Rdiff_c / Dw, dJ/drho += w_c dt u_c(1-u_c), dJ/dT += w_c R_c / N
// in FP64 per workgroup; every workgroup owns the same blocks in every step,
// so the accumulation order is fixed.  States u_k are recomputed per segment
// of K steps from checkpoints taken in the forward sweep.
#include <cmath>
#include <vector>

#include "gh_ctx.h"
#include "gh_rd_common.h"

namespace gh {
namespace {

__device__ inline int Q(int x, int y, int z) { return (z + 2) * 144 + (y + 2) * 12 + (x + 2); }  // [-2,10)^3

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

template <typename Real>
__global__ __launch_bounds__(512) void adj_step_kernel(int nslots, int per_xcd, const int32_t* __restrict__ nbr,
                                                       const Real* __restrict__ uk, const Real* __restrict__ uk1,
                                                       const double* __restrict__ lam1, double* __restrict__ lam0,
                                                       const Stat4<Real>* __restrict__ st,
                                                       const Real* __restrict__ sig, double rho, double ih2,
                                                       double dt, double invN, double invDw,
                                                       double* __restrict__ partial) {
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    __shared__ Real s_u[1728];
    __shared__ unsigned char s_o[1728];
    __shared__ double s_w[1000];
    __shared__ Real s_Dx[576], s_Dy[576], s_Dz[576];
    __shared__ double s_red[8];

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
    // face halos: depth 1 (state, flags, w, minus-face D) and depth 2 (state, flags)
    for (int h = t; h < 768; h += 512) {
        const int depth = h / 384, r = h % 384;
        const int f = r >> 6, a = r & 7, b = (r >> 3) & 7;
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
        const size_t nb = (size_t)nbr[6 * s + f] * BCELLS + src;
        const Real uh = uk[nb];
        const Stat4<Real> hs = st[nb];
        s_u[Q(qx, qy, qz)] = uh;
        s_o[Q(qx, qy, qz)] = uh < hs.tau;
        if (depth == 0) {
            const bool dm = uh > sig[nb];
            const bool mm = uk1[nb] < (Real)1;
            s_w[P(qx, qy, qz)] = (dm && mm) ? lam1[nb] : 0.0;
            if (f == FXM) s_Dx[b * 72 + a * 9] = hs.dx;
            else if (f == FYM) s_Dy[b * 72 + a] = hs.dy;
            else if (f == FZM) s_Dz[b * 8 + a] = hs.dz;
        }
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
    const double gd = block_sum(wc * dt * rdiff * invDw, s_red);
    const double gr = block_sum(wc * dt * u0 * (1. - u0), s_red);
    const double gt = block_sum(wc * R * invN, s_red);
    if (t == 0) {
        partial[3 * blockIdx.x + 0] += gd;
        partial[3 * blockIdx.x + 1] += gr;
        partial[3 * blockIdx.x + 2] += gt;
    }
}

// sum over cells of lam0 * u0 * (X - x0) / sigma^2 for X = ix, iy, iz
template <typename Real>
__global__ __launch_bounds__(256) void adj_ic_partials(size_t nc, int bpd, const double* __restrict__ lam,
                                                       const Real* __restrict__ u0, double x0, double y0, double z0,
                                                       double is2, double* __restrict__ partial) {
    double a[3] = {0, 0, 0};
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (size_t)gridDim.x * blockDim.x) {
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
        if (threadIdx.x == 0) partial[3 * blockIdx.x + q] = sq;
    }
}

template <typename Real>
int adjoint_t(Ctx& c, int64_t N, double T, int K, const double* lik, int stepPET, int stepTi, int use_roi,
              const double* ic, double sigma, double* loss, double* grad) {
    const size_t nc = c.ncells();
    const size_t nb = nc * sizeof(Real);
    const int64_t nck = (N + K - 1) / K;
    int rc;
    if ((rc = c.adjCk.ensure((size_t)nck * nb)) || (rc = c.adjSeg.ensure((size_t)(K + 1) * nb)) ||
        (rc = c.adjLam[0].ensure(nc * 8)) || (rc = c.adjLam[1].ensure(nc * 8)))
        return rc;
    const double dt = T / (double)N;
    char* ck = (char*)c.adjCk.p;
    char* seg = (char*)c.adjSeg.p;
    // forward sweep with checkpoints every K steps
    for (int64_t k = 0; k < N; ++k) {
        if (k % K == 0)
            GH_HIP_CHECK(hipMemcpyAsync(ck + (size_t)(k / K) * nb, c.phi[c.cur].p, nb, hipMemcpyDeviceToDevice,
                                        c.stream));
        if ((rc = rd_step_buffers(c, c.phi[c.cur].p, c.phi[1 - c.cur].p, dt))) return rc;
        c.cur = 1 - c.cur;
    }
    // loss at u_N and its gradient as the terminal adjoint
    int64_t npet = 0;
    if ((rc = lik_eval_current(c, lik, stepPET, stepTi, use_roi, loss, &npet))) return rc;
    int lc = 0;
    if ((rc = lik_seed_gradient(c, c.phi[c.cur].p, c.adjLam[lc].as<double>(), lik, stepPET, stepTi, use_roi)))
        return rc;
    // reverse sweep
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    const unsigned grid = 8 * per_xcd;
    if ((rc = c.adjPart.ensure(sizeof(double) * 3 * grid))) return rc;
    GH_HIP_CHECK(hipMemsetAsync(c.adjPart.p, 0, sizeof(double) * 3 * grid, c.stream));
    const double h = block_h_of(c);
    const double ih2 = 1. / (h * h);
    for (int64_t sgi = nck - 1; sgi >= 0; --sgi) {
        const int64_t k0 = sgi * K, k1 = std::min<int64_t>(N, k0 + K);
        GH_HIP_CHECK(hipMemcpyAsync(seg, ck + (size_t)sgi * nb, nb, hipMemcpyDeviceToDevice, c.stream));
        for (int64_t j = 0; j < k1 - k0; ++j)
            if ((rc = rd_step_buffers(c, seg + (size_t)j * nb, seg + (size_t)(j + 1) * nb, dt))) return rc;
        for (int64_t k = k1 - 1; k >= k0; --k) {
            hipLaunchKernelGGL(adj_step_kernel<Real>, dim3(grid), dim3(512), 0, c.stream, ns, per_xcd,
                               c.nbr.as<int32_t>(), (const Real*)(seg + (size_t)(k - k0) * nb),
                               (const Real*)(seg + (size_t)(k - k0 + 1) * nb), c.adjLam[lc].as<double>(),
                               c.adjLam[1 - lc].as<double>(), c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(),
                               c.rho, ih2, dt, 1.0 / (double)N, 1.0 / c.Dw, c.adjPart.as<double>());
            GH_HIP_CHECK(hipGetLastError());
            lc = 1 - lc;
        }
    }
    // initial-condition gradient from lambda_0 and u_0
    const unsigned igrid = 512;
    if ((rc = c.likPart.ensure(sizeof(double) * 4 * 1025))) return rc;
    hipLaunchKernelGGL(adj_ic_partials<Real>, dim3(igrid), dim3(256), 0, c.stream, nc, c.topo.uniform_bpd,
                       c.adjLam[lc].as<double>(), (const Real*)ck, ic[0], ic[1], ic[2], 1.0 / (sigma * sigma),
                       c.likPart.as<double>());
    GH_HIP_CHECK(hipGetLastError());
    std::vector<double> pp(3 * grid), ip(3 * igrid);
    GH_HIP_CHECK(hipMemcpyAsync(pp.data(), c.adjPart.p, pp.size() * 8, hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipMemcpyAsync(ip.data(), c.likPart.p, ip.size() * 8, hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int q = 0; q < 6; ++q) grad[q] = 0;
    for (unsigned w = 0; w < grid; ++w)
        for (int q = 0; q < 3; ++q) grad[q] += pp[3 * w + q];
    for (unsigned w = 0; w < igrid; ++w)
        for (int q = 0; q < 3; ++q) grad[3 + q] += ip[3 * w + q];
    return GH_OK;
}

}  // namespace
}  // namespace gh

using namespace gh;

extern "C" int gh_rd_adjoint(gh_ctx* ctx, int64_t nsteps, double T, int ckpt_every, const double* lik_params,
                             int stepPET, int stepTi, int use_roi, const double* ic_center, double ic_sigma,
                             double* loss, double* grad) {
    Ctx* c = as_ctx(ctx);
    if (!c || !lik_params || !ic_center || !loss || !grad) return set_error(GH_EINVAL, "null argument");
    if (!c->prepared) return set_error(GH_ESTATE, "gh_rd_adjoint: call gh_rd_prepare first");
    if (!c->lik_ready) return set_error(GH_ESTATE, "gh_rd_adjoint: call gh_lik_set_data first");
    if (c->topo.uniform_bpd <= 0) return set_error(GH_ESTATE, "gh_rd_adjoint: needs a uniform grid");
    if (nsteps < 1 || ckpt_every < 1 || ckpt_every > 4096 || !(T > 0) || !(ic_sigma > 0) || !(c->Dw > 0))
        return set_error(GH_EINVAL, "gh_rd_adjoint: bad argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return c->real_kind == GH_REAL_F32
               ? adjoint_t<float>(*c, nsteps, T, ckpt_every, lik_params, stepPET, stepTi, use_roi, ic_center,
                                  ic_sigma, loss, grad)
               : adjoint_t<double>(*c, nsteps, T, ckpt_every, lik_params, stepPET, stepTi, use_roi, ic_center,
                                   ic_sigma, loss, grad);
}
