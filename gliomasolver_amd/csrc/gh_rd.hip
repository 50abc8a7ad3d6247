// gh_rd.hip -- reaction-diffusion kernels for gfx950.
//
// Reference: Glioma/Operators/ReactionDiffusionOperator.h (rhs :98-139,
// harmonic mean :148-163, no-flux BC :166-181, UpdateTumor :215-222) driven by
// Glioma/Glioma_ReactionDiffusion.cpp:176-187.
//
// Arithmetic is bit-faithful to the reference for Real = float and double:
// the same float/double promotion points, no FMA contraction (this file is
// compiled with -ffp-contract=off and the pragma below), IEEE double division
// for the harmonic mean.
//
// Split of the work:
//   rd_statics_kernel  once per solve (Dw, Dg, anatomy fixed): the harmonic
//                      face diffusivity of each cell's +x,+y,+z face (the
//                      harmonic mean is symmetric, so each face is computed
//                      once), the threshold tau with "chf = phi+p_w+p_g < 0.1
//                      <=> phi < tau", and sigma with "p_w+p_g+phi > 0 <=>
//                      phi > sigma".  Both predicates are monotone in phi, so
//                      tau is found exactly by bisection over the ordered bit
//                      patterns of Real.
//   rd_step_kernel     every step: one workgroup per 8^3 block stages the
//                      block plus its six face halos (phi, outside flag, minus-
//                      face diffusivity) in LDS and evaluates the stencil, then
//                      either stores dphidt (two-pass API) or applies the
//                      Euler update and clamp in registers (fused step).
#include <cfloat>
#include <cmath>

#include "gh_ctx.h"
#include "gh_rd_common.h"

#pragma clang fp contract(off)

namespace gh {
namespace {

template <typename Real>
__global__ __launch_bounds__(256) void rd_statics_kernel(int nslots, const int32_t* __restrict__ nbr,
                                                         const Real* __restrict__ pw, const Real* __restrict__ pg,
                                                         Real Dw, Real Dg, Real* __restrict__ D,
                                                         Real* __restrict__ tau, Real* __restrict__ sig) {
    const size_t nc = (size_t)nslots * BCELLS;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nc) return;
    const int s = (int)(gid / BCELLS);
    const int l = (int)(gid % BCELLS);
    const int x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    const Real pwc = pw[gid], pgc = pg[gid];
    const Real dc = diffusivity(pwc, pgc, Dw, Dg);
    // +x, +y, +z neighbour cells (same slot or the same-level neighbour slot)
    size_t nx = (x < 7) ? gid + 1 : (size_t)nbr[6 * s + FXP] * BCELLS + (l - 7);
    size_t ny = (y < 7) ? gid + 8 : (size_t)nbr[6 * s + FYP] * BCELLS + (l - 56);
    size_t nz = (z < 7) ? gid + 64 : (size_t)nbr[6 * s + FZP] * BCELLS + (l - 448);
    D[gid] = harmonic(diffusivity(pw[nx], pg[nx], Dw, Dg), dc);
    D[nc + gid] = harmonic(diffusivity(pw[ny], pg[ny], Dw, Dg), dc);
    D[2 * nc + gid] = harmonic(diffusivity(pw[nz], pg[nz], Dw, Dg), dc);
    tau[gid] = chf_threshold(pwc, pgc);
    sig[gid] = domain_threshold(pwc, pgc);
}

__device__ inline int P(int x, int y, int z) { return (z + 1) * 100 + (y + 1) * 10 + (x + 1); }

// One workgroup of 512 lanes per block; lane = cell (x fastest).
// Slot mapping is XCD-aware: consecutive workgroups land on different XCDs,
// so workgroup w serves slot (w % 8) * per_xcd + w / 8 and each XCD sweeps a
// contiguous range of blocks whose halos stay in its own L2.
template <typename Real, bool FUSED>
__global__ __launch_bounds__(512) void rd_step_kernel(int nslots, int per_xcd, const int32_t* __restrict__ nbr,
                                                      const Real* __restrict__ phi, Real* __restrict__ out,
                                                      const Real* __restrict__ D, const Real* __restrict__ tau,
                                                      const Real* __restrict__ sig, Real rho, double ih2,
                                                      double dt) {
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    __shared__ Real s_phi[1000];
    __shared__ unsigned char s_out[1000];
    __shared__ Real s_Dx[576], s_Dy[576], s_Dz[576];

    const size_t nc = (size_t)nslots * BCELLS;
    const int t = threadIdx.x;
    const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
    const size_t g = (size_t)s * BCELLS + t;

    const Real pc = phi[g];
    const Real tc = tau[g];
    const Real sc = sig[g];
    const Real dxp = D[g], dyp = D[nc + g], dzp = D[2 * nc + g];
    s_phi[P(x, y, z)] = pc;
    s_out[P(x, y, z)] = pc < tc;
    s_Dx[z * 72 + y * 9 + x + 1] = dxp;
    s_Dy[z * 72 + (y + 1) * 8 + x] = dyp;
    s_Dz[(z + 1) * 64 + y * 8 + x] = dzp;

    if (t < 384) {  // six face halos of 64 cells, one lane each
        const int f = t >> 6, u = t & 7, v = (t >> 3) & 7;
        const size_t nb = (size_t)nbr[6 * s + f] * BCELLS;
        int src, dst;
        switch (f) {
            case FXM: src = 7 + 8 * u + 64 * v; dst = P(-1, u, v); break;
            case FXP: src = 0 + 8 * u + 64 * v; dst = P(8, u, v); break;
            case FYM: src = u + 8 * 7 + 64 * v; dst = P(u, -1, v); break;
            case FYP: src = u + 8 * 0 + 64 * v; dst = P(u, 8, v); break;
            case FZM: src = u + 8 * v + 64 * 7; dst = P(u, v, -1); break;
            default:  src = u + 8 * v + 64 * 0; dst = P(u, v, 8); break;
        }
        const Real ph = phi[nb + src];
        s_phi[dst] = ph;
        s_out[dst] = ph < tau[nb + src];
        if (f == FXM) s_Dx[v * 72 + u * 9] = D[nb + src];
        else if (f == FYM) s_Dy[v * 72 + u] = D[nc + nb + src];
        else if (f == FZM) s_Dz[v * 8 + u] = D[2 * nc + nb + src];
    }
    __syncthreads();

    Real d0 = s_Dx[z * 72 + y * 9 + x], d1 = dxp;
    Real d2 = s_Dy[z * 72 + y * 8 + x], d3 = dyp;
    Real d4 = s_Dz[z * 64 + y * 8 + x], d5 = dzp;
    const Real p0 = s_phi[P(x - 1, y, z)], p1 = s_phi[P(x + 1, y, z)];
    const Real p2 = s_phi[P(x, y - 1, z)], p3 = s_phi[P(x, y + 1, z)];
    const Real p4 = s_phi[P(x, y, z - 1)], p5 = s_phi[P(x, y, z + 1)];
    // _applyNoFluxBC: an outside neighbour doubles the opposite face (exact)
    if (s_out[P(x - 1, y, z)]) d1 = (Real)((double)d1 * 2.0);
    if (s_out[P(x + 1, y, z)]) d0 = (Real)((double)d0 * 2.0);
    if (s_out[P(x, y - 1, z)]) d3 = (Real)((double)d3 * 2.0);
    if (s_out[P(x, y + 1, z)]) d2 = (Real)((double)d2 * 2.0);
    if (s_out[P(x, y, z - 1)]) d5 = (Real)((double)d5 * 2.0);
    if (s_out[P(x, y, z + 1)]) d4 = (Real)((double)d4 * 2.0);
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
    Real dphidt = (Real)((fout + fin) + react);
    if (!(pc > sc)) dphidt = (Real)0.;  // outside the brain domain (:103, :137-138)
    if (FUSED) {
        // UpdateTumor: phi += dt*dphidt (double), then max(0,.), min(1,.)
        Real p = (Real)((double)pc + dt * (double)dphidt);
        p = ((Real)0. < p) ? p : (Real)0.;
        p = (p < (Real)1.) ? p : (Real)1.;
        out[g] = p;
    } else {
        out[g] = dphidt;
    }
}

template <typename Real>
__global__ __launch_bounds__(256) void rd_update_kernel(size_t nc, Real* __restrict__ phi,
                                                        const Real* __restrict__ dphidt, double dt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    Real p = (Real)((double)phi[i] + dt * (double)dphidt[i]);
    p = ((Real)0. < p) ? p : (Real)0.;
    p = (p < (Real)1.) ? p : (Real)1.;
    phi[i] = p;
}

inline double block_h(const Ctx& c) {
    // info.h[0] = 2^-level / 8 as Real (MRAGrid.inl:200-206); exact in float
    return std::ldexp(1.0, -c.topo.blocks[0].level) / BS;
}

template <typename Real>
int prepare_t(Ctx& c, double Dw, double Dg, double rho) {
    const int ns = c.topo.nblocks();
    const size_t nc = c.ncells();
    const unsigned grid = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(rd_statics_kernel<Real>, dim3(grid), dim3(256), 0, c.stream, ns, c.nbr.as<int32_t>(),
                       c.pw.as<Real>(), c.pg.as<Real>(), (Real)Dw, (Real)Dg, c.statD.as<Real>(),
                       c.statTau.as<Real>(), c.statSig.as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    c.Dw = Dw;
    c.Dg = Dg;
    c.rho = rho;
    c.prepared = true;
    return GH_OK;
}

template <typename Real, bool FUSED>
int step_t(Ctx& c, Real* out) {
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    const double h = block_h(c);
    const double ih2 = 1. / (h * h);
    hipLaunchKernelGGL((rd_step_kernel<Real, FUSED>), dim3(8 * per_xcd), dim3(512), 0, c.stream, ns, per_xcd,
                       c.nbr.as<int32_t>(), c.phi[c.cur].as<Real>(), out, c.statD.as<Real>(), c.statTau.as<Real>(),
                       c.statSig.as<Real>(), (Real)c.rho, ih2, c.dt_last);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

}  // namespace

int rd_prepare(Ctx& c, double Dw, double Dg, double rho) {
    if (c.topo.uniform_bpd <= 0) return set_error(GH_ESTATE, "rd_prepare: grid has blocks on several levels");
    return c.real_kind == GH_REAL_F32 ? prepare_t<float>(c, Dw, Dg, rho) : prepare_t<double>(c, Dw, Dg, rho);
}

int rd_rhs(Ctx& c) {
    return c.real_kind == GH_REAL_F32 ? step_t<float, false>(c, c.dphidt.as<float>())
                                      : step_t<double, false>(c, c.dphidt.as<double>());
}

int rd_update(Ctx& c, double dt) {
    const size_t nc = c.ncells();
    const unsigned grid = (unsigned)((nc + 255) / 256);
    if (c.real_kind == GH_REAL_F32)
        hipLaunchKernelGGL(rd_update_kernel<float>, dim3(grid), dim3(256), 0, c.stream, nc, c.phi[c.cur].as<float>(),
                           c.dphidt.as<float>(), dt);
    else
        hipLaunchKernelGGL(rd_update_kernel<double>, dim3(grid), dim3(256), 0, c.stream, nc,
                           c.phi[c.cur].as<double>(), c.dphidt.as<double>(), dt);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

int rd_steps(Ctx& c, int64_t nsteps, double dt) {
    c.dt_last = dt;
    for (int64_t k = 0; k < nsteps; ++k) {
        int rc = c.real_kind == GH_REAL_F32 ? step_t<float, true>(c, c.phi[1 - c.cur].as<float>())
                                            : step_t<double, true>(c, c.phi[1 - c.cur].as<double>());
        if (rc) return rc;
        c.cur = 1 - c.cur;
    }
    return GH_OK;
}

template <typename Real>
static void thresholds_host(int64_t n, const Real* pw, const Real* pg, Real* tau, Real* sig) {
    for (int64_t i = 0; i < n; ++i) {
        tau[i] = chf_threshold(pw[i], pg[i]);
        sig[i] = domain_threshold(pw[i], pg[i]);
    }
}

}  // namespace gh

extern "C" int gh_rd_thresholds(int real_kind, int64_t n, const void* p_w, const void* p_g, void* tau, void* sigma) {
    if (n < 0 || (n > 0 && (!p_w || !p_g || !tau || !sigma)))
        return gh::set_error(GH_EINVAL, "gh_rd_thresholds: bad argument");
    if (real_kind == GH_REAL_F32)
        gh::thresholds_host(n, (const float*)p_w, (const float*)p_g, (float*)tau, (float*)sigma);
    else if (real_kind == GH_REAL_F64)
        gh::thresholds_host(n, (const double*)p_w, (const double*)p_g, (double*)tau, (double*)sigma);
    else
        return gh::set_error(GH_EINVAL, "gh_rd_thresholds: bad real_kind");
    return GH_OK;
}
