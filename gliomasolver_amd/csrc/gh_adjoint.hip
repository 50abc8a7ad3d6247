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
// with R_c = Rdiff_c + rho u_c (1 - u_c): all three are linear in the sums
// A = sum w_c Rdiff_c and B = sum w_c u_c (1 - u_c), which are accumulated in
// FP64 per wave (dt, rho, N and Dw are constant over the solve and applied
// once at the end); every wave owns the same cells in every step, so the
// accumulation order is fixed.  States u_k are recomputed per segment of K
// steps from checkpoints taken in the forward sweep.
#include <cmath>
#include <vector>

#include "gh_ctx.h"
#include "gh_adjoint_common.h"

namespace gh {
namespace {

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
    // forward sweep with checkpoints every K steps; two steps per launch from
    // a clamped state (k >= 1) when no checkpoint falls in between
    const bool pairs = c.fast_ok && c.topo.uniform_bpd > 0;
    for (int64_t k = 0; k < N;) {
        if (k % K == 0)
            GH_HIP_CHECK(hipMemcpyAsync(ck + (size_t)(k / K) * nb, c.phi[c.cur].p, nb, hipMemcpyDeviceToDevice,
                                        c.stream));
        if (pairs && k >= 1 && k + 2 <= N && (k + 1) % K != 0) {
            if ((rc = rd_step2_buffers(c, c.phi[c.cur].p, c.phi[1 - c.cur].p, nullptr, dt))) return rc;
            k += 2;
        } else {
            if ((rc = rd_step_buffers(c, c.phi[c.cur].p, c.phi[1 - c.cur].p, dt))) return rc;
            k += 1;
        }
        c.cur = 1 - c.cur;
    }
    // loss at u_N and its gradient as the terminal adjoint
    int64_t npet = 0;
    if ((rc = lik_eval_current(c, lik, stepPET, stepTi, use_roi, loss, &npet))) return rc;
    int lc = 0;
    if ((rc = lik_seed_gradient(c, c.phi[c.cur].p, c.adjLam[lc].as<double>(), lik, stepPET, stepTi, use_roi)))
        return rc;
    hipLaunchKernelGGL(adj_mask_kernel<Real>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c.stream, nc,
                       c.adjLam[lc].as<double>(), (const Real*)c.phi[c.cur].p);
    GH_HIP_CHECK(hipGetLastError());
    // reverse sweep
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    const unsigned grid = 8 * per_xcd;
    if ((rc = c.adjPart.ensure(sizeof(double) * ADJ_PART * ns))) return rc;
    GH_HIP_CHECK(hipMemsetAsync(c.adjPart.p, 0, sizeof(double) * ADJ_PART * ns, c.stream));
    const double h = block_h_of(c);
    const double ih2 = 1. / (h * h);
    const UniformAdjParams prm{c.rho, ih2, dt, 1.0 / (double)N, 1.0 / c.Dw, c.topo.uniform_bpd,
                                bpd_magic(c.topo.uniform_bpd)};
    // blocks without diffusion take the reverse step's short path under the
    // two-step schedule (tile 0), whose plan holds their flags
    const unsigned char* zflux = (c.nItems2f > 0 && c.rd_bpw == 0) ? c.zeroFlux.as<unsigned char>() : nullptr;
    for (int64_t sgi = nck - 1; sgi >= 0; --sgi) {
        const int64_t k0 = sgi * K, k1 = std::min<int64_t>(N, k0 + K);
        GH_HIP_CHECK(hipMemcpyAsync(seg, ck + (size_t)sgi * nb, nb, hipMemcpyDeviceToDevice, c.stream));
        for (int64_t j = 0; j < k1 - k0;) {  // the segment's states, pairs from clamped states
            if (pairs && k0 + j >= 1 && j + 2 <= k1 - k0) {
                if ((rc = rd_step2_buffers(c, seg + (size_t)j * nb, seg + (size_t)(j + 2) * nb,
                                           seg + (size_t)(j + 1) * nb, dt)))
                    return rc;
                j += 2;
            } else {
                if ((rc = rd_step_buffers(c, seg + (size_t)j * nb, seg + (size_t)(j + 1) * nb, dt))) return rc;
                j += 1;
            }
        }
        for (int64_t k = k1 - 1; k >= k0; --k) {
            const Real* uk = (const Real*)(seg + (size_t)(k - k0) * nb);
            const int mask_out = k >= 1;
            if (c.fast_ok && k >= 1)  // u_k is a clamped state: the clamped-state records
                hipLaunchKernelGGL((adj_step_kernel<Real, UniformAdjParams, true>), dim3(grid), dim3(512), 0,
                                   c.stream, ns, per_xcd, uk, mask_out, c.adjLam[lc].as<double>(),
                                   c.adjLam[1 - lc].as<double>(), c.statF.as<Stat4<Real>>(), (const Real*)nullptr,
                                   prm, c.adjPart.as<double>(), zflux);
            else
                hipLaunchKernelGGL((adj_step_kernel<Real, UniformAdjParams>), dim3(grid), dim3(512), 0, c.stream,
                                   ns, per_xcd, uk, mask_out, c.adjLam[lc].as<double>(), c.adjLam[1 - lc].as<double>(),
                                   c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), prm, c.adjPart.as<double>(),
                                   zflux);
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
    std::vector<double> pp(ADJ_PART * (size_t)ns), ip(3 * igrid);
    GH_HIP_CHECK(hipMemcpyAsync(pp.data(), c.adjPart.p, pp.size() * 8, hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipMemcpyAsync(ip.data(), c.likPart.p, ip.size() * 8, hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int q = 0; q < 6; ++q) grad[q] = 0;
    double ab[2];
    sum_slot_partials(pp.data(), ns, ab);
    adj_param_grads(ab, prm.dt, prm.invDw, prm.invN, prm.rho, grad);
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
    if (!c->lik_valid())
        return set_error(GH_ESTATE, "gh_rd_adjoint: no observations for the current grid (call gh_lik_set_data)");
    if (c->topo.uniform_bpd <= 0) return set_error(GH_ESTATE, "gh_rd_adjoint: needs a uniform grid");
    if (nsteps < 1 || ckpt_every < 1 || ckpt_every > 4096 || !(T > 0) || !(ic_sigma > 0) || !(c->Dw > 0))
        return set_error(GH_EINVAL, "gh_rd_adjoint: bad argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    if (int rc = rd_ensure_prepared(*c)) return rc;
    return c->real_kind == GH_REAL_F32
               ? adjoint_t<float>(*c, nsteps, T, ckpt_every, lik_params, stepPET, stepTi, use_roi, ic_center,
                                  ic_sigma, loss, grad)
               : adjoint_t<double>(*c, nsteps, T, ckpt_every, lik_params, stepPET, stepTi, use_roi, ic_center,
                                   ic_sigma, loss, grad);
}
