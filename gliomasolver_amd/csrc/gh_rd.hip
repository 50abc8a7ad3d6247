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
                                                         Real Dw, Real Dg, Stat4<Real>* __restrict__ st,
                                                         Real* __restrict__ sig) {
    const size_t nc = (size_t)nslots * BCELLS;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nc) return;
    const int s = (int)(gid / BCELLS);
    const int l = (int)(gid % BCELLS);
    const int x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    const Real pwc = pw[gid], pgc = pg[gid];
    const Real dc = diffusivity(pwc, pgc, Dw, Dg);
    // +x, +y, +z neighbour cells (same slot or the same-level neighbour slot)
    const size_t nx = (x < 7) ? gid + 1 : (size_t)nbr[6 * s + FXP] * BCELLS + (l - 7);
    const size_t ny = (y < 7) ? gid + 8 : (size_t)nbr[6 * s + FYP] * BCELLS + (l - 56);
    const size_t nz = (z < 7) ? gid + 64 : (size_t)nbr[6 * s + FZP] * BCELLS + (l - 448);
    Stat4<Real> r;
    r.dx = harmonic(diffusivity(pw[nx], pg[nx], Dw, Dg), dc);
    r.dy = harmonic(diffusivity(pw[ny], pg[ny], Dw, Dg), dc);
    r.dz = harmonic(diffusivity(pw[nz], pg[nz], Dw, Dg), dc);
    r.tau = chf_threshold(pwc, pgc);
    st[gid] = r;
    sig[gid] = domain_threshold(pwc, pgc);
}

template <typename Real>
__global__ __launch_bounds__(256) void rd_update_kernel(size_t nc, Real* __restrict__ phi,
                                                        const Real* __restrict__ dphidt, double dt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    phi[i] = rd_update(phi[i], dphidt[i], dt);
}

inline double block_h(const Ctx& c) {
    // info.h[0] = 2^-level / 8 as Real (MRAGrid.inl:200-206); exact in float
    return std::ldexp(1.0, -c.topo.blocks[0].level) / BS;
}

// Work items of rd_step2f_kernel on a uniform grid: each block with diffusion
// is an item; runs of consecutive blocks without diffusion (zero_flux_kernel)
// are items of up to ZF_GROUP blocks, so those workgroups keep several blocks'
// loads in flight.  The eight XCDs get contiguous item ranges (halos stay in
// one L2) of equal estimated cost: a full block 1, a block without diffusion
// kZeroFluxCost.
template <typename Real>
int plan_step2f(Ctx& c) {
    const int ns = c.topo.nblocks();
    c.nZeroFlux = 0;
    c.nItems2f = 0;
    if (!c.fast_ok || c.topo.uniform_bpd <= 0 || ns == 0) return GH_OK;
    int rc;
    if ((rc = c.zeroFlux.ensure_grow(ns))) return rc;
    const UniformStepParams<Real> prm{(Real)c.rho, 0., 0., c.topo.uniform_bpd, bpd_magic(c.topo.uniform_bpd)};
    hipLaunchKernelGGL((zero_flux_kernel<Real, UniformStepParams<Real>>), dim3(ns), dim3(256), 0, c.stream, ns,
                       c.statF.as<Stat4<Real>>(), prm, c.zeroFlux.as<unsigned char>());
    GH_HIP_CHECK(hipGetLastError());
    std::vector<unsigned char> z(ns);
    GH_HIP_CHECK(hipMemcpyAsync(z.data(), c.zeroFlux.p, ns, hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    std::vector<int2> items;
    std::vector<double> cost;
    for (int s = 0; s < ns;) {
        if (!z[s]) {
            items.push_back(make_int2(s, 0));
            cost.push_back(1.0);
            ++s;
            continue;
        }
        int k = 0;
        while (s + k < ns && z[s + k] && k < ZF_GROUP) ++k;
        items.push_back(make_int2(s, k));
        cost.push_back(kZeroFluxCost * k);
        c.nZeroFlux += k;
        s += k;
    }
    const int ni = (int)items.size();
    double total = 0;
    for (double v : cost) total += v;
    // split point x: the first item whose prefix cost reaches x/8 of the total
    double acc = 0;
    int x = 1;
    c.split2f[0] = 0;
    for (int i = 0; i < ni && x < 8; ++i) {
        acc += cost[i];
        while (x < 8 && acc >= total * x / 8) c.split2f[x++] = i + 1;
    }
    for (; x <= 8; ++x) c.split2f[x] = ni;
    // within each XCD's range the blocks with diffusion go first (in slot
    // order, so neighbouring blocks still run close in time) and the short
    // groups of blocks without diffusion last, where they fill the tail
    for (int r = 0; r < 8; ++r)
        std::stable_partition(items.begin() + c.split2f[r], items.begin() + c.split2f[r + 1],
                              [](const int2& v) { return v.y == 0; });
    if ((rc = c.items2f.ensure_grow((size_t)ni * sizeof(int2)))) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(c.items2f.p, items.data(), (size_t)ni * sizeof(int2), hipMemcpyHostToDevice,
                                c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    c.nItems2f = ni;
    return GH_OK;
}

template <typename Real>
int prepare_t(Ctx& c, double Dw, double Dg, double rho) {
    const int ns = c.topo.nblocks();
    const size_t nc = c.ncells();
    const unsigned grid = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(rd_statics_kernel<Real>, dim3(grid), dim3(256), 0, c.stream, ns, c.nbr.as<int32_t>(),
                       c.pw.as<Real>(), c.pg.as<Real>(), (Real)Dw, (Real)Dg, c.statD.as<Stat4<Real>>(),
                       c.statSig.as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    int rc;
    if ((rc = build_stat_fast(c.stream, c.real_kind, nc, c.statD.p, c.statSig.p, c.statF, c.flag, &c.fast_ok)))
        return rc;
    if ((rc = plan_step2f<Real>(c))) return rc;
    c.Dw = Dw;
    c.Dg = Dg;
    c.rho = rho;
    c.prepared = true;
    c.constants_stale = false;
    return GH_OK;
}

template <typename Real, bool FUSED, int BPW>
void launch_step(Ctx& c, const Real* in, Real* out, double dt) {
    const int ns = c.topo.nblocks();
    const int groups = (ns + BPW - 1) / BPW;
    const int per_xcd = (groups + 7) / 8;
    const double h = block_h(c);
    const double ih2 = 1. / (h * h);
    const UniformStepParams<Real> prm{(Real)c.rho, ih2, dt, c.topo.uniform_bpd, bpd_magic(c.topo.uniform_bpd)};
    hipLaunchKernelGGL((rd_step_kernel<Real, FUSED, BPW, UniformStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                       c.stream, ns, per_xcd, in, out, c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), prm);
}

// two fused steps per launch (rd_step2_kernel), one block per workgroup
template <typename Real>
int launch_step2(Ctx& c, const Real* in, Real* out, double dt) {
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    const double h = block_h(c);
    const double ih2 = 1. / (h * h);
    const UniformStepParams<Real> prm{(Real)c.rho, ih2, dt, c.topo.uniform_bpd, bpd_magic(c.topo.uniform_bpd)};
    hipLaunchKernelGGL((rd_step2_kernel<Real, UniformStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0, c.stream,
                       ns, per_xcd, in, out, c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), prm);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

// two fused steps per launch on a clamped state (rd_step2f_kernel); mid, when
// not null, receives the state after the first step
template <typename Real>
int launch_step2f(Ctx& c, const Real* in, Real* out, double dt, Real* mid = nullptr) {
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    const double h = block_h(c);
    const double ih2 = 1. / (h * h);
    const UniformStepParams<Real> prm{(Real)c.rho, ih2, dt, c.topo.uniform_bpd, bpd_magic(c.topo.uniform_bpd)};
    (void)per_xcd;
    XcdSplit split = even_split(ns);
    const int2* items = nullptr;
    if (c.nItems2f > 0) {  // work items of the current prepare (plan_step2f)
        for (int i = 0; i <= 8; ++i) split.start[i] = c.split2f[i];
        items = c.items2f.as<int2>();
    }
    int widest = 0;
    for (int i = 0; i < 8; ++i) widest = std::max(widest, split.start[i + 1] - split.start[i]);
    hipLaunchKernelGGL((rd_step2f_kernel<Real, UniformStepParams<Real>>), dim3(8 * widest), dim3(512), 0, c.stream,
                       split, items, in, out, mid, c.statF.as<Stat4<Real>>(), (const unsigned char*)nullptr, prm);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

template <typename Real, bool FUSED>
int step_t(Ctx& c, const Real* in, Real* out, double dt) {
    switch (c.rd_bpw) {
        case 1: launch_step<Real, FUSED, 1>(c, in, out, dt); break;
        case 4: launch_step<Real, FUSED, 4>(c, in, out, dt); break;
        default: launch_step<Real, FUSED, 2>(c, in, out, dt); break;
    }
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

template <typename Real, bool FUSED>
int step_t(Ctx& c, Real* out) {
    return step_t<Real, FUSED>(c, c.phi[c.cur].as<Real>(), out, c.dt_last);
}

}  // namespace

int rd_ensure_prepared(Ctx& c) {
    if (!c.prepared || !c.constants_stale) return GH_OK;
    return rd_prepare(c, c.Dw, c.Dg, c.rho);
}

int rd_prepare(Ctx& c, double Dw, double Dg, double rho) {
    if (c.topo.uniform_bpd <= 0) {  // blocks on several levels: ghosts across level jumps
        int rc = rd_prepare_ghosted(c, Dw, Dg);
        if (rc) return rc;
        c.Dw = Dw;
        c.Dg = Dg;
        c.rho = rho;
        c.prepared = true;
        c.constants_stale = false;
        return GH_OK;
    }
    return c.real_kind == GH_REAL_F32 ? prepare_t<float>(c, Dw, Dg, rho) : prepare_t<double>(c, Dw, Dg, rho);
}

int rd_rhs(Ctx& c) {
    if (c.topo.uniform_bpd <= 0) return rd_step_ghosted(c, c.phi[c.cur].p, c.dphidt.p, 0.0, false);
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

// Times trial steps of each blocks-per-workgroup tile on the current state
// (saved in the dphidt scratch and restored after every trial) and keeps the
// fastest.
int rd_autotune(Ctx& c, double dt, int trial_steps, int* best_out) {
    static const int cand[4] = {1, 2, 4, 0};
    if (c.topo.uniform_bpd <= 0) {  // the multi-level kernel has a single tile
        if (best_out) *best_out = c.rd_bpw;
        return GH_OK;
    }
    const size_t nb = c.ncells() * c.real_size;
    GH_HIP_CHECK(hipMemcpyAsync(c.dphidt.p, c.phi[c.cur].p, nb, hipMemcpyDeviceToDevice, c.stream));
    const int cur0 = c.cur;
    float best_ms = 1e30f;
    int best = c.rd_bpw;
    for (int k = 0; k < 4; ++k) {
        c.rd_bpw = cand[k];
        int rc = rd_steps(c, 2, dt);  // warm
        if (rc) return rc;
        GH_HIP_CHECK(hipEventRecord(c.ev0, c.stream));
        if ((rc = rd_steps(c, trial_steps, dt))) return rc;
        GH_HIP_CHECK(hipEventRecord(c.ev1, c.stream));
        GH_HIP_CHECK(hipEventSynchronize(c.ev1));
        float ms = 0;
        GH_HIP_CHECK(hipEventElapsedTime(&ms, c.ev0, c.ev1));
        if (ms < best_ms) { best_ms = ms; best = cand[k]; }
        c.cur = cur0;
        GH_HIP_CHECK(hipMemcpyAsync(c.phi[c.cur].p, c.dphidt.p, nb, hipMemcpyDeviceToDevice, c.stream));
    }
    c.rd_bpw = best;
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (best_out) *best_out = best;
    return GH_OK;
}

double block_h_of(const Ctx& c) { return block_h(c); }

int build_stat_fast(hipStream_t stream, int real_kind, size_t n, const void* st, const void* sig, DevBuf& out,
                    DevBuf& flag, bool* ok) {
    const size_t rs = real_kind == GH_REAL_F32 ? 4 : 8;
    int rc;
    if ((rc = out.ensure_grow(std::max<size_t>(1, n) * 4 * rs)) || (rc = flag.ensure(sizeof(unsigned int)))) return rc;
    GH_HIP_CHECK(hipMemsetAsync(flag.p, 0, sizeof(unsigned int), stream));
    const unsigned grid = (unsigned)((n + 255) / 256);
    if (n > 0) {
        if (real_kind == GH_REAL_F32)
            hipLaunchKernelGGL(stat_fast_kernel<float>, dim3(grid), dim3(256), 0, stream, n, (const Stat4<float>*)st,
                               (const float*)sig, out.as<Stat4<float>>(), flag.as<unsigned int>());
        else
            hipLaunchKernelGGL(stat_fast_kernel<double>, dim3(grid), dim3(256), 0, stream, n, (const Stat4<double>*)st,
                               (const double*)sig, out.as<Stat4<double>>(), flag.as<unsigned int>());
        GH_HIP_CHECK(hipGetLastError());
    }
    unsigned int bad = 0;
    GH_HIP_CHECK(hipMemcpyAsync(&bad, flag.p, sizeof(bad), hipMemcpyDeviceToHost, stream));
    GH_HIP_CHECK(hipStreamSynchronize(stream));
    *ok = bad == 0;
    return GH_OK;
}

int rd_step2_buffers(Ctx& c, const void* in, void* out, void* mid, double dt) {
    if (!c.fast_ok || c.topo.uniform_bpd <= 0)
        return set_error(GH_ESTATE, "two-step launch: needs a uniform grid with p_w + p_g >= 0");
    return c.real_kind == GH_REAL_F32 ? launch_step2f<float>(c, (const float*)in, (float*)out, dt, (float*)mid)
                                      : launch_step2f<double>(c, (const double*)in, (double*)out, dt, (double*)mid);
}

int rd_step_buffers(Ctx& c, const void* in, void* out, double dt) {
    return c.real_kind == GH_REAL_F32 ? step_t<float, true>(c, (const float*)in, (float*)out, dt)
                                      : step_t<double, true>(c, (const double*)in, (double*)out, dt);
}

// Tile 0 takes two steps per launch.  A state written by the fused update is
// clamped, so after one single step from whatever state the call starts with,
// the pairs run rd_step2f_kernel when the anatomy allows it (fast_ok).
int rd_steps(Ctx& c, int64_t nsteps, double dt) {
    c.dt_last = dt;
    const bool ghosted = c.topo.uniform_bpd <= 0;
    const bool f32 = c.real_kind == GH_REAL_F32;
    int64_t k = 0;
    if (!ghosted && c.rd_bpw == 0) {
        const bool fast = c.fast_ok && nsteps >= 3;
        if (fast) {
            int rc = f32 ? step_t<float, true>(c, c.phi[1 - c.cur].as<float>())
                         : step_t<double, true>(c, c.phi[1 - c.cur].as<double>());
            if (rc) return rc;
            c.cur = 1 - c.cur;
            k = 1;
        }
        for (; k + 1 < nsteps; k += 2) {
            const int cur = c.cur;
            int rc = fast ? (f32 ? launch_step2f<float>(c, c.phi[cur].as<float>(), c.phi[1 - cur].as<float>(), dt)
                                 : launch_step2f<double>(c, c.phi[cur].as<double>(), c.phi[1 - cur].as<double>(), dt))
                          : (f32 ? launch_step2<float>(c, c.phi[cur].as<float>(), c.phi[1 - cur].as<float>(), dt)
                                 : launch_step2<double>(c, c.phi[cur].as<double>(), c.phi[1 - cur].as<double>(), dt));
            if (rc) return rc;
            c.cur = 1 - c.cur;
        }
    }
    for (; k < nsteps; ++k) {
        int rc = ghosted ? rd_step_ghosted(c, c.phi[c.cur].p, c.phi[1 - c.cur].p, dt, true)
                 : c.real_kind == GH_REAL_F32 ? step_t<float, true>(c, c.phi[1 - c.cur].as<float>())
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
