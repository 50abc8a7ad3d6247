// gh_batch.hip -- many independent solves of one TMCMC stage on one device.
//
// What the reference runs once per sample (tools/pi4u_lite/Inference/
// fitfun_glioma.c:118-152 -> runAll.sh -> `brain -UQ 1` then `likelihood`),
// here for a whole population at once: every sample has its own uniform
// periodic grid of bpd^3 blocks (bpd need not be a power of two), its own
// Dw, rho, Tend and tumour centre, its schedule (Glioma_ReactionDiffusion.cpp:
// 302-320) and the reference initial condition (:79-172) on the anatomy it
// names; several samples may share one anatomy and its observations.
//
// Layout: the samples' blocks are packed slot range after slot range, the
// samples sorted by step count (descending), so at step k the samples still
// running occupy a prefix of the slots and one launch of the single-solve step
// kernels (rd_step2_kernel / rd_step_kernel with per-slot parameters)
// advances all of them by two (or one) Euler steps.  Finished samples keep
// their final state in the buffer that was current when they ended.
// The arithmetic per cell is the single-solve path's, so every sample is
// bit-identical to solving it alone (gh_ic_from_anatomy + gh_rd_prepare +
// gh_rd_steps + gh_likelihood); the likelihood uses the same fixed 1024-way
// partition and reduction order per sample.
#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "gh_anatomy.h"
#include "gh_ctx.h"
#include "gh_lik_common.h"
#include "gh_adjoint_common.h"
#include "gh_rd_common.h"

#pragma clang fp contract(off)

namespace gh {
namespace {

struct BatchAnatomy {
    int bpd = 0;
    size_t nslots = 0;
    double L = 0;                // (Real)(0.1 * max(nx, ny, nz)), the characteristic length in cm
    DevBuf pw, pg;               // Real, block-major
    std::vector<unsigned char> tissue;  // host, Real per dense cell: voted pGM + pWM (gates the tumour)
    std::vector<unsigned char> gate;    // host, per dense cell: p_w + p_g > 0 (gates a Gaussian IC)
    DevBuf t1, fl, pet, roi;     // observations (block-major), ROI {cell, ix, iy, iz}
    int nroi = 0;
    bool has_obs = false;
};

// per running position (samples sorted by step count)
template <typename Real>
struct SlotRange {
    size_t slot0;    // first slot of the sample
    size_t aslot0;   // the anatomy's first slot is at aslot0 in its own arrays (always 0)
    Real rho, Dw, Dg;
    double ih2, dt;
    double rho_d, invN, invDw;  // adjoint: rho and 1/Dw as given, 1/N
    int bpd;                    // blocks per side of the sample's grid
    uint32_t magic;             // bpd_magic(bpd)
};

struct BatchSample {
    int anat;
    double Dw, Dg, rho, dt;  // schedule values (Real, held in double); adjoint: as given, dt = T/N
    int64_t nsteps;
    double ic[3];            // reference IC: centre in the unit cube; adjoint: Gaussian centre in cells
    double T;                // adjoint only
};

template <typename Real>
struct BatchAdjParams {
    const int32_t* __restrict__ slot_pos;
    const SlotRange<Real>* __restrict__ range;
    __device__ inline void get(int s, double& r, double& i, double& d, double& n, double& w) const {
        const SlotRange<Real>& q = range[slot_pos[s]];
        r = q.rho_d;
        i = q.ih2;
        d = q.dt;
        n = q.invN;
        w = q.invDw;
    }
    __device__ inline BlockPos pos(int s) const {
        const SlotRange<Real>& q = range[slot_pos[s]];
        return block_pos(s, (int)q.slot0, q.bpd, q.magic);
    }
};

// one sample's reverse-sweep seed and initial-condition gradient inputs
template <typename Real>
struct AdjJob {
    const Real* u;        // final state (seed) or initial state (IC gradient)
    double* lam;          // the sample's adjoint range
    const unsigned char* t1;
    const unsigned char* fl;
    const double* pet;
    size_t ncells;
    int bpd;
    LikParams p;
    double x0, y0, z0, is2;
};

template <typename Real>
__global__ __launch_bounds__(256) void batch_seed_kernel(const AdjJob<Real>* __restrict__ jobs, int full_bern) {
    const AdjJob<Real>& j = jobs[blockIdx.y];
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= j.ncells) return;
    Layout lay{0, 0, 0, j.bpd};
    lik_seed_cell<Real>(c, lay, j.u, j.t1, j.fl, j.pet, j.p, full_bern, j.lam);
}

// each sample's seed masked by [u_N < 1] (adj_mask_kernel per sample)
template <typename Real>
__global__ __launch_bounds__(256) void batch_mask_kernel(const AdjJob<Real>* __restrict__ jobs) {
    const AdjJob<Real>& j = jobs[blockIdx.y];
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < j.ncells) adj_mask_cell(c, j.lam, j.u);
}

template <typename Real>
__global__ __launch_bounds__(256) void batch_ic_partials(const AdjJob<Real>* __restrict__ jobs,
                                                         double* __restrict__ partial) {
    const AdjJob<Real>& j = jobs[blockIdx.y];
    adj_ic_body<Real>((int)blockIdx.x, (int)gridDim.x, j.ncells, j.bpd, j.lam, j.u, j.x0, j.y0, j.z0, j.is2,
                      partial + (size_t)blockIdx.y * 3 * gridDim.x);
}

template <typename Real>
struct BatchStepParams {
    const int32_t* __restrict__ slot_pos;
    const SlotRange<Real>* __restrict__ range;
    __device__ inline void get(int s, Real& r, double& i, double& d) const {
        const SlotRange<Real>& q = range[slot_pos[s]];
        r = q.rho;
        i = q.ih2;
        d = q.dt;
    }
    __device__ inline BlockPos pos(int s) const {
        const SlotRange<Real>& q = range[slot_pos[s]];
        return block_pos(s, (int)q.slot0, q.bpd, q.magic);
    }
};

// Statics of every slot (rd_statics_kernel per sample): anatomy cell =
// (slot - slot0) * 512 + lane, Dw, Dg of the sample.
template <typename Real>
__global__ __launch_bounds__(256) void batch_statics_kernel(size_t nslots, const int32_t* __restrict__ nbr,
                                                            const int32_t* __restrict__ slot_pos,
                                                            const SlotRange<Real>* __restrict__ range,
                                                            const Real* const* __restrict__ apw,
                                                            const Real* const* __restrict__ apg,
                                                            Stat4<Real>* __restrict__ st, Real* __restrict__ sig) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nslots * BCELLS) return;
    const size_t s = gid / BCELLS;
    const int l = (int)(gid % BCELLS);
    const int x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    const int pos = slot_pos[s];
    const SlotRange<Real>& q = range[pos];
    const Real* pw = apw[pos];
    const Real* pg = apg[pos];
    const size_t base = q.slot0 * BCELLS;
    const size_t a = gid - base;
    const Real pwc = pw[a], pgc = pg[a];
    const Real dc = diffusivity(pwc, pgc, q.Dw, q.Dg);
    const size_t nx = (x < 7) ? a + 1 : (size_t)nbr[6 * s + FXP] * BCELLS + (l - 7) - base;
    const size_t ny = (y < 7) ? a + 8 : (size_t)nbr[6 * s + FYP] * BCELLS + (l - 56) - base;
    const size_t nz = (z < 7) ? a + 64 : (size_t)nbr[6 * s + FZP] * BCELLS + (l - 448) - base;
    Stat4<Real> r;
    r.dx = harmonic(diffusivity(pw[nx], pg[nx], q.Dw, q.Dg), dc);
    r.dy = harmonic(diffusivity(pw[ny], pg[ny], q.Dw, q.Dg), dc);
    r.dz = harmonic(diffusivity(pw[nz], pg[nz], q.Dw, q.Dg), dc);
    r.tau = chf_threshold(pwc, pgc);
    st[gid] = r;
    sig[gid] = domain_threshold(pwc, pgc);
}

template <typename Real>
__global__ __launch_bounds__(256) void scatter_kernel(size_t n, const uint64_t* __restrict__ idx,
                                                      const Real* __restrict__ val, Real* __restrict__ phi) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) phi[idx[i]] = val[i];
}

// One sample's inputs to the likelihood, blockIdx.y = sample position.
template <typename T>
struct LikJob {
    const T* u;
    const unsigned char* t1;
    const unsigned char* fl;
    const double* pet;
    const int4* roi;
    int nroi;
    int bpd;
    size_t ncells;
    LikParams p;
};

template <typename T>
__global__ __launch_bounds__(LIK_THREADS) void batch_lik_partials_kernel(const LikJob<T>* __restrict__ jobs,
                                                                         double* __restrict__ partials) {
    const LikJob<T>& j = jobs[blockIdx.y];
    Layout lay{0, 0, 0, j.bpd};
    lik_partials_body<T>((int)blockIdx.x, (int)gridDim.x, j.ncells, lay, j.u, j.t1, j.fl, j.pet, j.roi, j.nroi, j.p,
                         partials + (size_t)blockIdx.y * 4 * LIK_GRID);
}

__global__ __launch_bounds__(256) void batch_lik_final_kernel(const double* __restrict__ partials,
                                                              double* __restrict__ out) {
    lik_final_body(LIK_GRID, partials + (size_t)blockIdx.x * 4 * LIK_GRID, out + 4 * (size_t)blockIdx.x);
}

// uniform periodic neighbours of block (bx,by,bz) of a bpd^3 grid, slot order
// bx + bpd*(by + bpd*bz), faces -x,+x,-y,+y,-z,+z
void periodic_neighbours(int bpd, size_t slot0, std::vector<int32_t>& nbr) {
    for (int bz = 0; bz < bpd; ++bz)
        for (int by = 0; by < bpd; ++by)
            for (int bx = 0; bx < bpd; ++bx) {
                auto id = [&](int x, int y, int z) {
                    x = (x + bpd) % bpd;
                    y = (y + bpd) % bpd;
                    z = (z + bpd) % bpd;
                    return (int32_t)(slot0 + (size_t)x + (size_t)bpd * ((size_t)y + (size_t)bpd * z));
                };
                nbr.push_back(id(bx - 1, by, bz));
                nbr.push_back(id(bx + 1, by, bz));
                nbr.push_back(id(bx, by - 1, bz));
                nbr.push_back(id(bx, by + 1, bz));
                nbr.push_back(id(bx, by, bz - 1));
                nbr.push_back(id(bx, by, bz + 1));
            }
}

// dense (x fastest, n = 8*bpd) -> block-major index of a uniform grid
inline size_t block_index(int bpd, size_t c) {
    const size_t n = (size_t)BS * bpd;
    const int i = (int)(c % n), j = (int)((c / n) % n), k = (int)(c / (n * n));
    const size_t slot = (size_t)(i / BS) + (size_t)bpd * ((size_t)(j / BS) + (size_t)bpd * (k / BS));
    return slot * BCELLS + (size_t)(i % BS) + BS * ((size_t)(j % BS) + BS * (size_t)(k % BS));
}

}  // namespace

struct Batch {
    int device = 0, real_kind = GH_REAL_F32;
    size_t real_size = 4;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::unique_ptr<BatchAnatomy>> anat;
    std::vector<BatchSample> smp;   // by sample index
    std::vector<int> order;         // position -> sample index
    std::vector<int> pos_of;        // sample index -> position
    std::vector<size_t> slot0;      // per position, plus the total
    DevBuf phi[2], statD, statSig, nbr, slotPos, range, apw, apg, likPart, likOut, likJobs;
    DevBuf statF, flag;  // rd_step2f_kernel records (valid iff fast_ok), count scratch
    bool fast_ok = false;
    DevBuf zeroFlux;                 // per slot: 1 = block without diffusion (zero_flux_kernel)
    DevBuf ck, tape, lam[2], apart, icpart, adjJobs;  // adjoint
    double ic_is2 = 1.0;                                 // adjoint: 1 / sigma^2 of the Gaussian IC
    std::vector<int> final_buf;                          // per position: phi buffer of the final state
    bool ran = false, adjoint = false;
    size_t nslots() const { return slot0.empty() ? 0 : slot0.back(); }
};

namespace {

Batch* as_batch(gh_batch* b) { return reinterpret_cast<Batch*>(b); }

template <typename Real>
int add_anatomy_t(Batch& b, BatchAnatomy& a, int nx, int ny, int nz, const double* gm, const double* wm,
                  const double* csf) {
    const int n = BS * a.bpd;
    const size_t nc = (size_t)n * n * n;
    std::vector<Real> pw(nc), pg(nc), pc(nc), tis(nc), bm(nc);
    anatomy_from_data<Real>(a.bpd, nx, ny, nz, gm, wm, csf, pw.data(), pg.data(), pc.data(), tis.data());
    a.tissue.resize(nc * sizeof(Real));
    std::memcpy(a.tissue.data(), tis.data(), nc * sizeof(Real));
    a.gate.resize(nc);
    for (size_t c = 0; c < nc; ++c) a.gate[c] = ((double)pw[c] + (double)pg[c] > 0.) ? 1 : 0;
    const int bmax = std::max(nx, std::max(ny, nz));
    a.L = (double)(Real)(bmax * 0.1);  // Glioma_ReactionDiffusion.cpp:95
    int rc;
    if ((rc = a.pw.ensure(nc * sizeof(Real))) || (rc = a.pg.ensure(nc * sizeof(Real)))) return rc;
    const std::vector<Real>* src[2] = {&pw, &pg};
    DevBuf* dst[2] = {&a.pw, &a.pg};
    for (int f = 0; f < 2; ++f) {
        for (size_t c = 0; c < nc; ++c) bm[block_index(a.bpd, c)] = (*src[f])[c];
        GH_HIP_CHECK(hipMemcpyAsync(dst[f]->p, bm.data(), nc * sizeof(Real), hipMemcpyHostToDevice, b.stream));
        GH_HIP_CHECK(hipStreamSynchronize(b.stream));
    }
    return GH_OK;
}

// Positions (samples by step count, descending), neighbour tables, per-slot
// parameters, statics, and phi[0] = 0.  b.smp must be complete.
template <typename Real>
int build_layout_t(Batch& b) {
    const int ns = (int)b.smp.size();
    b.order.resize(ns);
    std::iota(b.order.begin(), b.order.end(), 0);
    std::stable_sort(b.order.begin(), b.order.end(),
                     [&](int x, int y) { return b.smp[x].nsteps > b.smp[y].nsteps; });
    b.pos_of.assign(ns, 0);
    b.slot0.assign(ns + 1, 0);
    for (int p = 0; p < ns; ++p) {
        b.pos_of[b.order[p]] = p;
        b.slot0[p + 1] = b.slot0[p] + b.anat[b.smp[b.order[p]].anat]->nslots;
    }
    const size_t nsl = b.nslots();
    if (nsl > (size_t)INT32_MAX / 8) return set_error(GH_EINVAL, "gh_batch: too many blocks");
    const size_t nc = nsl * BCELLS;
    int rc;
    if ((rc = b.phi[0].ensure(nc * sizeof(Real))) || (rc = b.phi[1].ensure(nc * sizeof(Real))) ||
        (rc = b.statD.ensure(nc * sizeof(Stat4<Real>))) || (rc = b.statSig.ensure(nc * sizeof(Real))) ||
        (rc = b.nbr.ensure(6 * nsl * sizeof(int32_t))) || (rc = b.slotPos.ensure(nsl * sizeof(int32_t))) ||
        (rc = b.range.ensure(ns * sizeof(SlotRange<Real>))) || (rc = b.apw.ensure(ns * sizeof(void*))) ||
        (rc = b.apg.ensure(ns * sizeof(void*))))
        return rc;
    std::vector<int32_t> nbr, spos(nsl);
    nbr.reserve(6 * nsl);
    std::vector<SlotRange<Real>> rng(ns);
    std::vector<const void*> apw(ns), apg(ns);
    for (int p = 0; p < ns; ++p) {
        const BatchSample& q = b.smp[b.order[p]];
        const BatchAnatomy& a = *b.anat[q.anat];
        periodic_neighbours(a.bpd, b.slot0[p], nbr);
        for (size_t s = b.slot0[p]; s < b.slot0[p + 1]; ++s) spos[s] = p;
        const Real hr = (Real)(1.0 / (double)(BS * a.bpd));  // info.h[0] of a uniform grid
        rng[p].slot0 = b.slot0[p];
        rng[p].aslot0 = 0;
        rng[p].rho = (Real)q.rho;
        rng[p].Dw = (Real)q.Dw;
        rng[p].Dg = (Real)q.Dg;
        rng[p].ih2 = 1. / ((double)hr * (double)hr);
        rng[p].dt = q.dt;
        rng[p].rho_d = q.rho;
        rng[p].invN = 1.0 / (double)q.nsteps;
        rng[p].invDw = 1.0 / q.Dw;
        rng[p].bpd = a.bpd;
        rng[p].magic = bpd_magic(a.bpd);
        apw[p] = a.pw.p;
        apg[p] = a.pg.p;
    }
    GH_HIP_CHECK(hipMemcpyAsync(b.nbr.p, nbr.data(), nbr.size() * sizeof(int32_t), hipMemcpyHostToDevice, b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(b.slotPos.p, spos.data(), spos.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(b.range.p, rng.data(), rng.size() * sizeof(SlotRange<Real>), hipMemcpyHostToDevice,
                                b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(b.apw.p, apw.data(), ns * sizeof(void*), hipMemcpyHostToDevice, b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(b.apg.p, apg.data(), ns * sizeof(void*), hipMemcpyHostToDevice, b.stream));
    const unsigned g = (unsigned)((nc + 255) / 256);
    hipLaunchKernelGGL(batch_statics_kernel<Real>, dim3(g), dim3(256), 0, b.stream, nsl, b.nbr.as<int32_t>(),
                       b.slotPos.as<int32_t>(), b.range.as<SlotRange<Real>>(), b.apw.as<const Real* const>(),
                       b.apg.as<const Real* const>(), b.statD.as<Stat4<Real>>(), b.statSig.as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    if ((rc = build_stat_fast(b.stream, sizeof(Real) == 4 ? GH_REAL_F32 : GH_REAL_F64, nc, b.statD.p, b.statSig.p,
                              b.statF, b.flag, &b.fast_ok)))
        return rc;
    if (b.fast_ok && nsl > 0) {  // blocks without diffusion
        if ((rc = b.zeroFlux.ensure_grow(nsl))) return rc;
        const BatchStepParams<Real> prm{b.slotPos.as<int32_t>(), b.range.as<SlotRange<Real>>()};
        hipLaunchKernelGGL((zero_flux_kernel<Real, BatchStepParams<Real>>), dim3((unsigned)nsl), dim3(256), 0,
                           b.stream, (int)nsl, b.statF.as<Stat4<Real>>(), prm, b.zeroFlux.as<unsigned char>());
        GH_HIP_CHECK(hipGetLastError());
    }
    GH_HIP_CHECK(hipMemsetAsync(b.phi[0].p, 0, nc * sizeof(Real), b.stream));
    GH_HIP_CHECK(hipStreamSynchronize(b.stream));
    return GH_OK;
}

// phi[0][idx] = val for the nonzero initial cells of every sample
template <typename Real>
int scatter_ic(Batch& b, const std::vector<uint64_t>& idx, const std::vector<Real>& val) {
    if (idx.empty()) return GH_OK;
    DevBuf di, dv;
    int rc;
    if ((rc = di.ensure(idx.size() * sizeof(uint64_t))) || (rc = dv.ensure(val.size() * sizeof(Real)))) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(di.p, idx.data(), idx.size() * sizeof(uint64_t), hipMemcpyHostToDevice, b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(dv.p, val.data(), val.size() * sizeof(Real), hipMemcpyHostToDevice, b.stream));
    hipLaunchKernelGGL(scatter_kernel<Real>, dim3((unsigned)((idx.size() + 255) / 256)), dim3(256), 0, b.stream,
                       idx.size(), di.as<uint64_t>(), dv.as<Real>(), b.phi[0].as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    GH_HIP_CHECK(hipStreamSynchronize(b.stream));
    return GH_OK;
}

template <typename Real>
int set_samples_t(Batch& b, const double* params, double cfl) {
    const int ns = (int)b.smp.size();
    // schedules: the reference's Real arithmetic (gh_schedule)
    for (int i = 0; i < ns; ++i) {
        BatchSample& q = b.smp[i];
        const BatchAnatomy& a = *b.anat[q.anat];
        const double* pp = params + 6 * i;
        int64_t nsteps = 0;
        int rc = gh_schedule(b.real_kind, pp[0], a.L, pp[2], BS * a.bpd, cfl, &q.Dw, &q.Dg, &q.dt, &nsteps);
        if (rc) return rc;
        q.nsteps = nsteps;
        q.rho = (double)(Real)pp[1];
        q.T = pp[2];
        for (int d = 0; d < 3; ++d) q.ic[d] = pp[3 + d];
    }
    int rc = build_layout_t<Real>(b);
    if (rc) return rc;
    // the reference's tumour: a few hundred cells per sample
    std::vector<uint64_t> idx;
    std::vector<Real> val;
    std::vector<std::pair<size_t, Real>> cells;
    for (int p = 0; p < ns; ++p) {
        const BatchSample& q = b.smp[b.order[p]];
        const BatchAnatomy& a = *b.anat[q.anat];
        tumour_cells<Real>(a.bpd, reinterpret_cast<const Real*>(a.tissue.data()), q.ic, cells);
        for (const auto& e : cells) {
            idx.push_back(b.slot0[p] * BCELLS + block_index(a.bpd, e.first));
            val.push_back(e.second);
        }
    }
    if ((rc = scatter_ic<Real>(b, idx, val))) return rc;
    b.ran = false;
    b.adjoint = false;
    return GH_OK;
}

// The adjoint problem of gh_rd_adjoint per sample: Dw and rho as given
// (Dg = Dw/10), dt = T/N, Gaussian IC peak*exp(-|x-x0|^2/(2 sigma^2)) on
// cells with p_w + p_g > 0 (gh_gaussian_ic), evaluated in double exactly as
// there, only where it does not round to 0 in Real.
template <typename Real>
int set_adjoint_samples_t(Batch& b, const double* params, const int64_t* nsteps, double sigma, double peak) {
    const int ns = (int)b.smp.size();
    for (int i = 0; i < ns; ++i) {
        BatchSample& q = b.smp[i];
        const double* pp = params + 6 * i;
        q.Dw = pp[0];
        q.Dg = pp[0] / 10.0;
        q.rho = pp[1];
        q.T = pp[2];
        q.nsteps = nsteps[i];
        q.dt = pp[2] / (double)nsteps[i];
        for (int d = 0; d < 3; ++d) q.ic[d] = pp[3 + d];
    }
    int rc = build_layout_t<Real>(b);
    if (rc) return rc;
    const double is2 = 1.0 / (2.0 * sigma * sigma);
    // beyond exponent qmax, peak*exp(-q) is below half the smallest positive Real
    // (in logs: half the smallest double denormal is not representable)
    const double log_tiny = sizeof(Real) == 4 ? std::log(1.401298464324817e-45) : std::log(4.940656458412465e-324);
    const double qmax = std::log(std::fabs(peak)) - log_tiny + std::log(2.0) + 2.0;
    const double R = peak == 0 ? -1.0 : std::min(std::sqrt(std::max(qmax, 0.0) / is2) + 1.0, 1.0e6);
    std::vector<uint64_t> idx;
    std::vector<Real> val;
    for (int p = 0; p < ns; ++p) {
        const BatchSample& q = b.smp[b.order[p]];
        const BatchAnatomy& a = *b.anat[q.anat];
        const int n = BS * a.bpd;
        if (R < 0) continue;
        int lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = (int)std::max(0.0, std::floor(q.ic[d] - R));
            hi[d] = (int)std::min((double)(n - 1), std::ceil(q.ic[d] + R));
        }
        for (int k = lo[2]; k <= hi[2]; ++k)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i) {
                    const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                    if (!a.gate[c]) continue;
                    const double d2 = (i - q.ic[0]) * (i - q.ic[0]) + (j - q.ic[1]) * (j - q.ic[1]) +
                                      (k - q.ic[2]) * (k - q.ic[2]);
                    const Real v = (Real)(peak * std::exp(-d2 * is2));
                    if (v == (Real)0) continue;
                    idx.push_back(b.slot0[p] * BCELLS + block_index(a.bpd, c));
                    val.push_back(v);
                }
    }
    if ((rc = scatter_ic<Real>(b, idx, val))) return rc;
    b.ran = false;
    b.adjoint = true;
    return GH_OK;
}

// Launches of the forward sweep: one or two Euler steps over the prefix of
// running samples, reading phi[cur] and writing phi[1 - cur].  Two steps are
// taken when no running sample ends after the first of them and (with
// checkpoints every K) the state after the first is not a checkpoint.  A
// sample's final state is in the buffer current when it ends.
template <typename Real>
struct ForwardSweep {
    Batch& b;
    std::vector<int64_t> npos;  // step count per position (non-increasing)
    BatchStepParams<Real> prm;
    int cur = 0, active = 0;
    explicit ForwardSweep(Batch& bb)
        : b(bb), prm{bb.slotPos.as<int32_t>(), bb.range.as<SlotRange<Real>>()} {
        const int ns = (int)b.smp.size();
        npos.resize(ns);
        for (int p = 0; p < ns; ++p) npos[p] = b.smp[b.order[p]].nsteps;
        active = ns;
        b.final_buf.assign(ns, 0);
    }
    // retire the samples that have reached step k
    void retire(int64_t k) {
        while (active > 0 && npos[active - 1] <= k) b.final_buf[--active] = cur;
    }
    // one launch from step k; returns the steps taken.  From step 1 on every
    // running state was written by the fused update (clamped), so pairs take
    // rd_step2f_kernel when the anatomies allow it.
    int launch(int64_t k, int K) {
        const int nsl = (int)b.slot0[active];
        const bool two = npos[active - 1] >= k + 2 && (K <= 0 || (k + 1) % K != 0) && (k >= 1 || !b.fast_ok);
        const Real* in = b.phi[cur].as<Real>();
        Real* out = b.phi[1 - cur].as<Real>();
        if (two) {
            const int per_xcd = (nsl + 7) / 8;
            if (b.fast_ok)
                hipLaunchKernelGGL((rd_step2f_kernel<Real, BatchStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                                   b.stream, even_split(nsl), (const int2*)nullptr, in, out, (Real*)nullptr,
                                   b.statF.as<Stat4<Real>>(), b.zeroFlux.as<unsigned char>(), prm);
            else
                hipLaunchKernelGGL((rd_step2_kernel<Real, BatchStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                                   b.stream, nsl, per_xcd, in, out, b.statD.as<Stat4<Real>>(), b.statSig.as<Real>(),
                                   prm);
        } else {
            const int per_xcd = ((nsl + 1) / 2 + 7) / 8;  // two blocks per workgroup
            hipLaunchKernelGGL((rd_step_kernel<Real, true, 2, BatchStepParams<Real>>), dim3(8 * per_xcd), dim3(512),
                               0, b.stream, nsl, per_xcd, in, out, b.statD.as<Stat4<Real>>(), b.statSig.as<Real>(),
                               prm);
        }
        cur = 1 - cur;
        return two ? 2 : 1;
    }
};

template <typename Real>
int run_t(Batch& b, double* elapsed_ms) {
    ForwardSweep<Real> fw(b);
    const int64_t maxsteps = fw.npos.empty() ? 0 : fw.npos[0];
    GH_HIP_CHECK(hipEventRecord(b.ev0, b.stream));
    int64_t k = 0;
    while (true) {
        fw.retire(k);
        if (k >= maxsteps) break;
        k += fw.launch(k, 0);
    }
    GH_HIP_CHECK(hipGetLastError());
    GH_HIP_CHECK(hipEventRecord(b.ev1, b.stream));
    GH_HIP_CHECK(hipEventSynchronize(b.ev1));
    float ms = 0;
    GH_HIP_CHECK(hipEventElapsedTime(&ms, b.ev0, b.ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    b.ran = true;
    return GH_OK;
}

// the final state of the sample at position p
template <typename Real>
const Real* final_phi(const Batch& b, int p) {
    return b.phi[b.final_buf[p]].as<Real>() + b.slot0[p] * BCELLS;
}

template <typename Real>
int likelihood_t(Batch& b, const double* lik, int stepPET, int stepTi, int use_roi, double* out) {
    const int ns = (int)b.smp.size();
    std::vector<LikJob<Real>> jobs(ns);
    std::vector<LikParams> prm(ns);
    for (int p = 0; p < ns; ++p) {
        const int i = b.order[p];
        const BatchAnatomy& a = *b.anat[b.smp[i].anat];
        if (!a.has_obs) return set_error(GH_ESTATE, "gh_batch_likelihood: an anatomy has no observations");
        if (use_roi && !a.roi.p) return set_error(GH_ESTATE, "gh_batch_likelihood: an anatomy has no ROI");
        int rc = check_params(lik + 5 * i, stepPET, stepTi, prm[p]);
        if (rc) return rc;
        jobs[p] = LikJob<Real>{final_phi<Real>(b, p), a.t1.as<unsigned char>(), a.fl.as<unsigned char>(),
                               a.pet.as<double>(), use_roi ? a.roi.as<int4>() : nullptr, use_roi ? a.nroi : 0,
                               a.bpd, a.nslots * BCELLS, prm[p]};
    }
    int rc;
    if ((rc = b.likJobs.ensure(ns * sizeof(LikJob<Real>))) ||
        (rc = b.likPart.ensure((size_t)ns * 4 * LIK_GRID * sizeof(double))) ||
        (rc = b.likOut.ensure((size_t)ns * 4 * sizeof(double))))
        return rc;
    GH_HIP_CHECK(hipMemcpyAsync(b.likJobs.p, jobs.data(), ns * sizeof(LikJob<Real>), hipMemcpyHostToDevice, b.stream));
    for (int p0 = 0; p0 < ns; p0 += 65535) {  // gridDim.y limit
        const int cnt = std::min(65535, ns - p0);
        hipLaunchKernelGGL(batch_lik_partials_kernel<Real>, dim3(LIK_GRID, cnt), dim3(LIK_THREADS), 0, b.stream,
                           b.likJobs.as<LikJob<Real>>() + p0, b.likPart.as<double>() + (size_t)p0 * 4 * LIK_GRID);
        GH_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(batch_lik_final_kernel, dim3(ns), dim3(256), 0, b.stream, b.likPart.as<double>(),
                       b.likOut.as<double>());
    GH_HIP_CHECK(hipGetLastError());
    std::vector<double> sums((size_t)ns * 4);
    GH_HIP_CHECK(hipMemcpyAsync(sums.data(), b.likOut.p, sums.size() * sizeof(double), hipMemcpyDeviceToHost,
                                b.stream));
    GH_HIP_CHECK(hipStreamSynchronize(b.stream));
    for (int p = 0; p < ns; ++p) finish(&sums[4 * (size_t)p], prm[p], out + 4 * (size_t)b.order[p], nullptr);
    return GH_OK;
}


// Forward sweep with checkpoints every K steps, loss, seed, reverse sweep
// with per-segment recompute, initial-condition gradient -- gh_rd_adjoint for
// every sample, the samples still running at a step forming a slot prefix in
// both directions (the reverse of sample s starts at its own step N_s - 1,
// with its seed placed in the adjoint buffer of parity N_s).
template <typename Real>
int adjoint_t(Batch& b, int K, const double* lik, int stepPET, int stepTi, int use_roi, double* loss, double* grad,
              double* elapsed_ms) {
    const int ns = (int)b.smp.size();
    const int64_t Nmax = b.smp[b.order[0]].nsteps;
    std::vector<int64_t> npos(ns);
    for (int p = 0; p < ns; ++p) npos[p] = b.smp[b.order[p]].nsteps;
    auto active = [&](int64_t k) {  // positions with N > k (a prefix: npos is non-increasing)
        return (int)(std::upper_bound(npos.begin(), npos.end(), k, [](int64_t v, int64_t e) { return v >= e; }) -
                     npos.begin());
    };
    const size_t slot_bytes = BCELLS * sizeof(Real);
    const int64_t J = (Nmax + K - 1) / K;
    std::vector<size_t> ckoff(J + 1, 0);
    for (int64_t j = 0; j < J; ++j) ckoff[j + 1] = ckoff[j] + b.slot0[active(j * K)];
    const size_t nsl = b.nslots(), nc = nsl * BCELLS;
    constexpr int IC_GRID = 512;
    int rc;
    if ((rc = b.ck.ensure(ckoff[J] * slot_bytes)) || (rc = b.tape.ensure((size_t)(K + 1) * nsl * slot_bytes)) ||
        (rc = b.lam[0].ensure(nc * 8)) || (rc = b.lam[1].ensure(nc * 8)) || (rc = b.apart.ensure(ADJ_PART * nsl * 8)) ||
        (rc = b.icpart.ensure((size_t)ns * 3 * IC_GRID * 8)) || (rc = b.adjJobs.ensure(ns * sizeof(AdjJob<Real>))))
        return rc;
    const BatchStepParams<Real> fprm{b.slotPos.as<int32_t>(), b.range.as<SlotRange<Real>>()};
    const BatchAdjParams<Real> aprm{b.slotPos.as<int32_t>(), b.range.as<SlotRange<Real>>()};
    // blocks without diffusion take the reverse step's short path (flags exist when fast_ok)
    const unsigned char* zflux = b.fast_ok ? b.zeroFlux.as<unsigned char>() : nullptr;
    auto step = [&](const void* in, void* out, int nslots) {  // one Euler step (tape recompute)
        const int per_xcd = ((nslots + 1) / 2 + 7) / 8;
        hipLaunchKernelGGL((rd_step_kernel<Real, true, 2, BatchStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                           b.stream, nslots, per_xcd, (const Real*)in, (Real*)out, b.statD.as<Stat4<Real>>(),
                           b.statSig.as<Real>(), fprm);
    };
    char* ck = (char*)b.ck.p;
    char* tape = (char*)b.tape.p;
    GH_HIP_CHECK(hipEventRecord(b.ev0, b.stream));
    // forward, with a checkpoint of the running samples every K steps
    ForwardSweep<Real> fw(b);
    for (int64_t k = 0;;) {
        fw.retire(k);
        if (k >= Nmax) break;
        if (k % K == 0)
            GH_HIP_CHECK(hipMemcpyAsync(ck + ckoff[k / K] * slot_bytes, b.phi[fw.cur].p, b.slot0[fw.active] * slot_bytes,
                                        hipMemcpyDeviceToDevice, b.stream));
        k += fw.launch(k, K);
    }
    GH_HIP_CHECK(hipGetLastError());
    b.ran = true;
    // loss at every sample's final state
    std::vector<double> out((size_t)ns * 4);
    if ((rc = likelihood_t<Real>(b, lik, stepPET, stepTi, use_roi, out.data()))) return rc;
    // seeds dJ/du at u_N into the adjoint buffer of parity N
    std::vector<AdjJob<Real>> jobs(ns);
    size_t maxc = 0;
    for (int p = 0; p < ns; ++p) {
        const int i = b.order[p];
        const BatchAnatomy& a = *b.anat[b.smp[i].anat];
        LikParams lp;
        if ((rc = check_params(lik + 5 * i, stepPET, stepTi, lp))) return rc;
        jobs[p] = AdjJob<Real>{final_phi<Real>(b, p), b.lam[b.smp[i].nsteps & 1].as<double>() + b.slot0[p] * BCELLS,
                               a.t1.as<unsigned char>(), a.fl.as<unsigned char>(), a.pet.as<double>(),
                               a.nslots * BCELLS, a.bpd, lp, 0, 0, 0, 0};
        maxc = std::max(maxc, a.nslots * BCELLS);
    }
    GH_HIP_CHECK(hipMemcpyAsync(b.adjJobs.p, jobs.data(), ns * sizeof(AdjJob<Real>), hipMemcpyHostToDevice, b.stream));
    for (int p0 = 0; p0 < ns; p0 += 65535)
        hipLaunchKernelGGL(batch_seed_kernel<Real>, dim3((unsigned)((maxc + 255) / 256), std::min(65535, ns - p0)),
                           dim3(256), 0, b.stream, b.adjJobs.as<AdjJob<Real>>() + p0, use_roi ? 0 : 1);
    GH_HIP_CHECK(hipGetLastError());
    if (use_roi) {
        for (int p = 0; p < ns; ++p) {
            const int i = b.order[p];
            const BatchAnatomy& a = *b.anat[b.smp[i].anat];
            if (a.nroi == 0) continue;
            hipLaunchKernelGGL(lik_seed_roi_kernel<Real>, dim3((unsigned)((a.nroi + 255) / 256)), dim3(256), 0,
                               b.stream, a.nroi, a.roi.as<int4>(), jobs[p].u, a.t1.as<unsigned char>(),
                               a.fl.as<unsigned char>(), jobs[p].p, jobs[p].lam);
        }
        GH_HIP_CHECK(hipGetLastError());
    }
    for (int p0 = 0; p0 < ns; p0 += 65535)
        hipLaunchKernelGGL(batch_mask_kernel<Real>, dim3((unsigned)((maxc + 255) / 256), std::min(65535, ns - p0)),
                           dim3(256), 0, b.stream, b.adjJobs.as<AdjJob<Real>>() + p0);
    GH_HIP_CHECK(hipGetLastError());
    // reverse
    GH_HIP_CHECK(hipMemsetAsync(b.apart.p, 0, ADJ_PART * nsl * 8, b.stream));
    for (int64_t j = J - 1; j >= 0; --j) {
        const int64_t k0 = j * K, k1 = std::min<int64_t>(Nmax, k0 + K);
        GH_HIP_CHECK(hipMemcpyAsync(tape, ck + ckoff[j] * slot_bytes, b.slot0[active(k0)] * slot_bytes,
                                    hipMemcpyDeviceToDevice, b.stream));
        // recompute u_{k0+1} .. u_{k1} of the running samples, two steps per
        // launch (both states written) where the input is a clamped state
        // (k >= 1) and both fit the segment; a sample ending after the first of
        // the two gets a second state past its end that no reverse step reads
        for (int64_t k = k0; k < k1;) {
            char* in = tape + (size_t)(k - k0) * nsl * slot_bytes;
            const int nslots = (int)b.slot0[active(k)];
            if (b.fast_ok && k >= 1 && k + 2 <= k1) {
                const int per_xcd = (nslots + 7) / 8;
                hipLaunchKernelGGL((rd_step2f_kernel<Real, BatchStepParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                                   b.stream, even_split(nslots), (const int2*)nullptr, (const Real*)in,
                                   (Real*)(in + 2 * (size_t)nsl * slot_bytes), (Real*)(in + (size_t)nsl * slot_bytes),
                                   b.statF.as<Stat4<Real>>(), b.zeroFlux.as<unsigned char>(), fprm);
                k += 2;
            } else {
                step(in, in + (size_t)nsl * slot_bytes, nslots);
                k += 1;
            }
        }
        for (int64_t k = k1 - 1; k >= k0; --k) {
            const int nslots = (int)b.slot0[active(k)];
            const int per_xcd = (nslots + 7) / 8;
            const Real* uk = (const Real*)(tape + (size_t)(k - k0) * nsl * slot_bytes);
            const int mask_out = k >= 1;
            if (b.fast_ok && k >= 1)  // clamped states: the clamped-state records
                hipLaunchKernelGGL((adj_step_kernel<Real, BatchAdjParams<Real>, true>), dim3(8 * per_xcd), dim3(512),
                                   0, b.stream, nslots, per_xcd, uk, mask_out,
                                   b.lam[(k + 1) & 1].as<double>(),
                                   b.lam[k & 1].as<double>(), b.statF.as<Stat4<Real>>(), (const Real*)nullptr, aprm,
                                   b.apart.as<double>(), zflux);
            else
                hipLaunchKernelGGL((adj_step_kernel<Real, BatchAdjParams<Real>>), dim3(8 * per_xcd), dim3(512), 0,
                                   b.stream, nslots, per_xcd, uk, mask_out,
                                   b.lam[(k + 1) & 1].as<double>(),
                                   b.lam[k & 1].as<double>(), b.statD.as<Stat4<Real>>(), b.statSig.as<Real>(), aprm,
                                   b.apart.as<double>(), zflux);
        }
        GH_HIP_CHECK(hipGetLastError());
    }
    // initial-condition gradient: lambda_0 (buffer 0) against u_0 (checkpoint 0)
    for (int p = 0; p < ns; ++p) {
        const BatchSample& q = b.smp[b.order[p]];
        jobs[p].u = (const Real*)(ck) + b.slot0[p] * BCELLS;
        jobs[p].lam = b.lam[0].as<double>() + b.slot0[p] * BCELLS;
        jobs[p].x0 = q.ic[0];
        jobs[p].y0 = q.ic[1];
        jobs[p].z0 = q.ic[2];
        jobs[p].is2 = b.ic_is2;
    }
    GH_HIP_CHECK(hipMemcpyAsync(b.adjJobs.p, jobs.data(), ns * sizeof(AdjJob<Real>), hipMemcpyHostToDevice, b.stream));
    for (int p0 = 0; p0 < ns; p0 += 65535)
        hipLaunchKernelGGL(batch_ic_partials<Real>, dim3(IC_GRID, std::min(65535, ns - p0)), dim3(256), 0, b.stream,
                           b.adjJobs.as<AdjJob<Real>>() + p0, b.icpart.as<double>() + (size_t)p0 * 3 * IC_GRID);
    GH_HIP_CHECK(hipGetLastError());
    GH_HIP_CHECK(hipEventRecord(b.ev1, b.stream));
    std::vector<double> ap(ADJ_PART * nsl), ip((size_t)ns * 3 * IC_GRID);
    GH_HIP_CHECK(hipMemcpyAsync(ap.data(), b.apart.p, ap.size() * 8, hipMemcpyDeviceToHost, b.stream));
    GH_HIP_CHECK(hipMemcpyAsync(ip.data(), b.icpart.p, ip.size() * 8, hipMemcpyDeviceToHost, b.stream));
    GH_HIP_CHECK(hipStreamSynchronize(b.stream));
    float ms = 0;
    GH_HIP_CHECK(hipEventElapsedTime(&ms, b.ev0, b.ev1));
    if (elapsed_ms) *elapsed_ms = ms;
    for (int p = 0; p < ns; ++p) {
        const int i = b.order[p];
        double* gi = grad + 6 * (size_t)i;
        const BatchSample& q = b.smp[i];
        double ab[2];
        sum_slot_partials(ap.data() + ADJ_PART * b.slot0[p], (int)(b.slot0[p + 1] - b.slot0[p]), ab);
        adj_param_grads(ab, q.dt, 1.0 / q.Dw, 1.0 / (double)q.nsteps, q.rho, gi);
        gi[3] = gi[4] = gi[5] = 0;
        for (int w = 0; w < IC_GRID; ++w)
            for (int qd = 0; qd < 3; ++qd) gi[3 + qd] += ip[((size_t)p * IC_GRID + w) * 3 + qd];
        for (int qd = 0; qd < 4; ++qd) loss[4 * (size_t)i + qd] = out[4 * (size_t)i + qd];
    }
    return GH_OK;
}

}  // namespace
}  // namespace gh

using namespace gh;

extern "C" {

int gh_batch_create(int device, int real_kind, gh_batch** out) {
    if (!out) return set_error(GH_EINVAL, "gh_batch_create: out is NULL");
    *out = nullptr;
    if (real_kind != GH_REAL_F32 && real_kind != GH_REAL_F64) return set_error(GH_EINVAL, "bad real_kind");
    int ndev = 0;
    GH_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(GH_EINVAL, "gh_batch_create: no such device");
    GH_HIP_CHECK(hipSetDevice(device));
    std::unique_ptr<Batch> b(new Batch);
    b->device = device;
    b->real_kind = real_kind;
    b->real_size = real_kind == GH_REAL_F32 ? 4 : 8;
    GH_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    GH_HIP_CHECK(hipEventCreate(&b->ev0));
    GH_HIP_CHECK(hipEventCreate(&b->ev1));
    *out = reinterpret_cast<gh_batch*>(b.release());
    return GH_OK;
}

int gh_batch_destroy(gh_batch* bh) {
    Batch* b = as_batch(bh);
    if (!b) return GH_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& a : b->anat) {
        a->pw.release();
        a->pg.release();
        a->t1.release();
        a->fl.release();
        a->pet.release();
        a->roi.release();
    }
    DevBuf* bufs[] = {&b->phi[0], &b->phi[1], &b->statD,  &b->statSig, &b->nbr,    &b->slotPos, &b->range,
                      &b->apw,    &b->apg,    &b->likPart, &b->likOut, &b->likJobs, &b->ck,     &b->tape,
                      &b->lam[0], &b->lam[1], &b->apart,   &b->icpart, &b->adjJobs, &b->statF,  &b->flag};
    for (DevBuf* d : bufs) d->release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return GH_OK;
}

int gh_batch_add_anatomy(gh_batch* bh, int bpd, int nx, int ny, int nz, const double* gm, const double* wm,
                         const double* csf, int* anatomy_id) {
    Batch* b = as_batch(bh);
    if (!b || !gm || !wm || !csf || !anatomy_id) return set_error(GH_EINVAL, "gh_batch_add_anatomy: NULL argument");
    if (bpd < 1 || bpd > 64 || nx <= 0 || ny <= 0 || nz <= 0)
        return set_error(GH_EINVAL, "gh_batch_add_anatomy: 1 <= bpd <= 64 and positive data sizes");
    GH_HIP_CHECK(hipSetDevice(b->device));
    std::unique_ptr<BatchAnatomy> a(new BatchAnatomy);
    a->bpd = bpd;
    a->nslots = (size_t)bpd * bpd * bpd;
    int rc = b->real_kind == GH_REAL_F32 ? add_anatomy_t<float>(*b, *a, nx, ny, nz, gm, wm, csf)
                                         : add_anatomy_t<double>(*b, *a, nx, ny, nz, gm, wm, csf);
    if (rc) return rc;
    *anatomy_id = (int)b->anat.size();
    b->anat.push_back(std::move(a));
    return GH_OK;
}

int gh_batch_set_observations(gh_batch* bh, int anatomy_id, const double* t1c, const double* flair,
                              const double* pet, const double* roi, int nroi) {
    Batch* b = as_batch(bh);
    if (!b || !t1c || !flair || !pet) return set_error(GH_EINVAL, "gh_batch_set_observations: NULL argument");
    if (anatomy_id < 0 || anatomy_id >= (int)b->anat.size())
        return set_error(GH_EINVAL, "gh_batch_set_observations: no such anatomy");
    if (nroi < 0 || (nroi > 0 && !roi)) return set_error(GH_EINVAL, "gh_batch_set_observations: bad ROI");
    GH_HIP_CHECK(hipSetDevice(b->device));
    BatchAnatomy& a = *b->anat[anatomy_id];
    const int n = BS * a.bpd;
    const size_t nc = a.nslots * BCELLS;
    std::vector<unsigned char> t1(nc), fl(nc);
    std::vector<double> pv(nc);
    for (size_t c = 0; c < nc; ++c) {
        const size_t m = block_index(a.bpd, c);
        t1[m] = t1c[c] == 1;
        fl[m] = flair[c] == 1;
        pv[m] = pet[c];
    }
    int rc;
    if ((rc = a.t1.ensure(nc)) || (rc = a.fl.ensure(nc)) || (rc = a.pet.ensure(nc * sizeof(double)))) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(a.t1.p, t1.data(), nc, hipMemcpyHostToDevice, b->stream));
    GH_HIP_CHECK(hipMemcpyAsync(a.fl.p, fl.data(), nc, hipMemcpyHostToDevice, b->stream));
    GH_HIP_CHECK(hipMemcpyAsync(a.pet.p, pv.data(), nc * sizeof(double), hipMemcpyHostToDevice, b->stream));
    a.roi.release();
    a.nroi = 0;
    if (nroi > 0) {
        std::vector<int4> q(nroi);
        for (int i = 0; i < nroi; ++i) {
            const int ix = (int)roi[i], iy = (int)roi[i + nroi], iz = (int)roi[i + 2 * nroi];
            if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n)
                return set_error(GH_EINVAL, "gh_batch_set_observations: ROI point outside the grid");
            const size_t c = (size_t)ix + (size_t)n * ((size_t)iy + (size_t)n * iz);
            q[i] = make_int4((int)block_index(a.bpd, c), ix, iy, iz);
        }
        if ((rc = a.roi.ensure(nroi * sizeof(int4)))) return rc;
        GH_HIP_CHECK(hipMemcpyAsync(a.roi.p, q.data(), nroi * sizeof(int4), hipMemcpyHostToDevice, b->stream));
        a.nroi = nroi;
    }
    GH_HIP_CHECK(hipStreamSynchronize(b->stream));
    a.has_obs = true;
    return GH_OK;
}

int gh_batch_set_samples(gh_batch* bh, int nsamples, const int* anatomy_id, const double* params, double cfl) {
    Batch* b = as_batch(bh);
    if (!b || nsamples <= 0 || !anatomy_id || !params) return set_error(GH_EINVAL, "gh_batch_set_samples: bad argument");
    if (!(cfl > 0)) return set_error(GH_EINVAL, "gh_batch_set_samples: cfl must be > 0");
    for (int i = 0; i < nsamples; ++i) {
        if (anatomy_id[i] < 0 || anatomy_id[i] >= (int)b->anat.size())
            return set_error(GH_EINVAL, "gh_batch_set_samples: no such anatomy");
        const double* p = params + 6 * i;
        if (!(p[0] > 0) || !(p[1] >= 0) || !(p[2] >= 0) || !std::isfinite(p[0]) || !std::isfinite(p[1]) ||
            !std::isfinite(p[2]) || !std::isfinite(p[3]) || !std::isfinite(p[4]) || !std::isfinite(p[5]))
            return set_error(GH_EINVAL, "gh_batch_set_samples: need Dw > 0, rho >= 0, Tend >= 0 (all finite)");
        const double steps = p[2] * 6.0 * p[0] / (b->anat[anatomy_id[i]]->L * b->anat[anatomy_id[i]]->L) /
                             (cfl / ((double)BS * b->anat[anatomy_id[i]]->bpd * BS * b->anat[anatomy_id[i]]->bpd));
        if (steps > 1e8) return set_error(GH_EINVAL, "gh_batch_set_samples: more than 1e8 steps for a sample");
    }
    GH_HIP_CHECK(hipSetDevice(b->device));
    b->smp.assign(nsamples, BatchSample{});
    for (int i = 0; i < nsamples; ++i) b->smp[i].anat = anatomy_id[i];
    return b->real_kind == GH_REAL_F32 ? set_samples_t<float>(*b, params, cfl)
                                       : set_samples_t<double>(*b, params, cfl);
}

int gh_batch_run(gh_batch* bh, double* elapsed_ms) {
    Batch* b = as_batch(bh);
    if (!b) return set_error(GH_EINVAL, "null batch");
    if (b->smp.empty()) return set_error(GH_ESTATE, "gh_batch_run: no samples");
    if (b->ran) return set_error(GH_ESTATE, "gh_batch_run: already run; set the samples again");
    GH_HIP_CHECK(hipSetDevice(b->device));
    return b->real_kind == GH_REAL_F32 ? run_t<float>(*b, elapsed_ms) : run_t<double>(*b, elapsed_ms);
}

int gh_batch_likelihood(gh_batch* bh, const double* lik_params, int stepPET, int stepTi, int use_roi, double* out) {
    Batch* b = as_batch(bh);
    if (!b || !lik_params || !out) return set_error(GH_EINVAL, "gh_batch_likelihood: NULL argument");
    if (!b->ran) return set_error(GH_ESTATE, "gh_batch_likelihood: call gh_batch_run first");
    GH_HIP_CHECK(hipSetDevice(b->device));
    return b->real_kind == GH_REAL_F32 ? likelihood_t<float>(*b, lik_params, stepPET, stepTi, use_roi, out)
                                       : likelihood_t<double>(*b, lik_params, stepPET, stepTi, use_roi, out);
}

int gh_batch_set_adjoint_samples(gh_batch* bh, int nsamples, const int* anatomy_id, const double* params,
                                 const int64_t* nsteps, double ic_sigma, double ic_peak) {
    Batch* b = as_batch(bh);
    if (!b || nsamples <= 0 || !anatomy_id || !params || !nsteps)
        return set_error(GH_EINVAL, "gh_batch_set_adjoint_samples: bad argument");
    if (!(ic_sigma > 0) || !std::isfinite(ic_peak))
        return set_error(GH_EINVAL, "gh_batch_set_adjoint_samples: ic_sigma > 0 and a finite peak");
    for (int i = 0; i < nsamples; ++i) {
        if (anatomy_id[i] < 0 || anatomy_id[i] >= (int)b->anat.size())
            return set_error(GH_EINVAL, "gh_batch_set_adjoint_samples: no such anatomy");
        const double* p = params + 6 * i;
        if (!(p[0] > 0) || !(p[1] >= 0) || !(p[2] > 0) || !std::isfinite(p[0]) || !std::isfinite(p[1]) ||
            !std::isfinite(p[2]) || !std::isfinite(p[3]) || !std::isfinite(p[4]) || !std::isfinite(p[5]) ||
            nsteps[i] < 1 || nsteps[i] > 100000000)
            return set_error(GH_EINVAL, "gh_batch_set_adjoint_samples: need Dw > 0, rho >= 0, T > 0, 1 <= N <= 1e8");
    }
    GH_HIP_CHECK(hipSetDevice(b->device));
    b->smp.assign(nsamples, BatchSample{});
    for (int i = 0; i < nsamples; ++i) b->smp[i].anat = anatomy_id[i];
    b->ic_is2 = 1.0 / (ic_sigma * ic_sigma);
    return b->real_kind == GH_REAL_F32 ? set_adjoint_samples_t<float>(*b, params, nsteps, ic_sigma, ic_peak)
                                       : set_adjoint_samples_t<double>(*b, params, nsteps, ic_sigma, ic_peak);
}

int gh_batch_adjoint(gh_batch* bh, int ckpt_every, const double* lik_params, int stepPET, int stepTi, int use_roi,
                     double* loss, double* grad, double* elapsed_ms) {
    Batch* b = as_batch(bh);
    if (!b || !lik_params || !loss || !grad) return set_error(GH_EINVAL, "gh_batch_adjoint: NULL argument");
    if (!b->adjoint) return set_error(GH_ESTATE, "gh_batch_adjoint: set the samples with gh_batch_set_adjoint_samples");
    if (b->ran) return set_error(GH_ESTATE, "gh_batch_adjoint: already run; set the samples again");
    if (ckpt_every < 1 || ckpt_every > 4096) return set_error(GH_EINVAL, "gh_batch_adjoint: 1 <= ckpt_every <= 4096");
    GH_HIP_CHECK(hipSetDevice(b->device));
    return b->real_kind == GH_REAL_F32
               ? adjoint_t<float>(*b, ckpt_every, lik_params, stepPET, stepTi, use_roi, loss, grad, elapsed_ms)
               : adjoint_t<double>(*b, ckpt_every, lik_params, stepPET, stepTi, use_roi, loss, grad, elapsed_ms);
}

int gh_batch_sample_info(gh_batch* bh, int sample, double* Dw, double* Dg, double* dt, int64_t* nsteps) {
    Batch* b = as_batch(bh);
    if (!b || sample < 0 || sample >= (int)b->smp.size()) return set_error(GH_EINVAL, "gh_batch_sample_info: no such sample");
    const BatchSample& q = b->smp[sample];
    if (Dw) *Dw = q.Dw;
    if (Dg) *Dg = q.Dg;
    if (dt) *dt = q.dt;
    if (nsteps) *nsteps = q.nsteps;
    return GH_OK;
}

int gh_batch_get_phi(gh_batch* bh, int sample, void* dense) {
    Batch* b = as_batch(bh);
    if (!b || !dense || sample < 0 || sample >= (int)b->smp.size())
        return set_error(GH_EINVAL, "gh_batch_get_phi: bad argument");
    GH_HIP_CHECK(hipSetDevice(b->device));
    const int p = b->pos_of[sample];
    const BatchAnatomy& a = *b->anat[b->smp[sample].anat];
    const size_t nc = a.nslots * BCELLS;
    const int buf = b->ran ? b->final_buf[p] : 0;  // before the run: the initial state
    const unsigned char* src = (const unsigned char*)b->phi[buf].p + b->slot0[p] * BCELLS * b->real_size;
    std::vector<unsigned char> bm(nc * b->real_size);
    GH_HIP_CHECK(hipMemcpyAsync(bm.data(), src, bm.size(), hipMemcpyDeviceToHost, b->stream));
    GH_HIP_CHECK(hipStreamSynchronize(b->stream));
    for (size_t c = 0; c < nc; ++c)
        std::memcpy((unsigned char*)dense + c * b->real_size, bm.data() + block_index(a.bpd, c) * b->real_size,
                    b->real_size);
    return GH_OK;
}

int gh_batch_get_anatomy(gh_batch* bh, int anatomy_id, void* p_w, void* p_g) {
    Batch* b = as_batch(bh);
    if (!b || anatomy_id < 0 || anatomy_id >= (int)b->anat.size())
        return set_error(GH_EINVAL, "gh_batch_get_anatomy: no such anatomy");
    GH_HIP_CHECK(hipSetDevice(b->device));
    const BatchAnatomy& a = *b->anat[anatomy_id];
    const size_t nc = a.nslots * BCELLS;
    std::vector<unsigned char> bm(nc * b->real_size);
    const DevBuf* src[2] = {&a.pw, &a.pg};
    void* dst[2] = {p_w, p_g};
    for (int f = 0; f < 2; ++f) {
        if (!dst[f]) continue;
        GH_HIP_CHECK(hipMemcpyAsync(bm.data(), src[f]->p, bm.size(), hipMemcpyDeviceToHost, b->stream));
        GH_HIP_CHECK(hipStreamSynchronize(b->stream));
        for (size_t c = 0; c < nc; ++c)
            std::memcpy((unsigned char*)dst[f] + c * b->real_size, bm.data() + block_index(a.bpd, c) * b->real_size,
                        b->real_size);
    }
    return GH_OK;
}

}  // extern "C"
