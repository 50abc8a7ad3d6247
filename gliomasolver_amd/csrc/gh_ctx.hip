// gh_ctx.hip -- context, grid topology, field transfer and schedule entry
// points of the C ABI (include/gliomahip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "gh_anatomy.h"
#include "gh_ctx.h"

namespace gh {

static thread_local std::string g_err;

int set_error(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
const char* last_error() { return g_err.c_str(); }

int DevBuf::ensure(size_t nbytes) {
    if (nbytes <= bytes && p) return GH_OK;
    release();
    if (nbytes == 0) return GH_OK;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        p = nullptr;
        return set_error(GH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    bytes = nbytes;
    return GH_OK;
}
int DevBuf::ensure_grow(size_t nbytes) {
    if (nbytes <= bytes && p) return GH_OK;
    return ensure(nbytes + nbytes / 4);
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int PinnedBuf::ensure_grow(size_t nbytes) {
    if (nbytes <= bytes && p) return GH_OK;
    release();
    const size_t n = nbytes + nbytes / 4;
    GH_HIP_CHECK(hipHostMalloc(&p, std::max<size_t>(n, 256), hipHostMallocDefault));
    bytes = std::max<size_t>(n, 256);
    return GH_OK;
}
void PinnedBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
}

Ctx* as_ctx(gh_ctx* c) {
    Ctx* r = reinterpret_cast<Ctx*>(c);
    if (r) ++r->calls;  // every entry point on a context passes here once
    return r;
}

int Ctx::alloc_fields() {
    const size_t n = ncells() * real_size;
    int rc;
    if ((rc = phi[0].ensure(n)) || (rc = phi[1].ensure(n)) || (rc = dphidt.ensure(n)) || (rc = pw.ensure(n)) ||
        (rc = pg.ensure(n)) || (rc = pcsf.ensure(n)) || (rc = statD.ensure(4 * n)) ||
        (rc = statSig.ensure(n)))
        return rc;
    GH_HIP_CHECK(hipMemsetAsync(phi[0].p, 0, n, stream));
    GH_HIP_CHECK(hipMemsetAsync(phi[1].p, 0, n, stream));
    GH_HIP_CHECK(hipMemsetAsync(dphidt.p, 0, n, stream));
    GH_HIP_CHECK(hipMemsetAsync(pw.p, 0, n, stream));
    GH_HIP_CHECK(hipMemsetAsync(pg.p, 0, n, stream));
    GH_HIP_CHECK(hipMemsetAsync(pcsf.p, 0, n, stream));
    cur = 0;
    prepared = false;
    return GH_OK;
}

void Ctx::release_all() {
    DevBuf* bufs[] = {&dLevel,   &dNbr27,  &faceEnt,    &faceSrc,   &faceW,     &faceList,    &faceCount, &facePkBase, &scanBuf,     &fgPhi,       &fgTau,   &fgDm,
                      &fgPw,     &fgPg,    &labBuf,     &labSlots,  &labEnt,    &labSrc,      &labW,    &labPerm,  &labAllEnt,
                      &labAllSrc, &labAllW, &labAllSlots, &dDense, &dBlk, &dBad, &spare[0], &spare[1], &spare[2], &spare[3],
                      &labSplit[0], &labSplit[1], &labSplit[2], &labSplit[3], &remapBuf, &detBuf, &phi[0],  &phi[1],    &dphidt,      &pw,      &pg,
                      &pcsf,     &statD,   &statSig,    &statF,     &flag,      &zeroFlux,   &items2f,      &nbr,         &obsT1,   &obsFl,
                      &obsPet,   &obsRoi,  &likPart,    &adjCk,     &adjSeg,    &adjLam[0],   &adjLam[1], &adjPart};
    for (DevBuf* b : bufs) b->release();
    for (DevBuf& b : ckpt) b.release();
    faceScan.release();
    dHashKey.release();
    dHashVal.release();
    for (PinnedBuf* b : {&pinRemap, &pinDense, &pinBlk, &pinSplit, &pinSlots, &pinDet}) b->release();
    labPermOrder = 0;
}

// dense x-fastest n^3 <-> block-major [slot][z][y][x] for a uniform grid in
// slot order bx + bpd*(by + bpd*bz)
template <typename T>
static void dense_to_blocks(int n, const T* dense, T* bm) {
    const int bpd = n / BS;
    for (int bz = 0; bz < bpd; ++bz)
        for (int by = 0; by < bpd; ++by)
            for (int bx = 0; bx < bpd; ++bx) {
                T* dst = bm + ((size_t)bx + (size_t)bpd * (by + (size_t)bpd * bz)) * BCELLS;
                for (int z = 0; z < BS; ++z)
                    for (int y = 0; y < BS; ++y)
                        std::memcpy(dst + BS * (y + BS * z),
                                    dense + (size_t)BS * bx + (size_t)n * ((size_t)BS * by + y + (size_t)n * (BS * bz + z)),
                                    BS * sizeof(T));
            }
}
template <typename T>
static void blocks_to_dense(int n, const T* bm, T* dense) {
    const int bpd = n / BS;
    for (int bz = 0; bz < bpd; ++bz)
        for (int by = 0; by < bpd; ++by)
            for (int bx = 0; bx < bpd; ++bx) {
                const T* src = bm + ((size_t)bx + (size_t)bpd * (by + (size_t)bpd * bz)) * BCELLS;
                for (int z = 0; z < BS; ++z)
                    for (int y = 0; y < BS; ++y)
                        std::memcpy(dense + (size_t)BS * bx + (size_t)n * ((size_t)BS * by + y + (size_t)n * (BS * bz + z)),
                                    src + BS * (y + BS * z), BS * sizeof(T));
            }
}

static int check_dense(Ctx* c, int n) {
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.uniform_bpd <= 0) return set_error(GH_ESTATE, "dense transfer needs a uniform grid");
    if (n != BS * c->topo.uniform_bpd)
        return set_error(GH_EINVAL, "dense size n must equal 8*bpd of the uniform grid");
    return GH_OK;
}

static DevBuf* field_buf(Ctx* c, int field) {
    switch (field) {
        case GH_FIELD_PHI: return &c->phi[c->cur];
        case GH_FIELD_PW: return &c->pw;
        case GH_FIELD_PG: return &c->pg;
        case GH_FIELD_PCSF: return &c->pcsf;
        case GH_FIELD_DPHIDT: return &c->dphidt;
        default: return nullptr;
    }
}

// Glioma_ReactionDiffusion::_ic anatomy part (Glioma/Glioma_ReactionDiffusion.cpp:139-152):
// Real pGM,pWM,pCSF; tissue = (double)(Real)(pWM+pGM); majority vote; normalise.
template <typename Real>
static void anatomy(size_t nc, const double* gm, const double* wm, const double* csf, Real* p_w, Real* p_g,
                    Real* p_csf) {
    for (size_t c = 0; c < nc; ++c) {
        Real pGM = (Real)gm[c], pWM = (Real)wm[c], pCSF = (Real)csf[c];
        double tissue = (double)(Real)(pWM + pGM);
        pCSF = ((double)pCSF > tissue) ? (Real)1. : (Real)0.;
        pWM = ((double)pCSF > tissue) ? (Real)0. : pWM;
        pGM = ((double)pCSF > tissue) ? (Real)0. : pGM;
        tissue = (double)(Real)(pWM + pGM);
        p_w[c] = (tissue > 0.) ? (Real)((double)pWM / tissue) : (Real)0.;
        p_g[c] = (tissue > 0.) ? (Real)((double)pGM / tissue) : (Real)0.;
        p_csf[c] = pCSF;
    }
}

template <typename Real>
static int set_anatomy_t(Ctx* c, int n, const double* gm, const double* wm, const double* csf) {
    const size_t nc = (size_t)n * n * n;
    std::vector<Real> a(nc), b(nc), d(nc), bm(nc);
    anatomy<Real>(nc, gm, wm, csf, a.data(), b.data(), d.data());
    const std::vector<Real>* src[3] = {&a, &b, &d};
    DevBuf* dst[3] = {&c->pw, &c->pg, &c->pcsf};
    for (int f = 0; f < 3; ++f) {
        dense_to_blocks<Real>(n, src[f]->data(), bm.data());
        GH_HIP_CHECK(hipMemcpyAsync(dst[f]->p, bm.data(), nc * sizeof(Real), hipMemcpyHostToDevice, c->stream));
        GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    c->prepared = false;
    return GH_OK;
}

// Glioma_ReactionDiffusion::_ic (Glioma/Glioma_ReactionDiffusion.cpp:79-172) on
// a uniform grid for anatomy data of any size: the cell at x = origin + h*i
// (Real) samples voxel floor(x / (1/brainSizeMax)); majority vote and
// normalisation as above; with tumor_ic, the reference's smoothed ball of
// radius 0.005 and width 2/128 gated by the voted tissue fractions.
template <typename Real>
static int ic_from_anatomy_t(Ctx* c, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                             const double* tumor_ic) {
    const int bpd = c->topo.uniform_bpd;
    const size_t nc = c->ncells();
    const int n = BS * bpd;
    std::vector<Real> pw(nc), pg(nc), pc(nc), tis(nc), phi(nc, (Real)0);
    anatomy_from_data<Real>(bpd, nx, ny, nz, gm, wm, csf, pw.data(), pg.data(), pc.data(), tis.data());
    if (tumor_ic) {
        std::vector<std::pair<size_t, Real>> cells;
        tumour_cells<Real>(bpd, tis.data(), tumor_ic, cells);
        for (const auto& e : cells) phi[e.first] = e.second;
    }
    std::vector<Real> bm(nc);
    const std::vector<Real>* src[4] = {&pw, &pg, &pc, &phi};
    DevBuf* dst[4] = {&c->pw, &c->pg, &c->pcsf, &c->phi[c->cur]};
    for (int f = 0; f < (tumor_ic ? 4 : 3); ++f) {
        dense_to_blocks<Real>(n, src[f]->data(), bm.data());
        GH_HIP_CHECK(hipMemcpyAsync(dst[f]->p, bm.data(), nc * sizeof(Real), hipMemcpyHostToDevice, c->stream));
        GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    }
    c->prepared = false;
    return GH_OK;
}

// Grid(bpd,bpd,bpd): blocks at level log2(bpd), index (bx,by,bz); periodic
// same-level neighbours (MRAG/MRAGcore/MRAGrid.inl:37-39, 940-966).
static int init_uniform(Ctx* c, int bpd) {
    if (bpd < 1 || (bpd & (bpd - 1)) != 0 || bpd > 64)
        return set_error(GH_EINVAL, "bpd must be a power of two in [1,64]");
    int level = 0;
    while ((1 << level) < bpd) ++level;
    // slot order (level, z, y, x) = bx + bpd*(by + bpd*bz) on one level; the
    // old tree's tables may still be uploading
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->tree.init_uniform(level);
    if (c->tree.max_level < level) c->tree.max_level = level;
    int rc = apply_topology(*c);
    if (rc) return rc;
    return c->alloc_fields();
}

template <typename Real>
static void schedule_t(double Dw_phys, double L, double tend, int gpd, double cfl, double* Dw, double* Dg,
                       double* dt, int64_t* nsteps) {
    // Glioma_ReactionDiffusion.cpp:296-320, every quantity in Real
    Real dw = (Real)Dw_phys, LL = (Real)L, te = (Real)tend;
    dw = dw / (Real)(LL * LL);
    Real dg = (Real)(0.1 * (double)dw);
    Real h = (Real)(1. / (double)gpd);
    Real dmax = (dw >= dg) ? dw : dg;
    Real d = (Real)(cfl * (double)h * (double)h / (2. * 3 * (double)dmax));
    Real t = 0;
    int64_t k = 0;
    while (t <= te) {
        t = (Real)(t + d);
        ++k;
    }
    *Dw = dw;
    *Dg = dg;
    *dt = d;
    *nsteps = k;
}

}  // namespace gh

using namespace gh;

extern "C" {

const char* gh_last_error(void) { return gh::last_error(); }

int gh_ctx_create(int device, int real_kind, gh_ctx** out) {
    if (!out) return set_error(GH_EINVAL, "gh_ctx_create: out is NULL");
    *out = nullptr;
    if (real_kind != GH_REAL_F32 && real_kind != GH_REAL_F64)
        return set_error(GH_EINVAL, "gh_ctx_create: real_kind must be GH_REAL_F32 or GH_REAL_F64");
    int ndev = 0;
    GH_HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(GH_EINVAL, "gh_ctx_create: no such device");
    GH_HIP_CHECK(hipSetDevice(device));
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return set_error(GH_ENOMEM, "gh_ctx_create: out of host memory");
    c->device = device;
    c->real_kind = real_kind;
    c->real_size = real_kind == GH_REAL_F32 ? 4 : 8;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        delete c;
        return set_error(GH_EHIP, std::string("gh_ctx_create: ") + hipGetErrorString(e));
    }
    // device address of the allocation-free likelihood's scratch on this
    // device (once per device); a failure here leaves every other entry point
    // working and gh_likelihood_noalloc retries and reports it
    (void)lik_scratch_resolve(device);
    *out = reinterpret_cast<gh_ctx*>(c);
    return GH_OK;
}

int gh_ctx_destroy(gh_ctx* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return GH_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release_all();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evStage) (void)hipEventDestroy(c->evStage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return GH_OK;
}

int gh_sync(gh_ctx* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    return GH_OK;
}

int gh_grid_init_uniform(gh_ctx* ctx, int bpd) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return init_uniform(c, bpd);
}

int gh_grid_nblocks(gh_ctx* ctx, int* nblocks) {
    Ctx* c = as_ctx(ctx);
    if (!c || !nblocks) return set_error(GH_EINVAL, "null argument");
    *nblocks = c->topo.nblocks();
    return GH_OK;
}

int gh_grid_blocks(gh_ctx* ctx, int* level, int* index) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    for (int b = 0; b < c->topo.nblocks(); ++b) {
        if (level) level[b] = c->topo.blocks[b].level;
        if (index)
            for (int d = 0; d < 3; ++d) index[3 * b + d] = c->topo.blocks[b].index[d];
    }
    return GH_OK;
}

int gh_set_anatomy(gh_ctx* ctx, int n, const double* gm, const double* wm, const double* csf) {
    Ctx* c = as_ctx(ctx);
    int rc = check_dense(c, n);
    if (rc) return rc;
    if (!gm || !wm || !csf) return set_error(GH_EINVAL, "gh_set_anatomy: null array");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return c->real_kind == GH_REAL_F32 ? set_anatomy_t<float>(c, n, gm, wm, csf)
                                       : set_anatomy_t<double>(c, n, gm, wm, csf);
}

int gh_ic_from_anatomy(gh_ctx* ctx, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                       const double* tumor_ic) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.uniform_bpd <= 0) return set_error(GH_ESTATE, "gh_ic_from_anatomy: needs a uniform grid");
    if (nx <= 0 || ny <= 0 || nz <= 0 || !gm || !wm || !csf)
        return set_error(GH_EINVAL, "gh_ic_from_anatomy: bad argument");
    if (tumor_ic && !(std::isfinite(tumor_ic[0]) && std::isfinite(tumor_ic[1]) && std::isfinite(tumor_ic[2])))
        return set_error(GH_EINVAL, "gh_ic_from_anatomy: the tumour centre must be finite");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return c->real_kind == GH_REAL_F32 ? ic_from_anatomy_t<float>(c, nx, ny, nz, gm, wm, csf, tumor_ic)
                                       : ic_from_anatomy_t<double>(c, nx, ny, nz, gm, wm, csf, tumor_ic);
}

int gh_set_field_dense(gh_ctx* ctx, int field, int n, const void* values) {
    Ctx* c = as_ctx(ctx);
    int rc = check_dense(c, n);
    if (rc) return rc;
    DevBuf* b = field_buf(c, field);
    if (!b || !values) return set_error(GH_EINVAL, "gh_set_field_dense: bad field or null values");
    GH_HIP_CHECK(hipSetDevice(c->device));
    const size_t nc = c->ncells();
    std::vector<unsigned char> bm(nc * c->real_size);
    if (c->real_kind == GH_REAL_F32)
        dense_to_blocks<float>(n, (const float*)values, (float*)bm.data());
    else
        dense_to_blocks<double>(n, (const double*)values, (double*)bm.data());
    GH_HIP_CHECK(hipMemcpyAsync(b->p, bm.data(), bm.size(), hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (field == GH_FIELD_PW || field == GH_FIELD_PG) c->prepared = false;
    return GH_OK;
}

int gh_get_field_dense(gh_ctx* ctx, int field, int n, void* values) {
    Ctx* c = as_ctx(ctx);
    int rc = check_dense(c, n);
    if (rc) return rc;
    DevBuf* b = field_buf(c, field);
    if (!b || !values) return set_error(GH_EINVAL, "gh_get_field_dense: bad field or null values");
    GH_HIP_CHECK(hipSetDevice(c->device));
    const size_t nc = c->ncells();
    std::vector<unsigned char> bm(nc * c->real_size);
    GH_HIP_CHECK(hipMemcpyAsync(bm.data(), b->p, bm.size(), hipMemcpyDeviceToHost, c->stream));
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (c->real_kind == GH_REAL_F32)
        blocks_to_dense<float>(n, (const float*)bm.data(), (float*)values);
    else
        blocks_to_dense<double>(n, (const double*)bm.data(), (double*)values);
    return GH_OK;
}

int gh_schedule(int real_kind, double Dw_phys, double L, double tend, int gpd, double cfl, double* Dw, double* Dg,
                double* dt, int64_t* nsteps) {
    if (!Dw || !Dg || !dt || !nsteps) return set_error(GH_EINVAL, "gh_schedule: null output");
    if (gpd <= 0 || !(L > 0) || !(Dw_phys > 0) || !(cfl > 0) || !(tend >= 0) || !(tend < 1e7))
        return set_error(GH_EINVAL, "gh_schedule: bad argument");
    if (real_kind == GH_REAL_F32)
        schedule_t<float>(Dw_phys, L, tend, gpd, cfl, Dw, Dg, dt, nsteps);
    else if (real_kind == GH_REAL_F64)
        schedule_t<double>(Dw_phys, L, tend, gpd, cfl, Dw, Dg, dt, nsteps);
    else
        return set_error(GH_EINVAL, "gh_schedule: bad real_kind");
    return GH_OK;
}

int gh_rd_prepare(gh_ctx* ctx, double Dw, double Dg, double rho) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.nblocks() == 0) return set_error(GH_ESTATE, "gh_rd_prepare: no grid");
    // statics are invalidated by every change of p_w, p_g or the grid
    if (c->prepared && !c->constants_stale && c->Dw == Dw && c->Dg == Dg && c->rho == rho) return GH_OK;
    GH_HIP_CHECK(hipSetDevice(c->device));
    return rd_prepare(*c, Dw, Dg, rho);
}

int gh_rd_rhs(gh_ctx* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (!c->prepared) return set_error(GH_ESTATE, "gh_rd_rhs: call gh_rd_prepare first");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = rd_ensure_prepared(*c);
    if (rc) return rc;
    return rd_rhs(*c);
}

int gh_rd_update(gh_ctx* ctx, double dt) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.nblocks() == 0) return set_error(GH_ESTATE, "gh_rd_update: no grid");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return rd_update(*c, dt);
}

int gh_rd_steps(gh_ctx* ctx, int64_t nsteps, double dt, double* elapsed_ms) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (!c->prepared) return set_error(GH_ESTATE, "gh_rd_steps: call gh_rd_prepare first");
    if (nsteps < 0) return set_error(GH_EINVAL, "gh_rd_steps: nsteps < 0");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = rd_ensure_prepared(*c);
    if (rc) return rc;
    if (elapsed_ms) GH_HIP_CHECK(hipEventRecord(c->ev0, c->stream));
    rc = rd_steps(*c, nsteps, dt);
    if (rc) return rc;
    if (elapsed_ms) {
        GH_HIP_CHECK(hipEventRecord(c->ev1, c->stream));
        GH_HIP_CHECK(hipEventSynchronize(c->ev1));
        float ms = 0;
        GH_HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        *elapsed_ms = ms;
    }
    return GH_OK;
}

int gh_rd_set_tile(gh_ctx* ctx, int blocks_per_workgroup) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (blocks_per_workgroup != 0 && blocks_per_workgroup != 1 && blocks_per_workgroup != 2 &&
        blocks_per_workgroup != 4)
        return set_error(GH_EINVAL, "gh_rd_set_tile: 1, 2 or 4 blocks per workgroup, or 0 (two steps per launch)");
    c->rd_bpw = blocks_per_workgroup;
    return GH_OK;
}

int gh_rd_schedule_info(gh_ctx* ctx, int* tile, int* clamped_pairs) {
    Ctx* c = as_ctx(ctx);
    if (!c || !tile || !clamped_pairs) return set_error(GH_EINVAL, "gh_rd_schedule_info: NULL argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = rd_ensure_prepared(*c);
    if (rc) return rc;
    *tile = c->rd_bpw;
    *clamped_pairs = (c->prepared && c->fast_ok && c->topo.uniform_bpd > 0) ? 1 : 0;
    return GH_OK;
}

int gh_rd_zero_flux_blocks(gh_ctx* ctx, int* nblocks) {
    Ctx* c = as_ctx(ctx);
    if (!c || !nblocks) return set_error(GH_EINVAL, "gh_rd_zero_flux_blocks: NULL argument");
    if (!c->prepared) return set_error(GH_ESTATE, "gh_rd_zero_flux_blocks: call gh_rd_prepare first");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = rd_ensure_prepared(*c);
    if (rc) return rc;
    *nblocks = c->nZeroFlux;
    return GH_OK;
}

int gh_rd_autotune(gh_ctx* ctx, double dt, int trial_steps, int* best) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (!c->prepared) return set_error(GH_ESTATE, "gh_rd_autotune: call gh_rd_prepare first");
    if (trial_steps < 1 || trial_steps > 100000) return set_error(GH_EINVAL, "gh_rd_autotune: bad trial_steps");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = rd_ensure_prepared(*c);
    if (rc) return rc;
    return rd_autotune(*c, dt, trial_steps, best);
}

int gh_checkpoint_save(gh_ctx* ctx, int slot) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (slot < 0 || slot > 4096) return set_error(GH_EINVAL, "gh_checkpoint_save: slot out of range");
    if (c->topo.nblocks() == 0) return set_error(GH_ESTATE, "gh_checkpoint_save: no grid");
    GH_HIP_CHECK(hipSetDevice(c->device));
    if ((int)c->ckpt.size() <= slot) {
        c->ckpt.resize(slot + 1);
        c->ckptVersion.resize(slot + 1, ~(uint64_t)0);
    }
    const size_t nb = c->ncells() * c->real_size;
    int rc = c->ckpt[slot].ensure(nb);
    if (rc) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(c->ckpt[slot].p, c->phi[c->cur].p, nb, hipMemcpyDeviceToDevice, c->stream));
    c->ckptVersion[slot] = c->tree_version;
    return GH_OK;
}

int gh_checkpoint_load(gh_ctx* ctx, int slot) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    const size_t nb = c->ncells() * c->real_size;
    if (slot < 0 || slot >= (int)c->ckpt.size() || c->ckpt[slot].p == nullptr)
        return set_error(GH_EINVAL, "gh_checkpoint_load: slot was never saved");
    if (c->ckptVersion[slot] != c->tree_version || c->ckpt[slot].bytes < nb)
        return set_error(GH_ESTATE, "gh_checkpoint_load: the grid changed since the slot was saved");
    GH_HIP_CHECK(hipSetDevice(c->device));
    GH_HIP_CHECK(hipMemcpyAsync(c->phi[c->cur].p, c->ckpt[slot].p, nb, hipMemcpyDeviceToDevice, c->stream));
    return GH_OK;
}

}  // extern "C"
