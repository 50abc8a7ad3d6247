// gh_lik.hip -- TMCMC likelihood of a tumour field against imaging data.
//
// Reference: tools/pi4u_lite/Inference/Likelihood/source/HGG_Likelihood.cpp
//   PET term        :40-75   N = #{PET > 0 on the stepPET lattice},
//                            L = -N/2 log(2 pi s2) - 1/(2 s2) sum (u - b PET)^2
//   Bernoulli terms :77-156  alpha = 1/2 + 1/2 sgn(u-uc) (1 - exp(-w2/slope)),
//                            w2 = 1 if u > uc else (u-uc)^2; log alpha if y == 1
//                            else log(1-alpha); all voxels, or ROI points on
//                            the stepMRI lattice
//   run             :176-192 -logL = -(L_PET + L_T1 + L_T2)
// The reference accumulates in long double; here every per-voxel term is
// double and the sums are FP64 with a fixed-order two-stage reduction, so a
// result is reproducible run to run.
#include <mutex>

#include "gh_ctx.h"
#include "gh_lik_common.h"

namespace gh {

// Partials of gh_likelihood_noalloc: 4 sums per workgroup of the fixed
// LIK_GRID partition plus the 4 final sums (32.8 KB).  Each GPU holds its own
// copy until the process ends; in host code the name is only a stand-in, the
// device address comes from hipGetSymbolAddress (lik_scratch_resolve).
constexpr int LIK_SCRATCH_DOUBLES = 4 * (LIK_GRID + 1);
static __device__ __align__(256) double scratch[LIK_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};  // scratch's address on each device
}  // namespace

// Resolves scratch's address on `device`, which the caller has made current;
// once per device (gh_ctx_create calls it).
int lik_scratch_resolve(int device) {
    if (device < 0 || device >= kMaxDevices) return set_error(GH_EINVAL, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return GH_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_error(GH_EHIP, std::string("hipGetSymbolAddress(likelihood scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
    return GH_OK;
}

namespace {

// stage 1: per-workgroup partials {sse, npet, L_T1, L_T2}
template <typename T>
__global__ __launch_bounds__(LIK_THREADS) void lik_partials_kernel(size_t ncells, Layout lay, const T* __restrict__ u,
                                                                   const unsigned char* __restrict__ t1,
                                                                   const unsigned char* __restrict__ fl,
                                                                   const double* __restrict__ pet,
                                                                   const int4* __restrict__ roi, int nroi,
                                                                   LikParams p, double* __restrict__ partials) {
    lik_partials_body<T>((int)blockIdx.x, (int)gridDim.x, ncells, lay, u, t1, fl, pet, roi, nroi, p, partials);
}

// stage 2: one workgroup sums the partials in a fixed order
__global__ __launch_bounds__(256) void lik_final_kernel(int nparts, const double* __restrict__ partials,
                                                        double* __restrict__ out) {
    lik_final_body(nparts, partials, out);
}

// parts: 4 * (LIK_GRID + 1) doubles of device scratch (null: the context's
// own buffer, allocated on first use); every slot is written before it is read
template <typename T>
int run_lik(Ctx& c, size_t ncells, Layout lay, const T* u, const unsigned char* t1, const unsigned char* fl,
            const double* pet, const int4* roi, int nroi, const LikParams& p, double out[4], int64_t* npet,
            double* parts = nullptr) {
    if (!parts) {
        int rc = c.likPart.ensure(sizeof(double) * LIK_SCRATCH_DOUBLES);
        if (rc) return rc;
        parts = c.likPart.as<double>();
    }
    hipLaunchKernelGGL(lik_partials_kernel<T>, dim3(LIK_GRID), dim3(LIK_THREADS), 0, c.stream, ncells, lay, u, t1,
                       fl, pet, roi, nroi, p, parts);
    GH_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(lik_final_kernel, dim3(1), dim3(256), 0, c.stream, LIK_GRID, parts, parts + 4 * LIK_GRID);
    GH_HIP_CHECK(hipGetLastError());
    double sums[4];
    GH_HIP_CHECK(hipMemcpyAsync(sums, parts + 4 * LIK_GRID, sizeof(sums), hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    finish(sums, p, out, npet);
    return GH_OK;
}

// dense x-fastest -> block-major for a uniform grid (slot order bx + bpd*(by + bpd*bz))
template <typename T>
void to_blocks(int n, const T* dense, T* bm) {
    const int bpd = n / BS;
    for (size_t c = 0; c < (size_t)n * n * n; ++c) {
        const int ix = (int)(c % n), iy = (int)((c / n) % n), iz = (int)(c / ((size_t)n * n));
        const size_t s = (size_t)(ix / BS) + (size_t)bpd * ((iy / BS) + (size_t)bpd * (iz / BS));
        bm[s * BCELLS + (ix % BS) + BS * ((iy % BS) + BS * (iz % BS))] = dense[c];
    }
}

}  // namespace

namespace {
int lik_eval_into(Ctx& c, const double* params, int stepPET, int stepTi, int use_roi, double* out, int64_t* npet,
                  double* parts) {
    if (!c.lik_valid())
        return set_error(GH_ESTATE, "likelihood: no observations for the current grid (call gh_lik_set_data after "
                                    "the last grid change)");
    if (use_roi && c.nroi == 0) return set_error(GH_ESTATE, "likelihood: no ROI was given");
    LikParams p;
    int rc = check_params(params, stepPET, stepTi, p);
    if (rc) return rc;
    Layout lay{0, 0, 0, c.topo.uniform_bpd};
    const int4* roi = use_roi ? c.obsRoi.as<int4>() : nullptr;
    if (c.real_kind == GH_REAL_F32)
        return run_lik<float>(c, c.ncells(), lay, c.phi[c.cur].as<float>(), c.obsT1.as<unsigned char>(),
                              c.obsFl.as<unsigned char>(), c.obsPet.as<double>(), roi, c.nroi, p, out, npet, parts);
    return run_lik<double>(c, c.ncells(), lay, c.phi[c.cur].as<double>(), c.obsT1.as<unsigned char>(),
                           c.obsFl.as<unsigned char>(), c.obsPet.as<double>(), roi, c.nroi, p, out, npet, parts);
}
}  // namespace

int lik_eval_current(Ctx& c, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                     int64_t* npet) {
    return lik_eval_into(c, params, stepPET, stepTi, use_roi, out, npet, nullptr);
}

int lik_seed_gradient(Ctx& c, const void* u, double* g, const double* params, int stepPET, int stepTi,
                      int use_roi) {
    if (!c.lik_valid())
        return set_error(GH_ESTATE, "likelihood: no observations for the current grid (call gh_lik_set_data after "
                                    "the last grid change)");
    LikParams p;
    int rc = check_params(params, stepPET, stepTi, p);
    if (rc) return rc;
    if (use_roi && c.nroi == 0) return set_error(GH_ESTATE, "likelihood: no ROI was given");
    Layout lay{0, 0, 0, c.topo.uniform_bpd};
    const size_t nc = c.ncells();
    const unsigned grid = (unsigned)((nc + 255) / 256);
    if (c.real_kind == GH_REAL_F32)
        hipLaunchKernelGGL(lik_seed_kernel<float>, dim3(grid), dim3(256), 0, c.stream, nc, lay, (const float*)u,
                           c.obsT1.as<unsigned char>(), c.obsFl.as<unsigned char>(), c.obsPet.as<double>(), p,
                           use_roi ? 0 : 1, g);
    else
        hipLaunchKernelGGL(lik_seed_kernel<double>, dim3(grid), dim3(256), 0, c.stream, nc, lay, (const double*)u,
                           c.obsT1.as<unsigned char>(), c.obsFl.as<unsigned char>(), c.obsPet.as<double>(), p,
                           use_roi ? 0 : 1, g);
    GH_HIP_CHECK(hipGetLastError());
    if (use_roi) {
        const unsigned rg = (unsigned)((c.nroi + 255) / 256);
        if (c.real_kind == GH_REAL_F32)
            hipLaunchKernelGGL(lik_seed_roi_kernel<float>, dim3(rg), dim3(256), 0, c.stream, c.nroi,
                               c.obsRoi.as<int4>(), (const float*)u, c.obsT1.as<unsigned char>(),
                               c.obsFl.as<unsigned char>(), p, g);
        else
            hipLaunchKernelGGL(lik_seed_roi_kernel<double>, dim3(rg), dim3(256), 0, c.stream, c.nroi,
                               c.obsRoi.as<int4>(), (const double*)u, c.obsT1.as<unsigned char>(),
                               c.obsFl.as<unsigned char>(), p, g);
        GH_HIP_CHECK(hipGetLastError());
    }
    return GH_OK;
}

}  // namespace gh

using namespace gh;

extern "C" {

int gh_lik_set_data(gh_ctx* ctx, int n, const double* t1c, const double* flair, const double* pet, const double* roi,
                    int nroi) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.uniform_bpd <= 0 || n != BS * c->topo.uniform_bpd)
        return set_error(GH_ESTATE, "gh_lik_set_data: needs a uniform grid with n = 8*bpd");
    if (!t1c || !flair || !pet || nroi < 0 || (nroi > 0 && !roi))
        return set_error(GH_EINVAL, "gh_lik_set_data: null array");
    GH_HIP_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)n * n * n;
    std::vector<unsigned char> y1(nc), y2(nc), b1(nc), b2(nc);
    for (size_t i = 0; i < nc; ++i) {
        y1[i] = t1c[i] == 1;
        y2[i] = flair[i] == 1;
    }
    std::vector<double> pb(nc);
    to_blocks(n, y1.data(), b1.data());
    to_blocks(n, y2.data(), b2.data());
    to_blocks(n, pet, pb.data());
    int rc;
    if ((rc = c->obsT1.ensure(nc)) || (rc = c->obsFl.ensure(nc)) || (rc = c->obsPet.ensure(nc * sizeof(double))))
        return rc;
    GH_HIP_CHECK(hipMemcpyAsync(c->obsT1.p, b1.data(), nc, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipMemcpyAsync(c->obsFl.p, b2.data(), nc, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipMemcpyAsync(c->obsPet.p, pb.data(), nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->nroi = 0;
    if (nroi > 0) {
        // Matrix::D2D ROI(i, 0..2) as stored x-fastest: roi[i + nroi*col]
        std::vector<int4> q(nroi);
        const int bpd = c->topo.uniform_bpd;
        for (int i = 0; i < nroi; ++i) {
            const int ix = (int)roi[i], iy = (int)roi[i + nroi], iz = (int)roi[i + 2 * (size_t)nroi];
            if (ix < 0 || iy < 0 || iz < 0 || ix >= n || iy >= n || iz >= n)
                return set_error(GH_EINVAL, "gh_lik_set_data: ROI point outside the grid");
            const size_t s = (size_t)(ix / BS) + (size_t)bpd * ((iy / BS) + (size_t)bpd * (iz / BS));
            q[i] = make_int4((int)(s * BCELLS + (ix % BS) + BS * ((iy % BS) + BS * (iz % BS))), ix, iy, iz);
        }
        if ((rc = c->obsRoi.ensure(nroi * sizeof(int4)))) return rc;
        GH_HIP_CHECK(hipMemcpyAsync(c->obsRoi.p, q.data(), nroi * sizeof(int4), hipMemcpyHostToDevice, c->stream));
        c->nroi = nroi;
    }
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->lik_ready = true;
    c->lik_version = c->tree_version;
    return GH_OK;
}

int gh_likelihood(gh_ctx* ctx, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                  int64_t* npet) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out) return set_error(GH_EINVAL, "null argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    return lik_eval_current(*c, params, stepPET, stepTi, use_roi, out, npet);
}

int gh_likelihood_noalloc(gh_ctx* ctx, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                          int64_t* npet) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out) return set_error(GH_EINVAL, "null argument");
    // the partition is fixed at LIK_GRID workgroups whatever the grid size, so
    // the scratch bounds the partition, not the grid
    static_assert(LIK_SCRATCH_DOUBLES == 4 * (LIK_GRID + 1), "scratch holds the LIK_GRID partials + 4 sums");
    GH_HIP_CHECK(hipSetDevice(c->device));
    int rc = lik_scratch_resolve(c->device);  // a no-op once gh_ctx_create has resolved it
    if (rc) return rc;
    double* parts = g_scratch_ptr[c->device];
    return lik_eval_into(*c, params, stepPET, stepTi, use_roi, out, npet, parts);
}

int gh_likelihood_dense(gh_ctx* ctx, int nx, int ny, int nz, const double* model, const double* t1c,
                        const double* flair, const double* pet, const double* roi, int nroi, const double* params,
                        int stepPET, int stepTi, double* out, int64_t* npet) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out || !model || !t1c || !flair || !pet) return set_error(GH_EINVAL, "null argument");
    if (nx <= 0 || ny <= 0 || nz <= 0 || nroi < 0 || (nroi > 0 && !roi))
        return set_error(GH_EINVAL, "gh_likelihood_dense: bad size");
    LikParams p;
    int rc = check_params(params, stepPET, stepTi, p);
    if (rc) return rc;
    GH_HIP_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)nx * ny * nz;
    std::vector<unsigned char> y1(nc), y2(nc);
    for (size_t i = 0; i < nc; ++i) {
        y1[i] = t1c[i] == 1;
        y2[i] = flair[i] == 1;
    }
    std::vector<int4> q(nroi);
    for (int i = 0; i < nroi; ++i) {
        const int ix = (int)roi[i], iy = (int)roi[i + nroi], iz = (int)roi[i + 2 * (size_t)nroi];
        if (ix < 0 || iy < 0 || iz < 0 || ix >= nx || iy >= ny || iz >= nz)
            return set_error(GH_EINVAL, "gh_likelihood_dense: ROI point outside the grid");
        q[i] = make_int4((int)(ix + (size_t)nx * (iy + (size_t)ny * iz)), ix, iy, iz);
    }
    DevBuf dm, dy1, dy2, dp, dr;
    if ((rc = dm.ensure(nc * 8)) || (rc = dy1.ensure(nc)) || (rc = dy2.ensure(nc)) || (rc = dp.ensure(nc * 8)) ||
        (rc = dr.ensure((size_t)nroi * sizeof(int4)))) {
        dm.release(); dy1.release(); dy2.release(); dp.release(); dr.release();
        return rc;
    }
    GH_HIP_CHECK(hipMemcpyAsync(dm.p, model, nc * 8, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipMemcpyAsync(dy1.p, y1.data(), nc, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipMemcpyAsync(dy2.p, y2.data(), nc, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipMemcpyAsync(dp.p, pet, nc * 8, hipMemcpyHostToDevice, c->stream));
    if (nroi > 0)
        GH_HIP_CHECK(hipMemcpyAsync(dr.p, q.data(), nroi * sizeof(int4), hipMemcpyHostToDevice, c->stream));
    Layout lay{nx, ny, nz, 0};
    rc = run_lik<double>(*c, nc, lay, dm.as<double>(), dy1.as<unsigned char>(), dy2.as<unsigned char>(),
                         dp.as<double>(), nroi > 0 ? dr.as<int4>() : nullptr, nroi, p, out, npet);
    (void)hipStreamSynchronize(c->stream);
    dm.release(); dy1.release(); dy2.release(); dp.release(); dr.release();
    return rc;
}

}  // extern "C"
