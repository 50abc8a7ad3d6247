// gh_lik_common.h -- per-voxel terms and host-side combination of the
// HGG likelihood (tools/pi4u_lite/Inference/Likelihood/source/
// HGG_Likelihood.cpp:40-192), shared by the single-solve and batched kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gh_internal.h"

namespace gh {

struct LikParams {
    double s2, b, uc1, uc2, is2;
    int stepPET, stepTi;
};

__device__ inline double log_bernoulli(double u, bool y, double uc, double is2) {
    const double diff = u - uc;
    const double omega2 = (diff > 0.) ? 1. : diff * diff;
    const int sg = (diff < -0.0) ? -1 : (diff > 0.0);
    const double alpha = 0.5 + 0.5 * (double)sg * (1. - exp(-omega2 * is2));
    return y ? log(alpha) : log(1. - alpha);
}

__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

constexpr int LIK_THREADS = 256;
constexpr int LIK_GRID = 1024;  // fixed: the reduction order never changes

// cell index -> lattice coordinates for dense (x-fastest) or block-major
// uniform layouts
struct Layout {
    int nx, ny, nz;  // dense when bpd == 0
    int bpd;
    __device__ inline void coords(size_t c, int& ix, int& iy, int& iz) const {
        if (bpd == 0) {
            ix = (int)(c % nx);
            iy = (int)((c / nx) % ny);
            iz = (int)(c / ((size_t)nx * ny));
        } else {
            const int s = (int)(c / BCELLS), l = (int)(c % BCELLS);
            ix = (s % bpd) * BS + (l & 7);
            iy = ((s / bpd) % bpd) * BS + ((l >> 3) & 7);
            iz = (s / (bpd * bpd)) * BS + (l >> 6);
        }
    }
};

// stage 1: partials {sse, npet, L_T1, L_T2} of workgroup wg of nwg
// (LIK_THREADS lanes each); a fixed nwg fixes the summation order.
template <typename T>
__device__ inline void lik_partials_body(int wg, int nwg, size_t ncells, Layout lay, const T* __restrict__ u,
                                         const unsigned char* __restrict__ t1, const unsigned char* __restrict__ fl,
                                         const double* __restrict__ pet, const int4* __restrict__ roi, int nroi,
                                         const LikParams& p, double* __restrict__ partials) {
    double sse = 0, npet = 0, l1 = 0, l2 = 0;
    const size_t stride = (size_t)nwg * LIK_THREADS;
    for (size_t c = (size_t)wg * LIK_THREADS + threadIdx.x; c < ncells; c += stride) {
        int ix, iy, iz;
        lay.coords(c, ix, iy, iz);
        const double uc = (double)u[c];
        if ((ix % p.stepPET == 0) && (iy % p.stepPET == 0) && (iz % p.stepPET == 0)) {
            const double pv = pet[c];
            if (pv > 0.) {
                const double r = uc - p.b * pv;
                sse += r * r;
                npet += 1;
            }
        }
        if (roi == nullptr) {
            l1 += log_bernoulli(uc, t1[c] != 0, p.uc1, p.is2);
            l2 += log_bernoulli(uc, fl[c] != 0, p.uc2, p.is2);
        }
    }
    if (roi != nullptr) {
        for (size_t i = (size_t)wg * LIK_THREADS + threadIdx.x; i < (size_t)nroi; i += stride) {
            const int4 q = roi[i];  // {cell, ix, iy, iz}
            if ((q.y % p.stepTi == 0) && (q.z % p.stepTi == 0) && (q.w % p.stepTi == 0)) {
                const double uc = (double)u[q.x];
                l1 += log_bernoulli(uc, t1[q.x] != 0, p.uc1, p.is2);
                l2 += log_bernoulli(uc, fl[q.x] != 0, p.uc2, p.is2);
            }
        }
    }
    __shared__ double red[4][LIK_THREADS / 64];
    const int w = threadIdx.x / 64, lane = threadIdx.x % 64;
    sse = wave_sum(sse);
    npet = wave_sum(npet);
    l1 = wave_sum(l1);
    l2 = wave_sum(l2);
    if (lane == 0) {
        red[0][w] = sse;
        red[1][w] = npet;
        red[2][w] = l1;
        red[3][w] = l2;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int k = 0; k < LIK_THREADS / 64; ++k) s += red[threadIdx.x][k];
        partials[4 * wg + threadIdx.x] = s;
    }
}

// stage 2: one workgroup of 256 lanes sums the partials in a fixed order
__device__ inline void lik_final_body(int nparts, const double* __restrict__ partials, double* __restrict__ out) {
    __shared__ double red[4][256];
    for (int q = 0; q < 4; ++q) {
        double s = 0;
        for (int i = threadIdx.x; i < nparts; i += 256) s += partials[4 * i + q];
        red[q][threadIdx.x] = s;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[threadIdx.x] = red[threadIdx.x][0];
}

// d/du of log(alpha) (y == 1) or log(1 - alpha): 0 for u >= uc (alpha is
// constant above uc and flat at uc), alpha = e/2 with e = exp(-(u-uc)^2/slope) below
__device__ inline double dlog_bernoulli(double u, bool y, double uc, double is2) {
    const double diff = u - uc;
    if (!(diff < 0.)) return 0.;
    const double a = 0.5 * exp(-diff * diff * is2);
    const double da = a * (-2. * diff * is2);
    return y ? da / a : -da / (1. - a);
}

// g = dJ/du of J = -logL at cell c: PET lattice term and (without ROI) the Bernoulli terms
template <typename T>
__device__ inline void lik_seed_cell(size_t c, const Layout& lay, const T* __restrict__ u,
                                     const unsigned char* __restrict__ t1, const unsigned char* __restrict__ fl,
                                     const double* __restrict__ pet, const LikParams& p, int full_bern,
                                     double* __restrict__ g) {
    int ix, iy, iz;
    lay.coords(c, ix, iy, iz);
    const double uc = (double)u[c];
    double v = 0;
    if ((ix % p.stepPET == 0) && (iy % p.stepPET == 0) && (iz % p.stepPET == 0) && pet[c] > 0.)
        v += (uc - p.b * pet[c]) / p.s2;
    if (full_bern) v -= dlog_bernoulli(uc, t1[c] != 0, p.uc1, p.is2) + dlog_bernoulli(uc, fl[c] != 0, p.uc2, p.is2);
    g[c] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void lik_seed_kernel(size_t ncells, Layout lay, const T* __restrict__ u,
                                                       const unsigned char* __restrict__ t1,
                                                       const unsigned char* __restrict__ fl,
                                                       const double* __restrict__ pet, LikParams p, int full_bern,
                                                       double* __restrict__ g) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    lik_seed_cell<T>(c, lay, u, t1, fl, pet, p, full_bern, g);
}

// Bernoulli terms at ROI points on the stepMRI lattice (a point listed twice counts twice)
template <typename T>
__global__ __launch_bounds__(256) void lik_seed_roi_kernel(int nroi, const int4* __restrict__ roi,
                                                           const T* __restrict__ u,
                                                           const unsigned char* __restrict__ t1,
                                                           const unsigned char* __restrict__ fl, LikParams p,
                                                           double* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nroi) return;
    const int4 q = roi[i];
    if ((q.y % p.stepTi != 0) || (q.z % p.stepTi != 0) || (q.w % p.stepTi != 0)) return;
    const double uc = (double)u[q.x];
    atomicAdd(&g[q.x], -(dlog_bernoulli(uc, t1[q.x] != 0, p.uc1, p.is2) +
                         dlog_bernoulli(uc, fl[q.x] != 0, p.uc2, p.is2)));
}

inline int check_params(const double* params, int stepPET, int stepTi, LikParams& p) {
    if (!params) return set_error(GH_EINVAL, "likelihood: params is NULL");
    if (stepPET < 1 || stepTi < 1) return set_error(GH_EINVAL, "likelihood: steps must be >= 1");
    if (!(params[0] > 0) || !(params[4] != 0)) return set_error(GH_EINVAL, "likelihood: PETsigma2 > 0 and slope != 0");
    p.s2 = params[0];
    p.b = params[1];
    p.uc1 = params[2];
    p.uc2 = params[3];
    p.is2 = 1. / params[4];
    p.stepPET = stepPET;
    p.stepTi = stepTi;
    return GH_OK;
}

// combine on the host in long double, as HGG_Likelihood::run does
inline void finish(const double sums[4], const LikParams& p, double out[4], int64_t* npet) {
    const long double p1 = (long double)(-0.5 * (double)sums[1] * std::log(2. * M_PI * p.s2));
    const long double p2 = (long double)(-0.5 * (1. / p.s2)) * (long double)sums[0];
    const long double lpet = p1 + p2;
    const long double cost = lpet + (long double)sums[2] + (long double)sums[3];
    out[0] = (double)(-cost);
    out[1] = (double)lpet;
    out[2] = sums[2];
    out[3] = sums[3];
    if (npet) *npet = (int64_t)llround(sums[1]);
}

}  // namespace gh
