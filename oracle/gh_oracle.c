/*
 * gh_oracle.c -- CPU oracle (checker) for the GliomaSolver RD hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 *
 * Built with -ffp-contract=off and without -ffast-math so that float and
 * double operations round exactly where the reference C++ rounds them
 * (GCC -O3 on x86-64 SSE, no FMA: makefile/tools.make/make.linux:1).
 */
#include "gh_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* anatomy preprocessing                                                     */
/* Glioma/Glioma_ReactionDiffusion.cpp:110,139-152: Real pGM,pWM,pCSF; the    */
/* tissue sum is Real+Real promoted to double; majority vote then normalise.  */
/* ------------------------------------------------------------------------ */
#define ANATOMY_BODY(REAL)                                                         \
    for (int64_t c = 0; c < ncells; ++c) {                                         \
        REAL pGM = (REAL)gm[c], pWM = (REAL)wm[c], pCSF = (REAL)csf[c];            \
        double tissue = (double)(REAL)(pWM + pGM);                                 \
        pCSF = ((double)pCSF > tissue) ? (REAL)1. : (REAL)0.;                      \
        pWM = ((double)pCSF > tissue) ? (REAL)0. : pWM;                            \
        pGM = ((double)pCSF > tissue) ? (REAL)0. : pGM;                            \
        tissue = (double)(REAL)(pWM + pGM);                                        \
        p_w[c] = (tissue > 0.) ? (REAL)((double)pWM / tissue) : (REAL)0.;          \
        p_g[c] = (tissue > 0.) ? (REAL)((double)pGM / tissue) : (REAL)0.;          \
        p_csf[c] = pCSF;                                                           \
    }

void gho_anatomy_f32(int64_t ncells, const double* gm, const double* wm, const double* csf,
                     float* p_w, float* p_g, float* p_csf) { ANATOMY_BODY(float) }
void gho_anatomy_f64(int64_t ncells, const double* gm, const double* wm, const double* csf,
                     double* p_w, double* p_g, double* p_csf) { ANATOMY_BODY(double) }

/* ------------------------------------------------------------------------ */
/* initial condition: Glioma_ReactionDiffusion.cpp:79-172.                   */
/*   cell position x = origin + h*i (MRAG/MRAGcore/MRAGCommon.h:199-204),     */
/*   origin = block index * 1/bpd, h = (1/bpd)/8, all Real;                   */
/*   voxel = floor(x / (1./max(nx,ny,nz))) in double; cells outside stay 0;   */
/*   tumour: psi = (|x - ic| - 0.005) * (1./(2.*(1./128))) in Real;           */
/*   phi = 1 inside psi<-1 (tissue > 0.001), 0.5*(1-psi - sin(pi psi)/pi)     */
/*   on the smoothing shell (tissue > 0), else 0.  (1 - psi) is Real.         */
/* ------------------------------------------------------------------------ */
#define IC_BODY(REAL)                                                                        \
    const int n = 8 * bpd;                                                                   \
    const int bmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);                    \
    const double bh = 1.0 / (double)bmax;                                                    \
    const REAL dil = (REAL)(1.0 / bpd), h = dil / (REAL)8;                                   \
    const REAL rad = (REAL)0.005, sup = (REAL)2., hh = (REAL)(1. / 128);                    \
    const REAL iw = (REAL)(1. / (double)(REAL)(sup * hh));                                   \
    for (int k = 0; k < n; ++k)                                                              \
        for (int j = 0; j < n; ++j)                                                          \
            for (int i = 0; i < n; ++i) {                                                    \
                const int64_t c = i + (int64_t)n * (j + (int64_t)n * k);                     \
                p_w[c] = p_g[c] = p_csf[c] = phi[c] = (REAL)0;                               \
                REAL x[3];                                                                   \
                const int ii[3] = {i, j, k};                                                 \
                for (int d = 0; d < 3; ++d) {                                                \
                    REAL org = (REAL)(ii[d] / 8) * dil;                                      \
                    x[d] = org + h * (REAL)(ii[d] % 8);                                      \
                }                                                                            \
                const int mx = (int)floor((double)x[0] / bh);                                \
                const int my = (int)floor((double)x[1] / bh);                                \
                const int mz = (int)floor((double)x[2] / bh);                                \
                if (mx < 0 || mx >= nx || my < 0 || my >= ny || mz < 0 || mz >= nz) continue; \
                const int64_t v = mx + (int64_t)nx * (my + (int64_t)ny * mz);                \
                REAL pGM = (REAL)gm[v], pWM = (REAL)wm[v], pCSF = (REAL)csf[v];              \
                double tissue = (double)(REAL)(pWM + pGM);                                   \
                pCSF = ((double)pCSF > tissue) ? (REAL)1. : (REAL)0.;                        \
                pWM = ((double)pCSF > tissue) ? (REAL)0. : pWM;                              \
                pGM = ((double)pCSF > tissue) ? (REAL)0. : pGM;                              \
                tissue = (double)(REAL)(pWM + pGM);                                          \
                p_w[c] = (tissue > 0.) ? (REAL)((double)pWM / tissue) : (REAL)0.;            \
                p_g[c] = (tissue > 0.) ? (REAL)((double)pGM / tissue) : (REAL)0.;            \
                p_csf[c] = pCSF;                                                             \
                if (!tumor_ic) continue;                                                     \
                REAL p0 = x[0] - (REAL)tumor_ic[0], p1 = x[1] - (REAL)tumor_ic[1],           \
                     p2 = x[2] - (REAL)tumor_ic[2];                                          \
                REAL r2 = p0 * p0;                                                           \
                r2 = r2 + p1 * p1;                                                           \
                r2 = r2 + p2 * p2;                                                           \
                const REAL dist = (REAL)sqrt((double)r2);                                    \
                const REAL psi = (dist - rad) * iw;                                          \
                const REAL tis = pGM + pWM;                                                  \
                if (psi < -1 && tis > 0.001)                                                 \
                    phi[c] = (REAL)1.0;                                                      \
                else if (-1 <= psi && psi <= 1 && tis > 0)                                   \
                    phi[c] = (REAL)(0.5 * ((double)((REAL)1 - psi) - sin(M_PI * psi) / M_PI)); \
            }

void gho_ic_f32(int bpd, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                const double* tumor_ic, float* p_w, float* p_g, float* p_csf, float* phi) { IC_BODY(float) }
void gho_ic_f64(int bpd, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                const double* tumor_ic, double* p_w, double* p_g, double* p_csf, double* phi) { IC_BODY(double) }

/* ------------------------------------------------------------------------ */
/* schedule: Glioma_ReactionDiffusion.cpp:296-320                            */
/*   Dw = Dw/(L*L) (Real); Dg = 0.1*Dw; h = 1./(bs*bpd) (Real);              */
/*   dt = cfl*h*h/(2.*_DIM*max(Dw,Dg)) (Real); while(t<=tend){...; t+=dt;}    */
/* ------------------------------------------------------------------------ */
#define SCHEDULE_BODY(REAL)                                                        \
    REAL dw = (REAL)Dw_phys, LL = (REAL)L, te = (REAL)tend;                        \
    dw = dw / (REAL)(LL * LL);                                                     \
    REAL dg = (REAL)(0.1 * (double)dw);                                            \
    REAL h = (REAL)(1. / (double)gpd);                                             \
    REAL dmax = (dw >= dg) ? dw : dg;                                              \
    REAL d = (REAL)(cfl * (double)h * (double)h / (2. * 3 * (double)dmax));        \
    REAL t = 0;                                                                    \
    int64_t k = 0;                                                                 \
    while (t <= te) { t = (REAL)(t + d); ++k; }                                    \
    *Dw = dw; *Dg = dg; *dt = d; *nsteps = k;

void gho_schedule_f32(double Dw_phys, double L, double tend, int gpd, double cfl,
                      float* Dw, float* Dg, float* dt, int64_t* nsteps) { SCHEDULE_BODY(float) }
void gho_schedule_f64(double Dw_phys, double L, double tend, int gpd, double cfl,
                      double* Dw, double* Dg, double* dt, int64_t* nsteps) { SCHEDULE_BODY(double) }

/* ------------------------------------------------------------------------ */
/* RD step on a uniform periodic grid (MRAG grids are periodic:              */
/* MRAG/MRAGcore/MRAGrid.inl:37-39; same-level ghosts are exact copies).     */
/*                                                                          */
/* ReactionDiffusionOperator.h:98-139, promotion points (Real = float):      */
/*   df_i  = p_w*Dw + p_g*Dg                          Real                    */
/*   harm  : (df_i+df_loc < 1e-8f) ? 0 : 2.*df_i*df_loc/(Real)(df_i+df_loc)   */
/*           evaluated in double, stored to Real              (:148-163)    */
/*   chf_i = phi + p_w + p_g  (Real, left to right); chf_i < 0.1f doubles     */
/*           the coefficient of the opposite face              (:166-181)    */
/*   in    = ih2 * (Real)(sum_i df_i*phi_i)                     double        */
/*   out   = -((Real)(sum_i df_i) * phi_c * ih2)                double        */
/*   react = (Real)(rho*phi) * (1. - phi)                       double        */
/*   dphidt = (Real)(out + in + react)                                        */
/* UpdateTumor (:215-222): phi = (Real)(phi + dt*dphidt); clamp to [0,1].     */
/* ------------------------------------------------------------------------ */
#define RHS_BODY(REAL)                                                             \
    const double hh = (double)h;                                                   \
    const double ih2 = 1. / (hh * hh);                                             \
    const REAL eps_h = (REAL)1.0e-08, eps_bc = (REAL)0.1;                          \
    for (int k = 0; k < n; ++k)                                                    \
    for (int j = 0; j < n; ++j)                                                    \
    for (int i = 0; i < n; ++i) {                                                  \
        const int64_t c = i + (int64_t)n * (j + (int64_t)n * k);                   \
        const int64_t nb[6] = {                                                    \
            ((i + n - 1) % n) + (int64_t)n * (j + (int64_t)n * k),                 \
            ((i + 1) % n) + (int64_t)n * (j + (int64_t)n * k),                     \
            i + (int64_t)n * (((j + n - 1) % n) + (int64_t)n * k),                 \
            i + (int64_t)n * (((j + 1) % n) + (int64_t)n * k),                     \
            i + (int64_t)n * (j + (int64_t)n * ((k + n - 1) % n)),                 \
            i + (int64_t)n * (j + (int64_t)n * ((k + 1) % n))};                    \
        if ((double)(REAL)((REAL)(p_w[c] + p_g[c]) + phi[c]) > 0.) {               \
            REAL df_loc = (REAL)((REAL)(p_w[c] * Dw) + (REAL)(p_g[c] * Dg));       \
            REAL df[6], chf[6];                                                    \
            for (int f = 0; f < 6; ++f)                                            \
                df[f] = (REAL)((REAL)(p_w[nb[f]] * Dw) + (REAL)(p_g[nb[f]] * Dg)); \
            for (int f = 0; f < 6; ++f) {                                          \
                REAL s = (REAL)(df[f] + df_loc);                                   \
                df[f] = (s < eps_h) ? (REAL)0.                                     \
                      : (REAL)(((2. * (double)df[f]) * (double)df_loc) / (double)s); \
            }                                                                      \
            for (int f = 0; f < 6; ++f)                                            \
                chf[f] = (REAL)((REAL)(phi[nb[f]] + p_w[nb[f]]) + p_g[nb[f]]);     \
            if (chf[0] < eps_bc) df[1] = (REAL)((double)df[1] * 2.0);              \
            if (chf[1] < eps_bc) df[0] = (REAL)((double)df[0] * 2.0);              \
            if (chf[2] < eps_bc) df[3] = (REAL)((double)df[3] * 2.0);              \
            if (chf[3] < eps_bc) df[2] = (REAL)((double)df[2] * 2.0);              \
            if (chf[4] < eps_bc) df[5] = (REAL)((double)df[5] * 2.0);              \
            if (chf[5] < eps_bc) df[4] = (REAL)((double)df[4] * 2.0);              \
            REAL sin_ = (REAL)(df[0] * phi[nb[0]]);                                \
            for (int f = 1; f < 6; ++f) sin_ = (REAL)(sin_ + (REAL)(df[f] * phi[nb[f]])); \
            double fin = ih2 * (double)sin_;                                       \
            REAL sdf = df[0];                                                      \
            for (int f = 1; f < 6; ++f) sdf = (REAL)(sdf + df[f]);                 \
            double fout = -((double)(REAL)(sdf * phi[c]) * ih2);                   \
            double react = (double)(REAL)(rho * phi[c]) * (1. - (double)phi[c]);   \
            dphidt[c] = (REAL)((fout + fin) + react);                              \
        } else {                                                                   \
            dphidt[c] = (REAL)0.;                                                  \
        }                                                                          \
    }

#define UPDATE_BODY(REAL)                                                          \
    for (int64_t c = 0; c < (int64_t)n * n * n; ++c) {                             \
        REAL p = (REAL)((double)phi[c] + dt * (double)dphidt[c]);                  \
        p = ((REAL)0. < p) ? p : (REAL)0.;  /* std::max((Real)0., phi) */          \
        p = (p < (REAL)1.) ? p : (REAL)1.;  /* std::min((Real)1., phi) */          \
        phi[c] = p;                                                                \
    }

static void rhs_f32(int n, const float* phi, float* dphidt, const float* p_w, const float* p_g,
                    float Dw, float Dg, float rho, float h) { RHS_BODY(float) }
static void rhs_f64(int n, const double* phi, double* dphidt, const double* p_w, const double* p_g,
                    double Dw, double Dg, double rho, double h) { RHS_BODY(double) }

void gho_rd_steps_f32(int n, float* phi, float* dphidt, const float* p_w, const float* p_g,
                      float Dw, float Dg, float rho, float h, double dt, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; ++s) {
        rhs_f32(n, phi, dphidt, p_w, p_g, Dw, Dg, rho, h);
        UPDATE_BODY(float)
    }
}
void gho_rd_steps_f64(int n, double* phi, double* dphidt, const double* p_w, const double* p_g,
                      double Dw, double Dg, double rho, double h, double dt, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; ++s) {
        rhs_f64(n, phi, dphidt, p_w, p_g, Dw, Dg, rho, h);
        UPDATE_BODY(double)
    }
}

/* ------------------------------------------------------------------------ */
/* likelihood: HGG_Likelihood.cpp                                           */
/* ------------------------------------------------------------------------ */
static int sgn_d(double d) { /* :158-163 */
    if (d < -0.0) return -1;
    return d > 0.0;
}

long double gho_log_bernoulli(double u, double y, double uc, double slope) { /* :136-156 */
    double is2 = 1. / slope;
    double diff = u - uc;
    long double omega2 = (diff > 0.) ? 1.L : (long double)(diff * diff);
    long double alpha = 0.5L + (long double)(0.5 * sgn_d(diff)) * (1.L - expl(-omega2 * (long double)is2));
    return (y == 1) ? logl(alpha) : logl(1.L - alpha);
}

double gho_likelihood(int nx, int ny, int nz, const double* model, const double* pet,
                      const double* t1c, const double* flair, const double* roi, int nroi,
                      double PETsigma2, double PETscale, double T1uc, double T2uc, double slope,
                      int stepPET, int stepTi, double out[3], int64_t* npet) {
    /* _computePETLogLikelihood :40-75 */
    int N = 0;
    long double sum = 0.L;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                if ((ix % stepPET == 0) && (iy % stepPET == 0) && (iz % stepPET == 0)) {
                    int64_t c = ix + (int64_t)nx * (iy + (int64_t)ny * iz);
                    if (pet[c] > 0.) {
                        double r = model[c] - PETscale * pet[c];
                        sum += (long double)(r * r);
                        N++;
                    }
                }
            }
    long double p1 = (long double)(-0.5 * N * log(2. * M_PI * PETsigma2));
    long double p2 = (long double)(-0.5 * (1. / PETsigma2)) * sum;
    long double Lpet = p1 + p2;
    /* _computeTiLogLikelihood :77-128 */
    long double L[2];
    for (int t = 0; t < 2; ++t) {
        const double* data = (t == 0) ? t1c : flair;
        double uc = (t == 0) ? T1uc : T2uc;
        long double s = 0.L;
        if (roi == NULL) {
            for (int iz = 0; iz < nz; ++iz)
                for (int iy = 0; iy < ny; ++iy)
                    for (int ix = 0; ix < nx; ++ix) {
                        int64_t c = ix + (int64_t)nx * (iy + (int64_t)ny * iz);
                        s += gho_log_bernoulli(model[c], data[c], uc, slope);
                    }
        } else {
            /* Matrix::D2D ROI(i,0..2) is x-fastest over i: roi[i + nroi*col] */
            for (int i = 0; i < nroi; ++i) {
                int ix = (int)roi[i], iy = (int)roi[i + nroi], iz = (int)roi[i + 2 * nroi];
                if ((ix % stepTi == 0) && (iy % stepTi == 0) && (iz % stepTi == 0)) {
                    int64_t c = ix + (int64_t)nx * (iy + (int64_t)ny * iz);
                    s += gho_log_bernoulli(model[c], data[c], uc, slope);
                }
            }
        }
        L[t] = s;
    }
    out[0] = (double)Lpet;
    out[1] = (double)L[0];
    out[2] = (double)L[1];
    if (npet) *npet = N;
    long double cost = Lpet + L[0] + L[1]; /* run :176-192 */
    return (double)(-cost);
}

/* ------------------------------------------------------------------------ */
/* Matrix::D3D dump/load (Glioma/Matrix.h:625-665, header :72-89,           */
/* serialize type id :36-60): int32 {1234, 3, nx, ny, nz}, int32 type id     */
/* (0 = double), then nx*ny*nz values x-fastest.                             */
/* ------------------------------------------------------------------------ */
int gho_dat_write(const char* path, int nx, int ny, int nz, const double* data) {
    FILE* f = fopen(path, "wb");
    if (!f) return -1;
    int hdr[6] = {1234, 3, nx, ny, nz, 0};
    size_t n = (size_t)nx * ny * nz;
    int ok = fwrite(hdr, sizeof(int), 6, f) == 6 && fwrite(data, sizeof(double), n, f) == n;
    fclose(f);
    return ok ? 0 : -2;
}

int gho_dat_read_header(const char* path, int dims[3]) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int hdr[5];
    int ok = fread(hdr, sizeof(int), 5, f) == 5 && hdr[0] == 1234 && hdr[1] == 3;
    fclose(f);
    if (!ok) return -2;
    dims[0] = hdr[2]; dims[1] = hdr[3]; dims[2] = hdr[4];
    return 0;
}

int gho_dat_read(const char* path, int nx, int ny, int nz, double* data) {
    FILE* f = fopen(path, "rb");
    if (!f) return -1;
    int hdr[6];
    size_t n = (size_t)nx * ny * nz;
    int ok = fread(hdr, sizeof(int), 6, f) == 6 && hdr[0] == 1234 && hdr[1] == 3 &&
             hdr[2] == nx && hdr[3] == ny && hdr[4] == nz;
    if (ok) {
        if (hdr[5] == 0) ok = fread(data, sizeof(double), n, f) == n;
        else if (hdr[5] == 1) {
            float* tmp = (float*)malloc(n * sizeof(float));
            ok = tmp && fread(tmp, sizeof(float), n, f) == n;
            for (size_t i = 0; ok && i < n; ++i) data[i] = tmp[i];
            free(tmp);
        } else ok = 0;
    }
    fclose(f);
    return ok ? 0 : -2;
}

/* ------------------------------------------------------------------------ */
/* wavelet taps: MRAG/MRAGcore/MRAGWavelets_StaticData.cpp:13-21 (2nd),     */
/* :31-39 (4th).  Values are Real in the reference; returned here as double. */
/* ------------------------------------------------------------------------ */
int gho_wavelet_taps(int order, int filter, int support[2], double* taps) {
    static const double i2_Ha[1] = {1}, i2_Ga[3] = {-0.5, 1., -0.5}, i2_Hs[3] = {0.5, 1., 0.5}, i2_Gs[1] = {1};
    static const int i2_sup[4][2] = {{0, 1}, {-2, 1}, {-1, 2}, {-1, 0}};
    static const double i4_Ha[1] = {1};
    static const double i4_Ga[7] = {1 / 16., 0, -9 / 16., 1, -9 / 16., 0, 1 / 16.};
    static const double i4_Hs[7] = {-1 / 16., 0, 9 / 16., 1, 9 / 16., 0, -1 / 16.};
    static const double i4_Gs[1] = {1};
    static const int i4_sup[4][2] = {{0, 1}, {-4, 3}, {-3, 4}, {-1, 0}};
    const double* t2[4] = {i2_Ha, i2_Ga, i2_Hs, i2_Gs};
    const double* t4[4] = {i4_Ha, i4_Ga, i4_Hs, i4_Gs};
    if (filter < 0 || filter > 3 || (order != 2 && order != 4)) return -1;
    const int(*sup)[2] = (order == 2) ? i2_sup : i4_sup;
    const double* t = (order == 2) ? t2[filter] : t4[filter];
    support[0] = sup[filter][0];
    support[1] = sup[filter][1];
    int nt = support[1] - support[0];
    for (int i = 0; i < nt; ++i) taps[i] = t[i];
    return nt;
}

/* ------------------------------------------------------------------------ */
/* tangent-linear reference (FP64)                                           */
/* ------------------------------------------------------------------------ */
static double dlog_bern(double u, double y, double uc, double slope) {
    /* d/du of log(alpha) or log(1-alpha), alpha of HGG_Likelihood.cpp:149-152 */
    double diff = u - uc;
    if (!(diff < 0.)) return 0.; /* alpha is constant for u > uc; slope 0 at u == uc */
    double a = 0.5 * exp(-diff * diff / slope);
    double da = a * (-2. * diff / slope);
    return (y == 1) ? da / a : -da / (1. - a);
}

void gho_dloss_du(int n, const double* u, const double* t1c, const double* flair, const double* pet,
                  const double* lik, int stepPET, int stepTi, double* g) {
    (void)stepTi; /* full grid: every voxel counts for the Bernoulli terms */
    const double s2 = lik[0], b = lik[1], uc1 = lik[2], uc2 = lik[3], slope = lik[4];
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int64_t c = i + (int64_t)n * (j + (int64_t)n * k);
                double dl = dlog_bern(u[c], t1c[c], uc1, slope) + dlog_bern(u[c], flair[c], uc2, slope);
                if (i % stepPET == 0 && j % stepPET == 0 && k % stepPET == 0 && pet[c] > 0.)
                    dl += -(u[c] - b * pet[c]) / s2;
                g[c] = -dl; /* J = -logL */
            }
}

double gho_tlm_f64(int n, const double* phi0, const double* p_w, const double* p_g, double Dw, double Dg,
                   double rho, double h, double T, int64_t N, double sigma, const double* x0, const double* dir,
                   const double* t1c, const double* flair, const double* pet, const double* lik, int stepPET,
                   int stepTi, double* loss) {
    const int64_t nc = (int64_t)n * n * n;
    double* u = (double*)malloc(nc * sizeof(double));
    double* un = (double*)malloc(nc * sizeof(double));
    double* du = (double*)malloc(nc * sizeof(double));
    double* dun = (double*)malloc(nc * sizeof(double));
    double* dphidt = (double*)malloc(nc * sizeof(double));
    const double dt = T / (double)N, ddt = dir[2] / (double)N;
    const double ih2 = 1. / (h * h);
    memcpy(u, phi0, nc * sizeof(double));
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                int64_t c = i + (int64_t)n * (j + (int64_t)n * k);
                du[c] = phi0[c] * ((i - x0[0]) * dir[3] + (j - x0[1]) * dir[4] + (k - x0[2]) * dir[5]) /
                        (sigma * sigma);
            }
    for (int64_t s = 0; s < N; ++s) {
        memcpy(un, u, nc * sizeof(double));
        gho_rd_steps_f64(n, un, dphidt, p_w, p_g, Dw, Dg, rho, h, dt, 1);
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const int64_t c = i + (int64_t)n * (j + (int64_t)n * k);
                    const int64_t nb[6] = {
                        ((i + n - 1) % n) + (int64_t)n * (j + (int64_t)n * k),
                        ((i + 1) % n) + (int64_t)n * (j + (int64_t)n * k),
                        i + (int64_t)n * (((j + n - 1) % n) + (int64_t)n * k),
                        i + (int64_t)n * (((j + 1) % n) + (int64_t)n * k),
                        i + (int64_t)n * (j + (int64_t)n * ((k + n - 1) % n)),
                        i + (int64_t)n * (j + (int64_t)n * ((k + 1) % n))};
                    const int m = un[c] < 1.; /* upper clamp frozen; lower clamp inactive */
                    const int dom = (p_w[c] + p_g[c]) + u[c] > 0.;
                    if (!dom) { dun[c] = m ? du[c] : 0.; continue; }
                    double df_loc = p_w[c] * Dw + p_g[c] * Dg, df[6];
                    for (int f = 0; f < 6; ++f) {
                        double d = p_w[nb[f]] * Dw + p_g[nb[f]] * Dg, s_ = d + df_loc;
                        df[f] = (s_ < 1e-8) ? 0. : 2. * d * df_loc / s_;
                    }
                    int out[6];
                    for (int f = 0; f < 6; ++f) out[f] = (u[nb[f]] + p_w[nb[f]]) + p_g[nb[f]] < 0.1;
                    if (out[0]) df[1] *= 2; if (out[1]) df[0] *= 2;
                    if (out[2]) df[3] *= 2; if (out[3]) df[2] *= 2;
                    if (out[4]) df[5] *= 2; if (out[5]) df[4] *= 2;
                    double rdiff = 0., drdiff = 0.;
                    for (int f = 0; f < 6; ++f) {
                        rdiff += df[f] * (u[nb[f]] - u[c]);
                        drdiff += df[f] * (du[nb[f]] - du[c]);
                    }
                    rdiff *= ih2;
                    drdiff *= ih2;
                    const double R = rdiff + rho * u[c] * (1. - u[c]);
                    const double dR = drdiff + (dir[0] / Dw) * rdiff + rho * (1. - 2. * u[c]) * du[c] +
                                      dir[1] * u[c] * (1. - u[c]);
                    dun[c] = m ? du[c] + dt * dR + ddt * R : 0.;
                }
        double* t = u; u = un; un = t;
        t = du; du = dun; dun = t;
    }
    double comps[3];
    double J = gho_likelihood(n, n, n, u, pet, t1c, flair, NULL, 0, lik[0], lik[1], lik[2], lik[3], lik[4],
                              stepPET, stepTi, comps, NULL);
    if (loss) *loss = J;
    gho_dloss_du(n, u, t1c, flair, pet, lik, stepPET, stepTi, dun);
    double dJ = 0.;
    for (int64_t c = 0; c < nc; ++c) dJ += dun[c] * du[c];
    free(u); free(un); free(du); free(dun); free(dphidt);
    return dJ;
}
