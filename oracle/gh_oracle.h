/*
 * gh_oracle.h -- CPU oracle for the GliomaSolver reaction-diffusion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (gliomasolver_amd/libgliomahip.so) never links or calls it.
 *
 * Every function restates the reference algorithm in plain C and cites the
 * reference file:line it follows.  Arrays are x-fastest, index
 * i + n*(j + n*k), the layout of Matrix::D3D (Glioma/Matrix.h:593-601).
 */
#ifndef GH_ORACLE_H
#define GH_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- anatomy: Glioma/Glioma_ReactionDiffusion.cpp:139-152 (majority vote) */
void gho_anatomy_f32(int64_t ncells, const double* gm, const double* wm, const double* csf,
                     float* p_w, float* p_g, float* p_csf);
void gho_anatomy_f64(int64_t ncells, const double* gm, const double* wm, const double* csf,
                     double* p_w, double* p_g, double* p_csf);

/* ---- whole initial condition: Glioma/Glioma_ReactionDiffusion.cpp:79-172.
 * Grid of bpd^3 blocks of 8^3 cells; anatomy of nx*ny*nz voxels (x fastest);
 * outputs are dense n^3 arrays, n = 8*bpd.  tumor_ic NULL leaves phi at 0. */
void gho_ic_f32(int bpd, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                const double* tumor_ic, float* p_w, float* p_g, float* p_csf, float* phi);
void gho_ic_f64(int bpd, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                const double* tumor_ic, double* p_w, double* p_g, double* p_csf, double* phi);

/* ---- time step and step count: Glioma/Glioma_ReactionDiffusion.cpp:302-320 */
/* outputs: Dw, Dg (rescaled by L^2), dt, nsteps (iterations of while(t<=tend)) */
void gho_schedule_f32(double Dw_phys, double L, double tend, int gpd, double cfl,
                      float* Dw, float* Dg, float* dt, int64_t* nsteps);
void gho_schedule_f64(double Dw_phys, double L, double tend, int gpd, double cfl,
                      double* Dw, double* Dg, double* dt, int64_t* nsteps);

/* ---- uniform periodic grid, n^3 cells, spacing h (= info.h[0]).
 * nsteps two-pass Euler steps: rhs (ReactionDiffusionOperator.h:98-139)
 * then UpdateTumor (:215-222).  dphidt is an n^3 scratch array. */
void gho_rd_steps_f32(int n, float* phi, float* dphidt, const float* p_w, const float* p_g,
                      float Dw, float Dg, float rho, float h, double dt, int64_t nsteps);
void gho_rd_steps_f64(int n, double* phi, double* dphidt, const double* p_w, const double* p_g,
                      double Dw, double Dg, double rho, double h, double dt, int64_t nsteps);

/* ---- likelihood: tools/pi4u_lite/Inference/Likelihood/source/HGG_Likelihood.cpp:40-192.
 * roi may be NULL (bROI=0).  out[0]=L_PET, out[1]=L_T1, out[2]=L_T2 (log-likelihoods),
 * returns -(L_PET+L_T1+L_T2) as long double rounded to double; *npet gets N. */
double gho_likelihood(int nx, int ny, int nz, const double* model, const double* pet,
                      const double* t1c, const double* flair, const double* roi, int nroi,
                      double PETsigma2, double PETscale, double T1uc, double T2uc, double slope,
                      int stepPET, int stepTi, double out[3], int64_t* npet);
long double gho_log_bernoulli(double u, double y, double uc, double slope);

/* ---- Matrix::D3D .dat files: Glioma/Matrix.h:625-665 */
int gho_dat_write(const char* path, int nx, int ny, int nz, const double* data);
int gho_dat_read_header(const char* path, int dims[3]);
int gho_dat_read(const char* path, int nx, int ny, int nz, double* data);

/* ---- wavelets: MRAG/MRAGcore/MRAGWavelets_StaticData.cpp:13-39 */
/* order 2 or 4; filter in {0:Ha,1:Ga,2:Hs,3:Gs}; writes support[2] and taps; returns ntaps */
int gho_wavelet_taps(int order, int filter, int support[2], double* taps);

/* ---- tangent-linear reference (no reference counterpart: the reference has
 * no derivatives).  Differentiable model: N steps with dt = T/N of the FP64
 * operator above; the no-flux flags, the in-domain test and the upper clamp are
 * frozen at their values along the trajectory (a.e. derivative); the lower
 * clamp is taken as inactive.  IC u0 = the given phi0 (a Gaussian of width
 * sigma around x0 in cells, gated by tissue), so du0/dx0 = u0 (i - x0)/sigma^2.
 * J = -logL(u_N) with the likelihood above (full grid, no ROI).
 * dir = {dDw, drho, dT, dx0, dy0, dz0} (Dw = the operator's rescaled Dw with
 * Dg = 0.1 Dw).  Returns dJ along dir; *loss = J. */
double gho_tlm_f64(int n, const double* phi0, const double* p_w, const double* p_g, double Dw, double Dg,
                   double rho, double h, double T, int64_t N, double sigma, const double* x0, const double* dir,
                   const double* t1c, const double* flair, const double* pet, const double* lik, int stepPET,
                   int stepTi, double* loss);
/* ---- wavelet-adaptive grid (gh_oracle_amr.c).  leaves: nleaves rows of
 * (level, ix, iy, iz) in canonical order (level, iz, iy, ix) ascending; fields
 * block-major in that order.  order = 2 or 4 (interpolating wavelets). */
void gho_amr_lab_f32(int nleaves, const int* leaves, int order, const float* field, int leaf, float* lab);
void gho_amr_lab_f64(int nleaves, const int* leaves, int order, const double* field, int leaf, double* lab);
double gho_amr_fwt_f32(int order, const float* lab);
double gho_amr_fwt_f64(int order, const double* lab);
void gho_amr_details_f32(int nleaves, const int* leaves, int order, const float* field, int nlist, const int* list,
                         double* out);
void gho_amr_details_f64(int nleaves, const int* leaves, int order, const double* field, int nlist, const int* list,
                         double* out);
void gho_amr_split_f32(int order, const float* lab, int octant, float* child);
void gho_amr_split_f64(int order, const double* lab, int octant, double* child);
void gho_amr_rd_steps_f32(int nleaves, const int* leaves, int order, float* phi, const float* p_w,
                          const float* p_g, float Dw, float Dg, float rho, double dt, int64_t nsteps);
void gho_amr_rd_steps_f64(int nleaves, const int* leaves, int order, double* phi, const double* p_w,
                          const double* p_g, double Dw, double Dg, double rho, double dt, int64_t nsteps);
void gho_amr_resample(int nleaves, const int* leaves, const double* phi, int is_f32, int gpd, double* out);

int gho_amr_block_terms(int nleaves, const int* leaves, int order, int leaf, int is_f32, int cap, int* counts,
                        int* src, double* w);
/* ---- MRAG's own ghost construction (gh_oracle_mrag.c): instruction lists of
 * MRAG_BBInfoCreator evaluated as BlockLab::load does.  Return 0, or -1 when
 * the grid is not 2:1 balanced around the block or a ghost is left unfilled. */
int gho_mrag_lab_f32(int nleaves, const int* leaves, int order, const float* field, int leaf, float* lab);
int gho_mrag_lab_f64(int nleaves, const int* leaves, int order, const double* field, int leaf, double* lab);
int gho_mrag_block_terms(int nleaves, const int* leaves, int order, int leaf, int is_f32, int cap, int* counts,
                         int* src, double* w);

/* dJ/du of J = -logL (HGG_Likelihood.cpp:40-156 differentiated), full grid */
void gho_dloss_du(int n, const double* u, const double* t1c, const double* flair, const double* pet,
                  const double* lik, int stepPET, int stepTi, double* g);

#ifdef __cplusplus
}
#endif
#endif
