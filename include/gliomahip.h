/*
 * gliomahip.h -- C ABI of libgliomahip.so, the MI355X (gfx950) implementation
 * of GliomaSolver's reaction-diffusion hot path.
 *
 * Conventions
 *   - every entry point returns int status: GH_OK (0) or a negative GH_E* code;
 *     gh_last_error() returns the message of the calling thread's last failure.
 *     Nothing aborts across the ABI.
 *   - host pointers are borrowed for the duration of the call only.
 *   - "real" buffers hold float when the context was created with GH_REAL_F32
 *     and double with GH_REAL_F64 (the reference's Real, MRAG/MRAGcore/MRAGCommon.h:43-48).
 *   - dense arrays are x-fastest, index i + n*(j + n*k) (Matrix::D3D,
 *     Glioma/Matrix.h:593-601).  Block-major arrays are [block][z][y][x] with
 *     8^3 cells per block, blocks in the order returned by gh_grid_blocks().
 *   - one context per host thread; all device work is ordered on the context's
 *     HIP stream.
 */
#ifndef GLIOMAHIP_H
#define GLIOMAHIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define GH_OK 0
#define GH_EINVAL -1   /* bad argument or size */
#define GH_EHIP -2     /* HIP runtime failure */
#define GH_ESTATE -3   /* call out of order (e.g. no grid yet) */
#define GH_ENOMEM -4   /* device or host allocation failed */
#define GH_EIO -5      /* file could not be read or written */

#define GH_REAL_F32 0
#define GH_REAL_F64 1

typedef struct gh_ctx gh_ctx;

/* ---- context --------------------------------------------------------------
 * Replaces Environment::setup + the BlockCollection owned by Grid
 * (MRAG/MRAGcore/MRAGEnvironment.h:59-72, MRAG/MRAGcore/MRAGrid.inl:20-54). */
int gh_ctx_create(int device, int real_kind, gh_ctx** out);
int gh_ctx_destroy(gh_ctx* ctx);
const char* gh_last_error(void);
int gh_sync(gh_ctx* ctx);

/* ---- grid topology ---------------------------------------------------------
 * gh_grid_init_uniform: Grid<W,B>(bpd,bpd,bpd) -- all blocks at the finest
 * level log2(bpd) (MRAG/MRAGcore/MRAGrid.inl:92-111, MRAGridHelpers.h:9-50).
 * gh_grid_blocks: BlockInfo list (Grid::getBlocksInfo, MRAGrid.inl:190-219):
 * level[b], index[3b..3b+2]; h = 2^-level / 8, origin = index * 2^-level. */
int gh_grid_init_uniform(gh_ctx* ctx, int bpd);
int gh_grid_nblocks(gh_ctx* ctx, int* nblocks);
int gh_grid_blocks(gh_ctx* ctx, int* level, int* index);

/* ---- wavelet-adaptive grid ---------------------------------------------------
 * gh_grid_configure: wavelet order (2: Wavelets_Interp2ndOrder, 4:
 *   Wavelets_Interp4thOrder, MRAG/MRAGcore/MRAGWavelets_StaticData.cpp:13-39),
 *   maximum level (Refiner_SpaceExtension(resJump, maxLevel)) and a compression
 *   floor min_level (0 = none, the reference's behaviour).
 * gh_grid_init_leaves: any 2:1-free tiling by leaf blocks (level[i],
 *   index[3i..3i+2]); fields are zeroed.
 * gh_fwt_details: per block max |detail| of the projector phi
 *   (BlockFWT::multichannel_fwt + FWTReport::getOverAll_DetailMaxMag,
 *   MRAG/MRAGcore/MRAGBlockFWT.h:109-120,219-298,447-478).
 * gh_refine: Science::AutomaticRefinement<0,0>(grid, fwt, tol, maxLevel,
 *   max_loops) (MRAG/MRAGscience/MRAGScienceCore.h:42-108) with
 *   Refiner_SpaceExtension; *nrefined = refined blocks.
 * gh_compress: Science::AutomaticCompression<0,0>(grid, fwt, tol, max_loops)
 *   (:117-177); *ncollapsed = collapsed parents.  max_loops <= 0: until no change.
 *   As the next call after gh_refine on the context it reuses the details the
 *   refinement left valid (same result, fewer evaluations).
 * Both move phi, p_w, p_g, p_csf (split: Hs synthesis, MRAGBlockSplitter.h:24-99;
 * collapse: Ha injection, MRAGBlockCollapser.h:24-85) and recompute the
 * per-solve constants if gh_rd_prepare was called.
 * gh_set/get_field_blocks: block-major arrays in gh_grid_blocks() order.
 * gh_uq_resample: _dumpUQoutput (Glioma/Glioma_ReactionDiffusion.cpp:208-267),
 *   phi onto the uniform gpd^3 lattice (x-fastest, FP64). */
int gh_grid_configure(gh_ctx* ctx, int wavelet_order, int max_level, int min_level);
int gh_grid_init_leaves(gh_ctx* ctx, int nleaves, const int* level, const int* index);
int gh_fwt_details(gh_ctx* ctx, double* detail);
int gh_refine(gh_ctx* ctx, double tol, int max_loops, int* nrefined);
int gh_compress(gh_ctx* ctx, double tol, int max_loops, int* ncollapsed);
/* gh_grid_refine_blocks: Grid::refine(const set<int>& blocksToRefine)
 *   (MRAG/MRAGcore/MRAGrid.h:85, MRAGrid.inl:247-355) for a plan the caller
 *   made: the n listed blocks (ids in gh_grid_blocks() order, duplicates
 *   ignored) with the Refiner_SpaceExtension rules (space extension, maxLevel,
 *   2:1 chain reaction) are split; *nrefined = split blocks
 *   (RefinementResult::nCollapsedParentBlocks).
 * gh_grid_compress_blocks: Grid::compress(const set<int>&, int&
 *   nCollapsedBlocks) (MRAGrid.h:82, MRAGrid.inl:383-486): every parent whose
 *   eight children are listed leaves and pass the Compressor's level-jump test
 *   (MRAGCompressor.h:67-146, and the min_level floor) is collapsed;
 *   *ncollapsed = new parent blocks.  Both move the fields as gh_refine /
 *   gh_compress do; block ids change afterwards. */
int gh_grid_refine_blocks(gh_ctx* ctx, int n, const int* block_ids, int* nrefined);
int gh_grid_compress_blocks(gh_ctx* ctx, int n, const int* block_ids, int* ncollapsed);
/* gh_adapt_profile: host wall time (ms) spent by gh_refine / gh_compress
 * since the last reset, per phase -- the agents of MRAG's Profiler in
 * Science::AutomaticRefinement/Compression (MRAG/MRAGscience/
 * MRAGScienceCore.h:69-93,136-152: "AutoRef::FWT", "AutoRef::refine",
 * "AutoCompr::FWT", "AutoCompr::compress"), then the refine/compress time
 * split into plan, data movement (labs + remap), topology (neighbours, ghost
 * tables, uploads) and per-solve constants:
 * ms[8] = {ref_fwt, ref, comp_fwt, comp, plan, move, topology, statics}.
 * reset != 0 zeroes the counters after reading.  ms may be NULL. */
int gh_adapt_profile(gh_ctx* ctx, double* ms, int reset);
/* gh_adapt_counters: how the ghosts of the adaptation were built since the
 * last reset (no reference counterpart; for tuning and tests):
 * counts[4] = {lab batches evaluated on the device, lab batches from
 * host-built tables, face-ghost tables built on the device, on the host}.
 * Host tables are used only for ghosts that need more than two synthesis
 * levels; any depth of the tree is served on the device otherwise. */
int gh_adapt_counters(gh_ctx* ctx, int64_t* counts, int reset);
int gh_set_field_blocks(gh_ctx* ctx, int field, const void* values);
int gh_get_field_blocks(gh_ctx* ctx, int field, void* values);
int gh_uq_resample(gh_ctx* ctx, int gpd, double* out);
/* host-only plan queries on a leaf list (rows of level, ix, iy, iz in the
 * canonical order level, iz, iy, ix):
 * gh_plan_neighbours: leaves touching `leaf` (MRAGrid.inl:840-1149), rows of
 *   (leaf, wrapped) -- wrapped = a periodic image (capacity 128 rows).
 * gh_plan_refine: blocks refined for the flags (Refiner_SpaceExtension +
 *   Refiner::createPlan, MRAGRefiner_SpaceExtension.h:25-67, MRAGRefiner.h:43-197).
 * gh_plan_compress: parents collapsed for the flags (Compressor::createPlan,
 *   MRAGCompressor.h:67-146) as rows of (level, ix, iy, iz). */
int gh_plan_neighbours(int nleaves, const int* leaves, int leaf, int* nnbr, int* nbr);
int gh_plan_refine(int nleaves, const int* leaves, const unsigned char* flagged, int max_level, int* nvictims,
                   int* victims);
int gh_plan_compress(int nleaves, const int* leaves, const unsigned char* flagged, int min_level, int* nparents,
                     int* parents);

/* ---- fields ----------------------------------------------------------------
 * gh_set_anatomy: Glioma_ReactionDiffusion::_ic anatomy part
 * (Glioma/Glioma_ReactionDiffusion.cpp:79-152): n^3 GM/WM/CSF fractions as
 * read from GM.dat/WM.dat/CSF.dat, majority vote, normalised p_w, p_g, p_csf.
 * Requires a uniform grid with n = 8*bpd (voxel (i,j,k) -> cell (i,j,k)). */
int gh_set_anatomy(gh_ctx* ctx, int n, const double* gm, const double* wm, const double* csf);
/* gh_ic_from_anatomy: all of _ic (Glioma_ReactionDiffusion.cpp:79-172) for
 * anatomy arrays of any nx*ny*nz (1 mm voxels; the cell at x samples voxel
 * floor(x*max(nx,ny,nz)), cells outside the data stay 0) and, when tumor_ic
 * is not NULL, the reference's initial tumour (smoothed ball of radius 0.005
 * centred at tumor_ic in the unit cube, -icx/-icy/-icz or TumorIC.txt). */
int gh_ic_from_anatomy(gh_ctx* ctx, int nx, int ny, int nz, const double* gm, const double* wm,
                       const double* csf, const double* tumor_ic);
/* dense n^3 real arrays <-> the uniform grid's cells */
#define GH_FIELD_PHI 0
#define GH_FIELD_PW 1
#define GH_FIELD_PG 2
#define GH_FIELD_PCSF 3
#define GH_FIELD_DPHIDT 4
int gh_set_field_dense(gh_ctx* ctx, int field, int n, const void* values);
int gh_get_field_dense(gh_ctx* ctx, int field, int n, void* values);

/* ---- schedule --------------------------------------------------------------
 * Glioma_ReactionDiffusion::run (Glioma/Glioma_ReactionDiffusion.cpp:296-320):
 * Dw/(L*L), Dg = 0.1*Dw, dt = cfl*h^2/(6*max(Dw,Dg)) with h = 1/gpd, and the
 * number of iterations of while(t<=tend){t+=dt}, all in Real. */
int gh_schedule(int real_kind, double Dw_phys, double L, double tend, int gpd, double cfl,
                double* Dw, double* Dg, double* dt, int64_t* nsteps);

/* ---- reaction-diffusion ----------------------------------------------------
 * gh_rd_prepare: per-solve constants of ReactionDiffusionOperator(Dw,Dg,rho)
 * (Glioma/Operators/ReactionDiffusionOperator.h:29-35): harmonic face
 * diffusivities (:148-163) and the per-cell no-flux / in-domain thresholds
 * (:103, :166-181).  Must be called after the anatomy changes; returns at
 * once when called again with the same Dw, Dg, rho on unchanged p_w, p_g and
 * grid, so per-step callers pay nothing for it.
 * gh_rd_rhs: pipeline_process(vInfo, coll, bInfo, rhs) -> dphidt
 *   (Glioma_ReactionDiffusion.cpp:185, ReactionDiffusionOperator.h:98-139).
 * gh_rd_update: BlockProcessing::process(vInfo, coll, UpdateTumor(dt))
 *   (Glioma_ReactionDiffusion.cpp:186, ReactionDiffusionOperator.h:215-222).
 * gh_rd_steps: nsteps of _reactionDiffusionStep (Glioma_ReactionDiffusion.cpp:
 *   176-187, 317-322) fused into one kernel per step; bit-identical to
 *   alternating gh_rd_rhs and gh_rd_update.  If elapsed_ms is not NULL the
 *   call synchronises and returns the device time of the steps. */
int gh_rd_prepare(gh_ctx* ctx, double Dw, double Dg, double rho);
int gh_rd_rhs(gh_ctx* ctx);
int gh_rd_update(gh_ctx* ctx, double dt);
int gh_rd_steps(gh_ctx* ctx, int64_t nsteps, double dt, double* elapsed_ms);
/* gh_rd_set_tile / gh_rd_autotune (no reference counterpart): the uniform
 * step kernel's schedule -- 1, 2 or 4 blocks of 8^3 per workgroup and one
 * Euler step per launch, or 0: two Euler steps per launch (temporal blocking,
 * one block per workgroup); the autotuner times trial_steps of each on the
 * current state (restored afterwards) and keeps the fastest, returned in
 * *best.  Results do not depend on the schedule. */
int gh_rd_set_tile(gh_ctx* ctx, int blocks_per_workgroup);
int gh_rd_autotune(gh_ctx* ctx, double dt, int trial_steps, int* best);
/* gh_rd_schedule_info (no reference counterpart): the current schedule
 * (*tile as in gh_rd_set_tile) and whether two-steps-per-launch runs the
 * clamped-state kernel (*clamped_pairs = 1: p_w + p_g >= 0 in every cell, so
 * after one single step from the caller's state every pair of steps takes
 * it; 0: the general two-step kernel). */
int gh_rd_schedule_info(gh_ctx* ctx, int* tile, int* clamped_pairs);
/* gh_rd_zero_flux_blocks: no counterpart in the reference (a schedule
 *   diagnostic).  Blocks of a uniform grid whose cells carry only +0 face
 *   coefficients (no tissue, p_w*Dw + p_g*Dg == 0), which the two-step kernel
 *   advances by each cell's reaction update alone: the operator's flux sums
 *   are exact zeros there (ReactionDiffusionOperator.h:125-135), so the bits
 *   are the same.  Valid after gh_rd_prepare. */
int gh_rd_zero_flux_blocks(gh_ctx* ctx, int* nblocks);
/* gh_rd_thresholds (host-side, no device): the per-cell predicates that
 * gh_rd_prepare stores, for inspection and testing.  tau: the neighbour test
 * "phi + p_w + p_g < 0.1" of ReactionDiffusionOperator.h:115-120,166-181 holds
 * iff phi < tau; sigma: the domain test "p_w + p_g + phi > 0" (:103) holds iff
 * phi > sigma.  n cells of real_kind values. */
int gh_rd_thresholds(int real_kind, int64_t n, const void* p_w, const void* p_g, void* tau, void* sigma);

/* ---- likelihood ------------------------------------------------------------
 * HGG_Likelihood (tools/pi4u_lite/Inference/Likelihood/source/HGG_Likelihood.cpp:
 * 40-192) evaluated on the device in FP64.
 * gh_lik_set_data: the observations tumT1c.dat, tumFLAIR.dat, tumPET.dat (n^3,
 *   x-fastest, y == 1 marks a positive voxel) and optionally ROI.dat points
 *   (nroi rows, stored column-major roi[i + nroi*col], col = x,y,z), for a
 *   uniform grid with n = 8*bpd.
 * gh_likelihood: model u = the context's current phi (HGG_data.dat of a
 *   uniform grid).  params = {PETsigma2, PETscale, T1uc, T2uc, slope}
 *   (LikelihoodInput.txt order), stepPET / stepTi = -stepPET / -stepMRI,
 *   use_roi = -bROI.  out = {-logL, L_PET, L_T1, L_T2}; *npet = PET points.
 * gh_likelihood_dense: the same on host arrays of any nx*ny*nz (e.g. an
 *   HGG_data.dat read from disk). */
int gh_lik_set_data(gh_ctx* ctx, int n, const double* t1c, const double* flair, const double* pet,
                    const double* roi, int nroi);
int gh_likelihood(gh_ctx* ctx, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                  int64_t* npet);
int gh_likelihood_dense(gh_ctx* ctx, int nx, int ny, int nz, const double* model, const double* t1c,
                        const double* flair, const double* pet, const double* roi, int nroi,
                        const double* params, int stepPET, int stepTi, double* out, int64_t* npet);
/* gh_likelihood_noalloc: no counterpart in the reference.  gh_likelihood for
 *   callers that cannot allocate device memory (code under HIP graph capture,
 *   tight sampling loops, harnesses that forbid hipMalloc): same arguments,
 *   same result bit for bit, but the per-workgroup partials live in a
 *   file-scope device array of 4 * (1024 + 1) doubles (32.8 KB per GPU, held
 *   until the process ends) instead of a context buffer, and the call makes no
 *   hipMalloc/hipFree.  The partition is fixed at 1,024 workgroups for every
 *   grid size, so no input exceeds the array.  The array is shared by every
 *   context on a GPU: calls that overlap on one GPU (two threads, two
 *   streams) race and must be serialized by the caller. */
int gh_likelihood_noalloc(gh_ctx* ctx, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                          int64_t* npet);

/* ---- adjoint (no reference counterpart: the reference has no derivatives) --
 * gh_rd_adjoint: forward solve of nsteps fused steps with dt = T/nsteps from
 * the current phi (the IC), J = -logL(u_N) as gh_likelihood, then the
 * hand-written discrete adjoint backwards with checkpoint/recompute every
 * ckpt_every steps.  The no-flux flags, the in-domain test and the upper clamp
 * are frozen along the trajectory; the lower clamp is treated as inactive.
 * Requires gh_rd_prepare(Dw, Dg = Dw/10, rho) and gh_lik_set_data.  The IC is
 * taken to be a Gaussian of width ic_sigma (cells) centred at ic_center
 * (cells), so dJ/dx0 = sum lambda_0 u_0 (i - x0)/sigma^2.
 * loss[4] = {J, L_PET, L_T1, L_T2}; grad[6] = dJ/d{Dw (as passed to
 * gh_rd_prepare), rho, T, x0, y0, z0}, accumulated in FP64.  phi holds u_N on
 * return. */
int gh_rd_adjoint(gh_ctx* ctx, int64_t nsteps, double T, int ckpt_every, const double* lik_params,
                  int stepPET, int stepTi, int use_roi, const double* ic_center, double ic_sigma,
                  double* loss, double* grad);

/* ---- checkpoints (no reference counterpart: the reference keeps one copy of
 * the state) -- device-side copies of phi, used to restart solves from one
 * initial condition and for the adjoint's checkpoint/recompute.  Slots are
 * allocated on first save (0 <= slot <= 4096). */
int gh_checkpoint_save(gh_ctx* ctx, int slot);
int gh_checkpoint_load(gh_ctx* ctx, int slot);

/* ---- batched evaluation of a TMCMC population -------------------------------
 * Replaces the per-sample process chain of the reference's inference
 * (tools/pi4u_lite/Inference/fitfun_glioma.c:118-152: write
 * InputParameters.txt / TumorIC.txt / LikelihoodInput.txt, run runAll.sh =
 * `brain -UQ 1` + `likelihood`, read Likelihood.txt) by one device solve of
 * all samples.  Each sample is a uniform periodic grid of bpd^3 blocks
 * (bpd in [1,64], any value) -- the finest resolution of the reference's
 * adaptive run, at which HGG_data.dat is written -- with the schedule of
 * gh_schedule and the initial condition of gh_ic_from_anatomy; results are
 * bit-identical to solving each sample alone through gh_ctx.
 * gh_batch_add_anatomy: GM/WM/CSF data of any nx*ny*nz (the _ic mapping);
 *   L = 0.1*max(nx,ny,nz) in Real.  Returns its id.
 * gh_batch_set_observations: tumT1c / tumFLAIR / tumPET at the grid
 *   resolution n = 8*bpd (n^3, x fastest) and optional ROI points (as
 *   gh_lik_set_data).
 * gh_batch_set_samples: params = 6 doubles per sample {Dw (cm^2/day), rho
 *   (1/day), Tend (days), icx, icy, icz}; cfl 0.99 is the reference's.
 *   Allocates 2 states + statics per cell (28 B FP32, 56 B FP64); split a
 *   larger population into several batches.
 * gh_batch_run: all Euler steps of all samples; elapsed_ms = device time.
 * gh_batch_likelihood: lik_params = 5 doubles per sample (LikelihoodInput.txt
 *   order); out = 4 doubles per sample as gh_likelihood.
 * gh_batch_sample_info / gh_batch_get_phi: a sample's schedule and state
 *   (n^3 Real, x fastest).
 * gh_batch_get_anatomy: the p_w and p_g an anatomy's samples run on, as _ic
 *   left them on the grid (Glioma_ReactionDiffusion.cpp:103-150: cell
 *   position and voxel index in Real, so the FP32 and FP64 grids of one data
 *   set can differ where 8*bpd is not a power of two); n^3 Real each, either
 *   pointer may be NULL. */
typedef struct gh_batch gh_batch;
int gh_batch_create(int device, int real_kind, gh_batch** out);
int gh_batch_destroy(gh_batch* b);
int gh_batch_add_anatomy(gh_batch* b, int bpd, int nx, int ny, int nz, const double* gm, const double* wm,
                         const double* csf, int* anatomy_id);
int gh_batch_set_observations(gh_batch* b, int anatomy_id, const double* t1c, const double* flair,
                              const double* pet, const double* roi, int nroi);
int gh_batch_set_samples(gh_batch* b, int nsamples, const int* anatomy_id, const double* params, double cfl);
int gh_batch_run(gh_batch* b, double* elapsed_ms);
int gh_batch_likelihood(gh_batch* b, const double* lik_params, int stepPET, int stepTi, int use_roi,
                        double* out);
/* gh_batch_set_adjoint_samples: the problem of gh_rd_adjoint per sample --
 *   params = 6 doubles {Dw (already / L^2, Dg = Dw/10), rho, T, x0, y0, z0
 *   (Gaussian centre in cells)}, nsteps[i] Euler steps of dt = T/N, initial
 *   state ic_peak*exp(-|x-x0|^2/(2 ic_sigma^2)) where p_w + p_g > 0
 *   (gh_gaussian_ic).
 * gh_batch_adjoint: forward with checkpoints every ckpt_every steps, loss,
 *   reverse sweep with recompute, for all samples at once; loss = 4 doubles
 *   and grad = 6 doubles {dJ/dDw, rho, T, x0, y0, z0} per sample, each equal
 *   to gh_rd_adjoint on that sample alone.  Memory: the checkpoints
 *   (sum over samples of ceil(N/K) states), a (K+1)-state tape and two FP64
 *   adjoint states per cell. */
int gh_batch_set_adjoint_samples(gh_batch* b, int nsamples, const int* anatomy_id, const double* params,
                                 const int64_t* nsteps, double ic_sigma, double ic_peak);
int gh_batch_adjoint(gh_batch* b, int ckpt_every, const double* lik_params, int stepPET, int stepTi, int use_roi,
                     double* loss, double* grad, double* elapsed_ms);
int gh_batch_sample_info(gh_batch* b, int sample, double* Dw, double* Dg, double* dt, int64_t* nsteps);
int gh_batch_get_phi(gh_batch* b, int sample, void* dense);
int gh_batch_get_anatomy(gh_batch* b, int anatomy_id, void* p_w, void* p_g);

/* ---- TMCMC sampler (host) ----------------------------------------------------
 * The sampler of tools/pi4u_lite/Inference/engine_tmcmc.c (main loop
 * :1078-1290, prepare_newgen :754-1001, chaintask :683-738) and
 * tmcmc_stats.c (calculate_statistics :545-760) as a state machine that hands
 * out whole rounds of candidates, so one batched solve evaluates them:
 *   gh_tmcmc_create: nth parameters uniform in [lo, hi] (tmcmc_glioma.par
 *     B0..B10), PopSize, MaxStages, Min/MaxChainLength, TolCOV, bbeta, seed.
 *   gh_tmcmc_candidates: the points to evaluate now (*n of them, nth doubles
 *     each; points may be NULL to query n); *n == 0 once finished.
 *   gh_tmcmc_submit: their log-likelihoods (non-finite values become the
 *     reference's -1e12 penalty, fitfun_glioma.c:198-207), in the same order.
 *   gh_tmcmc_state: generation, its annealing exponent p_j, coefficient of
 *     variation, log evidence so far (sum of log-selections), finished flag.
 *   gh_tmcmc_population: the last completed generation (curgen_db).
 *   gh_tmcmc_next_exponent: p_{j+1} from log-likelihoods at p_j (stateless).
 * The random stream is this library's own (xoshiro256**), not GSL's. */
typedef struct gh_tmcmc gh_tmcmc;
int gh_tmcmc_create(int nth, const double* lo, const double* hi, int pop_size, int max_stages, int min_chain,
                    int max_chain, double tol_cov, double bbeta, uint64_t seed, gh_tmcmc** out);
int gh_tmcmc_destroy(gh_tmcmc* t);
int gh_tmcmc_candidates(gh_tmcmc* t, int* n, double* points);
int gh_tmcmc_submit(gh_tmcmc* t, int n, const double* loglik);
int gh_tmcmc_state(gh_tmcmc* t, int* gen, double* p, double* cov, double* log_evidence, int* done);
int gh_tmcmc_population(gh_tmcmc* t, int* n, double* points, double* loglik);
int gh_tmcmc_next_exponent(int n, const double* loglik, double p, double tol_cov, double* p_next, double* cov);
/* gh_glioma_params: fitfun_glioma.c:118-152 for n points of 11 log-space
 *   parameters: sample = {Dw, rho, Tend = sqrt(e^x2 / rho), icx, icy, icz}
 *   and lik = {PETsigma2 = (e^x6)^2, PETscale, T1uc, T2uc, slope}, each
 *   passed through the text the reference writes ("%.16lf", the centre
 *   "%.4f") and read back as brain (Real) or likelihood (double) would.
 * gh_glioma_fitness: -logL -> the value TMCMC receives: Likelihood.txt's
 *   "%Lf" read back, negated; non-finite becomes -1e12 (fitfun_glioma.c:
 *   190-210). */
int gh_glioma_params(int n, const double* x, int real_kind, double* sample, double* lik);
int gh_glioma_fitness(int n, const double* neg_loglik, double* F);

/* ---- synthetic inputs (no reference counterpart: the reference reads
 * patient .dat files) -------------------------------------------------------
 * gh_phantom: seeded brain phantom on an n^3 vertex grid -- outer CSF shell,
 * folded GM cortex, WM interior, two CSF ventricles, tanh phase fields of
 * width 1.5 cells.  seed_voxel receives a seeded WM voxel (i,j,k).
 * gh_gaussian_ic: phi = peak*exp(-|x-x0|^2/(2 sigma^2)) (cells), 0 where
 * p_w+p_g == 0; written as real_kind values. */
int gh_phantom(int n, uint64_t seed, double* gm, double* wm, double* csf, int* seed_voxel);
int gh_gaussian_ic(int real_kind, int n, const void* p_w, const void* p_g, const double* center,
                   double sigma, double peak, void* phi);

#ifdef __cplusplus
}
#endif
#endif
