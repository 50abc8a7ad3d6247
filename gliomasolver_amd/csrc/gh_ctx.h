// gh_ctx.h -- device context shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "gh_internal.h"
#include "gh_octree.h"

namespace gh {

#define GH_HIP_CHECK(expr)                                                                     \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return ::gh::set_error(GH_EHIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

// Device buffer (plain hipMalloc), move-only; the destructor frees it, so an
// early error return releases every local buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int ensure(size_t nbytes);
    // as ensure, but a reallocation takes 25% headroom (buffers of a growing
    // adaptive grid are not reallocated at every adaptation)
    int ensure_grow(size_t nbytes);
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Pinned host staging buffer (hipHostMalloc), move-only, freed by the
// destructor.  The adaptation's per-epoch tables go to and from the device
// through these, so each hipMemcpyAsync is one DMA (from pageable memory the
// runtime stages the copy: 100 KB 48 vs 10 us, 1.1 MB 61 vs 27 us on MI355X).
// A buffer is rewritten only after a synchronization that follows its last
// copy.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { release(); }
    int ensure_grow(size_t nbytes);  // 25% headroom on a reallocation
    void release();
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Per-context device state.  Field arrays are block-major [slot][z][y][x],
// 512 cells per slot, element type Real (float or double per real_kind).
struct Ctx {
    int device = 0;
    int real_kind = GH_REAL_F32;
    size_t real_size = 4;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // recorded after the last uploads from the pinned staging buffers (and from
    // the tree's hash table); waited for before any of them is rewritten, so a
    // tree change need not wait for its remap and uploads to finish
    hipEvent_t evStage = nullptr;
    Topology topo;
    Octree tree;  // host octree the topology is derived from

    // adaptive grid: per-slot level, same-level neighbours in 27 directions,
    // interpolated face ghosts of the RD stencil (jump faces only)
    DevBuf dLevel, dNbr27;
    DevBuf faceEnt, faceSrc, faceW;  // GhostEntry / int32 / Real
    DevBuf faceList, faceCount;      // device face-table build: jump faces, per-ghost term counts
    DevBuf faceScan;                 // per slot face + 1: jump flag, then its exclusive prefix sum
    DevBuf scanBuf;                  // face-table prefix sums: chunk sums, offsets, total
    int nFaceEnt = 0;
    // device-built face table in packets: per jump face, its 64 ghosts' term
    // lists interleaved (term k of ghost p at facePkBase[face] + 64 k + p for
    // k below its own term count in faceCount), so a wave's loads are contiguous
    DevBuf facePkBase;               // int32 per face + 1: packet starts (exclusive scan)
    DevBuf facePkOf;                 // int32 per slot face: its packet (index in the face list), -1 if none
    int nFacePk = 0;                 // faces in packets (0: the GhostEntry table is used)
    DevBuf fgPhi, fgTau, fgDm;       // per slot: 6*64 ghost phi, 6*64 ghost tau, 3*64 minus-face D
    DevBuf fgPw, fgPg;               // per slot: 6*64 ghost p_w, p_g (statics)
    // FWT / split labs (compact, one per listed slot)
    DevBuf labBuf, labSlots, labEnt, labSrc, labW;
    DevBuf labPerm;        // lab_kernel's cell order (int16, lab_order)
    int labPermOrder = 0;  // wavelet order labPerm was built for (0: none)
    // ghost table of the labs of every leaf, valid while tree_version == labAllVersion
    // (host-table path, used when the level maps below are not available)
    DevBuf labAllEnt, labAllSrc, labAllW, labAllSlots;
    int nLabAllEnt = 0;
    uint64_t tree_version = 0, labAllVersion = ~(uint64_t)0;
    // per-level block maps for device-side ghost evaluation (Octree::dense,
    // concatenated; nDense = 0: not available), block (level, x, y, z) per slot
    DevBuf dDense, dBlk, dBad;
    int nDense = 0;
    int topLevel = -1;           // finest leaf level
    DevBuf dHashKey, dHashVal;   // nodes of levels nDense..topLevel (Octree::hash_keys / hash_vals)
    int hashBits = 0;            // 0: no hashed levels
    double last_fwt_ms = 0;
    // adaptation scratch kept across epochs: spare field arrays swapped with
    // the live ones by apply_new_tree, split labs, remap list, FWT details
    DevBuf spare[4], labSplit[4], remapBuf, detBuf;
    PinnedBuf pinRemap, pinDense, pinBlk, pinSplit, pinSlots, pinDet;  // staging of the adaptation's host<->device tables
    // adaptation profile (gh_adapt_profile), host wall time in ms per phase
    double prof[8] = {0};
    // how ghosts were built (gh_adapt_counters): lab batches on the device,
    // lab batches from host tables, face tables on the device, from the host
    int64_t adapt_count[4] = {0};
    // FWT details (a flag and the value per slot) that gh_refine leaves valid
    // on the tree and data it ends with, for a gh_compress called right after
    // it: `calls` counts entry-point calls on this context (as_ctx), and the
    // details are used only when the compression is the next call after the
    // refinement's (carry_call), so that nothing can have changed them
    uint64_t calls = 0, carry_call = 0;
    std::vector<char> carry_known;
    std::vector<double> carry_det;

    DevBuf phi[2];      // ping-pong tumor density
    int cur = 0;        // phi[cur] holds the current state
    DevBuf dphidt;      // two-pass right-hand side
    DevBuf pw, pg, pcsf;
    DevBuf statD;       // 4 * ncells: per cell {D of the +x,+y,+z face, tau}: outside iff phi < tau
    DevBuf statSig;     // ncells: cell is in the domain iff phi > sigma
    DevBuf statF;       // 4 * ncells: {D+x, D+y, D+z, tau_enc} of rd_step2f_kernel (valid iff fast_ok)
    DevBuf flag;        // small scratch word for device-side counts
    bool fast_ok = false;  // every cell has p_w + p_g >= 0: clamped states may take rd_step2f_kernel
    DevBuf zeroFlux;       // per slot: 1 = block without diffusion (zero_flux_kernel), uniform grids
    int nZeroFlux = 0;     // blocks without diffusion
    DevBuf items2f;        // rd_step2f_kernel's work items {slot, blocks without diffusion} (plan_step2f)
    int nItems2f = 0;      // 0: one item per slot
    int split2f[9] = {0};  // per-XCD item ranges, balanced by estimated cost
    DevBuf nbr;         // 6 * nblocks int32
    std::vector<DevBuf> ckpt;  // device-side copies of phi (gh_checkpoint_*)
    std::vector<uint64_t> ckptVersion;  // tree_version each checkpoint was saved on
    // likelihood data (block-major): y == 1 flags of T1c/FLAIR, PET values, ROI {cell, ix, iy, iz}
    DevBuf obsT1, obsFl, obsPet, obsRoi, likPart;
    // adjoint: forward checkpoints, segment tape, adjoint ping-pong (FP64), parameter partials
    DevBuf adjCk, adjSeg, adjLam[2], adjPart;
    int nroi = 0;
    bool lik_ready = false;
    uint64_t lik_version = 0;  // tree_version the likelihood data was laid out for
    // likelihood data usable: set for this topology (apply_topology clears
    // lik_ready) on a uniform grid of the same size
    bool lik_valid() const { return lik_ready && lik_version == tree_version && topo.uniform_bpd > 0; }
    bool prepared = false;
    // set by a grid change: the face-ghost tables of the RD step, and (when
    // prepared) the per-solve constants, are rebuilt on their next use, so a
    // grid that the same adaptation pass changes again is never prepared
    bool faces_stale = false, constants_stale = false;
    double Dw = 0, Dg = 0, rho = 0;
    double dt_last = 0;  // dt of the running gh_rd_steps call
    int rd_bpw = 2;      // blocks per workgroup of the step kernel (1, 2 or 4)

    size_t ncells() const { return (size_t)topo.nblocks() * BCELLS; }
    int alloc_fields();
    void release_all();  // every device buffer of the context
};

Ctx* as_ctx(gh_ctx* c);

// RD launchers (gh_rd.hip)
int rd_prepare(Ctx& c, double Dw, double Dg, double rho);
// the constants of the current grid when a grid change left them stale
int rd_ensure_prepared(Ctx& c);
int rd_rhs(Ctx& c);
int rd_update(Ctx& c, double dt);
int rd_steps(Ctx& c, int64_t nsteps, double dt);
// one fused step from an explicit state buffer into another (same layout as phi)
int rd_step_buffers(Ctx& c, const void* in, void* out, double dt);
// two fused steps from a clamped state (every value >= +0, as written by the
// fused update) when c.fast_ok; mid (may be null) gets the state in between
int rd_step2_buffers(Ctx& c, const void* in, void* out, void* mid, double dt);
double block_h_of(const Ctx& c);
// the rd_step2f_kernel records of n cells from {D+, tau} and sigma into out;
// *ok = every cell has sigma <= 0 (p_w + p_g >= 0).  Synchronizes the stream.
int build_stat_fast(hipStream_t stream, int real_kind, size_t n, const void* st, const void* sig, DevBuf& out,
                    DevBuf& flag, bool* ok);
int rd_autotune(Ctx& c, double dt, int trial_steps, int* best);

// adaptive grid (gh_adapt.hip)
int apply_topology(Ctx& c);
int rd_prepare_ghosted(Ctx& c, double Dw, double Dg);
int rd_step_ghosted(Ctx& c, const void* in, void* out, double dt, bool fused);
// meanwhile (may be empty) runs on the host while the device evaluates the
// details (host work that does not need them: the tree-only parts of a plan)
int fwt_details(Ctx& c, const std::vector<int>& slots, std::vector<double>& det,
                const std::function<void()>& meanwhile = {});
// old_slot (may be null) receives, per new slot, the old slot of the same
// leaf or -1 for a new one (left untouched when nothing changes); ext (may be
// null) is the current tree's space_extension(); split_keys (may be null)
// receives the keys of the split leaves
int grid_refine(Ctx& c, const std::vector<char>& flagged, int* nrefined, std::vector<int>* old_slot = nullptr,
                const std::vector<char>* ext = nullptr, std::vector<uint64_t>* split_keys = nullptr);
// parents (may be null) receives the keys of the new parent leaves; cands
// (may be null) is the current tree's collapse_candidates()
int grid_compress(Ctx& c, const std::vector<char>& flagged, int* ncollapsed,
                  std::vector<uint64_t>* parents = nullptr, std::vector<int>* old_slot = nullptr,
                  const std::vector<Octree::CollapseCand>* cands = nullptr);

// likelihood helpers (gh_lik.hip): -logL of the current phi; dJ/du of a state buffer
int lik_eval_current(Ctx& c, const double* params, int stepPET, int stepTi, int use_roi, double* out,
                     int64_t* npet);
int lik_seed_gradient(Ctx& c, const void* u, double* g, const double* params, int stepPET, int stepTi,
                      int use_roi);
// resolves the device address of gh_likelihood_noalloc's scratch on `device`
// (current); once per device
int lik_scratch_resolve(int device);

}  // namespace gh
