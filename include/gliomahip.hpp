/*
 * gliomahip.hpp -- C++ host surface over the C ABI (gliomahip.h), shaped like
 * the reference's MRAG/Glioma types so that a Glioma_ReactionDiffusion-style
 * driver keeps its structure:
 *
 *   Grid<W,B>(bpd,bpd,bpd,maxStencil)            MRAG/MRAGcore/MRAGrid.h:57-91
 *   BlockInfo{blockID,index,level,origin,h}       MRAG/MRAGcore/MRAGCommon.h:166-260
 *   ReactionDiffusionOperator<_DIM>(Dw,Dg,rho)    Glioma/Operators/ReactionDiffusionOperator.h:27-47
 *   UpdateTumor<_DIM>(dt)                         Glioma/Operators/ReactionDiffusionOperator.h:189-200
 *   BlockProcessing::pipeline_process / process   MRAG/MRAGcore/MRAGBlockProcessing_TBB.h:447-512
 *   Science::AutomaticRefinement / Compression    MRAG/MRAGscience/MRAGScienceCore.h:42-45,117-120
 *   Matrix::D3D<double> .dat files                Glioma/Matrix.h:26-34,593-665
 *
 * The difference is ownership: the block fields live in device memory inside
 * the Grid's context; BlockInfo carries geometry only (no ptrBlock), and the
 * processing calls advance the WHOLE grid in one library call, which is the
 * only way the reference calls them (Glioma_ReactionDiffusion.cpp:176-187).
 * Passing a block list that is not the grid's full, current list is an error.
 *
 * Errors: every failing call throws gliomahip::Error carrying the status code
 * and gh_last_error() (the reference aborts instead).
 */
#ifndef GLIOMAHIP_HPP
#define GLIOMAHIP_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <set>
#include <vector>

#include "gliomahip.h"

namespace gliomahip {

class Error : public std::runtime_error {
public:
    int code;
    Error(int c, const std::string& m) : std::runtime_error("gliomahip error " + std::to_string(c) + ": " + m), code(c) {}
};

inline void check(int rc) {
    if (rc != GH_OK) throw Error(rc, gh_last_error());
}

/* ---------------------------------------------------------------------------
 * Matrix::D3D<double>: header {1234, 3, nx, ny, nz} (int), type id (int:
 * 0 double, 1 float, 2 int), nx*ny*nz values, x fastest (Glioma/Matrix.h:26-34,
 * 593-601, 625-665).  Values are held as double, like MatrixD3D.
 * ------------------------------------------------------------------------- */
class MatrixD3D {
public:
    MatrixD3D() {}
    MatrixD3D(int nx, int ny, int nz) : nx_(nx), ny_(ny), nz_(nz), data_((size_t)nx * ny * nz, 0.0) {}
    explicit MatrixD3D(const std::string& path) { load(path); }

    int getSizeX() const { return nx_; }
    int getSizeY() const { return ny_; }
    int getSizeZ() const { return nz_; }
    double operator()(int i, int j, int k) const { return data_[i + (size_t)nx_ * (j + (size_t)ny_ * k)]; }
    double& operator()(int i, int j, int k) { return data_[i + (size_t)nx_ * (j + (size_t)ny_ * k)]; }
    const double* data() const { return data_.data(); }
    double* data() { return data_.data(); }
    size_t size() const { return data_.size(); }

    void load(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f.is_open()) throw Error(GH_EIO, "cannot open " + path);
        int hdr[6];
        f.read((char*)hdr, sizeof(hdr));
        if (!f || hdr[0] != 1234 || hdr[1] != 3 || hdr[2] <= 0 || hdr[3] <= 0 || hdr[4] <= 0)
            throw Error(GH_EIO, "bad Matrix::D3D header in " + path);
        nx_ = hdr[2];
        ny_ = hdr[3];
        nz_ = hdr[4];
        const size_t n = (size_t)nx_ * ny_ * nz_;
        data_.assign(n, 0.0);
        if (hdr[5] == 0) {
            f.read((char*)data_.data(), n * sizeof(double));
        } else if (hdr[5] == 1) {
            std::vector<float> t(n);
            f.read((char*)t.data(), n * sizeof(float));
            for (size_t i = 0; i < n; ++i) data_[i] = t[i];
        } else if (hdr[5] == 2) {
            std::vector<int> t(n);
            f.read((char*)t.data(), n * sizeof(int));
            for (size_t i = 0; i < n; ++i) data_[i] = t[i];
        } else {
            throw Error(GH_EIO, "unknown Matrix type id in " + path);
        }
        if (!f) throw Error(GH_EIO, "truncated data in " + path);
    }

    void dump(const std::string& path) const {
        std::ofstream f(path, std::ios::binary);
        if (!f.is_open()) throw Error(GH_EIO, "cannot write " + path);
        const int hdr[6] = {1234, 3, nx_, ny_, nz_, 0};
        f.write((const char*)hdr, sizeof(hdr));
        f.write((const char*)data_.data(), data_.size() * sizeof(double));
        if (!f) throw Error(GH_EIO, "write failed: " + path);
    }

private:
    int nx_ = 0, ny_ = 0, nz_ = 0;
    std::vector<double> data_;
};

/* Matrix::D2D<double>: header {1234, 2, nx, ny}, type id, values x fastest
 * (Glioma/Matrix.h:26-34; ROI.dat holds N points as an N x 3 matrix). */
class MatrixD2D {
public:
    explicit MatrixD2D(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f.is_open()) throw Error(GH_EIO, "cannot open " + path);
        int hdr[5];
        f.read((char*)hdr, sizeof(hdr));
        if (!f || hdr[0] != 1234 || hdr[1] != 2 || hdr[2] < 0 || hdr[3] <= 0)
            throw Error(GH_EIO, "bad Matrix::D2D header in " + path);
        nx_ = hdr[2];
        ny_ = hdr[3];
        const size_t n = (size_t)nx_ * ny_;
        data_.assign(n, 0.0);
        if (hdr[4] == 0) {
            f.read((char*)data_.data(), n * sizeof(double));
        } else if (hdr[4] == 1) {
            std::vector<float> t(n);
            f.read((char*)t.data(), n * sizeof(float));
            for (size_t i = 0; i < n; ++i) data_[i] = t[i];
        } else if (hdr[4] == 2) {
            std::vector<int> t(n);
            f.read((char*)t.data(), n * sizeof(int));
            for (size_t i = 0; i < n; ++i) data_[i] = t[i];
        } else {
            throw Error(GH_EIO, "unknown Matrix type id in " + path);
        }
        if (!f) throw Error(GH_EIO, "truncated data in " + path);
    }
    int getSizeX() const { return nx_; }
    int getSizeY() const { return ny_; }
    double operator()(int i, int j) const { return data_[i + (size_t)nx_ * j]; }
    const double* data() const { return data_.data(); }

private:
    int nx_ = 0, ny_ = 0;
    std::vector<double> data_;
};

/* One device context (stream, fields, tables).  Not copyable. */
class Context {
public:
    explicit Context(int device = 0, int real_kind = GH_REAL_F32) { check(gh_ctx_create(device, real_kind, &h_)); }
    ~Context() {
        if (h_) gh_ctx_destroy(h_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    gh_ctx* get() const { return h_; }

private:
    gh_ctx* h_ = nullptr;
};

/* MRAG/MRAGcore/MRAGCommon.h:166-260: origin = index * 2^-level,
 * h = 2^-level / 8, pos = origin + h * i.  blockID is the block's slot in
 * the grid's canonical (level, z, y, x) order; it changes on adaptation. */
struct BlockInfo {
    int blockID;
    short int index[3];
    short int level;
    double origin[3];
    double h[3];

    template <typename T>
    void pos(T p[3], int ix, int iy, int iz) const {
        p[0] = origin[0] + h[0] * ix;
        p[1] = origin[1] + h[1] * iy;
        p[2] = origin[2] + h[2] * iz;
    }
};

/* Ghost tables live on the device; these are handles with the reference's names. */
class Grid;
struct BoundaryInfo {
    Grid* grid;
};
struct BlockCollection {
    Grid* grid;
};

/* Refiner_SpaceExtension(resJump, maxLevel) and Compressor(resJump)
 * (MRAG/MRAGscience/MRAGRefiner_SpaceExtension.h:25-31, MRAGCompressor.h:30-35).
 * Only resJump == 1 is supported, the reference's configuration
 * (Glioma/Glioma_Types.h:281-296). */
struct Refiner_SpaceExtension {
    int resJump, maxLevel;
    Refiner_SpaceExtension(int rj, int ml) : resJump(rj), maxLevel(ml) {}
};
struct Compressor {
    int resJump;
    explicit Compressor(int rj) : resJump(rj) {}
};
/* Outcome of Grid::refine / Grid::compress (MRAGRefinementPlan.h:86-104,
 * MRAGCompressionPlan.h:45-49). */
struct RefinementResult {
    int nChildrenBlocks = 0;
    int nCollapsedParentBlocks = 0;
    bool hasFailed() const { return false; }
};
struct CompressionResult {
    int nCollapsedBlocks = 0;
    int nNewBlocks = 0;
};
/* BlockFWT<W,B,RD_Projector_Wavelets>: details of phi (Glioma_Types.h:247-255). */
struct BlockFWT {};
struct Profiler {};

/*
 * Grid<W,B>(bpdx, bpdy, bpdz, maxStencil): bpd^3 blocks of 8^3 cells on level
 * log2(bpd) of the unit cube, periodic, interpolating wavelets of order 2
 * (Wavelets_Interp2ndOrder, the reference's _WAVELET_TYPE, MRAG/MRAGHeaders.h:14)
 * or 4.  maxStencil is accepted for signature compatibility; the lab width
 * follows from the wavelet order.
 */
class Grid {
public:
    Grid(int bpdx, int bpdy, int bpdz, const int maxStencil[2][3] = nullptr, int real_kind = GH_REAL_F32,
         int device = 0, int wavelet_order = 2, int max_level = -1)
        : ctx_(device, real_kind), bpd_(bpdx), real_kind_(real_kind) {
        (void)maxStencil;
        if (bpdx != bpdy || bpdy != bpdz) throw Error(GH_EINVAL, "Grid: blocks per dimension must be equal");
        int lvl = 0;
        while ((1 << lvl) < bpdx) ++lvl;
        max_level_ = max_level < 0 ? lvl : max_level;
        check(gh_grid_configure(ctx_.get(), wavelet_order, max_level_, 0));
        check(gh_grid_init_uniform(ctx_.get(), bpdx));
        order_ = wavelet_order;
        bi_.grid = this;
        bc_.grid = this;
    }

    gh_ctx* ctx() const { return ctx_.get(); }
    int realKind() const { return real_kind_; }
    int blocksPerDimension() const { return bpd_; }
    int maxLevel() const { return max_level_; }
    int waveletOrder() const { return order_; }

    std::vector<BlockInfo> getBlocksInfo() const {
        int n = 0;
        check(gh_grid_nblocks(ctx_.get(), &n));
        std::vector<int> lv(n), idx(3 * (size_t)n);
        check(gh_grid_blocks(ctx_.get(), lv.data(), idx.data()));
        std::vector<BlockInfo> out(n);
        for (int b = 0; b < n; ++b) {
            BlockInfo& i = out[b];
            i.blockID = b;
            i.level = (short)lv[b];
            const double ext = std::ldexp(1.0, -lv[b]);
            for (int a = 0; a < 3; ++a) {
                i.index[a] = (short)idx[3 * b + a];
                i.origin[a] = idx[3 * b + a] * ext;
                i.h[a] = ext / 8;
            }
        }
        return out;
    }
    int numberOfBlocks() const {
        int n = 0;
        check(gh_grid_nblocks(ctx_.get(), &n));
        return n;
    }
    BlockCollection& getBlockCollection() { return bc_; }
    BoundaryInfo& getBoundaryInfo() { return bi_; }
    void setRefiner(Refiner_SpaceExtension* r) {
        if (!r) return;
        if (r->resJump != 1) throw Error(GH_EINVAL, "Grid: only resJump 1 is supported");
        max_level_ = r->maxLevel;
        check(gh_grid_configure(ctx_.get(), order_, max_level_, 0));
    }
    void setCompressor(Compressor* c) {
        if (c && c->resJump != 1) throw Error(GH_EINVAL, "Grid: only resJump 1 is supported");
    }

    /* Grid::refine(const set<int>&) and Grid::compress(const set<int>&, int&)
     * (MRAG/MRAGcore/MRAGrid.h:82-85): apply a plan made by the caller over
     * blockIDs of getBlocksInfo(); the results carry the reference's counters
     * (MRAGRefinementPlan.h:86-104, MRAGCompressionPlan.h:45-49). */
    RefinementResult refine(const std::set<int>& blocksToRefine) {
        std::vector<int> ids(blocksToRefine.begin(), blocksToRefine.end());
        int n = 0;
        check(gh_grid_refine_blocks(ctx_.get(), (int)ids.size(), ids.data(), &n));
        RefinementResult r;
        r.nCollapsedParentBlocks = n;
        r.nChildrenBlocks = 8 * n;
        return r;
    }
    CompressionResult compress(const std::set<int>& blocksToCompress, int& nCollapsedBlocks) {
        std::vector<int> ids(blocksToCompress.begin(), blocksToCompress.end());
        int n = 0;
        check(gh_grid_compress_blocks(ctx_.get(), (int)ids.size(), ids.data(), &n));
        nCollapsedBlocks = n;
        CompressionResult r;
        r.nCollapsedBlocks = n;
        r.nNewBlocks = n;
        return r;
    }

    /* Field access at the uniform finest resolution (uniform grids only). */
    void setDense(int field, const void* values) {
        check(gh_set_field_dense(ctx_.get(), field, 8 * bpd_, values));
    }
    void getDense(int field, void* values) const { check(gh_get_field_dense(ctx_.get(), field, 8 * bpd_, values)); }

private:
    Context ctx_;
    int bpd_, real_kind_, order_ = 2, max_level_ = 0;
    BoundaryInfo bi_;
    BlockCollection bc_;
};

/* ReactionDiffusionOperator<_DIM>(Dw, Dg, rho): the 7-point harmonic-mean
 * Fisher-KPP right-hand side (ReactionDiffusionOperator.h:27-139).
 * Dw, Dg are already divided by L^2, as in Glioma_ReactionDiffusion.cpp:304-305. */
template <int nDim = 3>
struct ReactionDiffusionOperator {
    static_assert(nDim == 3, "the 3D operator only");
    int stencil_start[3] = {-1, -1, -1};
    int stencil_end[3] = {2, 2, 2};
    double Dw, Dg, rho;
    ReactionDiffusionOperator(double Dw_, double Dg_, double rho_) : Dw(Dw_), Dg(Dg_), rho(rho_) {}
};

/* UpdateTumor<_DIM>(dt): phi = clamp(phi + dt*dphidt, 0, 1) (ReactionDiffusionOperator.h:189-227). */
template <int nDim = 3>
struct UpdateTumor {
    static_assert(nDim == 3, "the 3D operator only");
    int stencil_start[3] = {0, 0, 0};
    int stencil_end[3] = {1, 1, 1};
    double dt;
    explicit UpdateTumor(double dt_) : dt(dt_) {}
};

/*
 * BlockProcessing with the reference's method names
 * (MRAGBlockProcessing_TBB.h:447-512, MRAGBlockProcessing_SingleCPU.h:306-326).
 * pipeline_process(vInfo, coll, bInfo, rhs): dphidt for every block (ghosts
 * filled on the device, across level jumps); process(vInfo, coll, update, g):
 * the Euler update.  The coefficient statics are recomputed when Dw, Dg or
 * the grid changed since the last call.
 */
class BlockProcessing_HIP {
public:
    template <int D>
    void pipeline_process(std::vector<BlockInfo>& vInfo, const BlockCollection& coll, BoundaryInfo& bi,
                          const ReactionDiffusionOperator<D>& rhs) {
        Grid& g = *coll.grid;
        if (bi.grid != coll.grid) throw Error(GH_EINVAL, "pipeline_process: BoundaryInfo of another grid");
        checkList(vInfo, g);
        check(gh_rd_prepare(g.ctx(), rhs.Dw, rhs.Dg, rhs.rho));
        check(gh_rd_rhs(g.ctx()));
    }
    template <int D>
    static void process(std::vector<BlockInfo>& vInfo, const BlockCollection& coll, const UpdateTumor<D>& up,
                        int nParallelGranularity = 1) {
        (void)nParallelGranularity;
        checkList(vInfo, *coll.grid);
        check(gh_rd_update(coll.grid->ctx(), up.dt));
    }

private:
    static void checkList(const std::vector<BlockInfo>& vInfo, const Grid& g) {
        if ((int)vInfo.size() != g.numberOfBlocks())
            throw Error(GH_EINVAL, "block list is not the grid's full current list");
    }
};
typedef BlockProcessing_HIP BlockProcessing;

namespace Science {
/* MRAGScienceCore.h:42-108: refine blocks whose detail >= tol, up to
 * iMaxLoops passes (<= 0: until nothing changes); blocks at iMaxLevel are
 * left alone.  Returns the number of blocks refined. */
template <int a = 0, int b = 0>
int AutomaticRefinement(Grid& g, BlockFWT&, double tol, int iMaxLevel, int iMaxLoops = -1, Profiler* = nullptr) {
    check(gh_grid_configure(g.ctx(), g.waveletOrder(), iMaxLevel < 0 ? g.maxLevel() : iMaxLevel, 0));
    int n = 0;
    check(gh_refine(g.ctx(), tol, iMaxLoops, &n));
    return n;
}
/* MRAGScienceCore.h:117-177: collapse families whose details are < tol. */
template <int a = 0, int b = 0>
int AutomaticCompression(Grid& g, BlockFWT&, double tol, int iMaxLoops = -1, Profiler* = nullptr) {
    int n = 0;
    check(gh_compress(g.ctx(), tol, iMaxLoops, &n));
    return n;
}
}  // namespace Science

/* Glioma_ReactionDiffusion::_ic (Glioma_ReactionDiffusion.cpp:79-172) on a
 * uniform Grid: anatomy from PatientFileName{GM,WM,CSF}.dat, L = 0.1 *
 * max(size) (cm), tumour at tumor_ic.  Returns L. */
inline double InitialCondition(Grid& g, const std::string& patientFileName, const double tumor_ic[3]) {
    MatrixD3D gm(patientFileName + "GM.dat"), wm(patientFileName + "WM.dat"), csf(patientFileName + "CSF.dat");
    if (wm.getSizeX() != gm.getSizeX() || wm.getSizeY() != gm.getSizeY() || wm.getSizeZ() != gm.getSizeZ() ||
        csf.getSizeX() != gm.getSizeX() || csf.getSizeY() != gm.getSizeY() || csf.getSizeZ() != gm.getSizeZ())
        throw Error(GH_EINVAL, "GM/WM/CSF sizes differ");
    check(gh_ic_from_anatomy(g.ctx(), gm.getSizeX(), gm.getSizeY(), gm.getSizeZ(), gm.data(), wm.data(), csf.data(),
                             tumor_ic));
    const int bmax = std::max(gm.getSizeX(), std::max(gm.getSizeY(), gm.getSizeZ()));
    // L = brainSizeMax * 0.1, stored in Real (Glioma_ReactionDiffusion.cpp:95)
    return g.realKind() == GH_REAL_F32 ? (double)(float)(bmax * 0.1) : bmax * 0.1;
}

/* _dumpUQoutput (Glioma_ReactionDiffusion.cpp:208-267): phi at the uniform
 * finest resolution gpd^3 as a Matrix::D3D<double>. */
inline MatrixD3D UQOutput(const Grid& g) {
    const int gpd = 8 * g.blocksPerDimension();
    MatrixD3D m(gpd, gpd, gpd);
    check(gh_uq_resample(g.ctx(), gpd, m.data()));
    return m;
}

}  // namespace gliomahip

#endif
