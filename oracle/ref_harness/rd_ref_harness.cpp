// rd_ref_harness.cpp -- drives the reference's own ReactionDiffusionOperator
// and UpdateTumor (Glioma/Operators/ReactionDiffusionOperator.h, compiled from
// the reference tree by oracle/Makefile) over a uniform periodic grid of 8^3
// blocks.  TEST INFRASTRUCTURE ONLY: its output is the golden answer that the
// oracle restatement (gh_oracle.c) and the HIP path are compared against, and
// its timed loop is the CPU baseline of the benchmark.
//
// What the harness supplies is the part of MRAG that a uniform grid reduces
// to, with MRAG's data shapes so that the timing is the reference's:
//   * Cell: the 27 Real fields of Glioma/Glioma_Types.h:21-201 in their order
//     (108 B in FP32), zero-initialised, with Cell * Real and Cell += Cell
//     applied field by field as there;
//   * BlockLab::load (MRAG/MRAGcore/MRAGBlockLab.h:154-301): the interior rows
//     copied with memcpy, then every ghost evaluated from its instruction list
//     -- on a uniform grid one instruction (the same-level source, weight
//     pool entry 1.0) -- as ghost = Cell(); ghost += value * w;
//   * BlockInfo with h = 1/(8*bpd) (MRAG/MRAGcore/MRAGrid.inl:200-214).
// Real is float unless -D_MRAG_REAL_=double (MRAG/MRAGcore/MRAGCommon.h:43-48).
//
// Input file (binary): int32 n, int64 nsteps, double Dw, Dg, rho, h, dt, then
// p_w[n^3], p_g[n^3], phi[n^3] as Real, x-fastest.  Output: phi[n^3] as Real.
// Optional third argument: worker threads.  Blocks are split into contiguous
// ranges per thread for both passes, as tbb::parallel_for does
// (MRAG/MRAGmultithreading/MRAGBlockProcessing_TBB.h:248-264); each thread owns
// one lab.  The step loop's wall time is printed as "elapsed_s=<t>".
#include "MRAGcore/MRAGCommon.h"
#undef min
#undef max
#include <algorithm>
#include <chrono>
#include <thread>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
using namespace std;
using namespace MRAG;

struct Cell {
    Real phi, dphidt;
    Real p_g, p_w, p_csf;
    Real wm, gm, csf;
    Real dwmdt, dgmdt, dcsfdt;
    Real ux, uy, uz;
    Real omega, domegadt;
    Real p, dpdt;
    Real f;
    Real chi;
    Real pff, dpffdt;
    Real exact;
    Real tmp;
    Real vp;
    Real mob, mu;
    Cell() { std::memset(this, 0, sizeof(Cell)); }
    void operator+=(const Cell& t) {
        Real* a = &phi;
        const Real* b = &t.phi;
        for (int k = 0; k < 27; ++k) a[k] += b[k];
    }
};
static_assert(sizeof(Cell) == 27 * sizeof(Real), "Cell holds the reference's 27 Real fields");
inline Cell operator*(const Cell& c, Real v) {
    Cell r;
    const Real* a = &c.phi;
    Real* o = &r.phi;
    for (int k = 0; k < 27; ++k) o[k] = a[k] * v;
    return r;
}

struct HBlock {
    static const int sizeX = 8, sizeY = 8, sizeZ = 8;
    Cell d[8][8][8];
    Cell& operator()(int ix, int iy = 0, int iz = 0) { return d[iz][iy][ix]; }
};

struct HLab {
    Cell d[10][10][10];
    Cell& operator()(int ix, int iy = 0, int iz = 0) { return d[iz + 1][iy + 1][ix + 1]; }
};

#include "Glioma/Operators/ReactionDiffusionOperator.h"

int main(int argc, char** argv) {
    if (argc != 3 && argc != 4) { fprintf(stderr, "usage: %s in.bin out.bin [threads]\n", argv[0]); return 2; }
    const int nthreads = argc == 4 ? std::max(1, atoi(argv[3])) : 1;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 3;
    int32_t n; int64_t nsteps; double Dw, Dg, rho, h, dt;
    if (fread(&n, 4, 1, f) != 1 || fread(&nsteps, 8, 1, f) != 1 || fread(&Dw, 8, 1, f) != 1 ||
        fread(&Dg, 8, 1, f) != 1 || fread(&rho, 8, 1, f) != 1 || fread(&h, 8, 1, f) != 1 ||
        fread(&dt, 8, 1, f) != 1) return 4;
    if (n % 8 != 0 || n <= 0) return 5;
    const size_t N = (size_t)n * n * n;
    vector<Real> pw(N), pg(N), phi(N);
    if (fread(pw.data(), sizeof(Real), N, f) != N || fread(pg.data(), sizeof(Real), N, f) != N ||
        fread(phi.data(), sizeof(Real), N, f) != N) return 6;
    fclose(f);

    const int bpd = n / 8;
    vector<HBlock> blocks((size_t)bpd * bpd * bpd);
    auto gidx = [n](int i, int j, int k) {
        i = (i + n) % n; j = (j + n) % n; k = (k + n) % n;
        return (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
    };
    for (int bz = 0; bz < bpd; ++bz) for (int by = 0; by < bpd; ++by) for (int bx = 0; bx < bpd; ++bx) {
        HBlock& b = blocks[bx + bpd * (by + bpd * bz)];
        for (int iz = 0; iz < 8; ++iz) for (int iy = 0; iy < 8; ++iy) for (int ix = 0; ix < 8; ++ix) {
            size_t g = gidx(8 * bx + ix, 8 * by + iy, 8 * bz + iz);
            b(ix, iy, iz).p_w = pw[g]; b(ix, iy, iz).p_g = pg[g]; b(ix, iy, iz).phi = phi[g];
        }
    }
    const ReactionDiffusionOperator<3> rhs((Real)Dw, (Real)Dg, (Real)rho);
    const UpdateTumor<3> update(dt);
    const Real hr = (Real)h;
    const Real weight_pool[1] = {(Real)1.0};  // BoundaryInfoBlock::weightsPool[0]
    const int nb = bpd * bpd * bpd;
    auto rhs_range = [&](int b0, int b1) {
        HLab* lab = new HLab;
        for (int b = b0; b < b1; ++b) {
            const int bx = b % bpd, by = (b / bpd) % bpd, bz = b / (bpd * bpd);
            // lab load: interior rows by memcpy, then each ghost from its
            // one-instruction list (same-level source, weight 1.0)
            for (int iz = 0; iz < 8; ++iz)
                for (int iy = 0; iy < 8; ++iy) std::memcpy(&(*lab)(0, iy, iz), &blocks[b](0, iy, iz), 8 * sizeof(Cell));
            for (int iz = -1; iz < 9; ++iz) for (int iy = -1; iy < 9; ++iy) for (int ix = -1; ix < 9; ++ix) {
                if (ix >= 0 && ix < 8 && iy >= 0 && iy < 8 && iz >= 0 && iz < 8) continue;
                int gi = 8 * bx + ix, gj = 8 * by + iy, gk = 8 * bz + iz;
                gi = (gi + n) % n; gj = (gj + n) % n; gk = (gk + n) % n;
                HBlock& src = blocks[(gi / 8) + bpd * ((gj / 8) + bpd * (gk / 8))];
                const Real w = weight_pool[0] * weight_pool[0] * weight_pool[0];
                Cell ghost = Cell();
                ghost += src(gi % 8, gj % 8, gk % 8) * w;
                (*lab)(ix, iy, iz) = ghost;
            }
            const int idx[3] = {bx, by, bz};
            const Real pos[3] = {(Real)(bx * 8 * h), (Real)(by * 8 * h), (Real)(bz * 8 * h)};
            const Real sp[3] = {hr, hr, hr};
            BlockInfo info(b, idx, 0, pos, sp);
            rhs(*lab, info, blocks[b]);
        }
        delete lab;
    };
    auto update_range = [&](int b0, int b1) {
        for (int b = b0; b < b1; ++b) {
            const int idx[3] = {0, 0, 0};
            BlockInfo info(b, idx, 0);
            update(info, blocks[b]);
        }
    };
    auto parallel = [&](auto&& body) {
        if (nthreads == 1) { body(0, nb); return; }
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; ++t) {
            const int b0 = (int)((long)nb * t / nthreads), b1 = (int)((long)nb * (t + 1) / nthreads);
            pool.emplace_back([&, b0, b1] { body(b0, b1); });
        }
        for (auto& th : pool) th.join();
    };
    const auto t0 = std::chrono::steady_clock::now();
    for (int64_t s = 0; s < nsteps; ++s) {
        parallel(rhs_range);
        parallel(update_range);
    }
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    printf("elapsed_s=%.9f\n", el);
    for (int bz = 0; bz < bpd; ++bz) for (int by = 0; by < bpd; ++by) for (int bx = 0; bx < bpd; ++bx) {
        HBlock& b = blocks[bx + bpd * (by + bpd * bz)];
        for (int iz = 0; iz < 8; ++iz) for (int iy = 0; iy < 8; ++iy) for (int ix = 0; ix < 8; ++ix)
            phi[gidx(8 * bx + ix, 8 * by + iy, 8 * bz + iz)] = b(ix, iy, iz).phi;
    }
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 7;
    if (fwrite(phi.data(), sizeof(Real), N, o) != N) return 8;
    fclose(o);
    return 0;
}
