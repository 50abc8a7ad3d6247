// brain: the reaction-diffusion driver of the reference's `brain -model RD`
// (Glioma/Glioma_main.cpp:52-54, Glioma/Glioma_ReactionDiffusion.cpp:16-66,
// 272-334) on the HIP library, with the same options and files:
//
//   -PatFileName P   anatomy P{GM,WM,CSF}.dat (Matrix::D3D)
//   -UQ 1            read TumorIC.txt (x y z) and InputParameters.txt
//                    (Dw rho Tend); write HGG_data.dat (phi at the finest
//                    uniform resolution) at the end
//   -Dw -rho -Tend   (defaults 0.0013, 0.025, 300) when -UQ 0
//   -icx -icy -icz   tumour centre (defaults 0.28, 0.67, 0.35) when -UQ 0
//   -adaptive 1      refine at every dump time and at the end
//   -dumpfreq F      dump/refinement period in days
//   -verbose, -vtk, -bDumpIC, -nthreads, -profiler: accepted; no VTK output
//                    is written by this build.
//
// Compile-time settings of the reference (Glioma_Types.h:265-298,
// makefile/tools.make/Makefile:17-36, MRAG/MRAGHeaders.h:14) are options here:
//   -bpd 16 (blocks per dimension), -maxLevel (4 for bpd 16, 3 for bpd 8,
//   else log2 bpd), -wavelets 2|4 (default 2), -precision float|double
//   (default float, the reference's Real), -device 0.
//
// -model UQprep runs the reference's UQ data preprocessing
// (Glioma/Glioma_UQ_DataPreprocessing.cpp, gh_uqprep.h) on the host:
// -PatFileName P, -bSynt 0|1, -bpd; writes tumPET/tumT1c/tumFLAIR.dat,
// ROI.dat, PriorRange.txt and LogPriorRange.txt (GenericPrior.txt in the
// working directory).
#include <chrono>
#include <cstdio>
#include <fstream>
#include <string>

#include "gh_args.h"
#include "gh_uqprep.h"
#include "gliomahip.hpp"

using namespace gliomahip;

namespace {

const double kRefinementTolerance = 1e-4;   // Glioma_Types.h:297
const double kCompressionTolerance = 1e-5;  // Glioma_Types.h:298

template <typename Real>
int run(const Args& a) {
    const int kind = sizeof(Real) == 4 ? GH_REAL_F32 : GH_REAL_F64;
    const int bpd = a.integer("-bpd", 16);
    int lvl = 0;
    while ((1 << lvl) < bpd) ++lvl;
    const int maxLevel = a.integer("-maxLevel", bpd == 16 ? 4 : (bpd == 8 ? 3 : lvl));
    const int order = a.integer("-wavelets", 2);
    const bool verbose = a.flag("-verbose", true);
    const bool uq = a.flag("-UQ", false);
    const bool adaptive = a.flag("-adaptive", true);
    const double dumpfreq = a.num("-dumpfreq", 0.0);
    const std::string pat = a.str("-PatFileName");
    if (a.flag("-vtk", true) && verbose) std::printf("note: VTK output is not written by this build\n");

    Refiner_SpaceExtension refiner(1, maxLevel);
    Compressor compressor(1);
    Grid grid(bpd, bpd, bpd, nullptr, kind, a.integer("-device", 0), order, maxLevel);
    grid.setRefiner(&refiner);
    grid.setCompressor(&compressor);
    BlockFWT fwt;

    Real ic[3];
    if (uq) {
        std::ifstream f("TumorIC.txt");
        if (!(f >> ic[0] >> ic[1] >> ic[2])) {
            std::printf("Aborting: missing input file TumorIC.txt \n");
            return 1;
        }
    } else {
        ic[0] = (Real)a.num("-icx", 0.28);
        ic[1] = (Real)a.num("-icy", 0.67);
        ic[2] = (Real)a.num("-icz", 0.35);
    }
    const double icd[3] = {(double)ic[0], (double)ic[1], (double)ic[2]};
    if (verbose) std::printf("Reading data from file: %s \n", pat.c_str());
    const double L = InitialCondition(grid, pat, icd);

    Real Dw, rho, tend;
    if (uq) {
        std::ifstream f("InputParameters.txt");
        if (!(f >> Dw >> rho >> tend)) {
            std::printf("Aborting: missing input file InputParameters.txt \n");
            return 1;
        }
    } else {
        Dw = (Real)a.num("-Dw", 0.0013);
        rho = (Real)a.num("-rho", 0.025);
        tend = (Real)a.num("-Tend", 300);
    }
    // Dw/(L*L), Dg = 0.1 Dw, dt = 0.99 h^2/(6 max(Dw,Dg)), all Real (:302-311)
    double sDw, sDg, sdt;
    int64_t nsteps_total;
    check(gh_schedule(kind, (double)Dw, L, (double)tend, 8 * bpd, 0.99, &sDw, &sDg, &sdt, &nsteps_total));
    const Real dt = (Real)sdt;
    if (verbose)
        std::printf("Dg=%e, Dw=%e, dt= %f, rho=%f , h=%f\n", sDg, sDw, (double)dt, (double)rho,
                    (double)(Real)(1. / (8 * bpd)));

    const auto t0 = std::chrono::steady_clock::now();
    Science::AutomaticRefinement<0, 0>(grid, fwt, kRefinementTolerance, maxLevel, 1);
    Science::AutomaticCompression<0, 0>(grid, fwt, kCompressionTolerance, -1);
    check(gh_rd_prepare(grid.ctx(), sDw, sDg, (double)rho));

    Real t = 0;
    double whenToWrite = dumpfreq;
    int64_t iterations = 0, cellUpdates = 0;
    int dumps = 0;
    while (t <= tend) {
        // the Euler steps up to the next dump time (or the end), in one call
        int64_t k = 0;
        bool dump = false;
        while (t <= tend) {
            t = (Real)(t + dt);
            ++k;
            if ((double)t >= whenToWrite) {
                dump = true;
                break;
            }
        }
        check(gh_rd_steps(grid.ctx(), k, (double)dt, nullptr));
        iterations += k;
        cellUpdates += k * (int64_t)grid.numberOfBlocks() * 512;
        if (dump) {
            if (adaptive) {
                Science::AutomaticRefinement<0, 0>(grid, fwt, kRefinementTolerance, maxLevel, 1);
                check(gh_rd_prepare(grid.ctx(), sDw, sDg, (double)rho));
            }
            ++dumps;
            whenToWrite = whenToWrite + dumpfreq;
        }
    }
    if (adaptive) Science::AutomaticRefinement<0, 0>(grid, fwt, kRefinementTolerance, maxLevel, 1);
    check(gh_sync(grid.ctx()));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (uq) UQOutput(grid).dump("HGG_data.dat");
    if (verbose) {
        std::printf("steps=%lld, dump times=%d, blocks=%d, cell-updates/s=%.3e, elapsed=%.3f s\n",
                    (long long)iterations, dumps, grid.numberOfBlocks(), cellUpdates / secs, secs);
        std::printf("\n\n Run Finished \n\n");
    }
    return 0;
}

// Glioma_UQ_DataPreprocessing::run (:539-548) with the constructor's choice
// of scans or ground truth (-bSynt, :17-52)
template <typename Real>
int run_uqprep(const Args& a) {
    const bool verbose = a.flag("-verbose", true);
    const std::string pat = a.str("-PatFileName");
    if (verbose) std::printf("[uqprep] patient files %s*\n", pat.c_str());
    const int points = uqprep::prepare<Real>(a.integer("-bpd", 16), pat, a.integer("-bSynt", 0) == 1, verbose);
    if (verbose) std::printf("[uqprep] done: %d ROI points\n", points);
    return 0;
}

}  // namespace

int main(int argc, const char** argv) {
    Args a(argc, argv);
    const std::string model = a.str("-model");
    const bool dbl = a.str("-precision", "float") == "double";
    if (model == "UQprep") {
        try {
            return dbl ? run_uqprep<double>(a) : run_uqprep<float>(a);
        } catch (const Error& e) {
            std::fprintf(stderr, "brain: %s\n", e.what());
            return 1;
        }
    }
    if (model != "RD") {
        std::fprintf(stderr, "brain: -model %s is not provided by this build (RD, UQprep)\n", model.c_str());
        return 2;
    }
    try {
        return dbl ? run<double>(a) : run<float>(a);
    } catch (const Error& e) {
        std::fprintf(stderr, "brain: %s\n", e.what());
        return 1;
    }
}
