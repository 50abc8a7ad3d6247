// rd_pipeline: the reference driver's structure written against the C++
// mirror types of gliomahip.hpp -- one ReactionDiffusionOperator pass through
// BlockProcessing::pipeline_process and one UpdateTumor pass through
// BlockProcessing::process per Euler step, as Glioma_ReactionDiffusion::
// _reactionDiffusionStep does (Glioma/Glioma_ReactionDiffusion.cpp:176-187).
// It takes brain's options and writes the same HGG_data.dat; it exists to show
// (and test) that a driver keeping the reference's per-step calls gets the
// same result as brain's batched gh_rd_steps.  Float precision only.
#include <cstdio>
#include <fstream>
#include <string>

#include "gh_args.h"
#include "gliomahip.hpp"

using namespace gliomahip;
typedef float Real;

static const double refinement_tolerance = 1e-4;
static const double compression_tolerance = 1e-5;

int main(int argc, const char** argv) {
    Args parser(argc, argv);
    try {
        const int blocksPerDimension = parser.integer("-bpd", 16);
        int lvl = 0;
        while ((1 << lvl) < blocksPerDimension) ++lvl;
        const int maxLevel = parser.integer("-maxLevel", blocksPerDimension == 16 ? 4 : (blocksPerDimension == 8 ? 3 : lvl));
        const bool bAdaptivity = parser.flag("-adaptive", true);
        const double whenToWriteOffset = parser.num("-dumpfreq", 0.0);

        Refiner_SpaceExtension refiner(1, maxLevel);
        Compressor compressor(1);
        Grid grid(blocksPerDimension, blocksPerDimension, blocksPerDimension, nullptr, GH_REAL_F32,
                  parser.integer("-device", 0), parser.integer("-wavelets", 2), maxLevel);
        grid.setRefiner(&refiner);
        grid.setCompressor(&compressor);
        BlockFWT blockfwt;
        BlockProcessing blockProcessing;

        Real tumor_ic[3];
        std::ifstream icf("TumorIC.txt");
        if (!(icf >> tumor_ic[0] >> tumor_ic[1] >> tumor_ic[2])) {
            std::printf("Aborting: missing input file TumorIC.txt \n");
            return 1;
        }
        const double icd[3] = {tumor_ic[0], tumor_ic[1], tumor_ic[2]};
        const Real L = (Real)InitialCondition(grid, parser.str("-PatFileName"), icd);

        Real Dw, rho, tend;
        std::ifstream pf("InputParameters.txt");
        if (!(pf >> Dw >> rho >> tend)) {
            std::printf("Aborting: missing input file InputParameters.txt \n");
            return 1;
        }
        Dw = Dw / (L * L);
        const Real Dg = (Real)(0.1 * Dw);
        Real t = 0.0;
        const Real h = (Real)(1. / (8 * blocksPerDimension));
        const Real dt = (Real)(0.99 * h * h / (2. * 3 * (Dw >= Dg ? Dw : Dg)));

        Science::AutomaticRefinement<0, 0>(grid, blockfwt, refinement_tolerance, maxLevel, 1);
        Science::AutomaticCompression<0, 0>(grid, blockfwt, compression_tolerance, -1);

        double whenToWrite = whenToWriteOffset;
        while (t <= tend) {
            std::vector<BlockInfo> vInfo = grid.getBlocksInfo();
            const BlockCollection& collection = grid.getBlockCollection();
            ReactionDiffusionOperator<3> rhs(Dw, Dg, rho);
            UpdateTumor<3> updateTumor(dt);
            blockProcessing.pipeline_process(vInfo, collection, grid.getBoundaryInfo(), rhs);
            BlockProcessing::process(vInfo, collection, updateTumor, 4);
            t += dt;
            if (t >= whenToWrite) {
                if (bAdaptivity) Science::AutomaticRefinement<0, 0>(grid, blockfwt, refinement_tolerance, maxLevel, 1);
                whenToWrite = whenToWrite + whenToWriteOffset;
            }
        }
        if (bAdaptivity) Science::AutomaticRefinement<0, 0>(grid, blockfwt, refinement_tolerance, maxLevel, 1);
        UQOutput(grid).dump("HGG_data.dat");
    } catch (const Error& e) {
        std::fprintf(stderr, "rd_pipeline: %s\n", e.what());
        return 1;
    }
    return 0;
}
