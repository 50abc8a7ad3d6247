// likelihood: the reference's likelihood executable
// (tools/pi4u_lite/Inference/Likelihood/source/Likelihood_main.cpp,
// HGG_Likelihood.cpp:11-38,165-192) evaluated on the device in FP64.
//
// Reads LikelihoodInput.txt (PETsigma2 PETscale T1uc T2uc slope),
// HGG_data.dat, tumPET.dat, tumT1c.dat, tumFLAIR.dat and, with -bROI 1,
// ROI.dat; options -stepPET, -stepMRI, -bROI as the reference; writes
// Likelihood.txt holding -logL ("%Lf \n").  -device selects the GPU.
#include <chrono>
#include <cstdio>
#include <fstream>

#include "gh_args.h"
#include "gliomahip.hpp"

using namespace gliomahip;

static bool same_size(const MatrixD3D& a, const MatrixD3D& b) {
    return a.getSizeX() == b.getSizeX() && a.getSizeY() == b.getSizeY() && a.getSizeZ() == b.getSizeZ();
}

int main(int argc, const char** argv) {
    std::printf("==========================================\n");
    std::printf("          Computing likelihood            \n");
    std::printf("==========================================\n");
    const auto t0 = std::chrono::steady_clock::now();
    Args a(argc, argv);
    double prm[5];
    {
        std::ifstream f("LikelihoodInput.txt");
        if (!(f >> prm[0] >> prm[1] >> prm[2] >> prm[3] >> prm[4])) {
            std::printf("Aborting: missing input file LikelihoodInput.txt \n");
            return 1;
        }
    }
    const bool roi = a.flag("-bROI", false);
    const int stepPET = a.integer("-stepPET", 1), stepTi = a.integer("-stepMRI", 1);
    std::printf("PET: PETsigma2=%f, PETscale=%f \n", prm[0], prm[1]);
    std::printf("MRI: T1uc=%f, T2uc =%f, slope=%f \n", prm[2], prm[3], prm[4]);
    std::printf("stepPET=%d, stepTi =%d \n", stepPET, stepTi);
    try {
        MatrixD3D model("HGG_data.dat"), pet("tumPET.dat"), t1c("tumT1c.dat"), flair("tumFLAIR.dat");
        if (!same_size(model, pet) || !same_size(model, t1c) || !same_size(model, flair))
            throw Error(GH_EINVAL, "HGG_data.dat and the observations differ in size");
        std::vector<double> pts;
        int npts = 0;
        if (roi) {
            MatrixD2D p("ROI.dat");
            if (p.getSizeY() < 3) throw Error(GH_EINVAL, "ROI.dat needs 3 columns");
            npts = p.getSizeX();
            pts.assign(p.data(), p.data() + (size_t)npts * 3);
        }
        Context ctx(a.integer("-device", 0), GH_REAL_F64);
        double out[4];
        int64_t npet = 0;
        check(gh_likelihood_dense(ctx.get(), model.getSizeX(), model.getSizeY(), model.getSizeZ(), model.data(),
                                  t1c.data(), flair.data(), pet.data(), roi ? pts.data() : nullptr, npts, prm,
                                  stepPET, stepTi, out, &npet));
        const long double Lpet = out[1], Lt1 = out[2], Lt2 = out[3];
        std::printf("L_Pet=%Lf, L_T1=%Lf, L_T2 = %Lf \n", Lpet, Lt1, Lt2);
        std::printf("LogLike = %Lf \n", -(long double)out[0]);
        FILE* f = std::fopen("Likelihood.txt", "w");
        if (!f) throw Error(GH_EIO, "cannot write Likelihood.txt");
        std::fprintf(f, "%Lf \n", (long double)out[0]);
        std::fclose(f);
    } catch (const Error& e) {
        std::fprintf(stderr, "likelihood: %s\n", e.what());
        return 1;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Likelihood: elapsed time=%f (sec) \n", secs);
    return 0;
}
