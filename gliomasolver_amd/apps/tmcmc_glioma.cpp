// tmcmc_glioma: Bayesian calibration of the tumour model with TMCMC, the
// reference's inference (tools/pi4u_lite/Inference/engine_tmcmc.c driving
// fitfun_glioma.c -> runAll.sh -> brain + likelihood per sample), with every
// generation's candidates solved together on the device (gh_batch) instead of
// one process chain per sample.
//
//   tmcmc_glioma -PatFileName P [-obs D] [-par tmcmc_glioma.par] [-bpd 16]
//                [-stepPET 3] [-stepMRI 1] [-bROI 0] [-precision float]
//                [-device 0] [-maxBatchCells 1073741824]
//
//   P{GM,WM,CSF}.dat   anatomy (any size, mapped as _ic does)
//   D/tumT1c.dat, D/tumFLAIR.dat, D/tumPET.dat (and D/ROI.dat with -bROI 1)
//                      observations at the grid resolution (8*bpd)^3
//   tmcmc_glioma.par   Nth, MaxStages, PopSize, Min/MaxChainLength, Bdef,
//                      B0..B10, TolCOV, bbeta, seed (engine_tmcmc.c:27-320)
// Writes curgen_db_NNN.txt per generation (engine_tmcmc.c dump_curgen_db:
// point and log-likelihood per line) and fitness.txt (log evidence).
// The forward model is the uniform finest-resolution solve of the sample's
// brain run (HGG_data.dat is written at that resolution).
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gh_args.h"
#include "gliomahip.hpp"

using namespace gliomahip;

namespace {

struct Par {
    int nth = 11, max_stages = 1000, pop = 1024, min_chain = 0, max_chain = 1;
    double tol_cov = 1.0, bbeta = 0.2;
    uint64_t seed = 280675;
    std::vector<double> lo, hi;
};

// "key value..." lines, '#' comments (engine_tmcmc.c read_data)
Par read_par(const std::string& path) {
    Par p;
    std::ifstream f(path);
    if (!f.is_open()) throw Error(GH_EIO, "cannot open " + path);
    std::vector<std::pair<int, std::pair<double, double>>> bounds;
    double dlo = 0, dhi = 1;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string key;
        if (!(is >> key)) continue;
        if (key == "Nth") is >> p.nth;
        else if (key == "MaxStages") is >> p.max_stages;
        else if (key == "PopSize") is >> p.pop;
        else if (key == "MinChainLength") is >> p.min_chain;
        else if (key == "MaxChainLength") { double v; is >> v; p.max_chain = (int)v; }
        else if (key == "TolCOV") is >> p.tol_cov;
        else if (key == "bbeta") is >> p.bbeta;
        else if (key == "seed") is >> p.seed;
        else if (key == "Bdef") is >> dlo >> dhi;
        else if (key.size() > 1 && key[0] == 'B' && std::isdigit((unsigned char)key[1])) {
            double a, b;
            is >> a >> b;
            bounds.push_back({std::atoi(key.c_str() + 1), {a, b}});
        }
    }
    p.lo.assign(p.nth, dlo);
    p.hi.assign(p.nth, dhi);
    for (auto& b : bounds)
        if (b.first >= 0 && b.first < p.nth) {
            p.lo[b.first] = b.second.first;
            p.hi[b.first] = b.second.second;
        }
    return p;
}

}  // namespace

int main(int argc, const char** argv) {
    Args a(argc, argv);
    try {
        const Par par = read_par(a.str("-par", "tmcmc_glioma.par"));
        if (par.nth != 11) throw Error(GH_EINVAL, "the glioma model has 11 parameters (Nth 11)");
        const int kind = a.str("-precision", "float") == "double" ? GH_REAL_F64 : GH_REAL_F32;
        const int bpd = a.integer("-bpd", 16);
        const int stepPET = a.integer("-stepPET", 3), stepTi = a.integer("-stepMRI", 1);
        const bool roi = a.flag("-bROI", false);
        const long long max_cells = (long long)a.num("-maxBatchCells", 1073741824.0);
        const std::string pat = a.str("-PatFileName"), obs = a.str("-obs", ".");

        MatrixD3D gm(pat + "GM.dat"), wm(pat + "WM.dat"), csf(pat + "CSF.dat");
        MatrixD3D t1(obs + "/tumT1c.dat"), fl(obs + "/tumFLAIR.dat"), pet(obs + "/tumPET.dat");
        const int n = 8 * bpd;
        if (t1.getSizeX() != n || t1.getSizeY() != n || t1.getSizeZ() != n)
            throw Error(GH_EINVAL, "observations must be (8*bpd)^3");
        std::vector<double> pts;
        int npts = 0;
        if (roi) {
            MatrixD2D r(obs + "/ROI.dat");
            npts = r.getSizeX();
            pts.assign(r.data(), r.data() + (size_t)npts * 3);
        }
        gh_batch* b = nullptr;
        check(gh_batch_create(a.integer("-device", 0), kind, &b));
        int aid = 0;
        check(gh_batch_add_anatomy(b, bpd, gm.getSizeX(), gm.getSizeY(), gm.getSizeZ(), gm.data(), wm.data(),
                                   csf.data(), &aid));
        check(gh_batch_set_observations(b, aid, t1.data(), fl.data(), pet.data(), roi ? pts.data() : nullptr, npts));

        gh_tmcmc* t = nullptr;
        check(gh_tmcmc_create(par.nth, par.lo.data(), par.hi.data(), par.pop, par.max_stages, par.min_chain,
                              par.max_chain, par.tol_cov, par.bbeta, par.seed, &t));
        const size_t per_batch = std::max<long long>(1, max_cells / ((long long)n * n * n));
        int last_gen = 0;
        long long evals = 0;
        double device_ms = 0;
        const auto t0 = std::chrono::steady_clock::now();
        while (true) {
            int m = 0;
            check(gh_tmcmc_candidates(t, &m, nullptr));
            if (m == 0) break;
            std::vector<double> x((size_t)m * 11), smp((size_t)m * 6), lik((size_t)m * 5), F(m), nll(m);
            check(gh_tmcmc_candidates(t, &m, x.data()));
            check(gh_glioma_params(m, x.data(), kind, smp.data(), lik.data()));
            for (int s0 = 0; s0 < m; s0 += (int)per_batch) {
                const int cnt = std::min<int>((int)per_batch, m - s0);
                std::vector<int> anat(cnt, aid);
                check(gh_batch_set_samples(b, cnt, anat.data(), smp.data() + 6 * (size_t)s0, 0.99));
                double ms = 0;
                check(gh_batch_run(b, &ms));
                device_ms += ms;
                std::vector<double> out((size_t)cnt * 4);
                check(gh_batch_likelihood(b, lik.data() + 5 * (size_t)s0, stepPET, stepTi, roi ? 1 : 0, out.data()));
                for (int i = 0; i < cnt; ++i) nll[s0 + i] = out[4 * (size_t)i];
            }
            check(gh_glioma_fitness(m, nll.data(), F.data()));
            evals += m;
            check(gh_tmcmc_submit(t, m, F.data()));
            int gen = 0, done = 0;
            double p = 0, cov = 0, ev = 0;
            check(gh_tmcmc_state(t, &gen, &p, &cov, &ev, &done));
            if (gen != last_gen || done) {
                int np = 0;
                check(gh_tmcmc_population(t, &np, nullptr, nullptr));
                std::vector<double> px((size_t)np * 11), pf(np);
                check(gh_tmcmc_population(t, &np, px.data(), pf.data()));
                char name[64];
                std::snprintf(name, sizeof(name), "curgen_db_%03d.txt", done ? gen : last_gen);
                FILE* f = std::fopen(name, "w");
                if (!f) throw Error(GH_EIO, std::string("cannot write ") + name);
                for (int i = 0; i < np; ++i) {
                    for (int k = 0; k < 11; ++k) std::fprintf(f, "%20.16lf ", px[(size_t)i * 11 + k]);
                    std::fprintf(f, "%20.16lf\n", pf[i]);
                }
                std::fclose(f);
                const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                std::printf("generation %d done: p=%.6f, log evidence so far %.6f, %lld evaluations, %.1f s "
                            "(device %.1f s)\n",
                            done ? gen : last_gen, p, ev, evals, secs, device_ms / 1e3);
                std::fflush(stdout);
                last_gen = gen;
            }
            if (done) break;
        }
        int gen = 0, done = 0;
        double p = 0, cov = 0, ev = 0;
        check(gh_tmcmc_state(t, &gen, &p, &cov, &ev, &done));
        FILE* f = std::fopen("fitness.txt", "w");
        if (f) {
            std::fprintf(f, "%lf\n", ev);
            std::fclose(f);
        }
        std::printf("logEvidence = %.6f after %d generations, %lld evaluations\n", ev, gen + 1, evals);
        gh_tmcmc_destroy(t);
        gh_batch_destroy(b);
    } catch (const Error& e) {
        std::fprintf(stderr, "tmcmc_glioma: %s\n", e.what());
        return 1;
    }
    return 0;
}
