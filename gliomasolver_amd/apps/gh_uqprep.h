// UQ data preprocessing of the reference's `brain -model UQprep`
// (Glioma/Glioma_UQ_DataPreprocessing.cpp), on a uniform grid of bpd^3 blocks
// of 8^3 cells:
//
//   _ic            (:66-171)  anatomy P{GM,WM,CSF}.dat -> p_w, p_g, p_csf with
//                             the CSF > 0.2 rule and normalisation by
//                             pWM + pGM + pCSF; observations
//                             P{Tum_T1c,Tum_FLAIR,Tum_FET}.dat -> T1c mask
//                             (> 0.1), FLAIR mask (> 0.1), PET
//   _ic_Synthetic  (:173-259) PHGG_data_GT.dat -> masks (> 0.7, > 0.25), PET =
//                             the density where either mask holds, plus
//                             N(0, 0.02^2) Box-Muller noise from drand48
//   _normalisePET  (:262-298) PET / max PET
//   _computePriorRange (:301-420) GenericPrior.txt -> PriorRange.txt,
//                             LogPriorRange.txt (IC box from the T1c centre of
//                             mass and volume, T range from the FLAIR volume)
//   _dumpUQdata    (:435-476) tumPET.dat, tumT1c.dat, tumFLAIR.dat
//   _dumpInferenceROI (:479-536) ROI.dat: cells with tissue within twice the
//                             FLAIR-equivalent radius of the centre of mass
//
// Cells are visited block after block in the grid's slot order (z, y, x of
// the block), cells z, y, x inside a block; sums accumulate in Real in that
// order, as the reference's loops over vInfo do.  The voxel mapping is the
// one of the RD initial condition (cell position origin + h*i in Real,
// voxel floor(x / (1/max(nx,ny,nz))) in double).  Not reproduced: VTK output
// and the reading of the FLAIR/T1c/FET scans that only feed it; the noise of
// the synthetic case follows the serial loop order (the reference's OpenMP
// loop draws in a thread-dependent order).
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "gliomahip.hpp"

namespace gliomahip {

class MatrixD2DWriter {
public:
    // Matrix::D2D<double> with nx rows of ny values (Glioma/Matrix.h:390-412)
    static void dump(const std::string& path, int nx, int ny, const std::vector<double>& v) {
        std::ofstream f(path, std::ios::binary);
        if (!f.is_open()) throw Error(GH_EIO, "cannot write " + path);
        const int hdr[5] = {1234, 2, nx, ny, 0};
        f.write((const char*)hdr, sizeof(hdr));
        if (!v.empty()) f.write((const char*)v.data(), v.size() * sizeof(double));
        if (!f) throw Error(GH_EIO, "write failed: " + path);
    }
};

template <typename Real>
class UQPreprocessing {
public:
    UQPreprocessing(int bpd, bool verbose) : bpd_(bpd), n_(8 * bpd), verbose_(verbose) {
        const size_t nc = (size_t)n_ * n_ * n_;
        pw_.assign(nc, 0);
        pg_.assign(nc, 0);
        pcsf_.assign(nc, 0);
        t1_.assign(nc, 0);
        fl_.assign(nc, 0);
        pet_.assign(nc, 0);
    }

    // _ic (:66-171)
    void ic(const std::string& pat) {
        MatrixD3D GM(pat + "GM.dat"), WM(pat + "WM.dat"), CSF(pat + "CSF.dat");
        MatrixD3D T1(pat + "Tum_T1c.dat"), FL(pat + "Tum_FLAIR.dat"), FET(pat + "Tum_FET.dat");
        const int nx = GM.getSizeX(), ny = GM.getSizeY(), nz = GM.getSizeZ();
        set_length(nx, ny, nz);
        map_cells(nx, ny, nz, [&](size_t c, int vx, int vy, int vz) {
            Real pGM = (Real)GM(vx, vy, vz), pWM = (Real)WM(vx, vy, vz), pCSF = (Real)CSF(vx, vy, vz);
            pWM = (pCSF > 0.2) ? (Real)0. : pWM;
            pGM = (pCSF > 0.2) ? (Real)0. : pGM;
            pCSF = (pCSF > 0.2) ? (Real)1. : pCSF;
            const double all = pWM + pGM + pCSF;
            pw_[c] = (all > 0.) ? (Real)(pWM / all) : (Real)0.;
            pg_[c] = (all > 0.) ? (Real)(pGM / all) : (Real)0.;
            pcsf_[c] = (all > 0.) ? (Real)(pCSF / all) : (Real)0.;
            const Real pT1 = (Real)T1(vx, vy, vz), pFL = (Real)FL(vx, vy, vz);
            t1_[c] = (pT1 > 0.1) ? (Real)1. : (Real)0.;
            fl_[c] = (pFL > 0.1) ? (Real)1. : (Real)0.;
            pet_[c] = (Real)FET(vx, vy, vz);
        });
    }

    // _ic_Synthetic (:173-259); drand48 in the serial loop order
    void ic_synthetic(const std::string& pat) {
        MatrixD3D GT(pat + "HGG_data_GT.dat");
        const int nx = GT.getSizeX(), ny = GT.getSizeY(), nz = GT.getSizeZ();
        set_length(nx, ny, nz);
        const Real ucFLAIR = (Real)0.25, ucT1 = (Real)0.7;
        const double alpha = 0.02;
        double xi_1 = 0, xi_2 = 0;
        long n = 0;
        map_cells(nx, ny, nz, [&](size_t c, int vx, int vy, int vz) {
            const Real data = (Real)GT(vx, vy, vz);
            t1_[c] = (data > ucT1) ? (Real)1. : (Real)0.;
            fl_[c] = (data > ucFLAIR) ? (Real)1. : (Real)0.;
            pet_[c] = (t1_[c] + fl_[c] >= 0.9) ? data : (Real)0.;
            if ((n % 2) == 0) {
                const double u1 = drand48(), u2 = drand48();
                xi_1 = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
                xi_2 = std::sqrt(-2.0 * std::log(u1)) * std::sin(2.0 * M_PI * u2);
            }
            if (pet_[c] > (Real)0.) {
                pet_[c] = (Real)std::max(0., (double)pet_[c] + alpha * xi_1);
                xi_1 = xi_2;
                ++n;
            }
        });
    }

    // _normalisePET (:262-298)
    void normalise_pet() {
        Real maxPET = 0.;
        for_cells([&](size_t c, int, int, int) { maxPET = std::max(maxPET, pet_[c]); });
        const Real scale = (Real)(1. / maxPET);
        for (Real& v : pet_) v = v * scale;
    }

    // _computePriorRange (:301-420)
    void prior_range(const std::string& generic = "GenericPrior.txt") {
        std::vector<float> prior;
        std::ifstream f(generic);
        if (!f.is_open()) throw Error(GH_EIO, "missing input file " + generic);
        float tmp;
        while (f.good()) {
            f >> tmp;
            if (f.eof()) break;
            prior.push_back(tmp);
        }
        if (prior.size() < 12) throw Error(GH_EIO, generic + ": expected at least 6 ranges");
        Real massT1 = 0., massT2 = 0., cmx = 0., cmy = 0., cmz = 0.;
        for_cells([&](size_t c, int i, int j, int k) {
            massT1 += t1_[c];
            massT2 += fl_[c];
            const double x[3] = {(double)coord(i), (double)coord(j), (double)coord(k)};
            cmx += x[0] * t1_[c];
            cmy += x[1] * t1_[c];
            cmz += x[2] * t1_[c];
        });
        cm_[0] = cmx / massT1;
        cm_[1] = cmy / massT1;
        cm_[2] = cmz / massT1;
        const Real h = (Real)(1.0 / bpd_) / (Real)8;  // info.h[0] of a uniform grid
        const Real h3 = h * h * h;
        const Real volT1 = massT1 * h3;
        volFLAIR_ = massT2 * h3;
        const Real T1radius3 = 3. * volT1 / (4. * M_PI);
        Real T1radius = std::pow((double)T1radius3, 1. / 3.);
        T1radius = T1radius * 1.25;
        const Real ICprior[6] = {cm_[0] - T1radius, cm_[0] + T1radius, cm_[1] - T1radius,
                                 cm_[1] + T1radius, cm_[2] - T1radius, cm_[2] + T1radius};
        const Real FLAIRradius3 = 3. * volFLAIR_ / (4. * M_PI);
        Real FLAIRradius = std::pow((double)FLAIRradius3, 1. / 3.);
        FLAIRradius = 1.25 * FLAIRradius * L_;
        const Real vmin = 2 * std::pow((double)(prior[0] * prior[2]), 1. / 2.);
        const Real vmax = 2 * std::pow((double)(prior[1] * prior[3]), 1. / 2.);
        Real Tmin = FLAIRradius / vmax;
        Real Tmax = (FLAIRradius + 2.) / vmin + 300.;
        if (verbose_) std::printf("Tmin =%f, Tmax=%f \n", Tmin, Tmax);
        Tmin = std::max((Real)30., Tmin);
        Tmax = std::min((Real)1800, Tmax);
        if (verbose_) std::printf("Restricted range: Tmin =%f, Tmax=%f \n", Tmin, Tmax);
        const Real cmin = Tmin * Tmin * prior[3];
        const Real cmax = Tmax * Tmax * prior[2];
        prior[4] = cmin;
        prior[5] = cmax;
        for (int q = 0; q < 6; ++q) prior[6 + q] = ICprior[q];
        if (verbose_) std::printf("FLAIRradius = %f, vmin =%f, vmax=%f \n", FLAIRradius, vmin, vmax);
        FILE* p = std::fopen("PriorRange.txt", "w");
        if (!p) throw Error(GH_EIO, "cannot write PriorRange.txt");
        for (size_t i = 0; i + 1 < prior.size(); i += 2) std::fprintf(p, "%.4f     %.4f\n", prior[i], prior[i + 1]);
        std::fclose(p);
        p = std::fopen("LogPriorRange.txt", "w");
        if (!p) throw Error(GH_EIO, "cannot write LogPriorRange.txt");
        for (size_t i = 0; i + 1 < prior.size(); i += 2)
            std::fprintf(p, "%.4f     %.4f\n", std::log(prior[i]), std::log(prior[i + 1]));
        std::fclose(p);
    }

    // _dumpUQdata (:435-476)
    void dump_uq_data() const {
        dump_field(pet_, "tumPET.dat");
        dump_field(t1_, "tumT1c.dat");
        dump_field(fl_, "tumFLAIR.dat");
    }

    // _dumpInferenceROI (:479-536)
    int dump_roi() const {
        const Real radius3 = 3. * volFLAIR_ / (4. * M_PI);
        Real radius = std::pow((double)radius3, 1. / 3.);
        radius = radius * 2.;
        if (verbose_) {
            std::printf("Center of mass =[ %f, %f, %f]\n", cm_[0], cm_[1], cm_[2]);
            std::printf("ROI radius %f\n", radius);
        }
        std::vector<int> pts;
        for_cells([&](size_t c, int i, int j, int k) {
            const Real p[3] = {coord(i) - cm_[0], coord(j) - cm_[1], coord(k) - cm_[2]};
            const Real dist = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            const Real tissue = pw_[c] + pg_[c];
            if (tissue > 0 && dist <= radius) {
                pts.push_back(i);
                pts.push_back(j);
                pts.push_back(k);
            }
        });
        const int np = (int)pts.size() / 3;
        std::vector<double> out((size_t)np * 3);
        for (int i = 0; i < np; ++i)
            for (int q = 0; q < 3; ++q) out[i + (size_t)np * q] = pts[3 * (size_t)i + q];
        MatrixD2DWriter::dump("ROI.dat", np, 3, out);
        return np;
    }

    Real L() const { return L_; }
    const Real* cm() const { return cm_; }

private:
    int bpd_, n_;
    bool verbose_;
    Real L_ = 1, cm_[3] = {0, 0, 0}, volFLAIR_ = 0;
    std::vector<Real> pw_, pg_, pcsf_, t1_, fl_, pet_;  // dense, x fastest

    Real coord(int i) const {  // BlockInfo::pos in Real (MRAG/MRAGcore/MRAGCommon.h:199-204)
        const Real dil = (Real)(1.0 / bpd_), h = dil / (Real)8;
        return (Real)((i / 8) * dil) + h * (Real)(i % 8);
    }
    void set_length(int nx, int ny, int nz) {
        const int m = std::max(nx, std::max(ny, nz));
        L_ = m * 0.1;
        if (verbose_) std::printf("dataSizeX=%i, dataSizeY=%i, dataSizeZ= %i \n", nx, ny, nz);
    }
    // every cell in block order: f(dense index, gix, giy, giz)
    template <typename F>
    void for_cells(F f) const {
        for (int bz = 0; bz < bpd_; ++bz)
            for (int by = 0; by < bpd_; ++by)
                for (int bx = 0; bx < bpd_; ++bx)
                    for (int iz = 0; iz < 8; ++iz)
                        for (int iy = 0; iy < 8; ++iy)
                            for (int ix = 0; ix < 8; ++ix) {
                                const int i = 8 * bx + ix, j = 8 * by + iy, k = 8 * bz + iz;
                                f((size_t)i + (size_t)n_ * ((size_t)j + (size_t)n_ * k), i, j, k);
                            }
    }
    // every cell whose voxel lies inside the data: f(dense index, vx, vy, vz)
    template <typename F>
    void map_cells(int nx, int ny, int nz, F f) {
        const double bh = 1.0 / (double)std::max(nx, std::max(ny, nz));
        for_cells([&](size_t c, int i, int j, int k) {
            const int vx = (int)std::floor((double)coord(i) / bh), vy = (int)std::floor((double)coord(j) / bh),
                      vz = (int)std::floor((double)coord(k) / bh);
            if ((vx >= 0 && vx < nx) && (vy >= 0 && vy < ny) && (vz >= 0 && vz < nz)) f(c, vx, vy, vz);
        });
    }
    void dump_field(const std::vector<Real>& v, const std::string& path) const {
        MatrixD3D m(n_, n_, n_);
        for (size_t c = 0; c < v.size(); ++c) m.data()[c] = (double)v[c];
        m.dump(path);
    }
};

}  // namespace gliomahip
