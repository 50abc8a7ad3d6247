// gh_uqprep.h -- inputs of a Bayesian calibration run, prepared on the host:
// observation volumes, the patient-specific prior box and the likelihood's
// region of interest.  It is the `brain -model UQprep` step of the reference
// (Glioma/Glioma_UQ_DataPreprocessing.cpp:17-548) on a uniform lattice of
// bpd^3 blocks of 8^3 cells.
//
// Stages (each a free function over plain data):
//   observe_scans / observe_ground_truth   volumes on the lattice
//        (reference :66-171 and :173-259)
//   scale_to_unit_peak                     PET / max PET      (:262-298)
//   LesionSummary::measure                 lesion counts and T1Gd centroid
//   derive_prior                           rows of PriorRange.txt (:301-420)
//   write_volume                           tum{PET,T1c,FLAIR}.dat (:435-476)
//   select_roi / write_roi                 ROI.dat            (:479-536)
//
// The output files are meant to equal the reference's bit for bit, so three
// things follow it exactly:
//   * the visiting order of the sums: block after block in slot order
//     (z, y, x of the block), cells z, y, x inside a block;
//   * the rounding points: which value is held in Real and which in double
//     (noted at each expression);
//   * the cell position origin + h*i in Real (MRAG/MRAGcore/MRAGCommon.h:
//     199-204) and the voxel floor(x / (1/max(nx,ny,nz))) in double.
// Not reproduced: VTK output and the raw FLAIR/T1c/FET scans that only feed
// it.  The synthetic noise follows the serial visiting order; the reference
// draws it inside an OpenMP loop, in a thread-dependent order.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "gliomahip.hpp"

namespace gliomahip {
namespace uqprep {

// A uniform lattice of bpd^3 blocks of 8^3 cells; dense arrays are x fastest.
struct Lattice {
    int bpd, side;
    explicit Lattice(int blocks_per_dim) : bpd(blocks_per_dim), side(8 * blocks_per_dim) {}
    size_t cells() const { return (size_t)side * side * side; }

    // cell centre along one axis, in Real, as BlockInfo::pos computes it
    template <typename Real>
    Real position(int i) const {
        const Real block_len = (Real)(1.0 / bpd), spacing = block_len / (Real)8;
        return (Real)((i / 8) * block_len) + spacing * (Real)(i % 8);
    }
    template <typename Real>
    Real spacing() const { return (Real)(1.0 / bpd) / (Real)8; }

    // f(dense index, i, j, k) over every cell in block-major visiting order
    template <typename F>
    void each_cell(F&& f) const {
        for (int b = 0; b < bpd * bpd * bpd; ++b) {
            const int x0 = 8 * (b % bpd), y0 = 8 * ((b / bpd) % bpd), z0 = 8 * (b / (bpd * bpd));
            for (int c = 0; c < 512; ++c) {
                const int i = x0 + (c & 7), j = y0 + ((c >> 3) & 7), k = z0 + (c >> 6);
                f((size_t)i + (size_t)side * ((size_t)j + (size_t)side * k), i, j, k);
            }
        }
    }
};

// Nearest-voxel lookup of a scan onto the lattice: per axis, the voxel of each
// lattice index (-1 outside the scan).
template <typename Real>
struct VoxelMap {
    std::array<std::vector<int>, 3> axis;
    VoxelMap(const Lattice& lat, int nx, int ny, int nz) {
        const int dims[3] = {nx, ny, nz};
        const double voxel = 1.0 / (double)std::max(nx, std::max(ny, nz));
        for (int a = 0; a < 3; ++a) {
            axis[a].resize(lat.side);
            for (int i = 0; i < lat.side; ++i) {
                const int v = (int)std::floor((double)lat.position<Real>(i) / voxel);
                axis[a][i] = (v >= 0 && v < dims[a]) ? v : -1;
            }
        }
    }
    bool inside(int i, int j, int k) const { return axis[0][i] >= 0 && axis[1][j] >= 0 && axis[2][k] >= 0; }
    double at(const MatrixD3D& m, int i, int j, int k) const { return m(axis[0][i], axis[1][j], axis[2][k]); }
};

// Everything the later stages read, one dense Real array per quantity.
template <typename Real>
struct Volumes {
    std::vector<Real> white, grey, fluid;      // tissue fractions
    std::vector<Real> enhancing, edema, pet;   // T1Gd mask, FLAIR mask, PET signal
    Real domain_cm = 1;                        // characteristic length L = 0.1 cm x max scan side
    explicit Volumes(const Lattice& lat)
        : white(lat.cells()), grey(lat.cells()), fluid(lat.cells()), enhancing(lat.cells()), edema(lat.cells()),
          pet(lat.cells()) {}
};

inline int longest_side(const MatrixD3D& m) { return std::max(m.getSizeX(), std::max(m.getSizeY(), m.getSizeZ())); }

// Patient scans (reference :66-171): fluid wins a voxel whose CSF fraction
// exceeds 0.2; the three fractions are divided by their Real sum (taken in
// double); lesion masks are the segmentations above 0.1; PET is the FET scan.
template <typename Real>
Volumes<Real> observe_scans(const Lattice& lat, const std::string& prefix) {
    const MatrixD3D gm(prefix + "GM.dat"), wm(prefix + "WM.dat"), csf(prefix + "CSF.dat");
    const MatrixD3D t1(prefix + "Tum_T1c.dat"), flair(prefix + "Tum_FLAIR.dat"), fet(prefix + "Tum_FET.dat");
    Volumes<Real> v(lat);
    v.domain_cm = (Real)(longest_side(gm) * 0.1);
    const VoxelMap<Real> map(lat, gm.getSizeX(), gm.getSizeY(), gm.getSizeZ());
    lat.each_cell([&](size_t c, int i, int j, int k) {
        if (!map.inside(i, j, k)) return;
        const Real f_in = (Real)map.at(csf, i, j, k);
        const bool fluid_voxel = (double)f_in > 0.2;
        const Real w = fluid_voxel ? (Real)0 : (Real)map.at(wm, i, j, k);
        const Real g = fluid_voxel ? (Real)0 : (Real)map.at(gm, i, j, k);
        const Real f = fluid_voxel ? (Real)1 : f_in;
        const double total = (double)(Real)(w + g + f);  // Real sum, compared and divided in double
        v.white[c] = total > 0. ? (Real)((double)w / total) : (Real)0;
        v.grey[c] = total > 0. ? (Real)((double)g / total) : (Real)0;
        v.fluid[c] = total > 0. ? (Real)((double)f / total) : (Real)0;
        v.enhancing[c] = (double)(Real)map.at(t1, i, j, k) > 0.1 ? (Real)1 : (Real)0;
        v.edema[c] = (double)(Real)map.at(flair, i, j, k) > 0.1 ? (Real)1 : (Real)0;
        v.pet[c] = (Real)map.at(fet, i, j, k);
    });
    return v;
}

// Standard normal deviates in pairs by the Box-Muller transform of two drand48
// draws.  The reference redraws a pair at every visited voxel while an even
// number of deviates has been handed out, and hands one out per noisy voxel.
class NormalPairs {
public:
    void at_voxel() {
        if (handed_ % 2 == 0) {
            const double u1 = drand48(), u2 = drand48();
            const double radius = std::sqrt(-2.0 * std::log(u1)), angle = 2.0 * M_PI * u2;
            pair_[0] = radius * std::cos(angle);
            pair_[1] = radius * std::sin(angle);
        }
    }
    double next() { return pair_[handed_++ % 2]; }

private:
    double pair_[2] = {0, 0};
    long handed_ = 0;
};

// Synthetic observations from a ground-truth density (reference :173-259):
// masks at the Real thresholds 0.7 (T1Gd) and 0.25 (FLAIR); PET is the density
// where a mask holds, plus N(0, 0.02^2) noise, floored at 0.  No tissue.
template <typename Real>
Volumes<Real> observe_ground_truth(const Lattice& lat, const std::string& prefix) {
    const MatrixD3D truth(prefix + "HGG_data_GT.dat");
    Volumes<Real> v(lat);
    v.domain_cm = (Real)(longest_side(truth) * 0.1);
    const VoxelMap<Real> map(lat, truth.getSizeX(), truth.getSizeY(), truth.getSizeZ());
    const Real enh_level = (Real)0.7, edema_level = (Real)0.25;
    const double noise_sd = 0.02;
    NormalPairs noise;
    lat.each_cell([&](size_t c, int i, int j, int k) {
        if (!map.inside(i, j, k)) return;
        const Real density = (Real)map.at(truth, i, j, k);
        v.enhancing[c] = density > enh_level ? (Real)1 : (Real)0;
        v.edema[c] = density > edema_level ? (Real)1 : (Real)0;
        const bool seen = (double)(Real)(v.enhancing[c] + v.edema[c]) >= 0.9;
        Real signal = seen ? density : (Real)0;
        noise.at_voxel();
        if (signal > (Real)0) signal = (Real)std::max(0., (double)signal + noise_sd * noise.next());
        v.pet[c] = signal;
    });
    return v;
}

// PET divided by its maximum (the reciprocal taken in double, stored in Real).
template <typename Real>
void scale_to_unit_peak(const Lattice& lat, std::vector<Real>& pet) {
    Real peak = 0;
    lat.each_cell([&](size_t c, int, int, int) { peak = std::max(peak, pet[c]); });
    const Real factor = (Real)(1. / peak);
    for (Real& p : pet) p = p * factor;
}

// Radius of the sphere of a given volume: the ratio in double rounded to Real,
// its cube root in double rounded to Real.
template <typename Real>
Real sphere_radius(Real volume) {
    const Real cubed = (Real)(3. * volume / (4. * M_PI));
    return (Real)std::pow((double)cubed, 1. / 3.);
}

// Lesion sizes and the T1Gd centroid.  Counts accumulate in Real; each
// centroid sum adds a double product (position in Real widened, times the
// mask) to a Real accumulator.
template <typename Real>
struct LesionSummary {
    Real enhancing_count = 0, edema_count = 0;
    Real centroid[3] = {0, 0, 0};
    Real enhancing_volume = 0, edema_volume = 0;

    static LesionSummary measure(const Lattice& lat, const Volumes<Real>& v) {
        LesionSummary s;
        Real moment[3] = {0, 0, 0};
        lat.each_cell([&](size_t c, int i, int j, int k) {
            s.enhancing_count += v.enhancing[c];
            s.edema_count += v.edema[c];
            const int ijk[3] = {i, j, k};
            for (int a = 0; a < 3; ++a)
                moment[a] = (Real)((double)moment[a] + (double)lat.position<Real>(ijk[a]) * (double)v.enhancing[c]);
        });
        for (int a = 0; a < 3; ++a) s.centroid[a] = moment[a] / s.enhancing_count;
        const Real h = lat.spacing<Real>(), cell_volume = h * h * h;
        s.enhancing_volume = s.enhancing_count * cell_volume;
        s.edema_volume = s.edema_count * cell_volume;
        return s;
    }
};

// Rows of (lower, upper) bounds in the order of GenericPrior.txt: Dw, rho,
// c = T^2 rho, icx, icy, icz, then the observation-model parameters.
struct PriorRows {
    std::vector<float> bound;  // 2 per row

    static PriorRows read(const std::string& path) {
        std::ifstream in(path);
        if (!in.is_open()) throw Error(GH_EIO, "missing input file " + path);
        PriorRows r;
        // a value that ends the file without trailing whitespace is dropped, as
        // the reference's read loop does (it tests eof after each extraction)
        for (float x; in >> x;) {
            if (in.eof()) break;
            r.bound.push_back(x);
        }
        if (r.bound.size() < 12) throw Error(GH_EIO, path + ": expected at least 6 ranges");
        return r;
    }
    void set(int row, float lo, float hi) {
        bound[2 * row] = lo;
        bound[2 * row + 1] = hi;
    }
    // "%.4f     %.4f" per row; the log file holds the float logarithms
    void write(const std::string& path, bool logarithms) const {
        FILE* f = std::fopen(path.c_str(), "w");
        if (!f) throw Error(GH_EIO, "cannot write " + path);
        for (size_t i = 0; i + 1 < bound.size(); i += 2) {
            const float lo = logarithms ? std::log(bound[i]) : bound[i];
            const float hi = logarithms ? std::log(bound[i + 1]) : bound[i + 1];
            std::fprintf(f, "%.4f     %.4f\n", lo, hi);
        }
        std::fclose(f);
    }
};

// Patient-specific rows of the prior (reference :301-420):
//  * seed box: the T1Gd centroid -/+ 1.25 x the T1Gd-equivalent radius;
//  * growth time: the FLAIR-equivalent radius r (x 1.25, in cm) crossed by the
//    fastest Fisher front 2 sqrt(D rho) gives the shortest age, r + 2 cm by the
//    slowest front plus 300 d the longest; clipped to [30, 1800] d and stored
//    as c = T^2 rho with the opposite rho bound.
template <typename Real>
void derive_prior(PriorRows& rows, const LesionSummary<Real>& s, Real domain_cm, bool verbose) {
    const Real half_side = (Real)((double)sphere_radius(s.enhancing_volume) * 1.25);
    const Real edema_r = (Real)(1.25 * (double)sphere_radius(s.edema_volume) * (double)domain_cm);
    const float* b = rows.bound.data();
    const Real slowest = (Real)(2 * std::pow(b[0] * b[2], 1. / 2.));  // float product, double root
    const Real fastest = (Real)(2 * std::pow(b[1] * b[3], 1. / 2.));
    Real t_lo = edema_r / fastest;
    Real t_hi = (Real)(((double)edema_r + 2.) / (double)slowest + 300.);
    if (verbose) std::printf("[uqprep] growth time from lesion size: %f .. %f d\n", t_lo, t_hi);
    t_lo = std::max((Real)30., t_lo);
    t_hi = std::min((Real)1800, t_hi);
    if (verbose)
        std::printf("[uqprep] growth time kept: %f .. %f d (edema radius %f cm, front speed %f .. %f)\n", t_lo, t_hi,
                    edema_r, slowest, fastest);
    const float rho_lo = b[2], rho_hi = b[3];
    rows.set(2, t_lo * t_lo * rho_hi, t_hi * t_hi * rho_lo);
    for (int a = 0; a < 3; ++a) rows.set(3 + a, s.centroid[a] - half_side, s.centroid[a] + half_side);
}

// Writes a dense volume as a gpd^3 Matrix::D3D.
template <typename Real>
void write_volume(const Lattice& lat, const std::vector<Real>& v, const std::string& path) {
    MatrixD3D m(lat.side, lat.side, lat.side);
    for (size_t c = 0; c < v.size(); ++c) m.data()[c] = (double)v[c];
    m.dump(path);
}

// Cells with tissue (white + grey > 0 in Real) within twice the
// FLAIR-equivalent radius of the T1Gd centroid (reference :479-536), in
// visiting order; distances in Real.
template <typename Real>
std::vector<std::array<int, 3>> select_roi(const Lattice& lat, const Volumes<Real>& v, const LesionSummary<Real>& s,
                                           bool verbose) {
    const Real reach = sphere_radius(s.edema_volume) * (Real)2;
    if (verbose)
        std::printf("[uqprep] region of interest: centre (%f, %f, %f), radius %f\n", s.centroid[0], s.centroid[1],
                    s.centroid[2], reach);
    std::vector<std::array<int, 3>> out;
    lat.each_cell([&](size_t c, int i, int j, int k) {
        const Real dx = lat.position<Real>(i) - s.centroid[0];
        const Real dy = lat.position<Real>(j) - s.centroid[1];
        const Real dz = lat.position<Real>(k) - s.centroid[2];
        const Real dist = std::sqrt((Real)((Real)(dx * dx + dy * dy) + dz * dz));
        if (v.white[c] + v.grey[c] > 0 && dist <= reach) out.push_back({i, j, k});
    });
    return out;
}

// ROI.dat: a Matrix::D2D of N rows (points) by 3 columns (ix, iy, iz), header
// {1234, 2, N, 3, 0}, doubles x fastest (Glioma/Matrix.h:390-412).
inline void write_roi(const std::vector<std::array<int, 3>>& pts, const std::string& path) {
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) throw Error(GH_EIO, "cannot write " + path);
    const int n = (int)pts.size();
    const int hdr[5] = {1234, 2, n, 3, 0};
    f.write((const char*)hdr, sizeof(hdr));
    std::vector<double> cols((size_t)n * 3);
    for (int p = 0; p < n; ++p)
        for (int a = 0; a < 3; ++a) cols[(size_t)p + (size_t)n * a] = pts[p][a];
    if (n) f.write((const char*)cols.data(), cols.size() * sizeof(double));
    if (!f) throw Error(GH_EIO, "write failed: " + path);
}

// The whole preprocessing step; returns the number of ROI points.
template <typename Real>
int prepare(int bpd, const std::string& prefix, bool synthetic, bool verbose) {
    const Lattice lat(bpd);
    Volumes<Real> v = synthetic ? observe_ground_truth<Real>(lat, prefix) : observe_scans<Real>(lat, prefix);
    scale_to_unit_peak(lat, v.pet);
    const LesionSummary<Real> s = LesionSummary<Real>::measure(lat, v);
    PriorRows rows = PriorRows::read("GenericPrior.txt");
    derive_prior(rows, s, v.domain_cm, verbose);
    rows.write("PriorRange.txt", false);
    rows.write("LogPriorRange.txt", true);
    write_volume(lat, v.pet, "tumPET.dat");
    write_volume(lat, v.enhancing, "tumT1c.dat");
    write_volume(lat, v.edema, "tumFLAIR.dat");
    const std::vector<std::array<int, 3>> roi = select_roi(lat, v, s, verbose);
    write_roi(roi, "ROI.dat");
    return (int)roi.size();
}

}  // namespace uqprep
}  // namespace gliomahip
