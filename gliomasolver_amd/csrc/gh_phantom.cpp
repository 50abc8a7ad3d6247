// gh_phantom.cpp -- seeded synthetic brain phantom and Gaussian initial
// condition.  Input synthesis only: the reference reads patient anatomy from
// GM.dat/WM.dat/CSF.dat (Glioma/Glioma_ReactionDiffusion.cpp:83-89); the
// phantom produces arrays of the same meaning on an n^3 vertex grid.
//
// Geometry (all lengths in cells):
//   head     ellipsoid, semi-axes n*(0.46,0.40,0.42) each scaled by 1+-3%
//   CSF      outer shell, depth 0 .. 2 + 2*max(0,F) (sulci reach 2 cells deeper)
//   GM       cortex below the pial surface, thickness 5 + G in [4,6]
//   WM       interior
//   CSF      two ventricle ellipsoids either side of the mid-plane
// F, G are products of sines of the polar/azimuthal angles with seeded
// frequencies.  Each interface is a phase field 0.5*(1+tanh(d/1.5)).
#include "gh_internal.h"
#include <cmath>
#include <vector>

namespace gh {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

static inline double phase(double d) { return 0.5 * (1.0 + std::tanh(d / 1.5)); }

int phantom(int n, uint64_t seed, double* gm, double* wm, double* csf, int* seed_voxel) {
    if (n < 8 || !gm || !wm || !csf) return set_error(GH_EINVAL, "gh_phantom: n >= 8 and non-NULL outputs required");
    uint64_t s = seed * 0x2545F4914F6CDD1DULL + 12345;
    const double c0[3] = {0.5 * (n - 1) + (uniform01(s) - 0.5), 0.5 * (n - 1) + (uniform01(s) - 0.5),
                          0.5 * (n - 1) + (uniform01(s) - 0.5)};
    const double base[3] = {0.46, 0.40, 0.42};
    double a[3];
    for (int d = 0; d < 3; ++d) a[d] = n * base[d] * (1.0 + 0.06 * (uniform01(s) - 0.5));
    const double amin = std::fmin(a[0], std::fmin(a[1], a[2]));
    const int w1 = 6 + (int)(4 * uniform01(s)), w2 = 5 + (int)(4 * uniform01(s));
    const int w3 = 3 + (int)(3 * uniform01(s)), w4 = 2 + (int)(3 * uniform01(s));
    const double ph0 = 6.283185307179586 * uniform01(s), ph1 = 6.283185307179586 * uniform01(s);
    // ventricles
    const double vs = 0.08 * n, vy = -0.02 * n, vz = 0.03 * n;
    const double va[3] = {0.05 * n, 0.12 * n, 0.06 * n};
    const double vmin = std::fmin(va[0], std::fmin(va[1], va[2]));

    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double p[3] = {i - c0[0], j - c0[1], k - c0[2]};
                double r2 = 0;
                for (int d = 0; d < 3; ++d) r2 += (p[d] / a[d]) * (p[d] / a[d]);
                const double r = std::sqrt(r2);
                const double d_out = (1.0 - r) * amin;  // > 0 inside the head
                const double rho = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) + 1e-12;
                const double theta = std::acos(std::fmax(-1.0, std::fmin(1.0, p[2] / rho)));
                const double phi = std::atan2(p[1], p[0]);
                const double F = std::sin(w1 * theta) * std::sin(w2 * phi + ph0);
                const double G = std::sin(w3 * theta + ph1) * std::cos(w4 * phi);
                const double depth_pial = 2.0 + 2.0 * std::fmax(0.0, F);
                const double depth_wm = depth_pial + 5.0 + G;
                double dv = 1e30;
                for (int side = -1; side <= 1; side += 2) {
                    const double q[3] = {p[0] - side * vs, p[1] - vy, p[2] - vz};
                    double rv2 = 0;
                    for (int d = 0; d < 3; ++d) rv2 += (q[d] / va[d]) * (q[d] / va[d]);
                    dv = std::fmin(dv, (std::sqrt(rv2) - 1.0) * vmin);  // > 0 outside
                }
                const double head = phase(d_out);
                const double brain = phase(d_out - depth_pial);
                const double wmf = phase(d_out - depth_wm);
                const double tissue = brain * phase(dv);
                const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                wm[c] = tissue * wmf;
                gm[c] = tissue * (1.0 - wmf);
                csf[c] = std::fmax(0.0, head - tissue);
            }
    if (seed_voxel) {
        // seeded WM voxel: uniform over voxels with WM fraction > 0.9
        std::vector<size_t> cand;
        for (size_t c = 0; c < (size_t)n * n * n; ++c)
            if (wm[c] > 0.9) cand.push_back(c);
        if (cand.empty()) return set_error(GH_EINVAL, "gh_phantom: no WM voxel (n too small)");
        const size_t c = cand[splitmix64(s) % cand.size()];
        seed_voxel[0] = (int)(c % n);
        seed_voxel[1] = (int)((c / n) % n);
        seed_voxel[2] = (int)(c / ((size_t)n * n));
    }
    return GH_OK;
}

template <typename Real>
static void gaussian(int n, const Real* pw, const Real* pg, const double* x0, double sigma, double peak, Real* phi) {
    const double is2 = 1.0 / (2.0 * sigma * sigma);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                const double d2 = (i - x0[0]) * (i - x0[0]) + (j - x0[1]) * (j - x0[1]) + (k - x0[2]) * (k - x0[2]);
                phi[c] = ((double)pw[c] + (double)pg[c] > 0.) ? (Real)(peak * std::exp(-d2 * is2)) : (Real)0;
            }
}

int gaussian_ic(int real_kind, int n, const void* p_w, const void* p_g, const double* center, double sigma,
                double peak, void* phi) {
    if (n <= 0 || !p_w || !p_g || !center || !phi || !(sigma > 0))
        return set_error(GH_EINVAL, "gh_gaussian_ic: bad argument");
    if (real_kind == GH_REAL_F32)
        gaussian(n, (const float*)p_w, (const float*)p_g, center, sigma, peak, (float*)phi);
    else if (real_kind == GH_REAL_F64)
        gaussian(n, (const double*)p_w, (const double*)p_g, center, sigma, peak, (double*)phi);
    else
        return set_error(GH_EINVAL, "gh_gaussian_ic: real_kind must be GH_REAL_F32 or GH_REAL_F64");
    return GH_OK;
}

}  // namespace gh

extern "C" int gh_phantom(int n, uint64_t seed, double* gm, double* wm, double* csf, int* seed_voxel) {
    return gh::phantom(n, seed, gm, wm, csf, seed_voxel);
}
extern "C" int gh_gaussian_ic(int real_kind, int n, const void* p_w, const void* p_g, const double* center,
                              double sigma, double peak, void* phi) {
    return gh::gaussian_ic(real_kind, n, p_w, p_g, center, sigma, peak, phi);
}
