// gh_anatomy.h -- host-side initial condition of Glioma_ReactionDiffusion::_ic
// (Glioma/Glioma_ReactionDiffusion.cpp:79-172) on a uniform grid of bpd^3
// blocks of 8^3 cells, dense x-fastest arrays of n = 8*bpd cells per side.
//
//   cell position  x = origin + h*i with origin = block*(1/bpd), h = (1/bpd)/8,
//                  all in Real (MRAG/MRAGcore/MRAGCommon.h:199-204)
//   voxel          floor(x / (1./max(nx,ny,nz))) in double; cells whose
//                  voxel lies outside the data are left at 0
//   tissue vote    CSF wins if pCSF > pWM+pGM, then p_w, p_g renormalised
//   tumour         psi = (|x - ic| - 0.005) * (1/(2*(1/128))) in Real;
//                  phi = 1 if psi < -1 and tissue > 0.001,
//                  0.5*((1 - psi) - sin(pi psi)/pi) if |psi| <= 1 and tissue > 0
//                  (the Real difference 1 - psi, the rest in double), else 0.
//                  The smoothing shell ends at |x - ic| = 0.005 + 1/64, so the
//                  tumour touches a few hundred cells at most.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "gh_internal.h"

namespace gh {

template <typename Real>
inline Real cell_coord(int i, int bpd) {
    const Real dil = (Real)(1.0 / bpd), h = dil / (Real)BS;
    return (Real)((i / BS) * dil) + h * (Real)(i % BS);
}

// p_w, p_g, p_csf and the voted tissue pGM + pWM (Real) of every cell.
template <typename Real>
void anatomy_from_data(int bpd, int nx, int ny, int nz, const double* gm, const double* wm, const double* csf,
                       Real* pw, Real* pg, Real* pc, Real* tissue) {
    const int n = BS * bpd;
    const int bmax = std::max(nx, std::max(ny, nz));
    const double bh = 1.0 / (double)bmax;
    std::vector<int> map[3];
    const int dims[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a) {  // the voxel index along each axis, -1 outside the data
        map[a].resize(n);
        for (int i = 0; i < n; ++i) {
            const int m = (int)std::floor((double)cell_coord<Real>(i, bpd) / bh);
            map[a][i] = (m >= 0 && m < dims[a]) ? m : -1;
        }
    }
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                pw[c] = pg[c] = pc[c] = (Real)0;
                if (tissue) tissue[c] = (Real)0;
                if (map[0][i] < 0 || map[1][j] < 0 || map[2][k] < 0) continue;
                const size_t v = (size_t)map[0][i] + (size_t)nx * ((size_t)map[1][j] + (size_t)ny * map[2][k]);
                Real pGM = (Real)gm[v], pWM = (Real)wm[v], pCSF = (Real)csf[v];
                double tis = (double)(Real)(pWM + pGM);
                pCSF = ((double)pCSF > tis) ? (Real)1. : (Real)0.;
                pWM = ((double)pCSF > tis) ? (Real)0. : pWM;
                pGM = ((double)pCSF > tis) ? (Real)0. : pGM;
                tis = (double)(Real)(pWM + pGM);
                pw[c] = (tis > 0.) ? (Real)((double)pWM / tis) : (Real)0.;
                pg[c] = (tis > 0.) ? (Real)((double)pGM / tis) : (Real)0.;
                pc[c] = pCSF;
                if (tissue) tissue[c] = pGM + pWM;
            }
}

// The tumour's nonzero cells: (dense cell index, phi).
template <typename Real>
void tumour_cells(int bpd, const Real* tissue, const double tumor_ic[3], std::vector<std::pair<size_t, Real>>& out) {
    out.clear();
    const int n = BS * bpd;
    const Real tumorRadius = (Real)0.005, smooth_sup = (Real)2., hh = (Real)(1. / 128);
    const Real iw = (Real)(1. / (smooth_sup * hh));
    const Real ic[3] = {(Real)tumor_ic[0], (Real)tumor_ic[1], (Real)tumor_ic[2]};
    // cells within 0.005 + 1/64 (+ a margin of two cells) of the centre
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const double r = 0.005 + 1.0 / 64 + 2.0 / n;
        lo[a] = std::max(0, (int)std::floor((tumor_ic[a] - r) * n));
        hi[a] = std::min(n - 1, (int)std::ceil((tumor_ic[a] + r) * n));
    }
    for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
            for (int i = lo[0]; i <= hi[0]; ++i) {
                const size_t c = (size_t)i + (size_t)n * ((size_t)j + (size_t)n * k);
                const Real p[3] = {cell_coord<Real>(i, bpd) - ic[0], cell_coord<Real>(j, bpd) - ic[1],
                                   cell_coord<Real>(k, bpd) - ic[2]};
                const Real dist = std::sqrt((Real)((Real)(p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]));
                const Real psi = (dist - tumorRadius) * iw;
                const Real t = tissue[c];
                Real phi = 0;
                if ((psi < -1) && (t > 0.001))
                    phi = (Real)1.0;
                else if (((-1 <= psi) && (psi <= 1)) && (t > 0))
                    phi = (Real)(1.0 * 0.5 * ((double)((Real)1 - psi) - std::sin(M_PI * (double)psi) / (M_PI)));
                if (phi != (Real)0) out.emplace_back(c, phi);
            }
}

}  // namespace gh
