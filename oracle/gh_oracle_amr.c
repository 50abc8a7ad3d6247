/*
 * gh_oracle_amr.c -- CPU oracle (checker) for the wavelet-adaptive grid.
 *
 * TEST INFRASTRUCTURE ONLY (see gh_oracle.h).
 *
 * Leaves are given as (level, ix, iy, iz) rows in the canonical order
 * (level, iz, iy, ix) ascending; block-major field arrays follow that order,
 * 512 cells per block, cell = x + 8y + 64z.  Ghost values follow the
 * interpolating-wavelet multiresolution of MRAG (MRAG/MRAGcore/
 * MRAGWavelets_StaticData.cpp:13-39, MRAG/MRAGcore/MRAGBlockLab.h:154-301):
 * a lattice point at level L covered by a leaf of level L is that cell; a
 * point inside a finer region is the coincident fine vertex (Ha = {1}); a
 * point inside a coarser region is Hs-synthesised from level L-1.  The value
 * is accumulated in Real, product by product, over its terms in recursion
 * order (synthesis taps z-major, then y, then x).
 */
#include "gh_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int n;
    int64_t* leaf; /* canonical keys of leaves, ascending */
    int ni;
    int64_t* inner; /* keys of nodes with children, ascending */
} tree_t;

static int64_t ck(int L, int x, int y, int z) {
    return ((int64_t)L << 48) | ((int64_t)z << 32) | ((int64_t)y << 16) | (int64_t)x;
}

static int cmp64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : (x > y);
}

static int find64(const int64_t* v, int n, int64_t k) {
    int lo = 0, hi = n - 1;
    while (lo <= hi) {
        int mid = (lo + hi) / 2;
        if (v[mid] == k) return mid;
        if (v[mid] < k) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

static void tree_build(tree_t* t, int n, const int* lv) {
    t->n = n;
    t->leaf = (int64_t*)malloc(sizeof(int64_t) * n);
    int cap = 16 * n + 16, m = 0;
    int64_t* in = (int64_t*)malloc(sizeof(int64_t) * cap);
    for (int i = 0; i < n; ++i) {
        int L = lv[4 * i], x = lv[4 * i + 1], y = lv[4 * i + 2], z = lv[4 * i + 3];
        t->leaf[i] = ck(L, x, y, z);
        for (int l = L - 1; l >= 0; --l) {
            if (m == cap) { cap *= 2; in = (int64_t*)realloc(in, sizeof(int64_t) * cap); }
            in[m++] = ck(l, x >> (L - l), y >> (L - l), z >> (L - l));
        }
    }
    qsort(in, m, sizeof(int64_t), cmp64);
    int u = 0;
    for (int i = 0; i < m; ++i)
        if (u == 0 || in[u - 1] != in[i]) in[u++] = in[i];
    t->inner = in;
    t->ni = u;
}

static void tree_free(tree_t* t) {
    free(t->leaf);
    free(t->inner);
}

static int64_t wrap64(int64_t v, int64_t n) {
    int64_t r = v % n;
    return r < 0 ? r + n : r;
}

typedef struct {
    int64_t src;
    double w;
} term_t;

typedef struct {
    term_t* v;
    int n, cap;
} terms_t;

static void push(terms_t* T, int64_t src, double w) {
    if (T->n == T->cap) {
        T->cap = T->cap ? 2 * T->cap : 64;
        T->v = (term_t*)realloc(T->v, sizeof(term_t) * T->cap);
    }
    T->v[T->n].src = src;
    T->v[T->n].w = w;
    T->n++;
}

/* value of lattice point q at level L as terms (source = rank*512 + cell) */
static void collect(const tree_t* t, int order, int L, int64_t qx, int64_t qy, int64_t qz, double w, terms_t* T) {
    const int64_t n = (int64_t)8 << L;
    qx = wrap64(qx, n); qy = wrap64(qy, n); qz = wrap64(qz, n);
    const int bx = (int)(qx >> 3), by = (int)(qy >> 3), bz = (int)(qz >> 3);
    const int r = find64(t->leaf, t->n, ck(L, bx, by, bz));
    if (r >= 0) {
        push(T, (int64_t)r * 512 + (qx & 7) + 8 * (qy & 7) + 64 * (qz & 7), w);
        return;
    }
    if (find64(t->inner, t->ni, ck(L, bx, by, bz)) >= 0) { /* finer region: coincident vertex */
        collect(t, order, L + 1, 2 * qx, 2 * qy, 2 * qz, w, T);
        return;
    }
    /* coarser region: Hs synthesis, odd points from Hs(+-1) (and Hs(+-3)) */
    int64_t s[3][4];
    double ws[3][4];
    int nt[3];
    const int64_t q[3] = {qx, qy, qz};
    for (int a = 0; a < 3; ++a) {
        if ((q[a] & 1) == 0) {
            s[a][0] = q[a] / 2; ws[a][0] = 1.0; nt[a] = 1;
        } else if (order == 2) {
            s[a][0] = (q[a] - 1) / 2; ws[a][0] = 0.5;
            s[a][1] = (q[a] + 1) / 2; ws[a][1] = 0.5;
            nt[a] = 2;
        } else {
            s[a][0] = (q[a] - 3) / 2; ws[a][0] = -1.0 / 16;
            s[a][1] = (q[a] - 1) / 2; ws[a][1] = 9.0 / 16;
            s[a][2] = (q[a] + 1) / 2; ws[a][2] = 9.0 / 16;
            s[a][3] = (q[a] + 3) / 2; ws[a][3] = -1.0 / 16;
            nt[a] = 4;
        }
    }
    for (int kz = 0; kz < nt[2]; ++kz)
        for (int ky = 0; ky < nt[1]; ++ky)
            for (int kx = 0; kx < nt[0]; ++kx)
                collect(t, order, L - 1, s[0][kx], s[1][ky], s[2][kz], w * (ws[2][kz] * ws[1][ky] * ws[0][kx]), T);
}

static int lab_lo(int order) { return order == 4 ? -2 : -1; }
static int lab_n(int order) { return order == 4 ? 13 : 10; }

/* lab of `leaf`; faces_only: just the interior and the six face layers at
 * distance 1 (what the 7-point stencil reads) */
#define LAB_FILL(NAME, REAL)                                                                         \
    static void NAME(const tree_t* t, const int* leaves, int order, const REAL* field, int leaf, REAL* lab, \
                     int faces_only, terms_t* T) {                                                   \
        const int L = leaves[4 * leaf], B[3] = {leaves[4 * leaf + 1], leaves[4 * leaf + 2], leaves[4 * leaf + 3]}; \
        const int lo = lab_lo(order), N = lab_n(order);                                              \
        for (int p = 0; p < N * N * N; ++p) {                                                        \
            const int i[3] = {p % N + lo, (p / N) % N + lo, p / (N * N) + lo};                       \
            if (faces_only) {                                                                        \
                int out = 0, far = 0;                                                                \
                for (int a = 0; a < 3; ++a) {                                                        \
                    if (i[a] < 0 || i[a] > 7) ++out;                                                 \
                    if (i[a] < -1 || i[a] > 8) far = 1;                                              \
                }                                                                                    \
                if (out > 1 || far) { lab[p] = (REAL)0; continue; }                                  \
            }                                                                                        \
            T->n = 0;                                                                                \
            collect(t, order, L, 8LL * B[0] + i[0], 8LL * B[1] + i[1], 8LL * B[2] + i[2], 1.0, T);   \
            REAL acc = (REAL)0;                                                                      \
            for (int k = 0; k < T->n; ++k) {                                                         \
                REAL prod = field[T->v[k].src] * (REAL)T->v[k].w;                                    \
                acc = acc + prod;                                                                    \
            }                                                                                        \
            lab[p] = acc;                                                                            \
        }                                                                                            \
    }
LAB_FILL(lab_fill_f32, float)
LAB_FILL(lab_fill_f64, double)

void gho_amr_lab_f32(int nleaves, const int* leaves, int order, const float* field, int leaf, float* lab) {
    tree_t t;
    terms_t T = {0, 0, 0};
    tree_build(&t, nleaves, leaves);
    lab_fill_f32(&t, leaves, order, field, leaf, lab, 0, &T);
    free(T.v);
    tree_free(&t);
}
void gho_amr_lab_f64(int nleaves, const int* leaves, int order, const double* field, int leaf, double* lab) {
    tree_t t;
    terms_t T = {0, 0, 0};
    tree_build(&t, nleaves, leaves);
    lab_fill_f64(&t, leaves, order, field, leaf, lab, 0, &T);
    free(T.v);
    tree_free(&t);
}

/* the terms of every lab position of `leaf` in recursion order (N^3
 * positions, x fastest; interior positions have none): counts[p], then up to
 * cap sources (leaf * 512 + cell) and weights rounded to FP32 (is_f32) or kept
 * in FP64 as the lab applies them, list after list; returns the total */
int gho_amr_block_terms(int nleaves, const int* leaves, int order, int leaf, int is_f32, int cap, int* counts,
                        int* src, double* w) {
    tree_t t;
    terms_t T = {0, 0, 0};
    tree_build(&t, nleaves, leaves);
    const int L = leaves[4 * leaf], lo = lab_lo(order), N = lab_n(order);
    int total = 0;
    for (int p = 0; p < N * N * N; ++p) {
        const int i[3] = {p % N + lo, (p / N) % N + lo, p / (N * N) + lo};
        counts[p] = 0;
        if (i[0] >= 0 && i[0] < 8 && i[1] >= 0 && i[1] < 8 && i[2] >= 0 && i[2] < 8) continue;
        T.n = 0;
        collect(&t, order, L, 8LL * leaves[4 * leaf + 1] + i[0], 8LL * leaves[4 * leaf + 2] + i[1],
                8LL * leaves[4 * leaf + 3] + i[2], 1.0, &T);
        counts[p] = T.n;
        for (int q = 0; q < T.n; ++q, ++total) {
            if (total >= cap) continue;
            src[total] = (int)T.v[q].src;
            w[total] = is_f32 ? (double)(float)T.v[q].w : T.v[q].w;
        }
    }
    free(T.v);
    tree_free(&t);
    return total;
}

/* BlockFWT::_filter / _fwt (MRAG/MRAGcore/MRAGBlockFWT.h:219-298) in Real;
 * report = max |value| over the wavelet codes 1..7 (:109-120) */
#define FWT_BODY(REAL)                                                                               \
    const int lo = lab_lo(order), N = lab_n(order);                                                  \
    int gs, ge;                                                                                      \
    REAL ga[7];                                                                                      \
    if (order == 2) { gs = -2; ge = 1; ga[0] = (REAL)-0.5; ga[1] = (REAL)1; ga[2] = (REAL)-0.5; }    \
    else {                                                                                           \
        static const double g[7] = {1 / 16., 0, -9 / 16., 1, -9 / 16., 0, 1 / 16.};                  \
        gs = -4; ge = 3;                                                                             \
        for (int k = 0; k < 7; ++k) ga[k] = (REAL)g[k];                                              \
    }                                                                                                \
    double mx = 0.0;                                                                                 \
    for (int dz = 0; dz < 4; ++dz) for (int dy = 0; dy < 4; ++dy) for (int dx = 0; dx < 4; ++dx)     \
    for (int code = 1; code < 8; ++code) {                                                           \
        const int bw[3] = {code & 1, (code >> 1) & 1, (code >> 2) & 1};                              \
        const int o[3] = {2 * dx, 2 * dy, 2 * dz};                                                   \
        int s[3], e[3];                                                                              \
        for (int a = 0; a < 3; ++a) { s[a] = bw[a] ? gs : 0; e[a] = bw[a] ? ge : 1; }                \
        REAL result = (REAL)0;                                                                       \
        for (int iz = s[2]; iz < e[2]; ++iz) {                                                       \
            const REAL wZ = bw[2] ? ga[iz - gs] : (REAL)1;                                           \
            for (int iy = s[1]; iy < e[1]; ++iy) {                                                   \
                const REAL wY = bw[1] ? ga[iy - gs] : (REAL)1;                                       \
                const REAL wZwY = wZ * wY;                                                           \
                for (int ix = s[0]; ix < e[0]; ++ix) {                                               \
                    const REAL wX = bw[0] ? ga[ix - gs] : (REAL)1;                                   \
                    const REAL w = wZwY * wX;                                                        \
                    const REAL v = lab[(o[2] - iz - lo) * N * N + (o[1] - iy - lo) * N + (o[0] - ix - lo)]; \
                    const REAL prod = v * w;                                                         \
                    result = result + prod;                                                          \
                }                                                                                    \
            }                                                                                        \
        }                                                                                            \
        const double a = fabs((double)result);                                                       \
        if (a > mx) mx = a;                                                                          \
    }                                                                                                \
    return mx;

double gho_amr_fwt_f32(int order, const float* lab) { FWT_BODY(float) }
double gho_amr_fwt_f64(int order, const double* lab) { FWT_BODY(double) }

/* FWT details (the report of BlockFWT, max |detail| over codes 1..7) of the
 * nlist listed leaves, the tree built once */
#define DETAILS_BODY(REAL, FILL, FWT)                                                               \
    tree_t t;                                                                                      \
    terms_t T = {0, 0, 0};                                                                         \
    tree_build(&t, nleaves, leaves);                                                               \
    const int N = lab_n(order);                                                                    \
    REAL* lab = (REAL*)malloc(sizeof(REAL) * N * N * N);                                           \
    for (int i = 0; i < nlist; ++i) {                                                              \
        FILL(&t, leaves, order, field, list[i], lab, 0, &T);                                       \
        out[i] = FWT(order, lab);                                                                  \
    }                                                                                              \
    free(lab);                                                                                     \
    free(T.v);                                                                                     \
    tree_free(&t);

void gho_amr_details_f32(int nleaves, const int* leaves, int order, const float* field, int nlist, const int* list,
                         double* out) {
    DETAILS_BODY(float, lab_fill_f32, gho_amr_fwt_f32)
}
void gho_amr_details_f64(int nleaves, const int* leaves, int order, const double* field, int nlist, const int* list,
                         double* out) {
    DETAILS_BODY(double, lab_fill_f64, gho_amr_fwt_f64)
}

/* BlockSplitter::_splitBlocks (MRAG/MRAGcore/MRAGBlockSplitter.h:24-99): the
 * child at `octant` of the parent whose lab is given; contributions in source
 * order (sz, sy, sx), weight (Real)(wZ*wY*Hs_x) with wZ, wY in double */
#define SPLIT_BODY(REAL)                                                                             \
    const int lo = lab_lo(order), N = lab_n(order);                                                  \
    int hs0, he0;                                                                                    \
    double hs[7];                                                                                    \
    if (order == 2) { hs0 = -1; he0 = 2; hs[0] = 0.5; hs[1] = 1; hs[2] = 0.5; }                      \
    else {                                                                                           \
        static const double h[7] = {-1 / 16., 0, 9 / 16., 1, 9 / 16., 0, -1 / 16.};                  \
        hs0 = -3; he0 = 4;                                                                           \
        for (int k = 0; k < 7; ++k) hs[k] = (double)(REAL)h[k];                                      \
    }                                                                                                \
    const int S0 = (int)ceil(hs0 * 0.5), E0 = 1 + (int)floor((he0 - 1) * 0.5 + 8);                   \
    for (int c = 0; c < 512; ++c) child[c] = (REAL)0;                                                \
    for (int sz = S0; sz < E0; ++sz) for (int sy = S0; sy < E0; ++sy) for (int sx = S0; sx < E0; ++sx) { \
        const int s[3] = {sx, sy, sz};                                                               \
        int is[3], ie[3];                                                                            \
        for (int a = 0; a < 3; ++a) {                                                                \
            is[a] = 2 * s[a] + 1 - he0; if (is[a] < 0) is[a] = 0;                                    \
            ie[a] = 2 * s[a] + 1 - hs0; if (ie[a] > 16) ie[a] = 16;                                  \
        }                                                                                            \
        const REAL v = lab[(sz - lo) * N * N + (sy - lo) * N + (sx - lo)];                           \
        for (int dz = is[2]; dz < ie[2]; ++dz) {                                                     \
            const double wZ = hs[2 * sz - dz - hs0];                                                 \
            for (int dy = is[1]; dy < ie[1]; ++dy) {                                                 \
                const double wZwY = wZ * hs[2 * sy - dy - hs0];                                      \
                for (int dx = is[0]; dx < ie[0]; ++dx) {                                             \
                    const int o[3] = {dx >> 3, dy >> 3, dz >> 3};                                    \
                    if ((o[0] | (o[1] << 1) | (o[2] << 2)) != octant) continue;                      \
                    const REAL w = (REAL)(wZwY * hs[2 * sx - dx - hs0]);                             \
                    const REAL prod = v * w;                                                         \
                    REAL* d = &child[(dx & 7) + 8 * (dy & 7) + 64 * (dz & 7)];                       \
                    *d = *d + prod;                                                                  \
                }                                                                                    \
            }                                                                                        \
        }                                                                                            \
    }

void gho_amr_split_f32(int order, const float* lab, int octant, float* child) { SPLIT_BODY(float) }
void gho_amr_split_f64(int order, const double* lab, int octant, double* child) { SPLIT_BODY(double) }

/* RD steps on a multi-level grid: labs of phi, p_w, p_g per leaf (face ghosts
 * from the multiresolution), operator of ReactionDiffusionOperator.h:98-139
 * with h = 2^-level/8 per block, then UpdateTumor (:215-222) */
#define AMR_RD_BODY(REAL, FILL)                                                                      \
    const int lo = lab_lo(order), N = lab_n(order), N3 = N * N * N;                                  \
    const int64_t nc = (int64_t)nleaves * 512;                                                       \
    tree_t tr;                                                                                       \
    terms_t TT = {0, 0, 0};                                                                          \
    tree_build(&tr, nleaves, leaves);                                                                \
    REAL* lp = (REAL*)malloc(sizeof(REAL) * N3);                                                     \
    REAL* lw = (REAL*)malloc(sizeof(REAL) * N3);                                                     \
    REAL* lg = (REAL*)malloc(sizeof(REAL) * N3);                                                     \
    REAL* dphi = (REAL*)malloc(sizeof(REAL) * nc);                                                   \
    const REAL eps_h = (REAL)1.0e-08, eps_bc = (REAL)0.1;                                            \
    for (int64_t st = 0; st < nsteps; ++st) {                                                        \
        for (int b = 0; b < nleaves; ++b) {                                                          \
            FILL(&tr, leaves, order, phi, b, lp, 1, &TT);                                            \
            FILL(&tr, leaves, order, p_w, b, lw, 1, &TT);                                            \
            FILL(&tr, leaves, order, p_g, b, lg, 1, &TT);                                            \
            const double hh = ldexp(1.0, -leaves[4 * b]) / 8;                                        \
            const double ih2 = 1. / (hh * hh);                                                       \
            for (int z = 0; z < 8; ++z) for (int y = 0; y < 8; ++y) for (int x = 0; x < 8; ++x) {    \
                const int c = (z - lo) * N * N + (y - lo) * N + (x - lo);                            \
                const int nb[6] = {c - 1, c + 1, c - N, c + N, c - N * N, c + N * N};                \
                REAL* out = &dphi[(int64_t)b * 512 + x + 8 * y + 64 * z];                            \
                if (!((double)(REAL)((REAL)(lw[c] + lg[c]) + lp[c]) > 0.)) { *out = (REAL)0; continue; } \
                REAL df_loc = (REAL)((REAL)(lw[c] * Dw) + (REAL)(lg[c] * Dg));                       \
                REAL df[6], chf[6];                                                                  \
                for (int f = 0; f < 6; ++f) {                                                        \
                    df[f] = (REAL)((REAL)(lw[nb[f]] * Dw) + (REAL)(lg[nb[f]] * Dg));                 \
                    REAL s = (REAL)(df[f] + df_loc);                                                 \
                    df[f] = (s < eps_h) ? (REAL)0. : (REAL)(((2. * (double)df[f]) * (double)df_loc) / (double)s); \
                    chf[f] = (REAL)((REAL)(lp[nb[f]] + lw[nb[f]]) + lg[nb[f]]);                      \
                }                                                                                    \
                if (chf[0] < eps_bc) df[1] = (REAL)((double)df[1] * 2.0);                            \
                if (chf[1] < eps_bc) df[0] = (REAL)((double)df[0] * 2.0);                            \
                if (chf[2] < eps_bc) df[3] = (REAL)((double)df[3] * 2.0);                            \
                if (chf[3] < eps_bc) df[2] = (REAL)((double)df[2] * 2.0);                            \
                if (chf[4] < eps_bc) df[5] = (REAL)((double)df[5] * 2.0);                            \
                if (chf[5] < eps_bc) df[4] = (REAL)((double)df[4] * 2.0);                            \
                REAL si = (REAL)(df[0] * lp[nb[0]]);                                                 \
                for (int f = 1; f < 6; ++f) si = (REAL)(si + (REAL)(df[f] * lp[nb[f]]));             \
                REAL sd = df[0];                                                                     \
                for (int f = 1; f < 6; ++f) sd = (REAL)(sd + df[f]);                                 \
                const double fin = ih2 * (double)si;                                                 \
                const double fout = -((double)(REAL)(sd * lp[c]) * ih2);                             \
                const double react = (double)(REAL)(rho * lp[c]) * (1. - (double)lp[c]);             \
                *out = (REAL)((fout + fin) + react);                                                 \
            }                                                                                        \
        }                                                                                            \
        for (int64_t c = 0; c < nc; ++c) {                                                           \
            REAL p = (REAL)((double)phi[c] + dt * (double)dphi[c]);                                  \
            p = ((REAL)0. < p) ? p : (REAL)0.;                                                       \
            p = (p < (REAL)1.) ? p : (REAL)1.;                                                       \
            phi[c] = p;                                                                              \
        }                                                                                            \
    }                                                                                                \
    free(lp); free(lw); free(lg); free(dphi); free(TT.v);                                            \
    tree_free(&tr);

void gho_amr_rd_steps_f32(int nleaves, const int* leaves, int order, float* phi, const float* p_w, const float* p_g,
                          float Dw, float Dg, float rho, double dt, int64_t nsteps) {
    AMR_RD_BODY(float, lab_fill_f32)
}
void gho_amr_rd_steps_f64(int nleaves, const int* leaves, int order, double* phi, const double* p_w,
                          const double* p_g, double Dw, double Dg, double rho, double dt, int64_t nsteps) {
    AMR_RD_BODY(double, lab_fill_f64)
}

/* _dumpUQoutput (Glioma/Glioma_ReactionDiffusion.cpp:208-267) with Real =
 * float positions (BlockInfo origin/h are Real): is_f32 selects that */
void gho_amr_resample(int nleaves, const int* leaves, const double* phi, int is_f32, int gpd, double* out) {
    const double hf = 1. / gpd, eps = hf * 0.5;
    memset(out, 0, sizeof(double) * (size_t)gpd * gpd * gpd);
    for (int b = 0; b < nleaves; ++b) {
        const int L = leaves[4 * b];
        for (int z = 0; z < 8; ++z) for (int y = 0; y < 8; ++y) for (int x = 0; x < 8; ++x) {
            double pos[3], h;
            const int idx[3] = {x, y, z};
            if (is_f32) {
                const float dil = (float)ldexp(1.0, -L), hr = dil / 8.f;
                for (int a = 0; a < 3; ++a) pos[a] = (double)((float)(leaves[4 * b + 1 + a] * dil) + hr * (float)idx[a]);
                h = hr;
            } else {
                const double dil = ldexp(1.0, -L), hr = dil / 8.;
                for (int a = 0; a < 3; ++a) pos[a] = leaves[4 * b + 1 + a] * dil + hr * idx[a];
                h = hr;
            }
            int m[3];
            for (int a = 0; a < 3; ++a) m[a] = (int)floor(pos[a] / hf);
            const int k = h < hf + eps ? 1 : (h < 2. * hf + eps ? 2 : (h < 3. * hf + eps ? 3 : 4));
            const double v = phi[(int64_t)b * 512 + x + 8 * y + 64 * z];
            for (int cz = 0; cz < k; ++cz) for (int cy = 0; cy < k; ++cy) for (int cx = 0; cx < k; ++cx) {
                const int X = m[0] + cx, Y = m[1] + cy, Z = m[2] + cz;
                if (X < gpd && Y < gpd && Z < gpd) out[X + (size_t)gpd * (Y + (size_t)gpd * Z)] = v;
            }
        }
    }
}
