/*
 * gh_oracle_mrag.c -- MRAG's own ghost construction, restated as a second,
 * independent checker of the adaptive lab.
 *
 * TEST INFRASTRUCTURE ONLY (see gh_oracle.h).
 *
 * gh_oracle_amr.c defines a ghost by recursion over the multiresolution
 * (depth first, synthesis taps z-major then y then x).  MRAG builds the same
 * ghosts differently: per block, an instruction list per lab ghost
 * (MRAG/MRAGcore/MRAG_BBInfoCreator.inl:1002-1052), evaluated by
 * BlockLab::load (MRAG/MRAGcore/MRAGBlockLab.h:256-295) as
 *
 *     ghost = 0;  for each instruction in list order:
 *         w = (Real)wx * (Real)wy * (Real)wz;  ghost += value * w;   (Real)
 *
 * The lists come from four places, restated here in the order MRAG runs them:
 *   1. a same-level neighbour: one instruction of weight 1 per covered ghost
 *      (_computeEasyGhosts_SameLevel, MRAG_BBInfoCreator.inl:43-152);
 *   2. a coarser neighbour: the ghosts whose Hs stencil lies inside it
 *      ("reconstructable", :193-205) take every coarse point of that
 *      neighbour whose product of 1-D Hs weights is nonzero, in the
 *      neighbour's (z, y, x) point order (:155-525); the other ghosts it
 *      covers are left over ("bastards", :223-260);
 *   3. a finer neighbour: the coincident fine point, Ha = {1}
 *      (_computeEasyGhosts_Analysis, :528-883);
 *   4. ghosts of codes no neighbour covers, and the left-overs: resolved one
 *      level at a time (_resolveBastardGhosts, :1055-1250;
 *      BastardGhost::generateRequests/collect, MRAG_BBInfoCreator.h:110-234):
 *      a point that no leaf of its own level holds asks for the Hs taps one
 *      level coarser (or the coincident point one level finer); its list is
 *      the depth-first walk over those requests, (z, y, x) at each level,
 *      skipping requests whose weight product is 0, with the 1-D weights
 *      multiplied per axis in double along the walk.
 * Points are located as MRAG_SmartBlockFinder.h:52-337 does: in a table of
 * b's neighbours over the ring of finest-level blocks around b, the point
 * clamped to that ring.
 *
 * A test that finds these lists (and the labs they give, in FP32 and FP64)
 * equal to gh_oracle_amr.c's recursion pins the build's accumulation order to
 * MRAG's.
 */
#include "gh_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int leaf;   /* canonical rank of the leaf */
    int level;  /* its level */
    int idx[3]; /* its index, unwrapped next to b (periodic images shifted) */
} mnode_t;

typedef struct {
    int src; /* leaf * 512 + cell */
    double w[3];
} minstr_t;

typedef struct {
    minstr_t* v;
    int n, cap;
} mlist_t;

typedef struct {
    int L, B[3], order, lo, N, send; /* block, wavelet order, lab [lo, lo + N), stencil_end */
    int S0, S1;                      /* Hs support [S0, S1) */
    mnode_t* nb;
    int nnb;
    int M, len; /* finest level of b and its neighbours; 2^(M - L) */
    int* ring;  /* (len + 2)^3 cells at level M around b: neighbour number, -1 = b */
    int err;
} mblock_t;

static void mpush(mlist_t* l, int src, const double w[3]) {
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 8;
        l->v = (minstr_t*)realloc(l->v, sizeof(minstr_t) * l->cap);
    }
    l->v[l->n].src = src;
    for (int a = 0; a < 3; ++a) l->v[l->n].w[a] = w[a];
    l->n++;
}

/* Hs tap at offset k (MRAGWavelets_StaticData.cpp:15,33), 0 outside */
static double hs_tap(int order, int k) {
    if (order == 2) return k == 0 ? 1.0 : ((k == 1 || k == -1) ? 0.5 : 0.0);
    switch (k) {
        case 0: return 1.0;
        case 1: case -1: return 9.0 / 16;
        case 3: case -3: return -1.0 / 16;
        default: return 0.0;
    }
}

/* leaves touching b's closed box, periodic images as separate nodes
 * (MRAG/MRAGcore/MRAGrid.inl:840-1149 neighbourhood) */
static void find_neighbours(mblock_t* m, int nleaves, const int* leaves, int self) {
    const int U = 12; /* common unit: level 12 */
    m->nb = (mnode_t*)malloc(sizeof(mnode_t) * (size_t)nleaves * 27);
    m->nnb = 0;
    const int64_t D = (int64_t)1 << U;
    int64_t blo[3], bhi[3];
    for (int a = 0; a < 3; ++a) {
        blo[a] = (int64_t)m->B[a] << (U - m->L);
        bhi[a] = blo[a] + ((int64_t)1 << (U - m->L));
    }
    for (int j = 0; j < nleaves; ++j) {
        const int l = leaves[4 * j];
        const int64_t sz = (int64_t)1 << (U - l);
        for (int s = 0; s < 27; ++s) {
            const int sh[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
            if (j == self && s == 13) continue;
            int touch = 1;
            for (int a = 0; a < 3 && touch; ++a) {
                const int64_t lo = ((int64_t)leaves[4 * j + 1 + a] << (U - l)) + sh[a] * D, hi = lo + sz;
                if (lo > bhi[a] || hi < blo[a]) touch = 0;
            }
            if (!touch) continue;
            mnode_t* n = &m->nb[m->nnb++];
            n->leaf = j;
            n->level = l;
            for (int a = 0; a < 3; ++a) n->idx[a] = leaves[4 * j + 1 + a] + sh[a] * (1 << l);
        }
    }
}

/* the neighbour table of MRAG_SmartBlockFinder.h:52-200 */
static void build_ring(mblock_t* m) {
    m->M = m->L;
    for (int i = 0; i < m->nnb; ++i)
        if (m->nb[i].level > m->M) m->M = m->nb[i].level;
    m->len = 1 << (m->M - m->L);
    const int W = m->len + 2;
    m->ring = (int*)malloc(sizeof(int) * W * W * W);
    for (int i = 0; i < W * W * W; ++i) m->ring[i] = -2; /* not covered */
    for (int i = 0; i < m->nnb; ++i) {
        const mnode_t* n = &m->nb[i];
        int s[3], e[3];
        for (int a = 0; a < 3; ++a) {
            const int sn = n->idx[a] * (1 << (m->M - n->level)) - m->B[a] * m->len;
            const int en = sn + (1 << (m->M - n->level));
            s[a] = sn > -1 ? sn : -1;
            e[a] = en < m->len + 1 ? en : m->len + 1;
        }
        for (int z = s[2]; z < e[2]; ++z)
            for (int y = s[1]; y < e[1]; ++y)
                for (int x = s[0]; x < e[0]; ++x) m->ring[(x + 1) + W * ((y + 1) + W * (z + 1))] = i;
    }
}

static int floor_div(int a, int b) { return (int)floor((double)a / (double)b); }

/* findLyingNode (MRAG_SmartBlockFinder.h:226-296): -1 = b itself */
static int lying_node(const mblock_t* m, int level, const int ri[3]) {
    int c[3];
    for (int a = 0; a < 3; ++a) {
        int v = floor_div(ri[a] * (1 << (m->M - level)), 8);
        if (v < -1) v = -1;
        if (v > m->len) v = m->len;
        c[a] = v;
    }
    if (c[0] >= 0 && c[0] < m->len && c[1] >= 0 && c[1] < m->len && c[2] >= 0 && c[2] < m->len) return -1;
    const int W = m->len + 2;
    return m->ring[(c[0] + 1) + W * ((c[1] + 1) + W * (c[2] + 1))];
}

static int node_level(const mblock_t* m, int k) { return k == -1 ? m->L : m->nb[k].level; }

/* findData (:298-337): the cell of a point whose lying node k has the
 * point's own level */
static int data_cell(mblock_t* m, int level, const int ri[3], int k) {
    int d[3];
    for (int a = 0; a < 3; ++a) {
        const int nidx = k == -1 ? m->B[a] : m->nb[k].idx[a];
        const double dilate = pow(0.5, m->L - level);
        d[a] = ri[a] - 8 * nidx + (int)floor(dilate * 8 * m->B[a]);
        if (d[a] < 0 || d[a] >= 8) m->err = 1; /* MRAG asserts here */
    }
    return (d[0] & 7) + 8 * (d[1] & 7) + 64 * (d[2] & 7);
}

/* BastardGhost: resolve a point and collect its list depth first */
static void resolve(mblock_t* m, int self, int level, const int ri[3], const double wacc[3], mlist_t* out, int depth) {
    if (depth > 8 || m->err) { m->err = 1; return; }
    const int k = lying_node(m, level, ri);
    if (k == -2) { m->err = 1; return; }
    if (node_level(m, k) == level) {
        const int leaf = k == -1 ? self : m->nb[k].leaf;
        mpush(out, leaf * 512 + data_cell(m, level, ri, k), wacc);
        return;
    }
    const int analysis = node_level(m, k) > level;
    int s[3], e[3];
    double vw[3][8];
    for (int a = 0; a < 3; ++a) {
        if (analysis) {
            s[a] = 2 * ri[a]; /* Ha support [0, 1) */
            e[a] = 2 * ri[a] + 1;
        } else {
            s[a] = (int)ceil((ri[a] + m->S0) * 0.5);
            e[a] = 1 + (int)floor((ri[a] + m->S1 - 1) * 0.5);
        }
        const int nw = e[a] - s[a];
        for (int i = s[a]; i < e[a]; ++i)
            vw[a][i - s[a]] = nw == 1 ? 1.0 : (analysis ? 1.0 : hs_tap(m->order, 2 * i - ri[a]));
    }
    int q[3];
    for (q[2] = s[2]; q[2] < e[2]; ++q[2])
        for (q[1] = s[1]; q[1] < e[1]; ++q[1])
            for (q[0] = s[0]; q[0] < e[0]; ++q[0]) {
                const double w = vw[0][q[0] - s[0]] * vw[1][q[1] - s[1]] * vw[2][q[2] - s[2]];
                if (w == 0.0) continue;
                const double wn[3] = {wacc[0] * vw[0][q[0] - s[0]], wacc[1] * vw[1][q[1] - s[1]],
                                      wacc[2] * vw[2][q[2] - s[2]]};
                resolve(m, self, level + (analysis ? 1 : -1), q, wn, out, depth + 1);
            }
}

/* all instruction lists of b's lab: lists[p], p over the N^3 lab positions
 * (x fastest); interior positions stay empty.  Returns 0, or -1 when the
 * grid has a neighbour two levels away or a point falls outside MRAG's table. */
static int build_lists(int nleaves, const int* leaves, int order, int self, mlist_t* lists, mblock_t* m) {
    memset(m, 0, sizeof(*m));
    m->L = leaves[4 * self];
    for (int a = 0; a < 3; ++a) m->B[a] = leaves[4 * self + 1 + a];
    m->order = order;
    m->lo = order == 4 ? -2 : -1;
    m->N = order == 4 ? 13 : 10;
    m->send = m->lo + m->N - 8 + 1;
    m->S0 = order == 4 ? -3 : -1;
    m->S1 = order == 4 ? 4 : 2;
    find_neighbours(m, nleaves, leaves, self);
    for (int i = 0; i < m->nnb; ++i)
        if (abs(m->nb[i].level - m->L) > 1) return -1;
    build_ring(m);
    const int lo = m->lo, N = m->N;
    char* pending = (char*)calloc((size_t)N * N * N, 1); /* bastards, by lab position */
    int covered[27] = {0};
    for (int i = 0; i < m->nnb; ++i) {
        const mnode_t* n = &m->nb[i];
        /* _computeNeighborCode (MRAG_BBInfoCreator.inl:20-39) */
        const double dil = pow(2., -(n->level - m->L));
        double p[3], pmax = 0;
        for (int a = 0; a < 3; ++a) {
            p[a] = dil * (n->idx[a] + 0.5) - (m->B[a] + 0.5);
            if (fabs(p[a]) > pmax) pmax = fabs(p[a]);
        }
        int dir[3];
        for (int a = 0; a < 3; ++a) dir[a] = (int)(p[a] / pmax);
        const int code = (dir[0] + 1) + 3 * (dir[1] + 1) + 9 * (dir[2] + 1);
        covered[code] = 1;
        int gs[3], ge[3]; /* the code's ghost box */
        for (int a = 0; a < 3; ++a) {
            gs[a] = dir[a] < 0 ? lo : (dir[a] == 0 ? 0 : 8);
            ge[a] = dir[a] < 0 ? 0 : (dir[a] == 0 ? 8 : 8 + m->send - 1);
        }
        int cs[3], ce[3]; /* n's coverage in b's index space */
        for (int a = 0; a < 3; ++a) {
            if (n->level >= m->L) {
                cs[a] = floor_div(8 * n->idx[a], 1 << (n->level - m->L)) - 8 * m->B[a];
                ce[a] = floor_div(8 * (n->idx[a] + 1) - 1, 1 << (n->level - m->L)) + 1 - 8 * m->B[a];
            } else {
                cs[a] = 8 * n->idx[a] * 2 - 8 * m->B[a];
                ce[a] = 8 * (n->idx[a] + 1) * 2 - 8 * m->B[a];
            }
        }
        int es[3], ee[3]; /* easy range */
        for (int a = 0; a < 3; ++a) {
            if (n->level < m->L) { /* reconstructable from n alone (:193-205) */
                const int rs = ((8 * n->idx[a] - m->S0 - 1) << 1) + m->S0 + 1 - 8 * m->B[a];
                const int re = ((8 * (n->idx[a] + 1) - m->S1 + 1) << 1) + m->S1 - 1 - 8 * m->B[a];
                es[a] = rs > gs[a] ? rs : gs[a];
                ee[a] = re < ge[a] ? re : ge[a];
            } else {
                es[a] = cs[a] > gs[a] ? cs[a] : gs[a];
                ee[a] = ce[a] < ge[a] ? ce[a] : ge[a];
            }
        }
        int d[3];
        for (d[2] = gs[2]; d[2] < ge[2]; ++d[2])
            for (d[1] = gs[1]; d[1] < ge[1]; ++d[1])
                for (d[0] = gs[0]; d[0] < ge[0]; ++d[0]) {
                    int inside = 1, easy = 1;
                    for (int a = 0; a < 3; ++a) {
                        if (d[a] < cs[a] || d[a] >= ce[a]) inside = 0;
                        if (d[a] < es[a] || d[a] >= ee[a]) easy = 0;
                    }
                    if (!inside) continue;
                    const int pidx = (d[0] - lo) + N * ((d[1] - lo) + N * (d[2] - lo));
                    if (!easy) { pending[pidx] = 1; continue; }
                    const double one[3] = {1, 1, 1};
                    if (n->level == m->L) {
                        int c[3];
                        for (int a = 0; a < 3; ++a) c[a] = d[a] - 8 * (n->idx[a] - m->B[a]);
                        mpush(&lists[pidx], n->leaf * 512 + c[0] + 8 * c[1] + 64 * c[2], one);
                    } else if (n->level > m->L) {
                        int c[3];
                        for (int a = 0; a < 3; ++a) c[a] = 2 * (8 * m->B[a] + d[a]) - 8 * n->idx[a];
                        mpush(&lists[pidx], n->leaf * 512 + c[0] + 8 * c[1] + 64 * c[2], one);
                    } else {
                        int s[3];
                        for (s[2] = 0; s[2] < 8; ++s[2])
                            for (s[1] = 0; s[1] < 8; ++s[1])
                                for (s[0] = 0; s[0] < 8; ++s[0]) {
                                    double w[3];
                                    for (int a = 0; a < 3; ++a)
                                        w[a] = (double)(float)hs_tap(order,
                                                                     2 * (8 * n->idx[a] + s[a]) - (8 * m->B[a] + d[a]));
                                    if (!(fabs(w[0] * w[1] * w[2]) > 0)) continue;
                                    mpush(&lists[pidx], n->leaf * 512 + s[0] + 8 * s[1] + 64 * s[2], w);
                                }
                    }
                }
    }
    /* codes no neighbour covers: all their ghosts are bastards (:1002-1031) */
    for (int code = 0; code < 27; ++code) {
        if (code == 13 || covered[code]) continue;
        const int dir[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        int gs[3], ge[3];
        for (int a = 0; a < 3; ++a) {
            gs[a] = dir[a] < 0 ? lo : (dir[a] == 0 ? 0 : 8);
            ge[a] = dir[a] < 0 ? 0 : (dir[a] == 0 ? 8 : 8 + m->send - 1);
        }
        for (int z = gs[2]; z < ge[2]; ++z)
            for (int y = gs[1]; y < ge[1]; ++y)
                for (int x = gs[0]; x < ge[0]; ++x) pending[(x - lo) + N * ((y - lo) + N * (z - lo))] = 1;
    }
    for (int p = 0; p < N * N * N; ++p) {
        if (!pending[p]) continue;
        if (lists[p].n) { m->err = 1; break; } /* a ghost claimed twice */
        const int ri[3] = {p % N + lo, (p / N) % N + lo, p / (N * N) + lo};
        const double one[3] = {1, 1, 1};
        resolve(m, self, m->L, ri, one, &lists[p], 0);
    }
    for (int p = 0; p < N * N * N && !m->err; ++p) {
        const int i[3] = {p % N + lo, (p / N) % N + lo, p / (N * N) + lo};
        const int interior = i[0] >= 0 && i[0] < 8 && i[1] >= 0 && i[1] < 8 && i[2] >= 0 && i[2] < 8;
        if (!interior && lists[p].n == 0) m->err = 1; /* a ghost nobody filled */
    }
    free(pending);
    return m->err ? -1 : 0;
}

static void free_lists(mlist_t* lists, int n, mblock_t* m) {
    for (int p = 0; p < n; ++p) free(lists[p].v);
    free(lists);
    free(m->nb);
    free(m->ring);
}

#define MRAG_LAB(NAME, REAL)                                                                          \
    int NAME(int nleaves, const int* leaves, int order, const REAL* field, int leaf, REAL* lab) {      \
        mblock_t m;                                                                                   \
        const int N = order == 4 ? 13 : 10, lo = order == 4 ? -2 : -1;                                \
        mlist_t* lists = (mlist_t*)calloc((size_t)N * N * N, sizeof(mlist_t));                        \
        const int rc = build_lists(nleaves, leaves, order, leaf, lists, &m);                          \
        for (int p = 0; p < N * N * N && rc == 0; ++p) {                                              \
            const int i[3] = {p % N + lo, (p / N) % N + lo, p / (N * N) + lo};                        \
            if (i[0] >= 0 && i[0] < 8 && i[1] >= 0 && i[1] < 8 && i[2] >= 0 && i[2] < 8) {            \
                lab[p] = field[(size_t)leaf * 512 + i[0] + 8 * i[1] + 64 * i[2]];                     \
                continue;                                                                             \
            }                                                                                         \
            REAL ghost = (REAL)0;                                                                     \
            for (int k = 0; k < lists[p].n; ++k) {                                                    \
                const minstr_t* t = &lists[p].v[k];                                                   \
                const REAL w = (REAL)((REAL)((REAL)t->w[0] * (REAL)t->w[1]) * (REAL)t->w[2]);         \
                const REAL prod = field[t->src] * w;                                                  \
                ghost = ghost + prod;                                                                 \
            }                                                                                         \
            lab[p] = ghost;                                                                           \
        }                                                                                             \
        free_lists(lists, N * N * N, &m);                                                             \
        return rc;                                                                                    \
    }
MRAG_LAB(gho_mrag_lab_f32, float)
MRAG_LAB(gho_mrag_lab_f64, double)

/* the instruction lists of every lab position of `leaf` (N^3 positions, x
 * fastest; interior positions have none): counts[p], then up to cap entries
 * of source (leaf * 512 + cell) and the Real weight (REAL)wx * (REAL)wy *
 * (REAL)wz as FP32 (is_f32) or FP64, list after list.  Returns the total
 * count, or -1. */
int gho_mrag_block_terms(int nleaves, const int* leaves, int order, int leaf, int is_f32, int cap, int* counts,
                         int* src, double* w) {
    mblock_t m;
    const int N = order == 4 ? 13 : 10;
    mlist_t* lists = (mlist_t*)calloc((size_t)N * N * N, sizeof(mlist_t));
    int total = build_lists(nleaves, leaves, order, leaf, lists, &m);
    if (total == 0) {
        for (int p = 0; p < N * N * N; ++p) {
            counts[p] = lists[p].n;
            for (int q = 0; q < lists[p].n; ++q, ++total) {
                if (total >= cap) continue;
                const double* t = lists[p].v[q].w;
                src[total] = lists[p].v[q].src;
                w[total] = is_f32 ? (double)((float)((float)t[0] * (float)t[1]) * (float)t[2]) : (t[0] * t[1]) * t[2];
            }
        }
    }
    free_lists(lists, N * N * N, &m);
    return total;
}
