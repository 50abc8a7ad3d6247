// gh_octree.cpp -- host octree of leaf blocks: topology, neighbourhood,
// refinement / compression plans, ghost recursion.  See gh_octree.h.
#include "gh_octree.h"

#include <algorithm>

namespace gh {

// periodic wrap into [0, n) for a power-of-two n (every lattice here is one)
static inline int wrapi(int64_t v, int64_t n) { return (int)(v & (n - 1)); }

Octree::Kind Octree::kind(int L, int x, int y, int z, int* slot) const {
    const int n = 1 << L;
    x = wrapi(x, n);
    y = wrapi(y, n);
    z = wrapi(z, n);
    if (L < dense_levels) {
        const int32_t v = dense[L][(size_t)x + ((size_t)y << L) + ((size_t)z << (2 * L))];
        if (v >= 0) {
            if (slot) *slot = v;
            return LEAF;
        }
        return v == -2 ? FINER : COARSER;
    }
    if (L > top_level || hash_bits == 0) return COARSER;  // no node below the finest leaf
    const uint64_t k = key(L, x, y, z);
    const size_t mask = ((size_t)1 << hash_bits) - 1;
    for (size_t h = node_hash(k, hash_bits);; h = (h + 1) & mask) {
        if (hash_keys[h] == kEmptyKey) return COARSER;
        if (hash_keys[h] != k) continue;
        const int32_t v = hash_vals[h];
        if (v >= 0) {
            if (slot) *slot = v;
            return LEAF;
        }
        return v == -2 ? FINER : COARSER;
    }
}

void Octree::init_uniform(int level) {
    leaves.clear();
    const int n = 1 << level;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) leaves.push_back(BlockRec{level, {x, y, z}});
    rebuild_index();
}

// (level, z, y, x) ascending by an LSD radix sort of the 31-bit keys
// level << 27 | z << 18 | y << 9 | x (levels < 16, coordinates < 512) in three
// passes of 11 bits
static void radix_sort_leaves(std::vector<BlockRec>& v) {
    const size_t nl = v.size();
    std::vector<uint32_t> a(nl), b(nl);
    for (size_t s = 0; s < nl; ++s) {
        const BlockRec& r = v[s];
        a[s] = (uint32_t)r.level << 27 | (uint32_t)r.index[2] << 18 | (uint32_t)r.index[1] << 9 | (uint32_t)r.index[0];
    }
    std::vector<uint32_t> cnt(2049);
    for (int pass = 0; pass < 3; ++pass) {
        const int sh = 11 * pass;
        std::fill(cnt.begin(), cnt.end(), 0u);
        for (uint32_t k : a) ++cnt[((k >> sh) & 2047u) + 1];
        for (int i = 0; i < 2048; ++i) cnt[i + 1] += cnt[i];
        for (uint32_t k : a) b[cnt[(k >> sh) & 2047u]++] = k;
        a.swap(b);
    }
    for (size_t s = 0; s < nl; ++s)
        v[s] = BlockRec{(int)(a[s] >> 27), {(int)(a[s] & 511u), (int)((a[s] >> 9) & 511u), (int)((a[s] >> 18) & 511u)}};
}

// canonical slot order: (level, z, y, x) ascending.  The leaves of levels up
// to 7 are marked in one bit set per level (occ[l]: 8^l bits, x fastest), read
// back in bit order; deeper leaves are radix sorted after them.  A leaf given
// twice leaves the whole list to the radix sort, so that it stays in the list,
// and the result is false.
static bool sort_leaves(std::vector<BlockRec>& leaves, int top, std::vector<uint64_t> (&occ)[8]) {
    const int nb = std::min(top, 7);
    for (int l = 0; l <= nb; ++l) occ[l].assign((((size_t)1 << (3 * l)) + 63) / 64, 0);
    std::vector<BlockRec> deep;
    uint64_t twice = 0;
    for (const BlockRec& b : leaves) {
        if (b.level > nb) {
            deep.push_back(b);
            continue;
        }
        const size_t i = (size_t)b.index[0] + ((size_t)b.index[1] << b.level) + ((size_t)b.index[2] << (2 * b.level));
        uint64_t& w = occ[b.level][i >> 6];
        const uint64_t bit = 1ull << (i & 63);
        twice |= w & bit;
        w |= bit;
    }
    if (twice) {
        radix_sort_leaves(leaves);
        return false;
    }
    radix_sort_leaves(deep);
    size_t s = 0;
    for (int l = 0; l <= nb; ++l) {
        const size_t m = ((size_t)1 << l) - 1;
        for (size_t wi = 0; wi < occ[l].size(); ++wi)
            for (uint64_t w = occ[l][wi]; w; w &= w - 1) {
                const size_t i = (wi << 6) + (size_t)__builtin_ctzll(w);
                leaves[s++] = BlockRec{l, {(int)(i & m), (int)((i >> l) & m), (int)(i >> (2 * l))}};
            }
    }
    for (const BlockRec& b : deep) leaves[s++] = b;
    return true;
}

// bit k of the result = bit 2k of v
static inline uint64_t even_bits(uint64_t v) {
    v &= 0x5555555555555555ull;
    v = (v | (v >> 1)) & 0x3333333333333333ull;
    v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
    v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
    return (v | (v >> 16)) & 0x00000000ffffffffull;
}

void Octree::rebuild_index() {
    for (int l = 0; l < MAXL; ++l) dense[l].clear();
    dense_levels = 0;
    top_level = -1;
    hash_bits = 0;
    hash_keys.clear();
    hash_vals.clear();
    interior_deep.clear();
    nested = false;
    row_levels = 0;
    if (leaves.empty()) return;
    int top = 0;
    for (const BlockRec& b : leaves) top = std::max(top, b.level);
    std::vector<uint64_t> occ[8];
    const bool marked = sort_leaves(leaves, top, occ);
    const int nd = std::min(top, DENSE_MAX);
    for (int l = 0; l <= nd; ++l) dense[l].assign((size_t)1 << (3 * l), -1);
    size_t deep_leaves = 0;  // leaves below the dense levels
    for (size_t s = 0; s < leaves.size(); ++s) {
        const BlockRec& b = leaves[s];
        if (b.level > nd) {
            ++deep_leaves;
            continue;
        }
        int32_t& e =
            dense[b.level][(size_t)b.index[0] + ((size_t)b.index[1] << b.level) + ((size_t)b.index[2] << (2 * b.level))];
        if (e != -1) nested = true;  // the same leaf twice
        e = (int32_t)s;
    }
    // the interior nodes are the strict ancestors of the leaves, marked -2; an
    // interior node that is also a leaf means nesting.  Nodes of levels >=
    // SCAN_MAX + 1 (and all of the hashed levels) are listed.
    std::vector<uint64_t>& deep = interior_deep;
    if (marked && top <= 6) {
        // level by level from the finest up, on rows of x bits (n <= 64): a
        // node is interior iff one of its 8 children is a leaf or interior
        // (no level here is listed: all are <= SCAN_MAX)
        auto occ_row = [&](int l, size_t r) -> uint64_t {
            if (l == 6) return occ[l][r];
            const size_t b = r << l;
            return (occ[l][b >> 6] >> (b & 63)) & ((1ull << (1 << l)) - 1);
        };
        std::vector<uint64_t> below((size_t)1 << (2 * top)), cur;  // rows of level l + 1: leaves and interior nodes
        for (int l = 0; l <= top; ++l) {
            leaf_rows[l].resize((size_t)1 << (2 * l));
            for (size_t r = 0; r < leaf_rows[l].size(); ++r) leaf_rows[l][r] = occ_row(l, r);
            node_rows[l].assign((size_t)1 << (2 * l), 0);
        }
        below = leaf_rows[top];
        for (int l = top - 1; l >= 0; --l) {
            const int n = 1 << l;
            const size_t dz = (size_t)1 << (l + 1);
            cur.assign((size_t)n * n, 0);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    const size_t r0 = (size_t)(2 * y) + (size_t)(2 * z) * dz;
                    uint64_t v = below[r0] | below[r0 + 1] | below[r0 + dz] | below[r0 + dz + 1];
                    v = even_bits(v | v >> 1);
                    const size_t r = (size_t)y + ((size_t)z << l);
                    for (uint64_t m = v; m; m &= m - 1) {
                        int32_t& e = dense[l][(r << l) + (size_t)__builtin_ctzll(m)];
                        if (e >= 0) nested = true;  // a leaf containing another leaf
                        e = -2;
                    }
                    node_rows[l][r] = v;
                    cur[r] = v | leaf_rows[l][r];
                }
            below.swap(cur);
        }
        row_levels = top + 1;
    } else {  // a walk up from each leaf that stops at a node marked before
        for (const BlockRec& b : leaves) {
            for (int l = b.level - 1; l >= 0; --l) {
                const int sh = b.level - l;
                const int x = b.index[0] >> sh, y = b.index[1] >> sh, z = b.index[2] >> sh;
                if (l > nd) {
                    deep.push_back(key(l, x, y, z));
                    continue;
                }
                int32_t& e = dense[l][(size_t)x + ((size_t)y << l) + ((size_t)z << (2 * l))];
                if (e == -2) break;
                if (e >= 0) nested = true;  // a leaf containing another leaf
                e = -2;
                if (l > SCAN_MAX) deep.push_back(key(l, x, y, z));
            }
        }
    }
    std::sort(deep.begin(), deep.end());
    deep.erase(std::unique(deep.begin(), deep.end()), deep.end());
    dense_levels = nd + 1;
    top_level = top;
    size_t deep_nodes = 0;  // interior nodes below the dense levels
    for (uint64_t k : deep) deep_nodes += (int)(k >> 48) > nd;
    if (deep_leaves + deep_nodes == 0) return;
    hash_bits = 4;
    while (((size_t)1 << hash_bits) < 2 * (deep_leaves + deep_nodes)) ++hash_bits;
    const size_t cap = (size_t)1 << hash_bits;
    hash_keys.assign(cap, kEmptyKey);
    hash_vals.assign(cap, -1);
    auto put = [&](uint64_t k, int32_t v) {
        size_t h = node_hash(k, hash_bits);
        while (hash_keys[h] != kEmptyKey) {
            if (hash_keys[h] == k) {  // a leaf that is also an interior node, or twice
                nested = true;
                return;
            }
            h = (h + 1) & (cap - 1);
        }
        hash_keys[h] = k;
        hash_vals[h] = v;
    };
    for (uint64_t k : deep)
        if ((int)(k >> 48) > nd) put(k, -2);
    for (size_t s = 0; s < leaves.size(); ++s)
        if (leaves[s].level > nd) put(key(leaves[s]), (int32_t)s);
}

// the canonical order puts the coarsest leaf first and the finest last
int Octree::levels_min() const { return leaves.empty() ? MAXL : leaves.front().level; }
int Octree::levels_max() const { return leaves.empty() ? 0 : leaves.back().level; }

// leaves of the subtree at (L, P) that touch the face/edge/corner facing
// direction -d (the side adjacent to the querying block)
static void touching_children(const Octree& t, int L, const int P[3], const int d[3], bool wrapped,
                              std::vector<Octree::Nbr>& out) {
    int slot;
    const Octree::Kind k = t.kind(L, P[0], P[1], P[2], &slot);
    if (k == Octree::LEAF) {
        out.push_back({slot, wrapped});
        return;
    }
    if (k != Octree::FINER) return;
    for (int o = 0; o < 8; ++o) {
        const int off[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            if (d[a] < 0 && off[a] != 1) ok = false;  // child on the max side faces the block
            if (d[a] > 0 && off[a] != 0) ok = false;
        }
        if (!ok) continue;
        const int C[3] = {2 * P[0] + off[0], 2 * P[1] + off[1], 2 * P[2] + off[2]};
        touching_children(t, L + 1, C, d, wrapped, out);
    }
}

std::vector<Octree::Nbr> Octree::neighbours(int s) const {
    const BlockRec& b = leaves[s];
    const int L = b.level, n = 1 << L;
    std::vector<Nbr> out;
    for (int code = 0; code < 27; ++code) {
        if (code == 13) continue;
        const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        int P[3];
        bool wrapped = false;
        for (int a = 0; a < 3; ++a) {
            P[a] = b.index[a] + d[a];
            if (P[a] < 0 || P[a] >= n) wrapped = true;
            P[a] = wrapi(P[a], n);
        }
        int slot;
        const Kind k = kind(L, P[0], P[1], P[2], &slot);
        if (k == LEAF) {
            out.push_back({slot, wrapped});
        } else if (k == FINER) {
            touching_children(*this, L, P, d, wrapped, out);
        } else {
            for (int l = L - 1; l >= 0; --l) {
                const int sh = L - l;
                if (kind(l, P[0] >> sh, P[1] >> sh, P[2] >> sh, &slot) == LEAF) {
                    // a coarser leaf reached through the wrap is a periodic image
                    // unless its box also touches the block directly
                    out.push_back({slot, wrapped});
                    break;
                }
            }
        }
    }
    std::sort(out.begin(), out.end(), [](const Nbr& a, const Nbr& b) {
        return a.slot != b.slot ? a.slot < b.slot : a.wrapped < b.wrapped;
    });
    out.erase(std::unique(out.begin(), out.end(),
                          [](const Nbr& a, const Nbr& b) { return a.slot == b.slot && a.wrapped == b.wrapped; }),
              out.end());
    return out;
}

void Octree::same_level_neighbours(int s, int out[27]) const {
    const BlockRec& b = leaves[s];
    const int L = b.level;
    if (L < dense_levels) {  // straight from the level's dense map
        const int m = (1 << L) - 1;
        const int32_t* map = dense[L].data();
        size_t ox[3], oy[3], oz[3];
        for (int d = 0; d < 3; ++d) {
            ox[d] = (size_t)((b.index[0] + d - 1) & m);
            oy[d] = (size_t)((b.index[1] + d - 1) & m) << L;
            oz[d] = (size_t)((b.index[2] + d - 1) & m) << (2 * L);
        }
        for (int dz = 0; dz < 3; ++dz)
            for (int dy = 0; dy < 3; ++dy)
                for (int dx = 0; dx < 3; ++dx) {
                    const int32_t v = map[ox[dx] + oy[dy] + oz[dz]];
                    out[dx + 3 * dy + 9 * dz] = v >= 0 ? v : -1;
                }
        return;
    }
    for (int code = 0; code < 27; ++code) {
        const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        int slot = -1;
        if (kind(b.level, b.index[0] + d[0], b.index[1] + d[1], b.index[2] + d[2], &slot) != LEAF) slot = -1;
        out[code] = slot;
    }
}

// Interp wavelet synthesis taps Hs at the fine point q: even q -> the coarse
// point q/2 with weight 1; odd q -> neighbours with Hs(+-1) (and Hs(+-3))
// (MRAG/MRAGcore/MRAGWavelets_StaticData.cpp:13-21, 31-39)
static int synth_taps(int order, int64_t q, int64_t s[4], double w[4]) {
    if ((q & 1) == 0) {
        s[0] = q / 2;
        w[0] = 1.0;
        return 1;
    }
    if (order == 2) {
        s[0] = (q - 1) / 2; w[0] = 0.5;
        s[1] = (q + 1) / 2; w[1] = 0.5;
        return 2;
    }
    s[0] = (q - 3) / 2; w[0] = -1.0 / 16;
    s[1] = (q - 1) / 2; w[1] = 9.0 / 16;
    s[2] = (q + 1) / 2; w[2] = 9.0 / 16;
    s[3] = (q + 3) / 2; w[3] = -1.0 / 16;
    return 4;
}

void Octree::value_terms(int L, const int64_t q_in[3], double w, std::vector<GhostTerm>& out) const {
    const int64_t n = (int64_t)BS << L;
    int64_t q[3];
    for (int a = 0; a < 3; ++a) q[a] = wrapi(q_in[a], n);
    int slot;
    const Kind k = kind(L, (int)(q[0] >> 3), (int)(q[1] >> 3), (int)(q[2] >> 3), &slot);
    if (k == LEAF) {
        out.push_back({slot * BCELLS + (int)((q[0] & 7) + 8 * (q[1] & 7) + 64 * (q[2] & 7)), w});
        return;
    }
    if (k == FINER) {  // Ha = {1}: the coincident fine vertex
        const int64_t f[3] = {2 * q[0], 2 * q[1], 2 * q[2]};
        value_terms(L + 1, f, w, out);
        return;
    }
    int64_t s[3][4];
    double ws[3][4];
    int nt[3];
    for (int a = 0; a < 3; ++a) nt[a] = synth_taps(order, q[a], s[a], ws[a]);
    for (int kz = 0; kz < nt[2]; ++kz)
        for (int ky = 0; ky < nt[1]; ++ky)
            for (int kx = 0; kx < nt[0]; ++kx) {
                const int64_t c[3] = {s[0][kx], s[1][ky], s[2][kz]};
                value_terms(L - 1, c, w * (ws[2][kz] * ws[1][ky] * ws[0][kx]), out);
            }
}

// Region of the 26-neighbourhood of leaf s in direction code: its index P at
// the leaf's level (wrapped periodically), whether the wrap was needed, and
// its kind (a leaf of the same level, subdivided, or inside a coarser leaf).
Octree::Kind Octree::neighbour_region(int s, int code, int P[3], bool& wrapped, int* slot) const {
    const BlockRec& b = leaves[s];
    const int L = b.level, n = 1 << L;
    const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
    wrapped = false;
    for (int a = 0; a < 3; ++a) {
        P[a] = b.index[a] + d[a];
        if (P[a] < 0 || P[a] >= n) wrapped = true;
        P[a] = wrapi(P[a], n);
    }
    return kind(L, P[0], P[1], P[2], slot);
}

// A subdivided neighbour region always holds finer leaves that touch the
// block (the leaves tiling it reach its boundary), and only a subdivided
// region holds leaves finer than the block: so s has a finer neighbour
// reached without the periodic wrap iff some unwrapped direction is FINER.
bool Octree::has_finer_direct_neighbour(int s) const {
    for (int code = 0; code < 27; ++code) {
        if (code == 13) continue;
        int P[3];
        bool wrapped;
        if (neighbour_region(s, code, P, wrapped) == FINER && !wrapped) return true;
    }
    return false;
}

// For each of the 27 direction codes: the children (octant bits x | y << 1 |
// z << 2) of the neighbour region in that direction that touch the node's
// box -- on the max side of an axis where the region lies below the node, on
// the min side where it lies above.  Faces have 4, edges 2, corners 1.
struct FacingChildren {
    int n[27];
    int oct[27][8];
    FacingChildren() {
        for (int code = 0; code < 27; ++code) {
            const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
            n[code] = 0;
            for (int o = 0; o < 8; ++o) {
                const int off[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
                bool faces = true;
                for (int a = 0; a < 3; ++a) {
                    if (d[a] < 0 && off[a] != 1) faces = false;
                    if (d[a] > 0 && off[a] != 0) faces = false;
                }
                if (faces) oct[code][n[code]++] = o;
            }
        }
    }
};
static const FacingChildren kFacing;

bool Octree::outside_within_one(int L, const int P[3]) const {
    const int n = 1 << L;
    if (L + 1 < dense_levels) {  // the same tests on the dense maps of levels L - 1, L and L + 1
        const int m = n - 1;
        const int32_t* map = dense[L].data();
        const int32_t* fine = dense[L + 1].data();
        for (int code = 0; code < 27; ++code) {
            if (code == 13) continue;
            const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
            const int Q[3] = {(P[0] + d[0]) & m, (P[1] + d[1]) & m, (P[2] + d[2]) & m};
            const int32_t v = map[(size_t)Q[0] + ((size_t)Q[1] << L) + ((size_t)Q[2] << (2 * L))];
            if (v >= 0) continue;  // a leaf
            if (v == -1) {         // inside a coarser leaf: it must be one level up
                if (L == 0 || dense[L - 1][(size_t)(Q[0] >> 1) + ((size_t)(Q[1] >> 1) << (L - 1)) +
                                           ((size_t)(Q[2] >> 1) << (2 * (L - 1)))] < 0)
                    return false;
                continue;
            }
            for (int i = 0; i < kFacing.n[code]; ++i) {  // subdivided: the children facing the node must be leaves
                const int o = kFacing.oct[code][i];
                if (fine[(size_t)(2 * Q[0] + (o & 1)) + ((size_t)(2 * Q[1] + ((o >> 1) & 1)) << (L + 1)) +
                         ((size_t)(2 * Q[2] + (o >> 2)) << (2 * (L + 1)))] < 0)
                    return false;
            }
        }
        return true;
    }
    for (int code = 0; code < 27; ++code) {
        if (code == 13) continue;
        const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        const int Q[3] = {wrapi(P[0] + d[0], n), wrapi(P[1] + d[1], n), wrapi(P[2] + d[2], n)};
        const Kind k = kind(L, Q[0], Q[1], Q[2]);
        if (k == LEAF) continue;
        if (k == COARSER) {  // the leaf covering Q must be one level up, not coarser
            if (L == 0 || kind(L - 1, Q[0] >> 1, Q[1] >> 1, Q[2] >> 1) != LEAF) return false;
            continue;
        }
        // subdivided: the children of Q that face the node must be leaves
        for (int o = 0; o < 8; ++o) {
            const int off[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
            bool faces = true;
            for (int a = 0; a < 3; ++a) {
                if (d[a] < 0 && off[a] != 1) faces = false;  // child on the max side faces the node
                if (d[a] > 0 && off[a] != 0) faces = false;
            }
            if (faces && kind(L + 1, 2 * Q[0] + off[0], 2 * Q[1] + off[1], 2 * Q[2] + off[2]) != LEAF) return false;
        }
    }
    return true;
}

// leaves of the subtree at (L, P) that touch the side facing -d and are at
// level >= min_level (touching_children's walk, filtered)
static void touching_deep(const Octree& t, int L, const int P[3], const int d[3], int min_level, std::vector<int>& out) {
    int slot;
    const Octree::Kind k = t.kind(L, P[0], P[1], P[2], &slot);
    if (k == Octree::LEAF) {
        if (L >= min_level) out.push_back(slot);
        return;
    }
    if (k != Octree::FINER) return;
    for (int o = 0; o < 8; ++o) {
        const int off[3] = {o & 1, (o >> 1) & 1, (o >> 2) & 1};
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            if (d[a] < 0 && off[a] != 1) ok = false;
            if (d[a] > 0 && off[a] != 0) ok = false;
        }
        if (ok) {
            const int C[3] = {2 * P[0] + off[0], 2 * P[1] + off[1], 2 * P[2] + off[2]};
            touching_deep(t, L + 1, C, d, min_level, out);
        }
    }
}

// The neighbours of leaf v (periodic images included) whose level differs by
// more than one from L + 1, L = v's level: the chain reaction's targets when v
// is refined.  A same-level leaf differs by one; a coarser leaf (level <= L - 1)
// always qualifies; inside a finer region only leaves of level >= L + 3 do.
void Octree::chain_targets(int v, std::vector<int>& out) const {
    const int L = leaves[v].level;
    if (L < dense_levels) {  // the same walk on the dense maps of levels L and L - 1
        const int m = (1 << L) - 1;
        const int32_t* map = dense[L].data();
        const BlockRec& b = leaves[v];
        for (int code = 0; code < 27; ++code) {
            if (code == 13) continue;
            const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
            const int P[3] = {(b.index[0] + d[0]) & m, (b.index[1] + d[1]) & m, (b.index[2] + d[2]) & m};
            const int32_t r = map[(size_t)P[0] + ((size_t)P[1] << L) + ((size_t)P[2] << (2 * L))];
            if (r >= 0) continue;  // a same-level leaf
            if (r == -2) {         // subdivided
                if (L + 3 <= top_level) touching_deep(*this, L, P, d, L + 3, out);
                continue;
            }
            const int32_t r1 = dense[L - 1][(size_t)(P[0] >> 1) + ((size_t)(P[1] >> 1) << (L - 1)) +
                                            ((size_t)(P[2] >> 1) << (2 * (L - 1)))];
            if (r1 >= 0) {
                out.push_back(r1);
                continue;
            }
            int slot;
            for (int l = L - 2; l >= 0; --l) {
                const int sh = L - l;
                if (kind(l, P[0] >> sh, P[1] >> sh, P[2] >> sh, &slot) == LEAF) {
                    out.push_back(slot);
                    break;
                }
            }
        }
        return;
    }
    for (int code = 0; code < 27; ++code) {
        if (code == 13) continue;
        const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
        int P[3], slot;
        bool wrapped;
        const Kind k = neighbour_region(v, code, P, wrapped, &slot);
        if (k == LEAF) continue;
        if (k == FINER) {
            if (L + 3 <= top_level) touching_deep(*this, L, P, d, L + 3, out);  // no leaf is finer than top_level
            continue;
        }
        for (int l = L - 1; l >= 0; --l) {
            const int sh = L - l;
            if (kind(l, P[0] >> sh, P[1] >> sh, P[2] >> sh, &slot) == LEAF) {
                out.push_back(slot);
                break;
            }
        }
    }
}

// space extension: a leaf with a finer non-periodic-image neighbour is refined
    // (MRAGRefiner_SpaceExtension.h:35-51).  Such a neighbour lies in a
    // subdivided region of the leaf's level next to it without the wrap
    // (has_finer_direct_neighbour), so the leaves to flag are the same-level
    // leaves next to each interior node, found from the interior nodes.
// Levels whose dense map is small are scanned for their subdivided nodes,
// the others visited from the listed interior nodes of the deeper levels.
std::vector<char> Octree::space_extension() const {
    std::vector<char> flag(leaves.size(), 0);
    auto extend = [&](int L, const int I[3]) {
        const int n = 1 << L;
        for (int code = 0; code < 27; ++code) {
            if (code == 13) continue;
            const int d[3] = {code % 3 - 1, (code / 3) % 3 - 1, code / 9 - 1};
            const int P[3] = {I[0] + d[0], I[1] + d[1], I[2] + d[2]};
            if (P[0] < 0 || P[0] >= n || P[1] < 0 || P[1] >= n || P[2] < 0 || P[2] >= n) continue;  // wrapped
            int slot;
            if (kind(L, P[0], P[1], P[2], &slot) == LEAF) flag[slot] = 1;
        }
    };
    // a scanned level as rows of x bits: the subdivided nodes, widened by one
    // along x, then along y and z (no wrap), meet the rows of leaves
    const int nscan = std::min(dense_levels, SCAN_MAX + 1);
    std::vector<uint64_t> sub, leaf;
    for (int L = 1; L < nscan && L < top_level; ++L) {  // (1 node on level 0; none subdivided on the finest)
        const int n = 1 << L;
        const uint64_t full = n == 64 ? ~0ull : (1ull << n) - 1;
        const int32_t* map = dense[L].data();
        sub.assign((size_t)n * n, 0);
        leaf.assign((size_t)n * n, 0);
        for (size_t r = 0; r < (size_t)n * n; ++r) {
            uint64_t sb = 0, lb = 0;
            for (int x = 0; x < n; ++x) {
                const int32_t v = map[(r << L) + x];
                sb |= (uint64_t)(v == -2) << x;
                lb |= (uint64_t)(v >= 0) << x;
            }
            sub[r] = (sb | sb << 1 | sb >> 1) & full;
            leaf[r] = lb;
        }
        for (int z = 0; z < n; ++z)
            for (int y = 0; y < n; ++y) {
                uint64_t near = 0;
                for (int zz = std::max(z - 1, 0); zz <= std::min(z + 1, n - 1); ++zz)
                    for (int yy = std::max(y - 1, 0); yy <= std::min(y + 1, n - 1); ++yy)
                        near |= sub[(size_t)yy + ((size_t)zz << L)];
                const size_t r = (size_t)y + ((size_t)z << L);
                for (uint64_t m = near & leaf[r]; m; m &= m - 1) flag[map[(r << L) + __builtin_ctzll(m)]] = 1;
            }
    }
    for (uint64_t k : interior_deep) {  // the levels not scanned (interior_deep holds levels >= SCAN_MAX + 1)
        const int L = (int)(k >> 48);
        if (L < nscan) continue;
        const int I[3] = {(int)((k >> 32) & 0xffff), (int)((k >> 16) & 0xffff), (int)(k & 0xffff)};
        extend(L, I);
    }
    return flag;
}

std::vector<int> Octree::refinement_victims(const std::vector<char>& flagged, const std::vector<char>* ext) const {
    const int nl = (int)leaves.size();
    std::vector<char> flag = ext ? *ext : space_extension();
    flag.resize(nl, 0);
    for (int s = 0; s < nl && s < (int)flagged.size(); ++s) flag[s] |= flagged[s];
    // nothing at the maximum level is refined (:54-63)
    for (int s = 0; s < nl; ++s)
        if (leaves[s].level >= max_level) flag[s] = 0;
    // chain reaction: neighbours that would end up more than one level coarser
    // than a refined block are refined too (MRAGRefiner.h:65-152)
    // (the targets of a block depend on the tree only, so the closure does not
    // depend on the order the front is processed in)
    std::vector<char> victim(nl, 0);
    std::vector<int> front, next;
    for (int s = 0; s < nl; ++s)
        if (flag[s]) {
            victim[s] = 1;
            front.push_back(s);
        }
    // a leaf of level L (1 <= L <= 6, L + 3 > top_level: no leaf 3 levels
    // finer) has targets only where a cell of level L next to it lies inside
    // a coarser leaf (dense map -1): rows of x bits of such cells answer that
    // for most leaves without the walk
    std::vector<uint64_t> coarse[7];
    auto no_targets = [&](int v) {
        const BlockRec& b = leaves[v];
        const int L = b.level;
        if (L < 1 || L > 6 || L >= dense_levels || L + 3 <= top_level) return false;
        const int n = 1 << L, m = n - 1;
        std::vector<uint64_t>& rows = coarse[L];
        if (rows.empty()) {
            rows.assign((size_t)n * n, 0);
            const int32_t* map = dense[L].data();
            for (size_t r = 0; r < (size_t)n * n; ++r) {
                uint64_t bits = 0;
                for (int x = 0; x < n; ++x) bits |= (uint64_t)(map[(r << L) + x] == -1) << x;
                rows[r] = bits;
            }
        }
        const uint64_t win = 1ull << ((b.index[0] - 1) & m) | 1ull << b.index[0] | 1ull << ((b.index[0] + 1) & m);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                if (rows[(size_t)((b.index[1] + dy) & m) + ((size_t)((b.index[2] + dz) & m) << L)] & win) return false;
        return true;
    };
    std::vector<int> tg;
    while (!front.empty()) {
        next.clear();
        for (int v : front) {
            if (no_targets(v)) continue;
            tg.clear();
            chain_targets(v, tg);
            for (int n : tg)
                if (!victim[n]) {
                    victim[n] = 1;
                    next.push_back(n);
                }
        }
        front.swap(next);
    }
    std::vector<int> out;
    for (int s = 0; s < nl; ++s)
        if (victim[s]) out.push_back(s);
    return out;
}

// The neighbourhood test of outside_within_one for a parent P at level L,
// restated on the cells of level l = L + 1: the 56 cells around P's children
// (the 4x4x4 box from 2P - 1, periodic, less the children themselves) are the
// children of P's 26 neighbours that face P, and the test fails exactly when
// one of them is subdivided or lies in a leaf coarser than L - 1.  With
// cov(c) = level of the leaf covering c, -1 when c is subdivided, a cell c of
// level l is bad when cov(c) < l - 2; P's own children are leaves, never bad,
// so P passes iff its whole box holds no bad cell.  The bad cells of each
// level 2..6 are kept as rows of x bits, 16 row tests per parent.
// bit k of the result = bits 4k .. 4k + 3 all equal to bit k of v (v < 2^16)
static inline uint64_t spread4(uint64_t v) {
    auto part1by1 = [](uint64_t x) {  // bit k -> bit 2k (x < 2^32)
        x &= 0x00000000ffffffffull;
        x = (x | (x << 16)) & 0x0000ffff0000ffffull;
        x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
        x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
        x = (x | (x << 2)) & 0x3333333333333333ull;
        return (x | (x << 1)) & 0x5555555555555555ull;
    };
    uint64_t t = part1by1(v);
    t |= t << 1;
    t = part1by1(t);
    return t | t << 1;
}

std::vector<Octree::CollapseCand> Octree::collapse_candidates() const {
    std::vector<CollapseCand> out;
    const int lrow = std::min(dense_levels - 1, 6);
    std::vector<uint64_t> bad_rows[7];
    if (row_levels == lrow + 1 && lrow >= 2) {
        // from the index's rows: a cell of level l is bad when it is interior
        // or its ancestor two levels up is neither a leaf nor interior (it
        // lies in a leaf coarser than l - 2)
        for (int l = 2; l <= lrow; ++l) {
            const int n = 1 << l;
            const uint64_t full = n == 64 ? ~0ull : (1ull << n) - 1;
            bad_rows[l].resize((size_t)n * n);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    const size_t up = (size_t)(y >> 2) + ((size_t)(z >> 2) << (l - 2));
                    const uint64_t present = leaf_rows[l - 2][up] | node_rows[l - 2][up];
                    const size_t r = (size_t)y + ((size_t)z << l);
                    bad_rows[l][r] = node_rows[l][r] | (~spread4(present) & full);
                }
            // widened to the box [x - 1, x + 2] x [y - 1, y + 2] x [z - 1, z + 2]
            // (periodic): a parent P of level l - 1 passes iff the bit of
            // (2P) is clear
            auto rot = [&](uint64_t v, int k) {  // rotate the n-bit row left by k (mod n)
                k = ((k % n) + n) % n;
                return k == 0 ? v : ((v << k) | (v >> (n - k))) & full;
            };
            std::vector<uint64_t> wx((size_t)n * n), wy((size_t)n * n);
            for (size_t r = 0; r < wx.size(); ++r) {
                const uint64_t v = bad_rows[l][r];
                wx[r] = v | rot(v, 1) | rot(v, -1) | rot(v, -2);
            }
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    uint64_t v = 0;
                    for (int d = -1; d <= 2; ++d) v |= wx[(size_t)((y + d) & (n - 1)) + ((size_t)z << l)];
                    wy[(size_t)y + ((size_t)z << l)] = v;
                }
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    uint64_t v = 0;
                    for (int d = -1; d <= 2; ++d) v |= wy[(size_t)y + ((size_t)((z + d) & (n - 1)) << l)];
                    bad_rows[l][(size_t)y + ((size_t)z << l)] = v;
                }
        }
        // parents of levels 1 .. lrow - 1 whose 8 children are leaves, from
        // the leaf rows of the level below; the root as before
        if (min_level <= 0) {
            CollapseCand c{key(0, 0, 0, 0), {}};
            bool ok = true;
            for (int o = 0; o < 8 && ok; ++o) ok = (c.kids[o] = dense[1][o]) >= 0;
            const int P[3] = {0, 0, 0};
            if (ok && outside_within_one(0, P)) out.push_back(c);
        }
        for (int L = std::max(1, min_level); L < lrow; ++L) {
            const int n = 1 << L, l = L + 1;
            const size_t dz = (size_t)1 << l;
            const uint64_t* rows = leaf_rows[l].data();
            const uint64_t* bad = bad_rows[l].data();
            const int32_t* map = dense[l].data();
            for (int pz = 0; pz < n; ++pz)
                for (int py = 0; py < n; ++py) {
                    const size_t r0 = (size_t)(2 * py) + (size_t)(2 * pz) * dz;
                    uint64_t t = rows[r0] & rows[r0 + 1] & rows[r0 + dz] & rows[r0 + dz + 1];
                    for (uint64_t c8 = even_bits(t & (t >> 1) & ~bad[r0]); c8; c8 &= c8 - 1) {
                        const int px = __builtin_ctzll(c8);
                        CollapseCand c{key(L, px, py, pz), {}};
                        for (int o = 0; o < 8; ++o)
                            c.kids[o] = map[(size_t)(2 * px + (o & 1)) + ((size_t)(2 * py + ((o >> 1) & 1)) << l) +
                                            ((size_t)(2 * pz + ((o >> 2) & 1)) << (2 * l))];
                        out.push_back(c);
                    }
                }
        }
        return out;
    }
    if (lrow >= 2) {
        std::vector<int8_t> prev(1, dense[0][0] >= 0 ? 0 : -1), cov;
        for (int l = 1; l <= lrow; ++l) {
            const int n = 1 << l;
            const int32_t* map = dense[l].data();
            cov.resize((size_t)n * n * n);
            if (l >= 2) bad_rows[l].assign((size_t)n * n, 0);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    const size_t row = (size_t)y + ((size_t)z << l);
                    const int8_t* up = prev.data() + ((size_t)(y >> 1) << (l - 1)) + ((size_t)(z >> 1) << (2 * (l - 1)));
                    uint64_t bits = 0;
                    for (int x = 0; x < n; ++x) {
                        const int32_t v = map[(row << l) + x];
                        const int8_t cv = v >= 0 ? (int8_t)l : v == -2 ? (int8_t)-1 : up[x >> 1];
                        cov[(row << l) + x] = cv;
                        bits |= (uint64_t)(cv < l - 2) << x;
                    }
                    if (l >= 2) bad_rows[l][row] = bits;
                }
            prev.swap(cov);
        }
    }
    // interior nodes whose 8 children are leaves -- each found once, from its
    // child of octant 0 (a leaf with even coordinates)
    for (const BlockRec& b : leaves) {
        if (b.level == 0 || ((b.index[0] | b.index[1] | b.index[2]) & 1) != 0) continue;
        const int L = b.level - 1;
        if (L < min_level) continue;
        const int px = b.index[0] >> 1, py = b.index[1] >> 1, pz = b.index[2] >> 1;
        CollapseCand c{key(L, px, py, pz), {}};
        bool ok = true;
        if (L + 1 < dense_levels) {  // the children straight from their level's map
            const int32_t* map = dense[L + 1].data();
            for (int o = 0; o < 8 && ok; ++o) {
                c.kids[o] = map[(size_t)(2 * px + (o & 1)) + ((size_t)(2 * py + ((o >> 1) & 1)) << (L + 1)) +
                                ((size_t)(2 * pz + ((o >> 2) & 1)) << (2 * (L + 1)))];
                ok = c.kids[o] >= 0;
            }
        } else {
            for (int o = 0; o < 8 && ok; ++o)
                ok = kind(L + 1, 2 * px + (o & 1), 2 * py + ((o >> 1) & 1), 2 * pz + ((o >> 2) & 1), &c.kids[o]) ==
                     LEAF;
        }
        if (!ok) continue;
        // every neighbour of every child within one level of the parent (MRAGCompressor.h:95-107)
        if (L >= 1 && L + 1 <= lrow) {
            const int l = L + 1, m = (1 << l) - 1;
            uint64_t win = 0;
            for (int k = 0; k < 4; ++k) win |= 1ull << ((2 * px - 1 + k) & m);
            const uint64_t* rows = bad_rows[l].data();
            for (int kz = 0; kz < 4 && ok; ++kz)
                for (int ky = 0; ky < 4 && ok; ++ky)
                    ok = (rows[(size_t)((2 * py - 1 + ky) & m) + ((size_t)((2 * pz - 1 + kz) & m) << l)] & win) == 0;
        } else {
            const int P[3] = {px, py, pz};
            ok = outside_within_one(L, P);
        }
        if (ok) out.push_back(c);
    }
    return out;
}

std::vector<Octree::CollapseCand> Octree::collapse_plan(const std::vector<char>& flagged,
                                                        const std::vector<CollapseCand>* cands) const {
    std::vector<CollapseCand> own;
    if (!cands) {
        own = collapse_candidates();
        cands = &own;
    }
    std::vector<CollapseCand> out;
    for (const CollapseCand& c : *cands) {
        bool ok = true;
        for (int o = 0; o < 8 && ok; ++o) ok = c.kids[o] < (int)flagged.size() && flagged[c.kids[o]];
        if (ok) out.push_back(c);
    }
    std::sort(out.begin(), out.end(), [](const CollapseCand& a, const CollapseCand& b) { return a.key < b.key; });
    return out;
}

std::vector<uint64_t> Octree::collapse_parents(const std::vector<char>& flagged,
                                               const std::vector<CollapseCand>* cands) const {
    std::vector<CollapseCand> plan = collapse_plan(flagged, cands);
    std::vector<uint64_t> out(plan.size());
    for (size_t i = 0; i < plan.size(); ++i) out[i] = plan[i].key;
    return out;
}

// octree from leaf rows (level, ix, iy, iz) given in canonical order
static int tree_from_rows(int nleaves, const int* rows, Octree& t) {
    if (nleaves <= 0 || !rows) return set_error(GH_EINVAL, "plan: no leaves");
    t.leaves.clear();
    for (int i = 0; i < nleaves; ++i) {
        const int L = rows[4 * i];
        if (L < 0 || L > 9) return set_error(GH_EINVAL, "plan: level out of range");
        BlockRec b{L, {rows[4 * i + 1], rows[4 * i + 2], rows[4 * i + 3]}};
        for (int a = 0; a < 3; ++a)
            if (b.index[a] < 0 || b.index[a] >= (1 << L)) return set_error(GH_EINVAL, "plan: index out of range");
        t.leaves.push_back(b);
    }
    std::vector<BlockRec> given = t.leaves;
    t.rebuild_index();
    if (t.nested) return set_error(GH_EINVAL, "plan: a leaf lies inside another leaf");
    for (int i = 0; i < nleaves; ++i)
        if (Octree::key(given[i]) != Octree::key(t.leaves[i]))
            return set_error(GH_EINVAL, "plan: leaves must be in canonical (level, z, y, x) order");
    return GH_OK;
}

}  // namespace gh

using namespace gh;

extern "C" int gh_plan_neighbours(int nleaves, const int* leaves, int leaf, int* nnbr, int* nbr) {
    Octree t;
    int rc = tree_from_rows(nleaves, leaves, t);
    if (rc) return rc;
    if (leaf < 0 || leaf >= nleaves || !nnbr || !nbr) return set_error(GH_EINVAL, "gh_plan_neighbours: bad argument");
    std::vector<Octree::Nbr> v = t.neighbours(leaf);
    *nnbr = (int)v.size();
    for (size_t i = 0; i < v.size() && i < 128; ++i) {
        nbr[2 * i] = v[i].slot;
        nbr[2 * i + 1] = v[i].wrapped;
    }
    return GH_OK;
}

extern "C" int gh_plan_refine(int nleaves, const int* leaves, const unsigned char* flagged, int max_level,
                              int* nvictims, int* victims) {
    Octree t;
    int rc = tree_from_rows(nleaves, leaves, t);
    if (rc) return rc;
    if (!flagged || !nvictims || !victims) return set_error(GH_EINVAL, "gh_plan_refine: null argument");
    t.max_level = max_level;
    std::vector<char> f(flagged, flagged + nleaves);
    std::vector<int> v = t.refinement_victims(f);
    *nvictims = (int)v.size();
    std::copy(v.begin(), v.end(), victims);
    return GH_OK;
}

extern "C" int gh_plan_compress(int nleaves, const int* leaves, const unsigned char* flagged, int min_level,
                                int* nparents, int* parents) {
    Octree t;
    int rc = tree_from_rows(nleaves, leaves, t);
    if (rc) return rc;
    if (!flagged || !nparents || !parents) return set_error(GH_EINVAL, "gh_plan_compress: null argument");
    t.min_level = min_level;
    std::vector<char> f(flagged, flagged + nleaves);
    std::vector<uint64_t> p = t.collapse_parents(f);
    *nparents = (int)p.size();
    for (size_t i = 0; i < p.size(); ++i) {
        parents[4 * i] = (int)(p[i] >> 48);
        parents[4 * i + 1] = (int)((p[i] >> 32) & 0xffff);
        parents[4 * i + 2] = (int)((p[i] >> 16) & 0xffff);
        parents[4 * i + 3] = (int)(p[i] & 0xffff);
    }
    return GH_OK;
}
