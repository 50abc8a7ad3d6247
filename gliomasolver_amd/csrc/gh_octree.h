// gh_octree.h -- host-side block octree of the wavelet-adaptive grid:
// topology, neighbourhood, refinement/compression plans and the ghost
// instruction tables consumed by the HIP kernels.
//
// Reference: MRAG/MRAGcore/MRAGrid.inl (hierarchy, periodic neighbourhood
// :840-1149, refine :247-355, compress :383-486), MRAG/MRAGcore/MRAGRefiner.h
// (:43-197), MRAG/MRAGscience/MRAGRefiner_SpaceExtension.h (:25-67),
// MRAG/MRAGcore/MRAGCompressor.h (:67-146), MRAG/MRAGscience/MRAGScienceCore.h
// (:42-177), MRAG/MRAGcore/MRAG_BBInfoCreator.inl (ghost instructions).
#pragma once
#include <cstdint>
#include <map>
#include <set>
#include <unordered_map>
#include <vector>

#include "gh_internal.h"

namespace gh {

// one contribution to a ghost value: value(src cell) * w
struct GhostTerm {
    int32_t src;  // slot * 512 + cell
    double w;
};

// a ghost whose value needs interpolation (neighbour not at the same level)
struct GhostEntry {
    int32_t dst;    // index into the output buffer
    int32_t start;  // first term
    int32_t count;  // number of terms
};

struct Octree {
    static constexpr int MAXL = 10;
    int order = 4;         // interpolating wavelet order: 2 or 4 (MRAGWavelets_Interp{2nd,4th}Order)
    int max_level = 4;     // Refiner_SpaceExtension maxLevel
    int min_level = 0;     // compression floor (0 = MRAG: no floor)
    std::vector<BlockRec> leaves;  // slot order; the tree is these leaves (interior nodes = their ancestors)
    // built by rebuild_index: keys of the interior nodes of levels >= SCAN_MAX + 1
    // (ascending), which the plans visit instead of scanning those levels' maps,
    // and whether the leaves nest (a leaf inside another, or a duplicate)
    static constexpr int SCAN_MAX = 5;  // levels whose maps (8^5 = 32,768 entries) the plans scan
    std::vector<uint64_t> interior_deep;
    bool nested = false;
    // for trees up to level 6 (row_levels = finest level + 1, else 0): per
    // level, rows of x bits (row y + z * 2^l) of its leaves and of its
    // interior nodes, which the compression candidates are read from
    int row_levels = 0;
    std::vector<uint64_t> leaf_rows[7], node_rows[7];
    // dense per-level form of the two maps for levels < dense_levels =
    // min(finest leaf level, DENSE_MAX) + 1 (built by rebuild_index): slot of a
    // leaf, -2 for a node with children, -1 for a point inside a coarser leaf
    static constexpr int DENSE_MAX = 7;
    int dense_levels = 0;
    int top_level = -1;  // finest leaf level (-1: no leaves)
    std::vector<int32_t> dense[MAXL];
    // the nodes of levels dense_levels..top_level (a grid finer than
    // DENSE_MAX) in an open-addressing table, for kind() and the device: key(node) at
    // slot node_hash(key) or the first free slot after it (linear probing),
    // value as in dense; empty slots hold kEmptyKey; 2^hash_bits slots, at
    // most half of them used
    static constexpr uint64_t kEmptyKey = ~0ull;
    int hash_bits = 0;
    std::vector<uint64_t> hash_keys;
    std::vector<int32_t> hash_vals;
    static uint32_t node_hash(uint64_t k, int bits) {
        return (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - bits));
    }

    static uint64_t key(int L, int x, int y, int z) {
        return ((uint64_t)L << 48) | ((uint64_t)(uint32_t)x << 32) | ((uint64_t)(uint32_t)y << 16) | (uint64_t)(uint32_t)z;
    }
    static uint64_t key(const BlockRec& b) { return key(b.level, b.index[0], b.index[1], b.index[2]); }

    enum Kind { LEAF = 0, FINER = 1, COARSER = 2 };
    // what covers block position (L, x, y, z) (wrapped periodically): a leaf, a
    // subdivided node, or nothing (a coarser leaf covers it)
    Kind kind(int L, int x, int y, int z, int* slot = nullptr) const;

    // a copy of the configuration without leaves or index: for building a
    // new tree that rebuild_index then indexes
    Octree shell() const {
        Octree o;
        o.order = order;
        o.max_level = max_level;
        o.min_level = min_level;
        return o;
    }

    void init_uniform(int level);
    void rebuild_index();  // canonical slot order (level, z, y, x) + maps; call after any change of `leaves`
    int levels_min() const;  // (leaves in canonical order: after rebuild_index)
    int levels_max() const;
    bool is_uniform() const { return levels_min() == levels_max(); }

    // neighbourhood of a leaf: leaves whose finest-level boxes touch its box,
    // with periodic images (MRAGrid.inl:840-1149, MRAGridHelpers.h:77-84).
    // wrapped = reached only through the periodic boundary (a "ghost node").
    struct Nbr {
        int slot;
        bool wrapped;
    };
    std::vector<Nbr> neighbours(int slot) const;
    Kind neighbour_region(int slot, int code, int P[3], bool& wrapped, int* region_slot = nullptr) const;
    bool has_finer_direct_neighbour(int slot) const;
    // every leaf touching node (L, P)'s box from outside (periodic images
    // included) is at level L - 1, L or L + 1: the same as every neighbour of
    // each of the node's eight children being within one level of L, since a
    // child's neighbours are its siblings (level L + 1) and leaves touching
    // the node's box, and every leaf touching the box touches a child
    bool outside_within_one(int L, const int P[3]) const;
    void chain_targets(int slot, std::vector<int>& out) const;
    // same-level neighbour slot in each of the 27 directions (-1 if none)
    void same_level_neighbours(int slot, int out[27]) const;

    // ghost value at level-L lattice point q as a linear combination of leaf
    // cell values: same-level leaf -> the cell; finer region -> injection
    // (Ha = {1}); coarser region -> Hs synthesis from level L-1.
    void value_terms(int L, const int64_t q[3], double w, std::vector<GhostTerm>& out) const;

    // --- plans (return the new leaf list; data movement is done by the caller)
    // Science::AutomaticRefinement's refine(set): space extension + 2:1 chain
    // reaction.  The space extension depends on the tree only: ext (may be
    // null) is space_extension() of this tree, computed beforehand.
    std::vector<char> space_extension() const;
    std::vector<int> refinement_victims(const std::vector<char>& flagged, const std::vector<char>* ext = nullptr) const;
    // Compressor::createPlan: parents whose 8 children are flagged leaves and whose
    // children's neighbours are all within one level of the parent.  The
    // candidates (parents of 8 leaves at or above min_level that pass the
    // level test) depend on the tree only: cands (may be null) is
    // collapse_candidates() of this tree, computed beforehand.
    struct CollapseCand {
        uint64_t key;  // the parent
        int kids[8];   // slots of its children by octant
    };
    std::vector<CollapseCand> collapse_candidates() const;
    // the chosen parents with their children, ascending keys
    std::vector<CollapseCand> collapse_plan(const std::vector<char>& flagged,
                                            const std::vector<CollapseCand>* cands = nullptr) const;
    std::vector<uint64_t> collapse_parents(const std::vector<char>& flagged,
                                           const std::vector<CollapseCand>* cands = nullptr) const;
};

// how a leaf of the new grid is produced from the old one
struct Remap {
    int kind;      // 0 copy, 1 split (child of an old leaf), 2 collapse (8 old leaves)
    int src;       // copy / split: old slot
    int octant;    // split: child position (x | y<<1 | z<<2)
    int kids[8];   // collapse: old slots of the children by octant
};

}  // namespace gh
