// gh_internal.h -- shared host-side declarations of libgliomahip.so.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gliomahip.h"

namespace gh {

// ---- errors: thread-local message, status codes of gliomahip.h -------------
int set_error(int code, const std::string& msg);
const char* last_error();

// ---- seeded RNG shared by the synthetic-input generators --------------------
uint64_t splitmix64(uint64_t& s);
double uniform01(uint64_t& s);

int phantom(int n, uint64_t seed, double* gm, double* wm, double* csf, int* seed_voxel);
int gaussian_ic(int real_kind, int n, const void* p_w, const void* p_g, const double* center, double sigma,
                double peak, void* phi);

constexpr int BS = 8;                 // block size (makefile/tools.make/Makefile:30, bs=8)
constexpr int BCELLS = BS * BS * BS;  // cells per block

// face order of ReactionDiffusionOperator.h:106-111: -x,+x,-y,+y,-z,+z
enum Face { FXM = 0, FXP = 1, FYM = 2, FYP = 3, FZM = 4, FZP = 5 };

// A leaf block (MRAG BlockInfo without the pointer: MRAG/MRAGcore/MRAGCommon.h:166-260).
struct BlockRec {
    int level;
    int index[3];
};

// Leaf-block topology in device slot order.
struct Topology {
    std::vector<BlockRec> blocks;
    // host copies of the device's neighbour tables, filled only when a
    // host-built ghost table needs them (empty otherwise)
    std::vector<int32_t> nbr;    // 6 per block: same-level face neighbour slot, -1 across a level jump
    std::vector<int32_t> nbr27;  // 27 per block: same-level neighbour slot per direction code, -1 if none
    int uniform_bpd = 0;       // >0 when every block sits on one level (bpd^3 blocks)
    int nblocks() const { return (int)blocks.size(); }
};

}  // namespace gh
