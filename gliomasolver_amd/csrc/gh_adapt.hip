// gh_adapt.hip -- wavelet-adaptive grid on the device: ghost fill across
// levels, FWT detail per block, refinement / compression with data movement,
// the RD step on a multi-level grid, and the uniform resample for the
// likelihood.
//
// Reference:
//   ghost fill       MRAG/MRAGcore/MRAGBlockLab.h:154-301 (ghost = sum w*value)
//   FWT details      MRAG/MRAGcore/MRAGBlockFWT.h:109-120, 219-298
//   refine/compress  MRAG/MRAGscience/MRAGScienceCore.h:42-177
//   split            MRAG/MRAGcore/MRAGBlockSplitter.h:24-99
//   collapse         MRAG/MRAGcore/MRAGBlockCollapser.h:24-85
//   RD on levels     Glioma/Operators/ReactionDiffusionOperator.h:47-139 (h per block)
//   UQ resample      Glioma/Glioma_ReactionDiffusion.cpp:208-267
//
// Ghost values: a ghost lattice point at the block's level is a same-level
// copy when a same-level leaf covers it, else a linear combination of leaf
// cells: injection from finer regions (Ha = {1}), Hs synthesis from coarser
// ones, recursively (Octree::value_terms on the host, ghost_visit on the
// device).  Terms are accumulated in Real, product by product, in recursion
// order (synthesis taps z-major, then y, then x); the weights are dyadic,
// exact in float and double.  The labs of the FWT and of the split, the
// neighbour tables and the RD face-ghost table of a multi-level grid are built
// on the device from per-level block maps (host-built tables only for ghosts
// deeper than the device recursion); the face ghosts are gathered every step.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>
#include <vector>

#include "gh_ctx.h"
#include "gh_rd_common.h"

#pragma clang fp contract(off)

namespace gh {
namespace {

// adds the wall time of its scope to one profile slot
struct ProfScope {
    double& slot;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit ProfScope(double& s) : slot(s) {}
    ~ProfScope() { slot += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};
enum ProfSlot { P_REF_FWT = 0, P_REF = 1, P_COMP_FWT = 2, P_COMP = 3, P_PLAN = 4, P_MOVE = 5, P_TOPO = 6, P_STAT = 7 };

// lab geometry: positions [LS, LS+LAB) per axis (MRAGrid.inl:761-789 with the
// RD stencil [-1,2)): order 4 -> [-2, 11), order 2 -> [-1, 9)
inline int lab_start(int order) { return order == 4 ? -2 : -1; }
inline int lab_size(int order) { return order == 4 ? 13 : 10; }

__device__ inline int dir_code(int ix, int iy, int iz) {
    const int dx = ix < 0 ? 0 : (ix < 8 ? 1 : 2), dy = iy < 0 ? 0 : (iy < 8 ? 1 : 2), dz = iz < 0 ? 0 : (iz < 8 ? 1 : 2);
    return dx + 3 * dy + 9 * dz;
}

// interior and same-level ghosts of the lab of each listed slot
template <typename Real>
__global__ __launch_bounds__(256) void lab_fill_kernel(int nlist, const int32_t* __restrict__ slots, int LS, int LAB,
                                                       const int32_t* __restrict__ nbr27, const Real* __restrict__ f,
                                                       Real* __restrict__ lab) {
    const int j = blockIdx.x;
    if (j >= nlist) return;
    const int s = slots[j];
    const int lab3 = LAB * LAB * LAB;
    for (int p = threadIdx.x; p < lab3; p += blockDim.x) {
        const int ix = p % LAB + LS, iy = (p / LAB) % LAB + LS, iz = p / (LAB * LAB) + LS;
        const int code = dir_code(ix, iy, iz);
        const int src = code == 13 ? s : nbr27[27 * s + code];
        if (src < 0) continue;  // interpolated ghost: written by ghost_entry_kernel
        const int cx = (ix + 8) & 7, cy = (iy + 8) & 7, cz = (iz + 8) & 7;
        lab[(size_t)j * lab3 + p] = f[(size_t)src * BCELLS + cx + 8 * cy + 64 * cz];
    }
}

// out[dst] = sum of f[src] * w over the entry's terms, in table order
template <typename Real>
__global__ __launch_bounds__(256) void ghost_entry_kernel(int n, const GhostEntry* __restrict__ ent,
                                                          const int32_t* __restrict__ tsrc,
                                                          const Real* __restrict__ tw, const Real* __restrict__ f,
                                                          Real* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GhostEntry e = ent[i];
    Real acc = (Real)0;
    for (int k = 0; k < e.count; ++k) {
        const Real prod = f[tsrc[e.start + k]] * tw[e.start + k];
        acc = acc + prod;
    }
    out[e.dst] = acc;
}

// Per-level block maps (Octree::dense on the device): map[dense_level_offset(l)
// + x + (y << l) + (z << 2l)] = slot of a leaf at (l, x, y, z), -2 for a node with
// children, -1 for a point inside a coarser leaf, for levels < nlev; levels
// nlev..top in the open-addressing table hkey/hval (Octree::hash_keys, 2^hbits
// slots, linear probing, absent = -1); levels > top hold no node.
struct LevelMaps {
    const int32_t* __restrict__ map;
    int nlev;
    int top;
    const uint64_t* __restrict__ hkey;
    const int32_t* __restrict__ hval;
    int hbits;
};

// the maps of levels 0, 1, ... lie one after another: level l starts at
// 8^0 + ... + 8^(l-1) = (8^l - 1) / 7 (l <= 7: no overflow)
__host__ __device__ inline int dense_level_offset(int l) { return (int)(((1u << (3 * l)) - 1u) / 7u); }

// value of block position (L, x, y, z) in the hashed levels (Octree::node_hash;
// the table is at most half full, so the probe reaches a free slot)
__device__ inline int32_t hashed_node(const LevelMaps& m, int L, int x, int y, int z) {
    const uint64_t k = ((uint64_t)L << 48) | ((uint64_t)(uint32_t)x << 32) | ((uint64_t)(uint32_t)y << 16) |
                       (uint64_t)(uint32_t)z;
    const uint32_t mask = (1u << m.hbits) - 1u;
    uint32_t h = (uint32_t)((k * 0x9E3779B97F4A7C15ull) >> (64 - m.hbits));
    for (;;) {
        const uint64_t e = m.hkey[h];
        if (e == k) return m.hval[h];
        if (e == Octree::kEmptyKey) return -1;
        h = (h + 1) & mask;
    }
}

inline LevelMaps level_maps(const Ctx& c) {
    LevelMaps m{};
    m.map = c.dDense.as<int32_t>();
    m.nlev = c.nDense;
    m.top = c.topLevel;
    m.hkey = c.dHashKey.as<uint64_t>();
    m.hval = c.dHashVal.as<int32_t>();
    m.hbits = c.hashBits;
    return m;
}

// Hs synthesis taps of fine point q (MRAGWavelets_StaticData.cpp:13-21,
// 31-39): even q -> coarse q/2 with weight 1; odd q -> (q-1)/2 + k, k < 2,
// weight 1/2 (order 2) or (q-3)/2 + k, k < 4, weights -1/16, 9/16, 9/16, -1/16
__device__ inline int synth_n(int order, int q) { return (q & 1) == 0 ? 1 : (order == 2 ? 2 : 4); }
__device__ inline int synth_src(int order, int q, int k) {
    return (q & 1) == 0 ? q / 2 : (order == 2 ? (q - 1) / 2 + k : (q - 3) / 2 + k);
}
__device__ inline double synth_w(int order, int q, int k) {
    return (q & 1) == 0 ? 1.0 : (order == 2 ? 0.5 : ((k == 0 || k == 3) ? -1.0 / 16 : 9.0 / 16));
}

// Calls visit(src, w) for every term of the value of lattice point q at
// level L times w, in the order of Octree::value_terms (src = slot * 512 +
// cell).  D more synthesis levels are allowed below this one; deeper
// recursion sets bad.
template <int D, typename Visit>
__device__ inline void ghost_visit(const LevelMaps& m, int order, int L, int qx, int qy, int qz, double w,
                                   Visit& visit, int& bad) {
    for (;;) {
        const int mask = (BS << L) - 1;  // periodic wrap, power-of-two lattice
        qx &= mask;
        qy &= mask;
        qz &= mask;
        if (L > m.top) break;  // inside a coarser leaf
        const int bx = (int)(qx >> 3), by = (int)(qy >> 3), bz = (int)(qz >> 3);
        const int32_t v = L < m.nlev ? m.map[dense_level_offset(L) + bx + (by << L) + (bz << (2 * L))] : hashed_node(m, L, bx, by, bz);
        if (v >= 0) {
            visit((int32_t)(v * BCELLS + (qx & 7) + 8 * (qy & 7) + 64 * (qz & 7)), w);
            return;
        }
        if (v != -2) break;  // coarser
        ++L;                 // finer region: the coincident fine vertex (Ha = {1})
        qx *= 2;
        qy *= 2;
        qz *= 2;
    }
    if constexpr (D == 0) {
        bad = 1;
    } else {
        const int nz = synth_n(order, qz), ny = synth_n(order, qy), nx = synth_n(order, qx);
#pragma unroll 1
        for (int kz = 0; kz < nz; ++kz)
#pragma unroll 1
            for (int ky = 0; ky < ny; ++ky)
#pragma unroll 1
                for (int kx = 0; kx < nx; ++kx)
                    ghost_visit<D - 1>(m, order, L - 1, synth_src(order, qx, kx), synth_src(order, qy, ky),
                                       synth_src(order, qz, kz),
                                       w * (synth_w(order, qz, kz) * synth_w(order, qy, ky) * synth_w(order, qx, kx)),
                                       visit, bad);
    }
}

// the ghost of face point (u, v) of face f of block b (level, x, y, z)
__device__ inline void face_point(int f, int u, int v, int i[3]) {
    switch (f) {
        case FXM: i[0] = -1; i[1] = u; i[2] = v; break;
        case FXP: i[0] = 8; i[1] = u; i[2] = v; break;
        case FYM: i[0] = u; i[1] = -1; i[2] = v; break;
        case FYP: i[0] = u; i[1] = 8; i[2] = v; break;
        case FZM: i[0] = u; i[1] = v; i[2] = -1; break;
        default:  i[0] = u; i[1] = v; i[2] = 8; break;
    }
}

// In-place exclusive prefix sum of n int32 counts on the device, in chunks of
// 1024: chunk totals, one workgroup scanning them (grand total to *total),
// then each chunk's own scan plus its offset.  Exact integer arithmetic.
template <typename T>
__device__ inline T block_exclusive_scan(T v, T* s_w, T& block_total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    if (w == 0) {
        T t = lane < nw ? s_w[lane] : (T)0;
        for (int o = 1; o < 64; o <<= 1) {
            const T y = __shfl_up(t, o, 64);
            if (lane >= o) t += y;
        }
        if (lane < nw) s_w[lane] = t;
    }
    __syncthreads();
    block_total = s_w[nw - 1];
    const T res = (w > 0 ? s_w[w - 1] : (T)0) + x - v;
    __syncthreads();  // s_w is reused by the caller
    return res;
}

__global__ __launch_bounds__(1024) void scan_chunk_sums(int n, const int32_t* __restrict__ cnt,
                                                        long long* __restrict__ sums) {
    __shared__ long long s_w[16];
    const int g = blockIdx.x * 1024 + threadIdx.x;
    long long tot;
    (void)block_exclusive_scan<long long>(g < n ? cnt[g] : 0, s_w, tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void scan_chunk_offsets(int nchunk, const long long* __restrict__ sums,
                                                           long long* __restrict__ offs,
                                                           long long* __restrict__ total) {
    __shared__ long long s_w[16];
    long long carry = 0;
    for (int b0 = 0; b0 < nchunk; b0 += 1024) {
        const int b = b0 + (int)threadIdx.x;
        long long tot;
        const long long e = block_exclusive_scan<long long>(b < nchunk ? sums[b] : 0, s_w, tot);
        if (b < nchunk) offs[b] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(1024) void scan_chunk_apply(int n, int32_t* __restrict__ cnt,
                                                         const long long* __restrict__ offs) {
    __shared__ long long s_w[16];
    const int g = blockIdx.x * 1024 + threadIdx.x;
    long long tot;
    const long long e = block_exclusive_scan<long long>(g < n ? cnt[g] : 0, s_w, tot);
    if (g < n) cnt[g] = (int32_t)(offs[blockIdx.x] + e);  // the caller rejects totals above INT32_MAX
}

// Same-level neighbours of every slot from the level maps, as
// Octree::same_level_neighbours: n27[27 s + k] = slot of the leaf at (L,
// index + d_k) (periodic) or -1; nbr6[6 s + f] = n27 entry of face f's code
// (FXM, FXP, FYM, FYP, FZM, FZP: 12, 14, 10, 16, 4, 22); level[s] = L.
__global__ __launch_bounds__(256) void topology_kernel(int n, const int4* __restrict__ blk, LevelMaps m,
                                                       int32_t* __restrict__ n27, int32_t* __restrict__ nbr6,
                                                       int32_t* __restrict__ level) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;  // n = 27 x blocks
    const int s = i / 27, k = i - 27 * s;
    const int4 b = blk[s];
    const int L = b.x, msk = (1 << L) - 1;
    const int bx = (b.y + k % 3 - 1) & msk, by = (b.z + (k / 3) % 3 - 1) & msk, bz = (b.w + k / 9 - 1) & msk;
    const int32_t v = L < m.nlev ? m.map[dense_level_offset(L) + bx + (by << L) + (bz << (2 * L))]
                                 : hashed_node(m, L, bx, by, bz);
    const int32_t r = v >= 0 ? v : -1;
    n27[i] = r;
    const int f = k == 12 ? 0 : k == 14 ? 1 : k == 10 ? 2 : k == 16 ? 3 : k == 4 ? 4 : k == 22 ? 5 : -1;
    if (f >= 0) nbr6[6 * s + f] = r;
    if (k == 13) level[s] = L;
}

// jump faces (no same-level neighbour): flag[i] = 1 for i < n, flag[n] = 0,
// then scanned in place into list positions
__global__ __launch_bounds__(256) void face_flag_kernel(int n, const int32_t* __restrict__ nbr6,
                                                        int32_t* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) flag[i] = i < n && nbr6[i] < 0 ? 1 : 0;
}

// the jump faces in slot-face order: faces[pos[i]] = i and pk_of[i] = pos[i]
// for a jump face i = 6 * slot + f, pk_of[i] = -1 for the others
__global__ __launch_bounds__(256) void face_list_kernel(int n, const int32_t* __restrict__ nbr6,
                                                        const int32_t* __restrict__ pos,
                                                        int32_t* __restrict__ faces, int32_t* __restrict__ pk_of) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (nbr6[i] < 0) {
        faces[pos[i]] = i;
        pk_of[i] = pos[i];
    } else {
        pk_of[i] = -1;
    }
}

// Face-ghost table on the device, pass 1: term count of ghost g = 64 * k + p
// of jump face k (faces[k] = 6 * slot + f), p = v * 8 + u.
// pkt (may be null): per face, 64 x the largest count of its 64 ghosts, which
// are the 64 lanes of one wave (n is a multiple of 64, blocks of 256 lanes)
__global__ __launch_bounds__(256) void face_count_kernel(int n, const int32_t* __restrict__ faces,
                                                         const int4* __restrict__ blk, LevelMaps m, int order,
                                                         int32_t* __restrict__ count, int32_t* __restrict__ pkt,
                                                         int* __restrict__ bad) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int sf = faces[g >> 6], s = sf / 6, f = sf % 6, p = g & 63;
    const int4 b = blk[s];
    int i[3];
    face_point(f, p & 7, p >> 3, i);
    int cnt = 0, my_bad = 0;
    auto visit = [&](int32_t, double) { ++cnt; };
    ghost_visit<2>(m, order, b.x, 8 * b.y + i[0], 8 * b.z + i[1], 8 * b.w + i[2], 1.0, visit, my_bad);
    count[g] = cnt;
    if (my_bad) atomicOr(bad, 1);
    if (pkt) {
        int mx = cnt;
        for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
        if (p == 0) pkt[g >> 6] = 64 * mx;
    }
}

// pass 2 in packets: the terms of ghost p of face k at base[k] + 64 j + p,
// j < its count (the rest of the packet's column is never read)
template <typename Real>
__global__ __launch_bounds__(256) void face_fill_packet_kernel(int n, const int32_t* __restrict__ faces,
                                                               const int4* __restrict__ blk, LevelMaps m, int order,
                                                               const int32_t* __restrict__ base,
                                                               int32_t* __restrict__ tsrc, Real* __restrict__ tw) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int k = g >> 6, sf = faces[k], s = sf / 6, f = sf % 6, p = g & 63;
    const int4 b = blk[s];
    int i[3];
    face_point(f, p & 7, p >> 3, i);
    const int b0 = base[k] + p;
    int j = 0, my_bad = 0;
    auto visit = [&](int32_t src, double w) {
        tsrc[b0 + 64 * j] = src;
        tw[b0 + 64 * j] = (Real)w;
        ++j;
    };
    ghost_visit<2>(m, order, b.x, 8 * b.y + i[0], 8 * b.z + i[1], 8 * b.w + i[2], 1.0, visit, my_bad);
}

// face ghosts from packets: out[64 * faces[k] + p] = sum of f[src] * w over
// ghost p's count[g] terms in order (the GhostEntry kernel's arithmetic), one
// wave per face, each term a contiguous 64-lane load; the loop bound is known
// up front, so the term loads of several iterations are in flight together
template <typename Real>
__global__ __launch_bounds__(256) void face_packet_kernel(int n, const int32_t* __restrict__ faces,
                                                          const int32_t* __restrict__ base,
                                                          const int32_t* __restrict__ count,
                                                          const int32_t* __restrict__ tsrc,
                                                          const Real* __restrict__ tw, const Real* __restrict__ f,
                                                          Real* __restrict__ out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int k = g >> 6, p = g & 63;
    const int b0 = base[k] + p, len = count[g];
    Real acc = (Real)0;
#pragma unroll 4
    for (int j = 0; j < len; ++j) {
        const Real prod = f[tsrc[b0 + 64 * j]] * tw[b0 + 64 * j];
        acc = acc + prod;
    }
    out[(size_t)faces[k] * 64 + p] = acc;
}

// BlockFWT::_fwt / _filter (MRAGBlockFWT.h:219-298): 64 coarse points x 8
// tensor filters per block; the block's detail is the largest |value| over
// the 7 wavelet codes (FWTReport::_getDetailMaxMag, :109-120; the code-0
// scaling values do not enter it and are not formed).
//
// The analysis taps Ga (MRAGWavelets_StaticData.cpp:14, 33) are compile-time
// constants.  A term whose Ga tap is 0 (order 4 has two) adds a signed zero
// -- the lab holds finite values -- which changes no partial sum except the
// sign of a zero, and so not |result|: those terms are left out (343 -> 125
// for code 7), the others are summed in the reference's order (z, y, x taps
// ascending).  The tap
// products wZ*wY*wX are exact in either precision (powers of two times 1, 9
// or 81, 729), so folding them at compile time gives the runtime values.
template <int ORDER>
struct FwtGeom;
template <>
struct FwtGeom<4> {  // Ga support [-4, 3): 1/16, 0, -9/16, 1, -9/16, 0, 1/16
    static constexpr int LS = -2, LAB = 13, NT = 5;
    static constexpr int tap(int k) { return k == 0 ? -4 : k == 1 ? -2 : k == 2 ? -1 : k == 3 ? 0 : 2; }
    static constexpr double w(int k) { return k == 0 ? 1 / 16. : k == 1 ? -9 / 16. : k == 2 ? 1. : k == 3 ? -9 / 16. : 1 / 16.; }
};
template <>
struct FwtGeom<2> {  // Ga support [-2, 1): -1/2, 1, -1/2
    static constexpr int LS = -1, LAB = 10, NT = 3;
    static constexpr int tap(int k) { return k - 2; }
    static constexpr double w(int k) { return k == 1 ? 1. : -0.5; }
};

// one wavelet code (BX, BY, BZ: the axes filtered with Ga, the others with
// Ha = {1}) at the coarse point whose lab position p points at
template <typename Real, int ORDER, int BX, int BY, int BZ>
__device__ __forceinline__ Real fwt_code(const Real* __restrict__ p) {
    using G = FwtGeom<ORDER>;
    constexpr int LAB = G::LAB;
    constexpr int nz = BZ ? G::NT : 1, ny = BY ? G::NT : 1, nx = BX ? G::NT : 1;
    Real r = (Real)0;
#pragma unroll
    for (int kz = 0; kz < nz; ++kz) {
        const Real wZ = BZ ? (Real)G::w(kz) : (Real)1;
        const int iz = BZ ? G::tap(kz) : 0;
#pragma unroll
        for (int ky = 0; ky < ny; ++ky) {
            const Real wY = BY ? (Real)G::w(ky) : (Real)1;
            const int iy = BY ? G::tap(ky) : 0;
            const Real wZwY = wZ * wY;
#pragma unroll
            for (int kx = 0; kx < nx; ++kx) {
                const Real wX = BX ? (Real)G::w(kx) : (Real)1;
                const int ix = BX ? G::tap(kx) : 0;
                const Real w = wZwY * wX;
                const Real prod = p[-(iz * LAB * LAB + iy * LAB + ix)] * w;
                r = r + prod;
            }
        }
    }
    return r;
}

// Full labs of the listed slots (slots == nullptr: every slot j), ghosts
// evaluated on the device; *bad is set when a ghost needs more than two
// synthesis levels (the caller then uses host tables).
// Lanes per lab: 512 while the labs are fewer than the CUs can hold several
// of, 256 beyond that (more labs in flight per CU hide the dependent map and
// field loads better; at least 256: the FWT epilogue takes four waves).
// Details of 1,114 / 7,225 / 13,455 blocks: 0.12 / 0.41 / 0.54 ms with 512,
// 0.15 / 0.38 / 0.47 ms with 256 lanes.
constexpr int kLabThreadsFewMax = 2048;  // largest list built with 512 lanes per lab
// up to four fields whose labs are built together: the ghost recursion (map
// lookups, terms, weights) runs once and each field accumulates its own sum
// with the single-field arithmetic, so every lab has the bits it has alone
template <typename Real>
struct LabFields {
    const Real* __restrict__ f[4];
    Real* __restrict__ lab[4];
};
// ORDER (2 or 4) fixes the lab geometry at compile time (the cell loops
// divide by the lab size).  FWT (NF = 1): the lab stays in LDS and the kernel
// ends with the block's FWT detail (fwt_code) in detail[j] instead of writing
// the lab out.
template <typename Real, int NF, int ORDER, bool FWT, int LAB_THREADS>
__global__ __launch_bounds__(LAB_THREADS) void lab_kernel(int nlist, const int32_t* __restrict__ slots,
                                                  const int32_t* __restrict__ nbr27,
                                                  const int4* __restrict__ blk, LevelMaps m,
                                                  const int16_t* __restrict__ perm, LabFields<Real> lf,
                                                  double* __restrict__ detail, int* __restrict__ bad) {
    static_assert(!FWT || NF == 1, "details of one field");
    static_assert(ORDER == 2 || ORDER == 4, "wavelet order");
    constexpr int order = ORDER, LS = ORDER == 4 ? -2 : -1, LAB = ORDER == 4 ? 13 : 10;
    const int j = blockIdx.x;
    if (j >= nlist) return;
    __shared__ Real s_lab[FWT ? LAB * LAB * LAB : 1];
    const int s = slots ? slots[j] : j;
    const int4 b = blk[s];
    constexpr int lab3 = LAB * LAB * LAB;
    int my_bad = 0;
    // Coarse-point cache.  A lab ghost inside a coarser leaf is the Hs
    // synthesis of level-(L-1) points; all of them lie in the PC^3 patch of
    // coarse points from 4*idx - PL on (PC = 9, PL = 2 for order 4; 6 and 1 for
    // order 2).  Most such points are a single leaf cell (or the coincident
    // vertex of a finer region); their values are resolved once per lab into
    // LDS, and a ghost's term then reads them there: the same cells, weights
    // and order of accumulation as the recursion, which the points that need
    // a further synthesis (flag 0) still take.
    const int L = b.x;
    constexpr int PC = ORDER == 4 ? 9 : 6, PL = ORDER == 4 ? 2 : 1;
    __shared__ Real s_cv[NF][729];
    __shared__ unsigned char s_cf[729];
    bool jump = false;
    for (int k = 0; k < 27; ++k) jump |= (k != 13 && nbr27[27 * s + k] < 0);
    const bool use_cache = jump && L >= 1;  // uniform per workgroup
    // what each of the 27 level-L regions around the block is: a leaf (its
    // slot), subdivided (-2) or inside a coarser leaf (-1) -- looked up once
    // here instead of for every ghost point
    __shared__ int32_t s_reg[27];
    if (jump && threadIdx.x < 27) {
        const int k = threadIdx.x, msk = (1 << L) - 1;
        const int bx = (b.y + k % 3 - 1) & msk, by = (b.z + (k / 3) % 3 - 1) & msk, bz = (b.w + k / 9 - 1) & msk;
        const int32_t v = L < m.nlev ? m.map[dense_level_offset(L) + bx + (by << L) + (bz << (2 * L))]
                                     : hashed_node(m, L, bx, by, bz);
        s_reg[k] = v >= 0 ? v : (v == -2 ? -2 : -1);
    }
    if (jump && !use_cache) __syncthreads();
    if (use_cache) {
        const int ox = 4 * b.y - PL, oy = 4 * b.z - PL, oz = 4 * b.w - PL;
        for (int q = threadIdx.x; q < PC * PC * PC; q += LAB_THREADS) {
            int Lc = L - 1, qx = ox + q % PC, qy = oy + (q / PC) % PC, qz = oz + q / (PC * PC);
            int32_t src = -1;
            for (;;) {  // ghost_visit's walk without synthesis
                const int mask = (BS << Lc) - 1;
                qx &= mask;
                qy &= mask;
                qz &= mask;
                if (Lc > m.top) break;
                const int bx = qx >> 3, by = qy >> 3, bz = qz >> 3;
                const int32_t v = Lc < m.nlev ? m.map[dense_level_offset(Lc) + bx + (by << Lc) + (bz << (2 * Lc))]
                                              : hashed_node(m, Lc, bx, by, bz);
                if (v >= 0) {
                    src = v * BCELLS + (qx & 7) + 8 * (qy & 7) + 64 * (qz & 7);
                    break;
                }
                if (v != -2) break;
                ++Lc;
                qx *= 2;
                qy *= 2;
                qz *= 2;
            }
            s_cf[q] = src >= 0 ? 1 : 0;
#pragma unroll
            for (int k = 0; k < NF; ++k) s_cv[k][q] = src >= 0 ? lf.f[k][src] : (Real)0;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < lab3; i += LAB_THREADS) {
        // a lab without level jumps is copies only: cells in storage order
        // (coalesced); else in lab_order(), so that waves see uniform work
        const int p = jump ? perm[i] : i;
        const int ix = p % LAB + LS, iy = (p / LAB) % LAB + LS, iz = p / (LAB * LAB) + LS;
        const int code = dir_code(ix, iy, iz);
        const int src = code == 13 ? s : nbr27[27 * s + code];
        Real v[NF];
        if (src >= 0) {
            const int cx = (ix + 8) & 7, cy = (iy + 8) & 7, cz = (iz + 8) & 7;
            const size_t c = (size_t)src * BCELLS + cx + 8 * cy + 64 * cz;
#pragma unroll
            for (int k = 0; k < NF; ++k) v[k] = lf.f[k][c];
        } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) v[k] = (Real)0;
            auto visit = [&](int32_t src, double w) {
                const Real wr = (Real)w;
#pragma unroll
                for (int k = 0; k < NF; ++k) {
                    const Real prod = lf.f[k][src] * wr;
                    v[k] = v[k] + prod;
                }
            };
            // the level-L region of this point: inside a coarser leaf?
            const int mask = (BS << L) - 1;
            const int qx = (8 * b.y + ix) & mask, qy = (8 * b.z + iy) & mask, qz = (8 * b.w + iz) & mask;
            const int32_t reg = s_reg[code];  // src < 0: the block has a level jump
            if (reg == -2) {  // subdivided: ghost_visit<2>'s first step, the coincident fine vertex
                ghost_visit<2>(m, order, L + 1, 2 * qx, 2 * qy, 2 * qz, 1.0, visit, my_bad);
            } else if (!(use_cache && reg == -1)) {
                ghost_visit<2>(m, order, L, qx, qy, qz, 1.0, visit, my_bad);
            } else {  // ghost_visit<2>'s synthesis step, coarse points from the cache
                const int nc = BS << (L - 1);  // coarse lattice size (a power of two: & wraps)
                constexpr int NT = ORDER == 4 ? 4 : 2;  // taps of an odd coordinate
                const int nz = synth_n(order, qz), ny = synth_n(order, qy), nx = synth_n(order, qx);
                // the x taps once per ghost (entries k >= nx unused)
                int sxk[NT], pxk[NT];
                double wxk[NT];
#pragma unroll
                for (int k = 0; k < NT; ++k) {
                    sxk[k] = synth_src(order, qx, k);
                    pxk[k] = (sxk[k] - (4 * b.y - PL)) & (nc - 1);
                    wxk[k] = synth_w(order, qx, k);
                }
#pragma unroll 1
                for (int kz = 0; kz < nz; ++kz) {
                    const int sz = synth_src(order, qz, kz);
                    const int pz = (sz - (4 * b.w - PL)) & (nc - 1);
                    const double wz = synth_w(order, qz, kz);
#pragma unroll 1
                    for (int ky = 0; ky < ny; ++ky) {
                        const int sy = synth_src(order, qy, ky);
                        const int py = (sy - (4 * b.z - PL)) & (nc - 1);
                        const double wzy = wz * synth_w(order, qy, ky);
                        const bool row_in = py < PC && pz < PC;
                        const int row = PC * (py + PC * pz);
#pragma unroll
                        for (int kx = 0; kx < NT; ++kx) {
                            if (kx >= nx) break;
                            const double w = 1.0 * (wzy * wxk[kx]);
                            const int pq = pxk[kx] + row;
                            if (row_in && pxk[kx] < PC && s_cf[pq]) {
                                const Real wr = (Real)w;
#pragma unroll
                                for (int k = 0; k < NF; ++k) {
                                    const Real prod = s_cv[k][pq] * wr;
                                    v[k] = v[k] + prod;
                                }
                            } else {
                                ghost_visit<1>(m, order, L - 1, sxk[kx], sy, sz, w, visit, my_bad);
                            }
                        }
                    }
                }
            }
        }
        if constexpr (FWT) {
            s_lab[p] = v[0];
        } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) lf.lab[k][(size_t)j * lab3 + p] = v[k];
        }
    }
    if (my_bad) atomicOr(bad, 1);
    if constexpr (FWT) {
        // one coarse point per lane: wave 0 forms code 7, wave 1 codes 3 and
        // 5, wave 2 codes 6 and 1, wave 3 codes 2 and 4 (125, 50, 30 and 10
        // terms at order 4)
        using G = FwtGeom<ORDER>;
        constexpr int FLAB = LAB;
        __shared__ double s_max[LAB_THREADS / 64];
        __syncthreads();
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int dx = lane & 3, dy = (lane >> 2) & 3, dz = lane >> 4;
        const Real* q = s_lab + (2 * dz - G::LS) * FLAB * FLAB + (2 * dy - G::LS) * FLAB + (2 * dx - G::LS);
        double dv = 0.0;
        if (wv == 0) {
            dv = (double)fabs(fwt_code<Real, ORDER, 1, 1, 1>(q));
        } else if (wv == 1) {
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 1, 1, 0>(q)));
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 1, 0, 1>(q)));
        } else if (wv == 2) {
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 0, 1, 1>(q)));
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 1, 0, 0>(q)));
        } else if (wv == 3) {
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 0, 1, 0>(q)));
            dv = fmax(dv, (double)fabs(fwt_code<Real, ORDER, 0, 0, 1>(q)));
        }
        for (int off = 32; off > 0; off >>= 1) dv = fmax(dv, __shfl_xor(dv, off, 64));
        if (lane == 0) s_max[wv] = dv;
        __syncthreads();
        if (threadIdx.x == 0) detail[j] = fmax(fmax(s_max[0], s_max[1]), fmax(s_max[2], s_max[3]));
    }
}

// Two blocks per 256-lane workgroup, two waves per block, one coarse point
// per lane: one wave forms code 7 (125 terms at order 4), the other codes 1-6
// (90), so that the waves of a workgroup finish together and a CU stays full
// (one wave per code left a CU running mostly the code-7 waves alone).
template <typename Real, int ORDER>
__global__ __launch_bounds__(256) void fwt_kernel(int nlist, const Real* __restrict__ lab,
                                                  double* __restrict__ detail) {
    using G = FwtGeom<ORDER>;
    constexpr int LAB = G::LAB, LS = G::LS, LAB3 = LAB * LAB * LAB;
    __shared__ Real s_lab[2 * LAB3];
    __shared__ double s_max[4];
    const int j0 = 2 * blockIdx.x;
    const int nb = min(2, nlist - j0);
    if (nb <= 0) return;
    for (int p = threadIdx.x; p < nb * LAB3; p += blockDim.x) s_lab[p] = lab[(size_t)j0 * LAB3 + p];
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = wv >> 1;
    const int dx = lane & 3, dy = (lane >> 2) & 3, dz = lane >> 4;
    const Real* p = s_lab + b * LAB3 + (2 * dz - LS) * LAB * LAB + (2 * dy - LS) * LAB + (2 * dx - LS);
    double v = 0.0;
    if (b < nb) {
        if ((wv & 1) == 0) {
            v = (double)fabs(fwt_code<Real, ORDER, 1, 1, 1>(p));
        } else {
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 1, 0, 0>(p)));
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 0, 1, 0>(p)));
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 1, 1, 0>(p)));
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 0, 0, 1>(p)));
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 1, 0, 1>(p)));
            v = fmax(v, (double)fabs(fwt_code<Real, ORDER, 0, 1, 1>(p)));
        }
    }
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if (lane == 0) s_max[wv] = v;
    __syncthreads();
    if (threadIdx.x < nb) detail[j0 + threadIdx.x] = fmax(s_max[2 * threadIdx.x], s_max[2 * threadIdx.x + 1]);
}

template <typename Real>
void launch_fwt(int order, int nlist, const Real* lab, double* detail, hipStream_t st) {
    const dim3 grid((unsigned)((nlist + 1) / 2));
    if (order == 4)
        hipLaunchKernelGGL((fwt_kernel<Real, 4>), grid, dim3(256), 0, st, nlist, lab, detail);
    else
        hipLaunchKernelGGL((fwt_kernel<Real, 2>), grid, dim3(256), 0, st, nlist, lab, detail);
}

// new leaf from the old grid: copy, split (Hs synthesis from the parent's
// lab, BlockSplitter::_splitBlocks) or collapse (Ha injection of the
// children, BlockCollapser::_collapseBlocks)
// the four fields of a remap (phi, p_w, p_g, p_csf): old arrays, their split
// labs and the new arrays
template <typename Real>
struct RemapFields {
    const Real* __restrict__ old[4];
    const Real* __restrict__ lab[4];
    Real* __restrict__ out[4];
};
// Split children: each destination cell d (0 <= d < 16 per axis in the
// parent's doubled lattice) sums the parent-lab sources s whose Hs tap 2s - d
// is nonzero, in increasing s -- the reference scatters every source of
// [S0, E0) (MRAGBlockSplitter.h:49-96) in increasing order, and the terms
// of its zero taps (order 4: 2s - d = +-2) add a signed zero to a sum that
// starts at +0 and so is never -0: they change nothing and are left out.
// The sources are synth_src / synth_w's (the same Hs taps): even d one
// (weight 1), odd d two (order 2) or four (order 4).  The lanes of a
// workgroup take the cells by parity class, one class per wave, so a wave's
// trip counts are uniform (1, 4, 16 or 64 terms at order 4).
template <typename Real>
__global__ __launch_bounds__(512) void remap_kernel(int nnew, const Remap* __restrict__ rm, RemapFields<Real> rf,
                                                    int LS, int LAB, int order) {
    const int ns = blockIdx.x;
    if (ns >= nnew) return;
    const Remap r = rm[ns];
    const int t = threadIdx.x;
    Real v[4];
    int cell = t;
    if (r.kind == 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) v[f] = rf.old[f][(size_t)r.src * BCELLS + t];
    } else if (r.kind == 2) {
        const int x = t & 7, y = (t >> 3) & 7, z = t >> 6;
        const int o = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2);
        const int c = 2 * (x & 3) + 8 * (2 * (y & 3)) + 64 * (2 * (z & 3));
#pragma unroll
        for (int f = 0; f < 4; ++f) v[f] = (Real)0 + rf.old[f][(size_t)r.kids[o] * BCELLS + c];
    } else {
        // the sources of this octant's cells: a W^3 window of the parent's lab
        // (W = 7 from 4 * octant - 1 at order 4, 5 from 4 * octant at order
        // 2), staged in LDS for the four fields
        __shared__ Real s_src[4 * 343];
        const int W = order == 4 ? 7 : 5;
        const int ws[3] = {4 * (r.octant & 1) - (order == 4 ? 1 : 0), 4 * ((r.octant >> 1) & 1) - (order == 4 ? 1 : 0),
                           4 * ((r.octant >> 2) & 1) - (order == 4 ? 1 : 0)};
        const size_t L0 = (size_t)r.src * LAB * LAB * LAB;
        const int W3 = W * W * W;
        for (int i = t; i < 4 * W3; i += 512) {
            const int f = i / W3, e = i - f * W3;
            const int ex = e % W, ey = (e / W) % W, ez = e / (W * W);
            s_src[i] = rf.lab[f][L0 + (size_t)((ws[2] + ez - LS) * LAB * LAB + (ws[1] + ey - LS) * LAB + (ws[0] + ex - LS))];
        }
        __syncthreads();
        const int w = t >> 6, l = t & 63;  // parity class (wave), cell of the class (lane)
        const int x = 2 * (l & 3) + (w & 1), y = 2 * ((l >> 2) & 3) + ((w >> 1) & 1), z = 2 * (l >> 4) + (w >> 2);
        cell = x + 8 * y + 64 * z;
        const int d[3] = {8 * (r.octant & 1) + x, 8 * ((r.octant >> 1) & 1) + y, 8 * ((r.octant >> 2) & 1) + z};
#pragma unroll
        for (int f = 0; f < 4; ++f) v[f] = (Real)0;
        const int nz = synth_n(order, d[2]), ny = synth_n(order, d[1]), nx = synth_n(order, d[0]);
        for (int kz = 0; kz < nz; ++kz) {
            const int sz = synth_src(order, d[2], kz);
            const double wZ = synth_w(order, d[2], kz);
            for (int ky = 0; ky < ny; ++ky) {
                const int sy = synth_src(order, d[1], ky);
                const double wZwY = wZ * synth_w(order, d[1], ky);
                for (int kx = 0; kx < nx; ++kx) {
                    const int sx = synth_src(order, d[0], kx);
                    const Real wt = (Real)(wZwY * synth_w(order, d[0], kx));
                    const int q = (sz - ws[2]) * W * W + (sy - ws[1]) * W + (sx - ws[0]);
#pragma unroll
                    for (int f = 0; f < 4; ++f) {
                        const Real prod = s_src[f * W3 + q] * wt;
                        v[f] = v[f] + prod;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) rf.out[f][(size_t)ns * BCELLS + cell] = v[f];
}

// per-cell constants on a multi-level grid: as rd_statics_kernel, with the
// face towards an interpolated ghost taken from the ghost's p_w, p_g; also the
// ghosts' tau and the harmonic diffusivity of the minus faces towards ghosts
template <typename Real>
__global__ __launch_bounds__(256) void statics_ghosted_kernel(int nslots, const int32_t* __restrict__ nbr,
                                                              const Real* __restrict__ pw, const Real* __restrict__ pg,
                                                              const Real* __restrict__ gpw,
                                                              const Real* __restrict__ gpg, Real Dw, Real Dg,
                                                              Stat4<Real>* __restrict__ st, Real* __restrict__ sig,
                                                              Real* __restrict__ fgTau, Real* __restrict__ fgDm) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)nslots * BCELLS) return;
    const int s = (int)(gid / BCELLS), l = (int)(gid % BCELLS);
    const int x = l & 7, y = (l >> 3) & 7, z = l >> 6;
    const Real pwc = pw[gid], pgc = pg[gid];
    const Real dc = diffusivity(pwc, pgc, Dw, Dg);
    auto nd = [&](int face, int inner_off, int wrap_off, int u, int v, bool at_edge) -> Real {
        if (!at_edge) return diffusivity(pw[gid + inner_off], pg[gid + inner_off], Dw, Dg);
        const int ns = nbr[6 * s + face];
        if (ns >= 0) {
            const size_t c = (size_t)ns * BCELLS + l + wrap_off;
            return diffusivity(pw[c], pg[c], Dw, Dg);
        }
        const size_t gi = (size_t)s * 384 + face * 64 + v * 8 + u;
        return diffusivity(gpw[gi], gpg[gi], Dw, Dg);
    };
    Stat4<Real> r;
    r.dx = harmonic(nd(FXP, 1, -7, y, z, x == 7), dc);
    r.dy = harmonic(nd(FYP, 8, -56, x, z, y == 7), dc);
    r.dz = harmonic(nd(FZP, 64, -448, x, y, z == 7), dc);
    r.tau = chf_threshold(pwc, pgc);
    st[gid] = r;
    sig[gid] = domain_threshold(pwc, pgc);
    // minus faces towards interpolated ghosts
    const int mfaces[3] = {FXM, FYM, FZM};
    const bool atm[3] = {x == 0, y == 0, z == 0};
    const int uu[3] = {y, x, x}, vv[3] = {z, z, y};
    for (int a = 0; a < 3; ++a)
        if (atm[a] && nbr[6 * s + mfaces[a]] < 0) {
            const size_t gi = (size_t)s * 384 + mfaces[a] * 64 + vv[a] * 8 + uu[a];
            fgDm[(size_t)s * 192 + a * 64 + vv[a] * 8 + uu[a]] = harmonic(diffusivity(gpw[gi], gpg[gi], Dw, Dg), dc);
        }
    // tau of the interpolated ghosts (one lane per ghost: the cells with x == 0 cover 64 per face)
    if (l < 384) {
        const int f = l >> 6;
        if (nbr[6 * s + f] < 0) {
            const size_t gi = (size_t)s * 384 + l;
            fgTau[gi] = chf_threshold(gpw[gi], gpg[gi]);
        }
    }
}

// Face-ghost packets as the step kernel reads them: packet of jump face
// (slot, f) = pkOf[6 * slot + f] (-1: no jump face), its terms as in
// face_packet_kernel
template <typename Real>
struct FacePackets {
    const int32_t* __restrict__ pkOf;
    const int32_t* __restrict__ base;
    const int32_t* __restrict__ count;
    const int32_t* __restrict__ tsrc;
    const Real* __restrict__ tw;
};

// RD step on a multi-level grid: one workgroup per block, h of the block's
// level; a face's halo comes from the same-level neighbour or, across a level
// jump, from the interpolated face ghosts -- precomputed in fgPhi, or (PK)
// summed here from the face's packet by the wave that owns the face, term by
// term in face_packet_kernel's order, so both give the same bits
template <typename Real, bool FUSED, bool PK, int NT>
__global__ __launch_bounds__(NT) void rd_step_ghosted_kernel(int nslots, int per_xcd, const int32_t* __restrict__ nbr,
                                                             const int32_t* __restrict__ level,
                                                             const Real* __restrict__ phi, Real* __restrict__ out,
                                                             const Stat4<Real>* __restrict__ st,
                                                             const Real* __restrict__ sig,
                                                             const Real* __restrict__ fgPhi,
                                                             const Real* __restrict__ fgTau,
                                                             const Real* __restrict__ fgDm, Real rho, double dt,
                                                             FacePackets<Real> fp) {
    static_assert(NT == 512 || NT == 256, "one or two cells per lane");
    constexpr int CPT = 512 / NT;  // cells per lane: i = t + c * NT
    const int s = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (s >= nslots) return;
    __shared__ Real s_phi[1000];
    __shared__ unsigned char s_out[1000];
    __shared__ Real s_Dx[576], s_Dy[576], s_Dz[576];
    const int t = threadIdx.x;
    Real pc[CPT], sc[CPT];
    Stat4<Real> sv[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int i = t + c * NT;
        const size_t g = (size_t)s * BCELLS + i;
        pc[c] = phi[g];
        sv[c] = st[g];
        sc[c] = sig[g];
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int i = t + c * NT;
        const int x = i & 7, y = (i >> 3) & 7, z = i >> 6;
        s_phi[P(x, y, z)] = pc[c];
        s_out[P(x, y, z)] = pc[c] < sv[c].tau;
        s_Dx[z * 72 + y * 9 + x + 1] = sv[c].dx;
        s_Dy[z * 72 + (y + 1) * 8 + x] = sv[c].dy;
        s_Dz[(z + 1) * 64 + y * 8 + x] = sv[c].dz;
    }
    // the 6 x 64 face halo points, one wave per face
#pragma unroll
    for (int c = 0; c < (384 + NT - 1) / NT; ++c) {
        const int fi = t + c * NT;
        if (fi >= 384) break;
        const int f = fi >> 6, u = fi & 7, v = (fi >> 3) & 7;
        int src, dst;
        switch (f) {
            case FXM: src = 7 + 8 * u + 64 * v; dst = P(-1, u, v); break;
            case FXP: src = 0 + 8 * u + 64 * v; dst = P(8, u, v); break;
            case FYM: src = u + 8 * 7 + 64 * v; dst = P(u, -1, v); break;
            case FYP: src = u + 8 * 0 + 64 * v; dst = P(u, 8, v); break;
            case FZM: src = u + 8 * v + 64 * 7; dst = P(u, v, -1); break;
            default:  src = u + 8 * v + 64 * 0; dst = P(u, v, 8); break;
        }
        const int ns = nbr[6 * s + f];
        Real ph, tau, dm = (Real)0;
        if (ns >= 0) {
            const size_t nb = (size_t)ns * BCELLS + src;
            const Stat4<Real> hs = st[nb];
            ph = phi[nb];
            tau = hs.tau;
            dm = f == FXM ? hs.dx : (f == FYM ? hs.dy : hs.dz);
        } else {
            if (PK) {
                const int k = fp.pkOf[6 * s + f];
                const int p = fi & 63;
                const int b0 = fp.base[k] + p, len = fp.count[64 * k + p];
                Real acc = (Real)0;
#pragma unroll 4
                for (int j = 0; j < len; ++j) {
                    const Real prod = phi[fp.tsrc[b0 + 64 * j]] * fp.tw[b0 + 64 * j];
                    acc = acc + prod;
                }
                ph = acc;
            } else {
                ph = fgPhi[(size_t)s * 384 + fi];
            }
            tau = fgTau[(size_t)s * 384 + fi];
            if (f == FXM || f == FYM || f == FZM) dm = fgDm[(size_t)s * 192 + (f >> 1) * 64 + v * 8 + u];
        }
        s_phi[dst] = ph;
        s_out[dst] = ph < tau;
        if (f == FXM) s_Dx[v * 72 + u * 9] = dm;
        else if (f == FYM) s_Dy[v * 72 + u] = dm;
        else if (f == FZM) s_Dz[v * 8 + u] = dm;
    }
    __syncthreads();
    const double h = ldexp(1.0, -level[s]) / 8;  // info.h[0] (MRAGrid.inl:200-206)
    const double ih2 = 1. / (h * h);
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int i = t + c * NT;
        const int x = i & 7, y = (i >> 3) & 7, z = i >> 6;
        Real d0 = s_Dx[z * 72 + y * 9 + x], d1 = sv[c].dx, d2 = s_Dy[z * 72 + y * 8 + x], d3 = sv[c].dy;
        Real d4 = s_Dz[z * 64 + y * 8 + x], d5 = sv[c].dz;
        if (s_out[P(x - 1, y, z)]) d1 = (Real)((double)d1 * 2.0);
        if (s_out[P(x + 1, y, z)]) d0 = (Real)((double)d0 * 2.0);
        if (s_out[P(x, y - 1, z)]) d3 = (Real)((double)d3 * 2.0);
        if (s_out[P(x, y + 1, z)]) d2 = (Real)((double)d2 * 2.0);
        if (s_out[P(x, y, z - 1)]) d5 = (Real)((double)d5 * 2.0);
        if (s_out[P(x, y, z + 1)]) d4 = (Real)((double)d4 * 2.0);
        const Real p0 = s_phi[P(x - 1, y, z)], p1 = s_phi[P(x + 1, y, z)], p2 = s_phi[P(x, y - 1, z)];
        const Real p3 = s_phi[P(x, y + 1, z)], p4 = s_phi[P(x, y, z - 1)], p5 = s_phi[P(x, y, z + 1)];
        Real sin_ = d0 * p0;
        sin_ = sin_ + d1 * p1;
        sin_ = sin_ + d2 * p2;
        sin_ = sin_ + d3 * p3;
        sin_ = sin_ + d4 * p4;
        sin_ = sin_ + d5 * p5;
        const double fin = ih2 * (double)sin_;
        Real sdf = d0 + d1;
        sdf = sdf + d2;
        sdf = sdf + d3;
        sdf = sdf + d4;
        sdf = sdf + d5;
        const Real sdp = sdf * pc[c];
        const double fout = -((double)sdp * ih2);
        const Real rp = rho * pc[c];
        const double react = (double)rp * (1. - (double)pc[c]);
        Real dphidt = (Real)((fout + fin) + react);
        if (!(pc[c] > sc[c])) dphidt = (Real)0.;
        const size_t g = (size_t)s * BCELLS + i;
        if (FUSED) {
            Real p = (Real)((double)pc[c] + dt * (double)dphidt);
            p = ((Real)0. < p) ? p : (Real)0.;
            p = (p < (Real)1.) ? p : (Real)1.;
            out[g] = p;
        } else {
            out[g] = dphidt;
        }
    }
}

// _dumpUQoutput (Glioma_ReactionDiffusion.cpp:208-267): leaf cells onto the
// finest uniform gpd^3 lattice; h = hf writes 1, 2hf writes 2^3, 3hf writes
// 3^3 and anything coarser writes 4^3 cells (the reference's quirk is kept)
template <typename Real, bool WRITE>
__global__ __launch_bounds__(512) void uq_resample_kernel(int nslots, const BlockRec* __restrict__ blocks,
                                                          const Real* __restrict__ phi, int gpd,
                                                          int* __restrict__ owner, double* __restrict__ out) {
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const BlockRec b = blocks[s];
    const int t = threadIdx.x;
    const int ix = t & 7, iy = (t >> 3) & 7, iz = t >> 6;
    const Real dil = (Real)ldexp(1.0, -b.level);
    const Real h = dil / (Real)8;
    const double hf = 1. / gpd, eps = hf * 0.5;
    const double xs[3] = {(double)((Real)(b.index[0] * dil) + h * (Real)ix),
                          (double)((Real)(b.index[1] * dil) + h * (Real)iy),
                          (double)((Real)(b.index[2] * dil) + h * (Real)iz)};
    const int m[3] = {(int)floor(xs[0] / hf), (int)floor(xs[1] / hf), (int)floor(xs[2] / hf)};
    const double hd = (double)h;
    const int k = hd < hf + eps ? 1 : (hd < 2. * hf + eps ? 2 : (hd < 3. * hf + eps ? 3 : 4));
    const int key = s * BCELLS + t;  // the reference's write order: blocks in order, cells z, y, x
    const double v = (double)phi[(size_t)s * BCELLS + t];
    for (int cz = 0; cz < k; ++cz)
        for (int cy = 0; cy < k; ++cy)
            for (int cx = 0; cx < k; ++cx) {
                const int X = m[0] + cx, Y = m[1] + cy, Z = m[2] + cz;
                if (X >= gpd || Y >= gpd || Z >= gpd) continue;
                const size_t o = (size_t)X + (size_t)gpd * (Y + (size_t)gpd * Z);
                if (!WRITE)
                    atomicMax(&owner[o], key);
                else if (owner[o] == key)
                    out[o] = v;
            }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// a table of ghost entries, each with its terms in recursion order
struct HostTable {
    std::vector<GhostEntry> ent;
    std::vector<int32_t> src;
    std::vector<double> w;
    void add(int32_t dst, const std::vector<GhostTerm>& terms) {
        ent.push_back(GhostEntry{dst, (int32_t)src.size(), (int32_t)terms.size()});
        for (auto& t : terms) {
            src.push_back(t.src);
            w.push_back(t.w);
        }
    }
};

// A table over items [0, n) built on up to 16 host threads: each thread fills
// the table of a contiguous item range with fill(i0, i1, table), and the parts
// are concatenated in item order, so the result equals a sequential build.
template <typename F>
HostTable parallel_table(int n, F fill) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({16, hw, n / 16}));
    HostTable h;
    if (nt == 1) {
        fill(0, n, h);
        return h;
    }
    std::vector<HostTable> part(nt);
    std::vector<std::exception_ptr> err(nt);
    std::vector<std::thread> th;
    for (int k = 0; k < nt; ++k)
        th.emplace_back([&, k] {
            try {
                fill((int)((int64_t)n * k / nt), (int)((int64_t)n * (k + 1) / nt), part[k]);
            } catch (...) {
                err[k] = std::current_exception();
            }
        });
    for (auto& t : th) t.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
    size_t ne = 0, nterm = 0;
    for (auto& p : part) {
        ne += p.ent.size();
        nterm += p.src.size();
    }
    h.ent.reserve(ne);
    h.src.reserve(nterm);
    h.w.reserve(nterm);
    for (auto& p : part) {
        const int32_t off = (int32_t)h.src.size();
        for (GhostEntry e : p.ent) {
            e.start += off;
            h.ent.push_back(e);
        }
        h.src.insert(h.src.end(), p.src.begin(), p.src.end());
        h.w.insert(h.w.end(), p.w.begin(), p.w.end());
    }
    return h;
}

template <typename Real>
int upload_table(Ctx& c, const HostTable& h, DevBuf& ent, DevBuf& src, DevBuf& w) {
    int rc;
    if ((rc = ent.ensure_grow(std::max<size_t>(1, h.ent.size()) * sizeof(GhostEntry))) ||
        (rc = src.ensure_grow(std::max<size_t>(1, h.src.size()) * sizeof(int32_t))) ||
        (rc = w.ensure_grow(std::max<size_t>(1, h.w.size()) * sizeof(Real))))
        return rc;
    std::vector<Real> wr(h.w.size());
    for (size_t i = 0; i < wr.size(); ++i) wr[i] = (Real)h.w[i];
    if (!h.ent.empty()) {
        GH_HIP_CHECK(hipMemcpyAsync(ent.p, h.ent.data(), h.ent.size() * sizeof(GhostEntry), hipMemcpyHostToDevice,
                                    c.stream));
        GH_HIP_CHECK(hipMemcpyAsync(src.p, h.src.data(), h.src.size() * 4, hipMemcpyHostToDevice, c.stream));
        GH_HIP_CHECK(hipMemcpyAsync(w.p, wr.data(), wr.size() * sizeof(Real), hipMemcpyHostToDevice, c.stream));
    }
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    return GH_OK;
}

// face ghosts of the RD stencil for faces without a same-level neighbour
HostTable build_face_table(const Octree& t, const std::vector<int32_t>& nbr) {
    return parallel_table((int)t.leaves.size(), [&](int s0, int s1, HostTable& h) {
        std::vector<GhostTerm> terms;
        for (int s = s0; s < s1; ++s) {
            const BlockRec& b = t.leaves[s];
            for (int f = 0; f < 6; ++f) {
                if (nbr[6 * s + f] >= 0) continue;
                for (int v = 0; v < 8; ++v)
                    for (int u = 0; u < 8; ++u) {
                        int i[3];
                        switch (f) {
                            case FXM: i[0] = -1; i[1] = u; i[2] = v; break;
                            case FXP: i[0] = 8; i[1] = u; i[2] = v; break;
                            case FYM: i[0] = u; i[1] = -1; i[2] = v; break;
                            case FYP: i[0] = u; i[1] = 8; i[2] = v; break;
                            case FZM: i[0] = u; i[1] = v; i[2] = -1; break;
                            default:  i[0] = u; i[1] = v; i[2] = 8; break;
                        }
                        const int64_t q[3] = {8LL * b.index[0] + i[0], 8LL * b.index[1] + i[1],
                                              8LL * b.index[2] + i[2]};
                        terms.clear();
                        t.value_terms(b.level, q, 1.0, terms);
                        h.add(s * 384 + f * 64 + v * 8 + u, terms);
                    }
            }
        }
    });
}

// interpolated ghosts of the full labs of the listed slots (compact index j)
HostTable build_lab_table(const Octree& t, const std::vector<int>& slots, const std::vector<int32_t>& nbr27,
                          int order) {
    const int LS = lab_start(order), LAB = lab_size(order), lab3 = LAB * LAB * LAB;
    return parallel_table((int)slots.size(), [&](int j0, int j1, HostTable& h) {
        std::vector<GhostTerm> terms;
        for (int j = j0; j < j1; ++j) {
            const int s = slots[j];
            const BlockRec& b = t.leaves[s];
            for (int p = 0; p < lab3; ++p) {
                const int i[3] = {p % LAB + LS, (p / LAB) % LAB + LS, p / (LAB * LAB) + LS};
                const int code = (i[0] < 0 ? 0 : (i[0] < 8 ? 1 : 2)) + 3 * (i[1] < 0 ? 0 : (i[1] < 8 ? 1 : 2)) +
                                 9 * (i[2] < 0 ? 0 : (i[2] < 8 ? 1 : 2));
                if (code == 13 || nbr27[27 * s + code] >= 0) continue;
                const int64_t q[3] = {8LL * b.index[0] + i[0], 8LL * b.index[1] + i[1], 8LL * b.index[2] + i[2]};
                terms.clear();
                t.value_terms(b.level, q, 1.0, terms);
                h.add(j * lab3 + p, terms);
            }
        }
    });
}

// The face-ghost table of the RD step built on the device (same entries and
// term order as build_face_table): the jump faces listed, a counting pass,
// a prefix sum, a filling pass.  *done = false when a ghost is deeper than
// the device recursion (the caller then builds it on the host).
int device_face_table(Ctx& c, bool& done) {
    done = false;
    c.nFacePk = 0;
    c.nFaceEnt = 0;
    const int nb = c.topo.nblocks();
    const int n6 = 6 * nb;
    int rc;
    DevBuf& dfaces = c.faceList;  // kept across calls: no allocation per adaptation
    if ((rc = c.faceScan.ensure_grow(((size_t)n6 + 1) * 4)) || (rc = dfaces.ensure_grow(std::max(1, n6) * 4)) ||
        (rc = c.facePkOf.ensure_grow(std::max(1, n6) * 4)))
        return rc;
    // the jump faces in slot-face order, listed on the device from the
    // neighbour table: flags, their exclusive prefix sum, the list
    int32_t* pos = c.faceScan.as<int32_t>();
    {
        const int np6 = n6 + 1;
        const int nch = (np6 + 1023) / 1024;
        if ((rc = c.scanBuf.ensure_grow(((size_t)2 * nch + 1) * sizeof(long long)))) return rc;
        long long* sums = c.scanBuf.as<long long>();
        hipLaunchKernelGGL(face_flag_kernel, dim3((unsigned)((np6 + 255) / 256)), dim3(256), 0, c.stream, n6,
                           c.nbr.as<int32_t>(), pos);
        hipLaunchKernelGGL(scan_chunk_sums, dim3(nch), dim3(1024), 0, c.stream, np6, pos, sums);
        hipLaunchKernelGGL(scan_chunk_offsets, dim3(1), dim3(1024), 0, c.stream, nch, sums, sums + nch,
                           sums + 2 * nch);
        hipLaunchKernelGGL(scan_chunk_apply, dim3(nch), dim3(1024), 0, c.stream, np6, pos, sums + nch);
        if (n6 > 0)
            hipLaunchKernelGGL(face_list_kernel, dim3((unsigned)((n6 + 255) / 256)), dim3(256), 0, c.stream, n6,
                               c.nbr.as<int32_t>(), pos, dfaces.as<int32_t>(), c.facePkOf.as<int32_t>());
        GH_HIP_CHECK(hipGetLastError());
    }
    int nf = 0;
    GH_HIP_CHECK(hipMemcpyAsync(&nf, pos + n6, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    const int n = nf * 64;
    if (n == 0) {
        done = true;
        return GH_OK;
    }
    const LevelMaps m = level_maps(c);
    const int order = c.tree.order;
    DevBuf& dcount = c.faceCount;
    if ((rc = dcount.ensure_grow((size_t)n * 4)) || (rc = c.facePkBase.ensure_grow(((size_t)nf + 1) * 4))) return rc;
    GH_HIP_CHECK(hipMemsetAsync(c.dBad.p, 0, sizeof(int), c.stream));
    GH_HIP_CHECK(hipMemsetAsync(c.facePkBase.as<int32_t>() + nf, 0, sizeof(int32_t), c.stream));
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(face_count_kernel, dim3(grid), dim3(256), 0, c.stream, n, dfaces.as<int32_t>(),
                       c.dBlk.as<int4>(), m, order, dcount.as<int32_t>(), c.facePkBase.as<int32_t>(),
                       c.dBad.as<int>());
    GH_HIP_CHECK(hipGetLastError());
    // packet starts: exclusive prefix sum of the packet sizes, in place, the
    // total in the extra last entry
    const int np = nf + 1;
    const int nchunk = (np + 1023) / 1024;
    if ((rc = c.scanBuf.ensure_grow(((size_t)2 * nchunk + 1) * sizeof(long long)))) return rc;
    long long* sums = c.scanBuf.as<long long>();
    int32_t* pk = c.facePkBase.as<int32_t>();
    hipLaunchKernelGGL(scan_chunk_sums, dim3(nchunk), dim3(1024), 0, c.stream, np, pk, sums);
    hipLaunchKernelGGL(scan_chunk_offsets, dim3(1), dim3(1024), 0, c.stream, nchunk, sums, sums + nchunk,
                       sums + 2 * nchunk);
    hipLaunchKernelGGL(scan_chunk_apply, dim3(nchunk), dim3(1024), 0, c.stream, np, pk, sums + nchunk);
    GH_HIP_CHECK(hipGetLastError());
    int bad = 1;
    long long total = 0;
    GH_HIP_CHECK(hipMemcpyAsync(&total, sums + 2 * nchunk, sizeof(long long), hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipMemcpyAsync(&bad, c.dBad.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (bad) return GH_OK;
    if (total > INT32_MAX) return set_error(GH_EINVAL, "face ghost table: too many terms");
    const size_t rs = c.real_size;
    if ((rc = c.faceSrc.ensure_grow(std::max<int64_t>(1, total) * 4)) ||
        (rc = c.faceW.ensure_grow(std::max<int64_t>(1, total) * rs)))
        return rc;
    if (c.real_kind == GH_REAL_F32)
        hipLaunchKernelGGL(face_fill_packet_kernel<float>, dim3(grid), dim3(256), 0, c.stream, n,
                           dfaces.as<int32_t>(), c.dBlk.as<int4>(), m, order, pk, c.faceSrc.as<int32_t>(),
                           c.faceW.as<float>());
    else
        hipLaunchKernelGGL(face_fill_packet_kernel<double>, dim3(grid), dim3(256), 0, c.stream, n,
                           dfaces.as<int32_t>(), c.dBlk.as<int4>(), m, order, pk, c.faceSrc.as<int32_t>(),
                           c.faceW.as<double>());
    GH_HIP_CHECK(hipGetLastError());
    c.nFacePk = nf;
    done = true;
    return GH_OK;
}

}  // namespace

// the staging buffers and the tree's hash table may be rewritten once the
// uploads recorded by stage_mark have completed
int stage_wait(Ctx& c) {
    if (c.evStage) GH_HIP_CHECK(hipEventSynchronize(c.evStage));
    return GH_OK;
}
int stage_mark(Ctx& c) {
    if (!c.evStage) GH_HIP_CHECK(hipEventCreateWithFlags(&c.evStage, hipEventDisableTiming));
    GH_HIP_CHECK(hipEventRecord(c.evStage, c.stream));
    return GH_OK;
}

// topology derived from the octree: slots, level maps, block coordinates
// uploaded; same-level neighbours and levels built on the device from the
// maps (topology_kernel); the per-slot ghost buffers sized
int apply_topology(Ctx& c) {
    Octree& t = c.tree;
    ++c.tree_version;  // invalidates the cached lab tables
    c.lik_ready = false;  // observations were laid out for the old topology
    Topology topo;
    topo.blocks = t.leaves;
    const int nb = (int)t.leaves.size();
    topo.uniform_bpd = t.is_uniform() ? (1 << t.levels_min()) : 0;
    c.topo = std::move(topo);
    int rc;
    if ((rc = c.nbr.ensure_grow(std::max(1, 6 * nb) * 4)) || (rc = c.dNbr27.ensure_grow(std::max(1, 27 * nb) * 4)) ||
        (rc = c.dLevel.ensure_grow(std::max(1, nb) * 4)))
        return rc;
    // level maps and block coordinates for the device-side lab ghosts, staged
    // in pinned memory (read by the uploads until stage_mark's event passes)
    if ((rc = stage_wait(c))) return rc;
    c.nDense = t.dense_levels;
    c.topLevel = t.top_level;
    c.hashBits = t.hash_bits;
    const size_t ndense = c.nDense > 0 ? (size_t)dense_level_offset(c.nDense) : 0;
    if ((rc = c.pinDense.ensure_grow(std::max<size_t>(1, ndense) * 4)) ||
        (rc = c.pinBlk.ensure_grow(std::max(1, nb) * sizeof(int4))))
        return rc;
    int32_t* dense = c.pinDense.as<int32_t>();
    for (int l = 0; l < c.nDense; ++l) {
        if (t.dense[l].size() != ((size_t)1 << (3 * l))) return set_error(GH_EINVAL, "level maps: layout");
        std::memcpy(dense + dense_level_offset(l), t.dense[l].data(), t.dense[l].size() * 4);
    }
    if (c.hashBits > 0) {  // the tree outlives the asynchronous copies (apply_new_tree waits for them)
        int rc2;
        if ((rc2 = c.dHashKey.ensure_grow(t.hash_keys.size() * 8)) ||
            (rc2 = c.dHashVal.ensure_grow(t.hash_vals.size() * 4)))
            return rc2;
        GH_HIP_CHECK(hipMemcpyAsync(c.dHashKey.p, t.hash_keys.data(), t.hash_keys.size() * 8, hipMemcpyHostToDevice,
                                    c.stream));
        GH_HIP_CHECK(hipMemcpyAsync(c.dHashVal.p, t.hash_vals.data(), t.hash_vals.size() * 4, hipMemcpyHostToDevice,
                                    c.stream));
    }
    int4* blk = c.pinBlk.as<int4>();
    for (int s = 0; s < nb; ++s)
        blk[s] = make_int4(t.leaves[s].level, t.leaves[s].index[0], t.leaves[s].index[1], t.leaves[s].index[2]);
    if ((rc = c.dDense.ensure_grow(std::max<size_t>(1, ndense) * 4)) ||
        (rc = c.dBlk.ensure_grow(std::max(1, nb) * sizeof(int4))) || (rc = c.dBad.ensure_grow(sizeof(int))))
        return rc;
    if (ndense > 0)
        GH_HIP_CHECK(hipMemcpyAsync(c.dDense.p, dense, ndense * 4, hipMemcpyHostToDevice, c.stream));
    if (nb > 0) GH_HIP_CHECK(hipMemcpyAsync(c.dBlk.p, blk, (size_t)nb * sizeof(int4), hipMemcpyHostToDevice, c.stream));
    if (nb > 0) {
        hipLaunchKernelGGL(topology_kernel, dim3((unsigned)((27 * nb + 255) / 256)), dim3(256), 0, c.stream, 27 * nb,
                           c.dBlk.as<int4>(), level_maps(c), c.dNbr27.as<int32_t>(), c.nbr.as<int32_t>(),
                           c.dLevel.as<int32_t>());
        GH_HIP_CHECK(hipGetLastError());
    }
    c.faces_stale = true;  // built by ensure_faces when an RD step or the constants need them
    c.nFacePk = 0;
    c.nFaceEnt = 0;
    const size_t rs = c.real_size;
    rc = c.fgPhi.ensure_grow(std::max(1, nb) * 384 * rs);
    if (!rc) rc = c.fgTau.ensure_grow(std::max(1, nb) * 384 * rs);
    if (!rc) rc = c.fgDm.ensure_grow(std::max(1, nb) * 192 * rs);
    if (!rc) rc = c.fgPw.ensure_grow(std::max(1, nb) * 384 * rs);
    if (!rc) rc = c.fgPg.ensure_grow(std::max(1, nb) * 384 * rs);
    if (rc) return rc;
    // no synchronization: the next reader of the device tables is on the same
    // stream, and the staging buffers are rewritten only after stage_wait
    return stage_mark(c);
}

// the host copies of the neighbour tables (c.topo.nbr, c.topo.nbr27), for a
// host-built ghost table; the device's own are built by apply_topology
void host_neighbours(Ctx& c) {
    const int nb = c.topo.nblocks();
    if ((int)c.topo.nbr27.size() == 27 * nb && (int)c.topo.nbr.size() == 6 * nb) return;
    static const int face_code[6] = {12, 14, 10, 16, 4, 22};
    c.topo.nbr.assign(6 * (size_t)nb, -1);
    c.topo.nbr27.assign(27 * (size_t)nb, -1);
    for (int s = 0; s < nb; ++s) {
        int o[27];
        c.tree.same_level_neighbours(s, o);
        for (int k = 0; k < 27; ++k) c.topo.nbr27[27 * (size_t)s + k] = o[k];
        for (int f = 0; f < 6; ++f) c.topo.nbr[6 * (size_t)s + f] = o[face_code[f]];
    }
}

namespace {

// the face-ghost table of the current grid (device-built, or host-built when
// a ghost needs deeper recursion), once per grid change
int ensure_faces(Ctx& c) {
    if (!c.faces_stale) return GH_OK;
    bool on_device = false;
    int rc;
    if (c.nDense > 0 && (rc = device_face_table(c, on_device))) return rc;
    ++c.adapt_count[on_device ? 2 : 3];
    if (!on_device) {
        c.nFacePk = 0;
        host_neighbours(c);
        HostTable ft = build_face_table(c.tree, c.topo.nbr);
        c.nFaceEnt = (int)ft.ent.size();
        rc = c.real_kind == GH_REAL_F32 ? upload_table<float>(c, ft, c.faceEnt, c.faceSrc, c.faceW)
                                        : upload_table<double>(c, ft, c.faceEnt, c.faceSrc, c.faceW);
        if (rc) return rc;
        GH_HIP_CHECK(hipStreamSynchronize(c.stream));  // the table's host vectors are being uploaded
    }
    c.faces_stale = false;
    return GH_OK;
}

template <typename Real>
int face_ghosts(Ctx& c, const Real* field, Real* out) {
    if (c.nFacePk > 0) {
        const int n = c.nFacePk * 64;
        hipLaunchKernelGGL(face_packet_kernel<Real>, dim3((n + 255) / 256), dim3(256), 0, c.stream, n,
                           c.faceList.as<int32_t>(), c.facePkBase.as<int32_t>(), c.faceCount.as<int32_t>(),
                           c.faceSrc.as<int32_t>(), c.faceW.as<Real>(), field, out);
        GH_HIP_CHECK(hipGetLastError());
        return GH_OK;
    }
    if (c.nFaceEnt == 0) return GH_OK;
    hipLaunchKernelGGL(ghost_entry_kernel<Real>, dim3((c.nFaceEnt + 255) / 256), dim3(256), 0, c.stream, c.nFaceEnt,
                       c.faceEnt.as<GhostEntry>(), c.faceSrc.as<int32_t>(), c.faceW.as<Real>(), field, out);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

template <typename Real>
int prepare_ghosted_t(Ctx& c, double Dw, double Dg) {
    int rc;
    if ((rc = ensure_faces(c)) || (rc = face_ghosts<Real>(c, c.pw.as<Real>(), c.fgPw.as<Real>())) ||
        (rc = face_ghosts<Real>(c, c.pg.as<Real>(), c.fgPg.as<Real>())))
        return rc;
    const int ns = c.topo.nblocks();
    const size_t nc = c.ncells();
    hipLaunchKernelGGL(statics_ghosted_kernel<Real>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c.stream, ns,
                       c.nbr.as<int32_t>(), c.pw.as<Real>(), c.pg.as<Real>(), c.fgPw.as<Real>(), c.fgPg.as<Real>(),
                       (Real)Dw, (Real)Dg, c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), c.fgTau.as<Real>(),
                       c.fgDm.as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

constexpr int kFusedGhostMaxBlocks = 1450;  // largest grid whose step kernel gathers its own ghosts

// The step kernel with 512 lanes per block when it gathers the face ghosts
// itself (one wave per face: with 256 lanes two waves would gather two faces
// each, in turn: 1,114 blocks 15.9 -> 23.0 us per step), else with 256 lanes
// of two cells each (more blocks in flight per CU: 7,225 / 13,455 blocks 45.3
// / 91.6 -> 44.6 / 86.1 us per step with the separate gather).
template <typename Real, bool FUSED, bool PK>
void launch_step_ghosted(Ctx& c, int ns, int per_xcd, const Real* in, Real* out, double dt, const Real* fg,
                         const FacePackets<Real>& fp) {
    if (!PK)
        hipLaunchKernelGGL((rd_step_ghosted_kernel<Real, FUSED, PK, 256>), dim3(8 * per_xcd), dim3(256), 0, c.stream,
                           ns, per_xcd, c.nbr.as<int32_t>(), c.dLevel.as<int32_t>(), in, out,
                           c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), fg, c.fgTau.as<Real>(),
                           c.fgDm.as<Real>(), (Real)c.rho, dt, fp);
    else
        hipLaunchKernelGGL((rd_step_ghosted_kernel<Real, FUSED, PK, 512>), dim3(8 * per_xcd), dim3(512), 0, c.stream,
                           ns, per_xcd, c.nbr.as<int32_t>(), c.dLevel.as<int32_t>(), in, out,
                           c.statD.as<Stat4<Real>>(), c.statSig.as<Real>(), fg, c.fgTau.as<Real>(),
                           c.fgDm.as<Real>(), (Real)c.rho, dt, fp);
}

template <typename Real, bool FUSED>
int step_ghosted_t(Ctx& c, const Real* in, Real* out, double dt) {
    int rc0 = ensure_faces(c);
    if (rc0) return rc0;
    const int ns = c.topo.nblocks();
    const int per_xcd = (ns + 7) / 8;
    // Small grids: ghosts summed inside the step kernel (one launch per step).
    // Larger grids: a packet kernel first, whose one-thread-per-ghost gather
    // hides the scattered reads better than the step kernel's ghost threads,
    // which hold up their whole workgroup at the tile barrier (measured on
    // MI355X at 1,114 blocks: 16.0 fused vs 17.2 us per step; at 7,225: 67.1
    // vs 45.4 us; the crossover of the two straight lines is ~1,450 blocks).
    if (c.nFacePk > 0 && ns <= kFusedGhostMaxBlocks) {
        const FacePackets<Real> fp{c.facePkOf.as<int32_t>(), c.facePkBase.as<int32_t>(), c.faceCount.as<int32_t>(),
                                   c.faceSrc.as<int32_t>(), c.faceW.as<Real>()};
        launch_step_ghosted<Real, FUSED, true>(c, ns, per_xcd, in, out, dt, nullptr, fp);
        GH_HIP_CHECK(hipGetLastError());
        return GH_OK;
    }
    int rc = face_ghosts<Real>(c, in, c.fgPhi.as<Real>());
    if (rc) return rc;
    launch_step_ghosted<Real, FUSED, false>(c, ns, per_xcd, in, out, dt, c.fgPhi.as<Real>(), FacePackets<Real>{});
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

// full labs of `slots` for `field` into `out` (compact, one lab per listed
// slot, slot list already on the device in dslots); the interpolated ghosts
// come from the uploaded lab table (nent entries in ent/src/w)
template <typename Real>
int build_labs(Ctx& c, int nslots, const DevBuf& dslots, const Real* field, int nent, const DevBuf& ent,
               const DevBuf& src, const DevBuf& w, DevBuf& out) {
    const int order = c.tree.order;
    const int LS = lab_start(order), LAB = lab_size(order), lab3 = LAB * LAB * LAB;
    int rc;
    if ((rc = out.ensure_grow(std::max<size_t>(1, (size_t)nslots) * lab3 * sizeof(Real)))) return rc;
    if (nslots == 0) return GH_OK;
    hipLaunchKernelGGL(lab_fill_kernel<Real>, dim3((unsigned)nslots), dim3(256), 0, c.stream, nslots,
                       dslots.as<int32_t>(), LS, LAB, c.dNbr27.as<int32_t>(), field, out.as<Real>());
    GH_HIP_CHECK(hipGetLastError());
    if (nent > 0) {
        hipLaunchKernelGGL(ghost_entry_kernel<Real>, dim3((unsigned)((nent + 255) / 256)), dim3(256), 0, c.stream, nent,
                           ent.as<GhostEntry>(), src.as<int32_t>(), w.as<Real>(), field, out.as<Real>());
        GH_HIP_CHECK(hipGetLastError());
    }
    return GH_OK;
}

// labs of the listed slots with their own freshly built ghost table
template <typename Real>
int build_labs_for(Ctx& c, const std::vector<int>& slots, const Real* field, const HostTable& tab, DevBuf& out) {
    int rc;
    if ((rc = c.labSlots.ensure_grow(std::max<size_t>(1, slots.size()) * 4))) return rc;
    if (!slots.empty())
        GH_HIP_CHECK(hipMemcpyAsync(c.labSlots.p, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, c.stream));
    return build_labs<Real>(c, (int)slots.size(), c.labSlots, field, (int)tab.ent.size(), c.labEnt, c.labSrc, c.labW,
                            out);
}

// The order in which lab_kernel visits the lab cells: by region of the
// 27-neighbourhood (same-level copies, finer and coarser neighbours differ
// per region), then by the parity of the cell's coordinates (a ghost inside a
// coarser leaf sums 1, 2, 4, ... 64 terms by parity), then by position; so
// the lanes of a wave mostly take the same path with the same trip counts.
int lab_order(Ctx& c, int order) {
    if (c.labPermOrder == order) return GH_OK;
    const int LS = lab_start(order), LAB = lab_size(order), lab3 = LAB * LAB * LAB;
    std::vector<std::pair<int, int>> key(lab3);
    for (int p = 0; p < lab3; ++p) {
        const int i[3] = {p % LAB + LS, (p / LAB) % LAB + LS, p / (LAB * LAB) + LS};
        int code = 0, par = 0;
        for (int a = 2; a >= 0; --a) code = 3 * code + (i[a] < 0 ? 0 : (i[a] < 8 ? 1 : 2));
        for (int a = 0; a < 3; ++a) par |= (i[a] & 1) << a;
        key[p] = {code * 8 + par, p};
    }
    std::sort(key.begin(), key.end());
    std::vector<int16_t> perm(lab3);
    for (int q = 0; q < lab3; ++q) perm[q] = (int16_t)key[q].second;
    int rc;
    if ((rc = c.labPerm.ensure((size_t)lab3 * sizeof(int16_t)))) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(c.labPerm.p, perm.data(), perm.size() * sizeof(int16_t), hipMemcpyHostToDevice,
                                c.stream));
    GH_HIP_CHECK(hipStreamSynchronize(c.stream));  // perm is read by the copy
    c.labPermOrder = order;
    return GH_OK;
}

// lab_kernel for the wavelet order and the list size
template <typename Real, int NF, bool FWT>
void launch_labs(Ctx& c, int order, int nlist, const int32_t* dslots, const LevelMaps& m, const LabFields<Real>& lf,
                 double* detail) {
    const bool few = nlist <= kLabThreadsFewMax;
    auto go = [&](auto kern, int threads) {
        hipLaunchKernelGGL(kern, dim3((unsigned)nlist), dim3(threads), 0, c.stream, nlist, dslots,
                           c.dNbr27.as<int32_t>(), c.dBlk.as<int4>(), m, c.labPerm.as<int16_t>(), lf, detail,
                           c.dBad.as<int>());
    };
    if (order == 4) {
        if (few) go(lab_kernel<Real, NF, 4, FWT, 512>, 512);
        else go(lab_kernel<Real, NF, 4, FWT, 256>, 256);
    } else {
        if (few) go(lab_kernel<Real, NF, 2, FWT, 512>, 512);
        else go(lab_kernel<Real, NF, 2, FWT, 256>, 256);
    }
}

// Labs of the four fields (phi, p_w, p_g, p_csf) of nlist slots (dslots on
// the device, or nullptr: slots 0..nlist-1) with the ghosts evaluated on the
// device; may set c.dBad (a ghost deeper than the kernel's recursion: use
// host tables instead).
template <typename Real>
int device_labs(Ctx& c, int nlist, const int32_t* dslots, int nf, const Real* const* fields, DevBuf* const* outs) {
    const int order = c.tree.order;
    const int LAB = lab_size(order), lab3 = LAB * LAB * LAB;
    if (nf != 4) return set_error(GH_EINVAL, "device_labs: 4 fields");
    int rc;
    LabFields<Real> lf{};
    for (int k = 0; k < nf; ++k) {
        if ((rc = outs[k]->ensure_grow(std::max<size_t>(1, (size_t)nlist) * lab3 * sizeof(Real)))) return rc;
        lf.f[k] = fields[k];
        lf.lab[k] = outs[k]->as<Real>();
    }
    if (nlist == 0) return GH_OK;
    if ((rc = lab_order(c, order))) return rc;
    const LevelMaps m = level_maps(c);
    launch_labs<Real, 4, false>(c, order, nlist, dslots, m, lf, nullptr);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

// FWT details of nlist slots (dslots on the device) from labs built in LDS
// and never stored; clears and then may set c.dBad as device_labs
template <typename Real>
int device_details(Ctx& c, int nlist, const int32_t* dslots, const Real* field, double* detail) {
    const int order = c.tree.order;
    int rc;
    if (nlist == 0) return GH_OK;
    if ((rc = lab_order(c, order))) return rc;
    const LevelMaps m = level_maps(c);
    LabFields<Real> lf{};
    lf.f[0] = field;
    launch_labs<Real, 1, true>(c, order, nlist, dslots, m, lf, detail);
    GH_HIP_CHECK(hipGetLastError());
    return GH_OK;
}

// synchronizes; true if no lab since the last clear needed host tables
bool device_labs_ok(Ctx& c) {
    int bad = 1;
    if (hipMemcpyAsync(&bad, c.dBad.p, sizeof(int), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return false;
    return bad == 0;
}

int clear_bad(Ctx& c) {
    GH_HIP_CHECK(hipMemsetAsync(c.dBad.p, 0, sizeof(int), c.stream));
    return GH_OK;
}

// Details of the listed slots: labs of those slots evaluated on the device
// when per-level maps exist, else labs of every leaf from a ghost table cached
// per tree version (an adaptation epoch that changes nothing rebuilds nothing
// on the host); a lab does not depend on which other labs are built, so both
// give the details of labs built for the list alone.
template <typename Real>
int fwt_details_t(Ctx& c, const std::vector<int>& slots, std::vector<double>& det,
                  const std::function<void()>& meanwhile) {
    det.assign(slots.size(), 0.0);
    bool ran = false;
    auto run_meanwhile = [&] {
        if (!ran && meanwhile) meanwhile();
        ran = true;
    };
    if (slots.empty()) {
        run_meanwhile();
        return GH_OK;
    }
    const int order = c.tree.order;
    const int n = c.topo.nblocks();
    int rc;
    DevBuf& d = c.detBuf;
    if ((rc = d.ensure_grow((size_t)n * sizeof(double)))) return rc;
    if (c.nDense > 0) {
        // labs and details of the listed slots only; the recursion check is
        // read back together with the details (one synchronization)
        const int m = (int)slots.size();
        // pinSlots and pinDet are used here only, each call synchronizing
        // before it returns
        if ((rc = c.labSlots.ensure_grow((size_t)m * 4)) || (rc = clear_bad(c)) ||
            (rc = c.pinSlots.ensure_grow((size_t)m * 4)) || (rc = c.pinDet.ensure_grow((size_t)m * 8 + 8)))
            return rc;
        std::memcpy(c.pinSlots.p, slots.data(), (size_t)m * 4);
        GH_HIP_CHECK(hipMemcpyAsync(c.labSlots.p, c.pinSlots.p, (size_t)m * 4, hipMemcpyHostToDevice, c.stream));
        if ((rc = device_details<Real>(c, m, c.labSlots.as<int32_t>(), c.phi[c.cur].as<Real>(), d.as<double>())))
            return rc;
        // details, then the recursion flag after them, into the pinned buffer
        double* pd = c.pinDet.as<double>();
        int* pbad = reinterpret_cast<int*>(pd + m);
        *pbad = 1;
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(pd, d.p, (size_t)m * 8, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(pbad, c.dBad.p, sizeof(int), hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) run_meanwhile();
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
        if (e != hipSuccess) return set_error(GH_EHIP, std::string("fwt: ") + hipGetErrorString(e));
        std::memcpy(det.data(), pd, (size_t)m * 8);
        const int bad = *pbad;
        if (bad == 0) {
            ++c.adapt_count[0];
            return GH_OK;
        }
    }
    ++c.adapt_count[1];
    run_meanwhile();
    // host-built lab table of every slot (cached per tree version)
    std::vector<int> all;  // source of an async upload: lives until the sync below
    if (c.labAllVersion != c.tree_version) {
        all.resize(n);
        for (int i = 0; i < n; ++i) all[i] = i;
        host_neighbours(c);
        HostTable tab = build_lab_table(c.tree, all, c.topo.nbr27, order);
        if ((rc = upload_table<Real>(c, tab, c.labAllEnt, c.labAllSrc, c.labAllW)) ||
            (rc = c.labAllSlots.ensure_grow(std::max<size_t>(1, all.size()) * 4)))
            return rc;
        GH_HIP_CHECK(hipMemcpyAsync(c.labAllSlots.p, all.data(), all.size() * 4, hipMemcpyHostToDevice, c.stream));
        c.nLabAllEnt = (int)tab.ent.size();
        c.labAllVersion = c.tree_version;
    }
    if ((rc = build_labs<Real>(c, n, c.labAllSlots, c.phi[c.cur].as<Real>(), c.nLabAllEnt, c.labAllEnt, c.labAllSrc,
                               c.labAllW, c.labBuf)))
        return rc;
    launch_fwt<Real>(order, n, c.labBuf.as<Real>(), d.as<double>(), c.stream);
    std::vector<double> all_det(n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(all_det.data(), d.p, all_det.size() * 8, hipMemcpyDeviceToHost, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) return set_error(GH_EHIP, std::string("fwt: ") + hipGetErrorString(e));
    for (size_t i = 0; i < slots.size(); ++i) det[i] = all_det[slots[i]];
    return GH_OK;
}

// replace the grid by `nt`: every new leaf is a copy, a split child (from the
// old parent's lab) or a collapse of 8 old leaves; field arrays are rebuilt
// on the device and the per-solve constants recomputed
template <typename Real>
// rm: one Remap per new leaf, in c.pinRemap (pinned; read by the upload
// until apply_topology synchronizes)
int apply_new_tree(Ctx& c, Octree& nt, const Remap* rm, const std::vector<int>& split_parents) {
    const int order = c.tree.order;
    const int LS = lab_start(order), LAB = lab_size(order);
    int rc;
    DevBuf* fields[4] = {&c.phi[c.cur], &c.pw, &c.pg, &c.pcsf};
    DevBuf* labs = c.labSplit;
    const auto t_move = std::chrono::steady_clock::now();
    if (!split_parents.empty()) {
        bool on_device = false;
        if (c.nDense > 0) {
            if ((rc = c.labSlots.ensure_grow(split_parents.size() * 4)) || (rc = clear_bad(c)) ||
                (rc = c.pinSplit.ensure_grow(split_parents.size() * 4)))
                return rc;
            std::memcpy(c.pinSplit.p, split_parents.data(), split_parents.size() * 4);
            GH_HIP_CHECK(hipMemcpyAsync(c.labSlots.p, c.pinSplit.p, split_parents.size() * 4, hipMemcpyHostToDevice,
                                        c.stream));
            const Real* fld[4];
            DevBuf* out[4];
            for (int f = 0; f < 4; ++f) {
                fld[f] = fields[f]->as<Real>();
                out[f] = &labs[f];
            }
            if ((rc = device_labs<Real>(c, (int)split_parents.size(), c.labSlots.as<int32_t>(), 4, fld, out))) return rc;
            on_device = device_labs_ok(c);
        }
        ++c.adapt_count[on_device ? 0 : 1];
        if (!on_device) {
            host_neighbours(c);
            HostTable tab = build_lab_table(c.tree, split_parents, c.topo.nbr27, order);
            if ((rc = upload_table<Real>(c, tab, c.labEnt, c.labSrc, c.labW))) return rc;
            for (int f = 0; f < 4; ++f)
                if ((rc = build_labs_for<Real>(c, split_parents, fields[f]->as<Real>(), tab, labs[f]))) return rc;
        }
    }
    const int nnew = (int)nt.leaves.size();
    DevBuf& drm = c.remapBuf;
    if ((rc = drm.ensure_grow((size_t)nnew * sizeof(Remap)))) return rc;
    GH_HIP_CHECK(hipMemcpyAsync(drm.p, rm, (size_t)nnew * sizeof(Remap), hipMemcpyHostToDevice, c.stream));
    DevBuf* nf = c.spare;  // the new arrays go to the spares, which then swap with the live ones
    RemapFields<Real> rf{};
    for (int f = 0; f < 4; ++f) {
        if ((rc = nf[f].ensure_grow((size_t)nnew * BCELLS * sizeof(Real)))) return rc;
        rf.old[f] = fields[f]->as<Real>();
        rf.lab[f] = labs[f].as<Real>();
        rf.out[f] = nf[f].as<Real>();
    }
    hipLaunchKernelGGL(remap_kernel<Real>, dim3(nnew), dim3(512), 0, c.stream, nnew, drm.as<Remap>(), rf, LS, LAB,
                       order);
    GH_HIP_CHECK(hipGetLastError());
    // rm (in c.pinRemap) and the split list (copied to c.pinSplit) are read by
    // the uploads until the event apply_topology records has passed
    for (int f = 0; f < 4; ++f) std::swap(*fields[f], nf[f]);
    c.prof[P_MOVE] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_move).count();
    {
        ProfScope topo(c.prof[P_TOPO]);
        if ((rc = stage_wait(c))) return rc;  // the old tree's hash table may still be uploading
        c.tree = std::move(nt);                 // nt is the caller's scratch tree
        if ((rc = apply_topology(c))) return rc;
    }
    // the other state buffers follow the new size
    const size_t n = c.ncells() * c.real_size;
    if ((rc = c.phi[1 - c.cur].ensure_grow(n)) || (rc = c.dphidt.ensure_grow(n)) || (rc = c.statD.ensure_grow(4 * n)) ||
        (rc = c.statSig.ensure_grow(n)))
        return rc;
    if (c.prepared) c.constants_stale = true;  // rebuilt by the next RD call (rd_ensure_prepared)
    return GH_OK;
}

}  // namespace

int rd_prepare_ghosted(Ctx& c, double Dw, double Dg) {
    return c.real_kind == GH_REAL_F32 ? prepare_ghosted_t<float>(c, Dw, Dg) : prepare_ghosted_t<double>(c, Dw, Dg);
}

int rd_step_ghosted(Ctx& c, const void* in, void* out, double dt, bool fused) {
    if (c.real_kind == GH_REAL_F32)
        return fused ? step_ghosted_t<float, true>(c, (const float*)in, (float*)out, dt)
                     : step_ghosted_t<float, false>(c, (const float*)in, (float*)out, dt);
    return fused ? step_ghosted_t<double, true>(c, (const double*)in, (double*)out, dt)
                 : step_ghosted_t<double, false>(c, (const double*)in, (double*)out, dt);
}

int fwt_details(Ctx& c, const std::vector<int>& slots, std::vector<double>& det,
                const std::function<void()>& meanwhile) {
    return c.real_kind == GH_REAL_F32 ? fwt_details_t<float>(c, slots, det, meanwhile)
                                      : fwt_details_t<double>(c, slots, det, meanwhile);
}

// Grid::refine with Refiner_SpaceExtension (MRAGrid.inl:247-355)
int grid_refine(Ctx& c, const std::vector<char>& flagged, int* nrefined, std::vector<int>* old_slot,
                const std::vector<char>* ext, std::vector<uint64_t>* split_keys) {
    const auto t_plan = std::chrono::steady_clock::now();
    const Octree& t = c.tree;
    std::vector<int> victims = t.refinement_victims(flagged, ext);
    if (nrefined) *nrefined = (int)victims.size();
    if (victims.empty()) return GH_OK;
    if (split_keys) {
        split_keys->resize(victims.size());
        for (size_t j = 0; j < victims.size(); ++j) (*split_keys)[j] = Octree::key(t.leaves[victims[j]]);
    }
    Octree nt = t.shell();
    nt.leaves.reserve(t.leaves.size() + 7 * victims.size());
    std::vector<char> is_victim(t.leaves.size(), 0);
    for (int v : victims) is_victim[v] = 1;
    std::vector<int> victim_index(t.leaves.size(), -1);  // old slot -> compact parent index
    for (int j = 0; j < (int)victims.size(); ++j) victim_index[victims[j]] = j;
    for (int s = 0; s < (int)t.leaves.size(); ++s)
        if (!is_victim[s]) nt.leaves.push_back(t.leaves[s]);
    for (int j = 0; j < (int)victims.size(); ++j) {
        const BlockRec& b = t.leaves[victims[j]];
        for (int o = 0; o < 8; ++o) {
            BlockRec ch{b.level + 1, {2 * b.index[0] + (o & 1), 2 * b.index[1] + ((o >> 1) & 1),
                                      2 * b.index[2] + ((o >> 2) & 1)}};
            nt.leaves.push_back(ch);
        }
    }
    nt.rebuild_index();
    const size_t nrm = nt.leaves.size();
    int rc;
    if ((rc = stage_wait(c)) || (rc = c.pinRemap.ensure_grow(nrm * sizeof(Remap)))) return rc;
    Remap* rm = c.pinRemap.as<Remap>();
    // the leaves not split keep their (canonical) order: new slots and old
    // slots not split pair up in one pass; the rest are children
    int j = 0;
    const int nold = (int)t.leaves.size();
    auto same = [](const BlockRec& a, const BlockRec& b) {
        return a.level == b.level && a.index[0] == b.index[0] && a.index[1] == b.index[1] && a.index[2] == b.index[2];
    };
    for (int s = 0; s < (int)nrm; ++s) {
        const BlockRec& nb = nt.leaves[s];
        while (j < nold && is_victim[j]) ++j;
        Remap r{};
        if (j < nold && same(nb, t.leaves[j])) {
            r.kind = 0;
            r.src = j++;
        } else {  // a child of a split leaf: the parent is the old leaf one level up
            int ps = -1;
            if (nb.level == 0 || t.kind(nb.level - 1, nb.index[0] >> 1, nb.index[1] >> 1, nb.index[2] >> 1, &ps) !=
                                     Octree::LEAF || victim_index[ps] < 0)
                return set_error(GH_EINVAL, "grid_refine: inconsistent plan");
            r.kind = 1;
            r.src = victim_index[ps];
            r.octant = (nb.index[0] & 1) | (nb.index[1] & 1) << 1 | (nb.index[2] & 1) << 2;
        }
        rm[s] = r;
    }
    while (j < nold && is_victim[j]) ++j;
    if (j != nold) return set_error(GH_EINVAL, "grid_refine: inconsistent plan");
    if (old_slot) {
        old_slot->resize(nrm);
        for (size_t s = 0; s < nrm; ++s) (*old_slot)[s] = rm[s].kind == 0 ? rm[s].src : -1;
    }
    c.prof[P_PLAN] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    return c.real_kind == GH_REAL_F32 ? apply_new_tree<float>(c, nt, rm, victims)
                                      : apply_new_tree<double>(c, nt, rm, victims);
}

// Grid::compress with Compressor (MRAGrid.inl:383-486)
int grid_compress(Ctx& c, const std::vector<char>& flagged, int* ncollapsed, std::vector<uint64_t>* parents_out,
                  std::vector<int>* old_slot, const std::vector<Octree::CollapseCand>* cands) {
    const auto t_plan = std::chrono::steady_clock::now();
    const Octree& t = c.tree;
    const std::vector<Octree::CollapseCand> plan = t.collapse_plan(flagged, cands);  // ascending keys
    std::vector<uint64_t> parents(plan.size());
    for (size_t i = 0; i < plan.size(); ++i) parents[i] = plan[i].key;
    if (ncollapsed) *ncollapsed = (int)parents.size();
    if (parents_out) *parents_out = parents;
    if (parents.empty()) return GH_OK;
    Octree nt = t.shell();
    std::vector<char> gone(t.leaves.size(), 0);
    for (const Octree::CollapseCand& p : plan)
        for (int o = 0; o < 8; ++o) gone[p.kids[o]] = 1;
    nt.leaves.reserve(t.leaves.size() - 7 * parents.size());
    for (int s = 0; s < (int)t.leaves.size(); ++s)
        if (!gone[s]) nt.leaves.push_back(t.leaves[s]);
    for (uint64_t pk : parents)
        nt.leaves.push_back(BlockRec{(int)(pk >> 48), {(int)((pk >> 32) & 0xffff), (int)((pk >> 16) & 0xffff),
                                                       (int)(pk & 0xffff)}});
    nt.rebuild_index();
    const size_t nrm = nt.leaves.size();
    int rc;
    if ((rc = stage_wait(c)) || (rc = c.pinRemap.ensure_grow(nrm * sizeof(Remap)))) return rc;
    Remap* rm = c.pinRemap.as<Remap>();
    // the surviving leaves keep their (canonical) order: new slots and old
    // slots not collapsed pair up in one pass; the rest are the parents
    int j = 0;
    const int nold = (int)t.leaves.size();
    auto same = [](const BlockRec& a, const BlockRec& b) {
        return a.level == b.level && a.index[0] == b.index[0] && a.index[1] == b.index[1] && a.index[2] == b.index[2];
    };
    for (int s = 0; s < (int)nrm; ++s) {
        const BlockRec& nb = nt.leaves[s];
        while (j < nold && gone[j]) ++j;
        Remap r{};
        if (j < nold && same(nb, t.leaves[j])) {
            r.kind = 0;
            r.src = j++;
        } else {
            r.kind = 2;
            const uint64_t k = Octree::key(nb);
            const auto it = std::lower_bound(parents.begin(), parents.end(), k);
            if (it == parents.end() || *it != k) return set_error(GH_EINVAL, "grid_compress: inconsistent plan");
            const Octree::CollapseCand& p = plan[it - parents.begin()];
            for (int o = 0; o < 8; ++o) r.kids[o] = p.kids[o];
        }
        rm[s] = r;
    }
    while (j < nold && gone[j]) ++j;
    if (j != nold) return set_error(GH_EINVAL, "grid_compress: inconsistent plan");
    if (old_slot) {
        old_slot->resize(nrm);
        for (size_t s = 0; s < nrm; ++s) (*old_slot)[s] = rm[s].kind == 0 ? rm[s].src : -1;
    }
    c.prof[P_PLAN] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    return c.real_kind == GH_REAL_F32 ? apply_new_tree<float>(c, nt, rm, {})
                                      : apply_new_tree<double>(c, nt, rm, {});
}

}  // namespace gh

using namespace gh;

extern "C" {

int gh_grid_configure(gh_ctx* ctx, int wavelet_order, int max_level, int min_level) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if ((wavelet_order != 2 && wavelet_order != 4) || max_level < 0 || max_level > 9 || min_level < 0 ||
        min_level > max_level)
        return set_error(GH_EINVAL, "gh_grid_configure: order 2|4, 0 <= min_level <= max_level <= 9");
    if (c->tree.order != wavelet_order) c->labAllVersion = ~(uint64_t)0;  // labs are sized by the order
    c->tree.order = wavelet_order;
    c->tree.max_level = max_level;
    c->tree.min_level = min_level;
    return GH_OK;
}

int gh_grid_init_leaves(gh_ctx* ctx, int nleaves, const int* level, const int* index) {
    Ctx* c = as_ctx(ctx);
    if (!c || nleaves <= 0 || !level || !index) return set_error(GH_EINVAL, "gh_grid_init_leaves: bad argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    Octree t = c->tree.shell();
    double vol = 0;
    for (int i = 0; i < nleaves; ++i) {
        const int L = level[i];
        if (L < 0 || L > 9) return set_error(GH_EINVAL, "gh_grid_init_leaves: level out of range");
        for (int a = 0; a < 3; ++a)
            if (index[3 * i + a] < 0 || index[3 * i + a] >= (1 << L))
                return set_error(GH_EINVAL, "gh_grid_init_leaves: index out of range");
        t.leaves.push_back(BlockRec{L, {index[3 * i], index[3 * i + 1], index[3 * i + 2]}});
        vol += std::ldexp(1.0, -3 * L);
    }
    t.rebuild_index();
    for (int s = 1; s < nleaves; ++s)  // canonical order puts equal leaves next to each other
        if (Octree::key(t.leaves[s]) == Octree::key(t.leaves[s - 1]))
            return set_error(GH_EINVAL, "gh_grid_init_leaves: duplicate leaf");
    if (t.nested) return set_error(GH_EINVAL, "gh_grid_init_leaves: a leaf contains another leaf");
    if (vol != 1.0) return set_error(GH_EINVAL, "gh_grid_init_leaves: leaves do not tile the domain");
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));  // the old tree's tables may still be uploading
    c->tree = t;
    int rc = apply_topology(*c);
    if (rc) return rc;
    return c->alloc_fields();
}

int gh_fwt_details(gh_ctx* ctx, double* detail) {
    Ctx* c = as_ctx(ctx);
    if (!c || !detail) return set_error(GH_EINVAL, "null argument");
    GH_HIP_CHECK(hipSetDevice(c->device));
    std::vector<int> all(c->topo.nblocks());
    for (size_t i = 0; i < all.size(); ++i) all[i] = (int)i;
    std::vector<double> d;
    int rc = fwt_details(*c, all, d);
    if (rc) return rc;
    std::copy(d.begin(), d.end(), detail);
    return GH_OK;
}

// Which blocks a set of changed boxes can reach.  A block's lab -- hence its
// FWT details -- reads the leaves within its own box widened by two of its
// block widths on every side (periodically): the lab spans [-2, 11) cells of
// the block's level, and a ghost in a coarser region adds at most two
// synthesis levels of 3 and 6 of those cells (order 4; less for order 2), 12 <
// 16 cells in all; a finer region answers with the coincident point.  A
// compression changes the tree and the data only inside the collapsed
// parents' boxes, a refinement only inside the split blocks' boxes, so a
// block whose widened box misses all of them keeps its lab bit for bit.
// The boxes are rasterized at level R = min(finest, 7).  Up to R = 6 as rows
// of x bits: the changed cells of level R are OR-reduced to every coarser
// level L and widened there by two cells each way (periodic), so that a block
// of level L is one bit test.  At R = 7 with a summed-volume table, blocks
// finer than R tested conservatively.
class ChangedBoxes {
  public:
    ChangedBoxes(int finest, const std::vector<uint64_t>& boxes) : R_(std::min(finest, 7)), n_(1 << R_) {
        if (R_ <= 6) {
            build_rows(boxes);
            return;
        }
        const size_t m = (size_t)(n_ + 1);
        sum_.assign(m * m * m, 0);
        for (uint64_t k : boxes) {
            int lo[3], hi[3];
            raster(k, lo, hi);
            for (int z = lo[2]; z < hi[2]; ++z)
                for (int y = lo[1]; y < hi[1]; ++y)
                    for (int x = lo[0]; x < hi[0]; ++x) at(x + 1, y + 1, z + 1) = 1;
        }
        for (int z = 1; z <= n_; ++z)  // summed-volume table
            for (int y = 1; y <= n_; ++y)
                for (int x = 1; x <= n_; ++x)
                    at(x, y, z) += at(x - 1, y, z) + at(x, y - 1, z) + at(x, y, z - 1) - at(x - 1, y - 1, z) -
                                   at(x - 1, y, z - 1) - at(x, y - 1, z - 1) + at(x - 1, y - 1, z - 1);
    }
    // true if block (L, idx) widened by two of its widths meets a changed box
    bool reaches(const BlockRec& b) const {
        if (R_ <= 6) {  // rows of the block's level (none finer than R: conservatively reached)
            const int L = b.level;
            if (L > R_) return true;
            return (wide_[L][(size_t)b.index[1] + ((size_t)b.index[2] << L)] >> b.index[0]) & 1;
        }
        int lo[2][3], hi[2][3], nseg[3];
        for (int a = 0; a < 3; ++a) {
            int l, h;  // [l, h) at level R, before the periodic wrap
            if (b.level <= R_) {
                l = (b.index[a] - 2) << (R_ - b.level);
                h = (b.index[a] + 3) << (R_ - b.level);
            } else {
                const int sh = b.level - R_;
                l = (b.index[a] - 2) >> sh;                  // floor
                h = (b.index[a] + 3 + (1 << sh) - 1) >> sh;  // ceil
            }
            const int len = h - l;
            nseg[a] = 1;
            if (len >= n_) {
                lo[0][a] = 0;
                hi[0][a] = n_;
                continue;
            }
            l = ((l % n_) + n_) % n_;
            h = l + len;
            lo[0][a] = l;
            hi[0][a] = std::min(h, n_);
            if (h > n_) {
                lo[1][a] = 0;
                hi[1][a] = h - n_;
                nseg[a] = 2;
            }
        }
        for (int i = 0; i < nseg[0]; ++i)
            for (int j = 0; j < nseg[1]; ++j)
                for (int k = 0; k < nseg[2]; ++k)
                    if (box_sum(lo[i][0], hi[i][0], lo[j][1], hi[j][1], lo[k][2], hi[k][2]) > 0) return true;
        return false;
    }

  private:
    int R_, n_;
    std::vector<int32_t> sum_;
    std::vector<uint64_t> wide_[7];  // per level L <= R: rows of x bits, the cells a changed box can reach
    // cells [lo, hi) of level R covered by box k (the containing cell for a box finer than R)
    void raster(uint64_t k, int lo[3], int hi[3]) const {
        const int L = (int)(k >> 48);
        const int idx[3] = {(int)((k >> 32) & 0xffff), (int)((k >> 16) & 0xffff), (int)(k & 0xffff)};
        for (int a = 0; a < 3; ++a) {
            if (L <= R_) {
                lo[a] = idx[a] << (R_ - L);
                hi[a] = (idx[a] + 1) << (R_ - L);
            } else {
                lo[a] = idx[a] >> (L - R_);
                hi[a] = lo[a] + 1;
            }
        }
    }
    static uint64_t even_bits(uint64_t v) {  // bit k of the result = bit 2k of v
        v &= 0x5555555555555555ull;
        v = (v | (v >> 1)) & 0x3333333333333333ull;
        v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
        v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
        v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
        return (v | (v >> 16)) & 0x00000000ffffffffull;
    }
    void build_rows(const std::vector<uint64_t>& boxes) {
        std::vector<uint64_t> cells((size_t)n_ * n_, 0);  // the changed cells of level R
        for (uint64_t k : boxes) {
            int lo[3], hi[3];
            raster(k, lo, hi);
            const uint64_t span = (hi[0] - lo[0] == 64 ? ~0ull : ((1ull << (hi[0] - lo[0])) - 1)) << lo[0];
            for (int z = lo[2]; z < hi[2]; ++z)
                for (int y = lo[1]; y < hi[1]; ++y) cells[(size_t)y + ((size_t)z << R_)] |= span;
        }
        std::vector<uint64_t> wx, wy;
        for (int L = R_; L >= 0; --L) {
            const int n = 1 << L;
            const uint64_t full = n == 64 ? ~0ull : (1ull << n) - 1;
            if (L < R_) {  // OR-reduce the level L + 1 rows
                std::vector<uint64_t> up((size_t)n * n);
                const size_t dz = (size_t)1 << (L + 1);
                for (int z = 0; z < n; ++z)
                    for (int y = 0; y < n; ++y) {
                        const size_t r0 = (size_t)(2 * y) + (size_t)(2 * z) * dz;
                        const uint64_t v = cells[r0] | cells[r0 + 1] | cells[r0 + dz] | cells[r0 + dz + 1];
                        up[(size_t)y + ((size_t)z << L)] = even_bits(v | v >> 1);
                    }
                cells.swap(up);
            }
            // widen by two cells each way, periodically: x by rotations, then y, then z
            auto rot = [&](uint64_t v, int k) {  // rotate the n-bit row v left by k (mod n)
                k = ((k % n) + n) % n;
                return k == 0 ? v : ((v << k) | (v >> (n - k))) & full;
            };
            wx.assign((size_t)n * n, 0);
            for (size_t r = 0; r < (size_t)n * n; ++r) {
                const uint64_t v = cells[r];
                wx[r] = v | rot(v, 1) | rot(v, 2) | rot(v, -1) | rot(v, -2);
            }
            wy.assign((size_t)n * n, 0);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    uint64_t v = 0;
                    for (int d = -2; d <= 2; ++d) v |= wx[(size_t)((y + d) & (n - 1)) + ((size_t)z << L)];
                    wy[(size_t)y + ((size_t)z << L)] = v;
                }
            std::vector<uint64_t>& w = wide_[L];
            w.assign((size_t)n * n, 0);
            for (int z = 0; z < n; ++z)
                for (int y = 0; y < n; ++y) {
                    uint64_t v = 0;
                    for (int d = -2; d <= 2; ++d) v |= wy[(size_t)y + ((size_t)((z + d) & (n - 1)) << L)];
                    w[(size_t)y + ((size_t)z << L)] = v;
                }
        }
    }
    int32_t& at(int x, int y, int z) { return sum_[(size_t)x + (size_t)(n_ + 1) * ((size_t)y + (size_t)(n_ + 1) * z)]; }
    int32_t get(int x, int y, int z) const {
        return sum_[(size_t)x + (size_t)(n_ + 1) * ((size_t)y + (size_t)(n_ + 1) * z)];
    }
    int32_t box_sum(int x0, int x1, int y0, int y1, int z0, int z1) const {
        return get(x1, y1, z1) - get(x0, y1, z1) - get(x1, y0, z1) - get(x1, y1, z0) + get(x0, y0, z1) +
               get(x0, y1, z0) + get(x1, y0, z0) - get(x0, y0, z0);
    }
};

// Science::AutomaticRefinement<0,0>(grid, fwt, tol, maxLevel, loops) (MRAGScienceCore.h:42-108)
int gh_refine(gh_ctx* ctx, double tol, int max_loops, int* nrefined) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    GH_HIP_CHECK(hipSetDevice(c->device));
    // blocks found nice (detail < tol) stay nice while they exist: a flag per
    // slot, carried to the surviving leaves' new slots after each refinement
    std::vector<char> nice(c->topo.nblocks(), 0), carried;
    std::vector<int> old_slot;
    // the details of the last evaluation, per slot of the current tree, where
    // still valid (for the compression that may follow: see Ctx::carry_call)
    c->carry_known.clear();
    c->carry_det.clear();
    std::vector<char> known;
    std::vector<double> kdet;
    int loops = 0, total = 0;
    while (true) {
        std::vector<int> cand;
        for (int s = 0; s < c->topo.nblocks(); ++s) {
            const BlockRec& b = c->tree.leaves[s];
            if (c->tree.max_level >= 0 && c->tree.max_level <= b.level) continue;
            if (!nice[s]) cand.push_back(s);
        }
        if (cand.empty()) break;
        std::vector<double> det;
        std::vector<char> ext;  // the space extension, found while the device evaluates the details
        int rc;
        bool host_tables = false;
        {
            ProfScope fwt(c->prof[P_REF_FWT]);
            const int64_t host_before = c->adapt_count[1];
            if ((rc = fwt_details(*c, cand, det, [&] { ext = c->tree.space_extension(); }))) return rc;
            host_tables = c->adapt_count[1] != host_before;
        }
        known.assign(c->topo.nblocks(), 0);
        kdet.assign(c->topo.nblocks(), 0.0);
        if (!host_tables)  // (host-built labs are not bounded by the two-block reach)
            for (size_t i = 0; i < cand.size(); ++i) {
                known[cand[i]] = 1;
                kdet[cand[i]] = det[i];
            }
        std::vector<char> flagged(c->topo.nblocks(), 0);
        int nflag = 0;
        for (size_t i = 0; i < cand.size(); ++i) {
            if (det[i] >= tol) {
                flagged[cand[i]] = 1;
                ++nflag;
            } else {
                nice[cand[i]] = 1;
            }
        }
        if (nflag == 0) break;
        int nr = 0;
        old_slot.clear();
        std::vector<uint64_t> split;
        const int finest = c->tree.top_level;
        {
            ProfScope ref(c->prof[P_REF]);
            if ((rc = grid_refine(*c, flagged, &nr, &old_slot, &ext, &split))) return rc;
        }
        if (!old_slot.empty()) {
            carried.assign(old_slot.size(), 0);
            for (size_t s = 0; s < old_slot.size(); ++s)
                if (old_slot[s] >= 0) carried[s] = nice[old_slot[s]];
            nice.swap(carried);
            // details the split boxes cannot reach stay valid on the new tree
            const ChangedBoxes changed(finest, split);
            std::vector<char> known2(old_slot.size(), 0);
            std::vector<double> kdet2(old_slot.size(), 0.0);
            for (size_t s = 0; s < old_slot.size(); ++s) {
                const int o = old_slot[s];
                if (o >= 0 && known[o] && !changed.reaches(c->tree.leaves[s])) {
                    known2[s] = 1;
                    kdet2[s] = kdet[o];
                }
            }
            known.swap(known2);
            kdet.swap(kdet2);
        }
        total += nr;
        ++loops;
        if (max_loops > 0 && loops >= max_loops) break;
    }
    if (nrefined) *nrefined = total;
    if (known.size() == (size_t)c->topo.nblocks()) {
        c->carry_known.swap(known);
        c->carry_det.swap(kdet);
        c->carry_call = c->calls;
    }
    return GH_OK;
}

// Science::AutomaticCompression<0,0>(grid, fwt, tol, loops) (MRAGScienceCore.h:117-177).
// Every loop needs the details of every block not yet found nice; a block
// whose lab no collapse of the previous loop could reach (ChangedBoxes) keeps
// the detail computed then -- the same lab, so the same value -- and only the
// others are evaluated again (all of them when the previous loop's labs came
// from host tables, whose recursion is not bounded by two levels).
int gh_compress(gh_ctx* ctx, double tol, int max_loops, int* ncollapsed) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    GH_HIP_CHECK(hipSetDevice(c->device));
    // per slot: found nice; a detail still valid on the current tree (known);
    // both carried to the surviving leaves' new slots after each compression
    std::vector<char> nice(c->topo.nblocks(), 0), known(c->topo.nblocks(), 0), nice2, known2;
    std::vector<double> kdet(c->topo.nblocks(), 0.0), kdet2;
    // a refinement called just before (no other call on this context in
    // between) leaves the details still valid on this tree and data
    if (c->carry_call + 1 == c->calls && c->carry_known.size() == known.size()) {
        known.swap(c->carry_known);
        kdet.swap(c->carry_det);
    }
    c->carry_known.clear();
    c->carry_det.clear();
    std::vector<int> old_slot;
    int loops = 0, total = 0;
    while (true) {
        std::vector<int> cand, need;
        std::vector<size_t> need_at;
        std::vector<double> det;
        for (int s = 0; s < c->topo.nblocks(); ++s)
            if (!nice[s]) cand.push_back(s);
        det.assign(cand.size(), 0.0);
        for (size_t i = 0; i < cand.size(); ++i) {
            if (known[cand[i]]) {
                det[i] = kdet[cand[i]];
            } else {
                need.push_back(cand[i]);
                need_at.push_back(i);
            }
        }
        int rc;
        bool host_tables = false;
        std::vector<Octree::CollapseCand> pcand;  // found while the device evaluates the details
        {
            ProfScope fwt(c->prof[P_COMP_FWT]);
            std::vector<double> dn;
            const int64_t host_before = c->adapt_count[1];
            if ((rc = fwt_details(*c, need, dn, [&] { pcand = c->tree.collapse_candidates(); }))) return rc;
            host_tables = c->adapt_count[1] != host_before;
            for (size_t j = 0; j < need.size(); ++j) det[need_at[j]] = dn[j];
        }
        std::vector<char> flagged(c->topo.nblocks(), 0);
        int nflag = 0;
        for (size_t i = 0; i < cand.size(); ++i) {
            if (det[i] < tol) {
                flagged[cand[i]] = 1;
                ++nflag;
            } else {
                nice[cand[i]] = 1;
            }
        }
        if (nflag == 0) break;
        int nc = 0;
        std::vector<uint64_t> parents;
        const int finest = c->tree.top_level;
        old_slot.clear();
        {
            ProfScope comp(c->prof[P_COMP]);
            if ((rc = grid_compress(*c, flagged, &nc, &parents, &old_slot, &pcand))) return rc;
        }
        if (nc == 0) break;
        total += nc;
        ++loops;
        if (max_loops > 0 && loops >= max_loops) break;
        // this loop's details of the candidates, kept where no collapse reaches
        // the block's lab; the other blocks (and all of them after host-built
        // labs) are evaluated again
        std::vector<char> fresh(known.size(), 0);
        for (size_t i = 0; i < cand.size(); ++i) {
            fresh[cand[i]] = 1;
            kdet[cand[i]] = det[i];
        }
        const size_t nn = old_slot.size();
        nice2.assign(nn, 0);
        known2.assign(nn, 0);
        kdet2.assign(nn, 0.0);
        std::unique_ptr<ChangedBoxes> changed;
        if (!host_tables) changed.reset(new ChangedBoxes(finest, parents));
        for (size_t s = 0; s < nn; ++s) {
            const int o = old_slot[s];
            if (o < 0) continue;
            nice2[s] = nice[o];
            if (changed && fresh[o] && !changed->reaches(c->tree.leaves[s])) {
                known2[s] = 1;
                kdet2[s] = kdet[o];
            }
        }
        nice.swap(nice2);
        known.swap(known2);
        kdet.swap(kdet2);
    }
    if (ncollapsed) *ncollapsed = total;
    // the grid a compression leaves is the one that steps next: its constants
    // (and those of any refinement before it) are built here, once
    ProfScope stat(c->prof[P_STAT]);
    return rd_ensure_prepared(*c);
}

// the listed blocks as per-slot flags; ids outside [0, nblocks) are refused
static int flags_of(Ctx* c, int n, const int* ids, std::vector<char>& flagged) {
    if (n < 0 || (n > 0 && !ids)) return set_error(GH_EINVAL, "block list: bad size or null pointer");
    const int nb = c->topo.nblocks();
    flagged.assign(nb, 0);
    for (int i = 0; i < n; ++i) {
        if (ids[i] < 0 || ids[i] >= nb) return set_error(GH_EINVAL, "block list: block id out of range");
        flagged[ids[i]] = 1;
    }
    return GH_OK;
}

// Grid::refine(const set<int>&) (MRAGrid.inl:247-355)
int gh_grid_refine_blocks(gh_ctx* ctx, int n, const int* block_ids, int* nrefined) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.nblocks() == 0) return set_error(GH_ESTATE, "gh_grid_refine_blocks: no grid");
    std::vector<char> flagged;
    int rc = flags_of(c, n, block_ids, flagged);
    if (rc) return rc;
    GH_HIP_CHECK(hipSetDevice(c->device));
    int nr = 0;
    {
        ProfScope ref(c->prof[P_REF]);
        if ((rc = grid_refine(*c, flagged, &nr))) return rc;
    }
    if (nrefined) *nrefined = nr;
    return GH_OK;
}

// Grid::compress(const set<int>&, int& nCollapsedBlocks) (MRAGrid.inl:383-486)
int gh_grid_compress_blocks(gh_ctx* ctx, int n, const int* block_ids, int* ncollapsed) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (c->topo.nblocks() == 0) return set_error(GH_ESTATE, "gh_grid_compress_blocks: no grid");
    std::vector<char> flagged;
    int rc = flags_of(c, n, block_ids, flagged);
    if (rc) return rc;
    GH_HIP_CHECK(hipSetDevice(c->device));
    int nc = 0;
    {
        ProfScope comp(c->prof[P_COMP]);
        if ((rc = grid_compress(*c, flagged, &nc))) return rc;
    }
    if (ncollapsed) *ncollapsed = nc;
    return GH_OK;
}

int gh_adapt_counters(gh_ctx* ctx, int64_t* counts, int reset) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (counts)
        for (int i = 0; i < 4; ++i) counts[i] = c->adapt_count[i];
    if (reset)
        for (int64_t& v : c->adapt_count) v = 0;
    return GH_OK;
}

int gh_adapt_profile(gh_ctx* ctx, double* ms, int reset) {
    Ctx* c = as_ctx(ctx);
    if (!c) return set_error(GH_EINVAL, "null context");
    if (ms)
        for (int i = 0; i < 8; ++i) ms[i] = c->prof[i];
    if (reset)
        for (double& v : c->prof) v = 0;
    return GH_OK;
}

int gh_set_field_blocks(gh_ctx* ctx, int field, const void* values) {
    Ctx* c = as_ctx(ctx);
    if (!c || !values) return set_error(GH_EINVAL, "null argument");
    DevBuf* b = field == GH_FIELD_PHI ? &c->phi[c->cur]
              : field == GH_FIELD_PW  ? &c->pw
              : field == GH_FIELD_PG  ? &c->pg
              : field == GH_FIELD_PCSF ? &c->pcsf
              : field == GH_FIELD_DPHIDT ? &c->dphidt : nullptr;
    if (!b) return set_error(GH_EINVAL, "gh_set_field_blocks: bad field");
    GH_HIP_CHECK(hipSetDevice(c->device));
    GH_HIP_CHECK(hipMemcpyAsync(b->p, values, c->ncells() * c->real_size, hipMemcpyHostToDevice, c->stream));
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (field == GH_FIELD_PW || field == GH_FIELD_PG) c->prepared = false;
    return GH_OK;
}

int gh_get_field_blocks(gh_ctx* ctx, int field, void* values) {
    Ctx* c = as_ctx(ctx);
    if (!c || !values) return set_error(GH_EINVAL, "null argument");
    DevBuf* b = field == GH_FIELD_PHI ? &c->phi[c->cur]
              : field == GH_FIELD_PW  ? &c->pw
              : field == GH_FIELD_PG  ? &c->pg
              : field == GH_FIELD_PCSF ? &c->pcsf
              : field == GH_FIELD_DPHIDT ? &c->dphidt : nullptr;
    if (!b) return set_error(GH_EINVAL, "gh_get_field_blocks: bad field");
    GH_HIP_CHECK(hipSetDevice(c->device));
    GH_HIP_CHECK(hipMemcpyAsync(values, b->p, c->ncells() * c->real_size, hipMemcpyDeviceToHost, c->stream));
    GH_HIP_CHECK(hipStreamSynchronize(c->stream));
    return GH_OK;
}

int gh_uq_resample(gh_ctx* ctx, int gpd, double* out) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out || gpd <= 0 || gpd > 2048) return set_error(GH_EINVAL, "gh_uq_resample: bad argument");
    if ((size_t)c->topo.nblocks() * BCELLS >= (size_t)INT32_MAX)
        return set_error(GH_EINVAL, "gh_uq_resample: grid too large");
    GH_HIP_CHECK(hipSetDevice(c->device));
    const size_t n = (size_t)gpd * gpd * gpd;
    DevBuf d, own, blk;
    int rc;
    if ((rc = d.ensure(n * 8)) || (rc = own.ensure(n * sizeof(int))) ||
        (rc = blk.ensure(c->topo.blocks.size() * sizeof(BlockRec))))
        return rc;
    GH_HIP_CHECK(hipMemsetAsync(d.p, 0, n * 8, c->stream));
    GH_HIP_CHECK(hipMemsetAsync(own.p, 0xff, n * sizeof(int), c->stream));  // -1: no writer
    GH_HIP_CHECK(hipMemcpyAsync(blk.p, c->topo.blocks.data(), c->topo.blocks.size() * sizeof(BlockRec),
                                hipMemcpyHostToDevice, c->stream));
    // where several cells land on one voxel (blocks finer than 1/gpd) the
    // last writer of the reference's loop order wins: owner pass, then write pass
    const dim3 g(c->topo.nblocks()), b(512);
    const int ns = c->topo.nblocks();
    if (c->real_kind == GH_REAL_F32) {
        hipLaunchKernelGGL((uq_resample_kernel<float, false>), g, b, 0, c->stream, ns, blk.as<BlockRec>(),
                           c->phi[c->cur].as<float>(), gpd, own.as<int>(), d.as<double>());
        hipLaunchKernelGGL((uq_resample_kernel<float, true>), g, b, 0, c->stream, ns, blk.as<BlockRec>(),
                           c->phi[c->cur].as<float>(), gpd, own.as<int>(), d.as<double>());
    } else {
        hipLaunchKernelGGL((uq_resample_kernel<double, false>), g, b, 0, c->stream, ns, blk.as<BlockRec>(),
                           c->phi[c->cur].as<double>(), gpd, own.as<int>(), d.as<double>());
        hipLaunchKernelGGL((uq_resample_kernel<double, true>), g, b, 0, c->stream, ns, blk.as<BlockRec>(),
                           c->phi[c->cur].as<double>(), gpd, own.as<int>(), d.as<double>());
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d.p, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    d.release();
    own.release();
    blk.release();
    if (e != hipSuccess) return set_error(GH_EHIP, std::string("gh_uq_resample: ") + hipGetErrorString(e));
    return GH_OK;
}

}  // extern "C"

