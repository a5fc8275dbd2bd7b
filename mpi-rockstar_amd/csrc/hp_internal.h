// hp_internal.h -- device layouts and helpers shared by the halopot kernels.
//
// Everything here is compiled for gfx950 (wave64) with -ffp-contract=off:
// float and double expressions that must round exactly like the reference's
// C99 code (bounding boxes, split planes, centres of mass, dmin, acceptance)
// are written as plain operators and are never fused; the hot pair loops use
// explicit fmaf where fusing is wanted.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/halopot.h"

namespace hp {

constexpr int kLeafPoints = 10;   // POINTS_PER_LEAF, src/potential.c:57
constexpr int kDegenerate = 20;   // 2 * POINTS_PER_LEAF, src/potential.c:145
constexpr int kWave = 64;
constexpr int kGroup = 64;        // sink particles per walk wavefront (one per lane)
constexpr int kTile = 1024;       // points per workgroup tile of a big node
constexpr int kSmallMax = 1024;   // largest subtree finished by one workgroup
constexpr int kSmallChunks = kSmallMax / 64;
constexpr int kSrcBatch = 8;      // leaf points the walk reads per scalar batch
// Call-size buckets of the launch setting: calls of at most 1,024, 8,192,
// 65,536 and more records (HP_TUNE_BUCKETS in the header).
constexpr int kTuneBuckets = 4;
__host__ __device__ __forceinline__ int tune_bucket(int64_t count) {
    return count <= 1024 ? 0 : (count <= 8192 ? 1 : (count <= 65536 ? 2 : 3));
}
constexpr int kFatLeaf = 64;      // leaves above this size get a wavefront
constexpr int kTileThreads = 256;
constexpr int kMaxLevels = 4096;  // depth_end entries kept per build
// The level buffer: depth_end[kMaxLevels], then the split levels' ranges.
// The big tasks of split level L are [lvl_big[L - 1], lvl_big[L]) and their
// tiles [lvl_tile[L - 1], lvl_tile[L]); entry 0 is 0.
constexpr int kLevelWords = 3 * kMaxLevels + 4;
__host__ __device__ inline int32_t *lvl_big(int32_t *lv) { return lv + kMaxLevels; }
__host__ __device__ inline int32_t *lvl_tile(int32_t *lv) { return lv + 2 * kMaxLevels + 2; }
__host__ __device__ inline const int32_t *lvl_tile(const int32_t *lv) { return lv + 2 * kMaxLevels + 2; }

// Tree node, 64 bytes, read by the walk with scalar loads.
//   [first, first + count) is the node's range in tree order (batch-global);
//   left is the first child (the right child is left + 1) or -1 for a leaf;
//   skip is the node that follows the subtree in depth-first order (-1: end).
// What the walk reads of a node comes first (NodeHead, 32 bytes), so that a
// pair of siblings is read as two 32-byte loads from one 128-byte line.
struct __attribute__((aligned(16))) Node {
    float com[3];
    float mass;
    float dmin2;     // float(dmin) * float(dmin), src/potential.c:128
    int32_t first;
    int32_t count;
    int32_t left;
    float lo[3];
    float hi[3];
    int32_t skip;
    float dmin;
};
struct NodeHead {
    float com[3];
    float mass;
    float dmin2;
    int32_t first;
    int32_t count;
    int32_t left;
};
static_assert(sizeof(Node) == 64 && sizeof(NodeHead) == 32 && offsetof(Node, com) == 0 &&
                  offsetof(Node, mass) == offsetof(NodeHead, mass) &&
                  offsetof(Node, dmin2) == offsetof(NodeHead, dmin2) &&
                  offsetof(Node, first) == offsetof(NodeHead, first) &&
                  offsetof(Node, count) == offsetof(NodeHead, count) &&
                  offsetof(Node, left) == offsetof(NodeHead, left),
              "node layout");

// A node waiting to be split at the current level.
struct Task {
    int32_t node;
    int32_t dim;
    float lim;
    int32_t tile0;   // first tile (big tasks only)
    int32_t ntiles;
    int32_t depth;
    int32_t pad[2];
};

// Outcome of a big task's count pass, read by its scatter pass.
struct BigSplit {
    int32_t nleft;
    int32_t left;      // child id, or -1 when the split failed (node is a leaf)
    int32_t leafL;     // left child is a leaf (points go straight to tree order)
    int32_t leafR;
};

// Per-tile partial of a big task: left count and the bounding boxes of the
// left and right parts.
struct TilePart {
    int32_t nleft;
    int32_t pad[3];
    float lo[2][3];
    float hi[2][3];
};

// One walk wavefront: sink records [p0, min(p0 + kGroup, pend)) of the halo
// whose root node is `root`.
struct Group {
    int32_t root;
    int32_t p0;
    int32_t pend;
    int32_t pad;
};

// Device-side counters of one build.
enum Counter {
    kNodeCount = 0,
    kSmallNext = 1,   // wavefront subtree tasks (<= wave_max points), append-only
    kBigNext = 2,     // big tasks pushed for the next level (append-only)
    kTileNext = 3,    // tiles of big tasks (append-only)
    kMaxDepth = 4,
    kLeafCount = 5,
    kErrFlags = 6,
    kBigDmin = 7,     // nodes whose dmin is reduced by tiles
    kFatCount = 8,    // leaves of more than kFatLeaf points (failed splits)
    // level state, advanced on the device between levels
    kGlobalNodes = 9, // nodes made by the split levels (subtree nodes follow)
    // 10-13 unused (the split levels' ranges are in the level buffer)
    kLevel = 14,      // levels advanced so far
    kActive = 15,     // the current level has big tasks
    kBadRoots = 16,   // calls with non-finite points, listed for their own pass
    kGroupNext = 17,  // workgroup subtree tasks (wave_max < n <= big_threshold), append-only
    kClassNext = 18,    // [18, 21): workgroup subtree tasks of each kernel size
    kClassCursor = 21,  // [21, 24): next task of each size to start (dynamic queues)
    kWaveIdBase = 24,   // first node id of the wavefront subtrees' block
    kNumCounters = 25
};

// kErrDeep: the split levels recorded in a graph did not finish the trees;
// every kernel that writes the caller's records, halos or bound flags then
// returns at once, so the caller's buffers are untouched and the step can be
// run again.
enum ErrFlag { kErrWalkCap = 1, kErrBadTree = 2, kErrCapacity = 4, kErrDeep = 8 };

// ----------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// number of set bits of m below this lane
__device__ __forceinline__ int rank_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wavefront reductions on DPP lane moves (no LDS round trips): swap within
// pairs and quads, mirror within half rows and rows (every lane then holds
// its row's result), then combine the four row results read from lanes 0,
// 16, 32 and 48.  The result is wavefront-uniform.  Call with all 64 lanes
// active, in wavefront-uniform control flow.
template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit lanes");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, dpp_i32<CTRL>(__builtin_bit_cast(int, v)));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)dpp_i32<CTRL>((int)(uint32_t)u);
        const uint32_t hi = (uint32_t)dpp_i32<CTRL>((int)(uint32_t)(u >> 32));
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

template <typename T>
__device__ __forceinline__ T read_lane(T v, int l) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    } else {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
        return __builtin_bit_cast(T, ((uint64_t)hi << 32) | lo);
    }
}

template <typename T, typename Op>
__device__ __forceinline__ T wave_reduce(T v, Op op) {
    v = op(v, dpp_move<0xB1>(v));   // quad_perm [1, 0, 3, 2]
    v = op(v, dpp_move<0x4E>(v));   // quad_perm [2, 3, 0, 1]
    v = op(v, dpp_move<0x141>(v));  // row_half_mirror
    v = op(v, dpp_move<0x140>(v));  // row_mirror
    return op(op(read_lane(v, 0), read_lane(v, 16)), op(read_lane(v, 32), read_lane(v, 48)));
}

// Inclusive prefix sum over the lanes of a wavefront, on DPP moves: row_shr
// 1, 2, 4, 8 within rows of 16, then row_bcast 15 and 31 across rows.  Call
// with all 64 lanes active.
template <typename T>
__device__ __forceinline__ T wave_scan_add(T v) {
    const int lane = lane_id(), r = lane & 15;
    T u;
    u = dpp_move<0x111>(v);
    if (r >= 1) v = v + u;
    u = dpp_move<0x112>(v);
    if (r >= 2) v = v + u;
    u = dpp_move<0x114>(v);
    if (r >= 4) v = v + u;
    u = dpp_move<0x118>(v);
    if (r >= 8) v = v + u;
    u = dpp_move<0x142>(v);  // row_bcast:15, rows 1 and 3
    if ((lane >> 4) & 1) v = v + u;
    u = dpp_move<0x143>(v);  // row_bcast:31, rows 2 and 3
    if (lane >= 32) v = v + u;
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
    return wave_reduce(v, [](T a, T b) { return b < a ? b : a; });
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
    return wave_reduce(v, [](T a, T b) { return b > a ? b : a; });
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    return wave_reduce(v, [](T a, T b) { return a + b; });
}

// A load through the constant address space: at a wavefront-uniform
// address it becomes a scalar load.  Only for data that nothing in the same
// launch writes (the scalar cache is invalidated between launches).
template <typename T>
__device__ __forceinline__ T load_const(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    typedef const __attribute__((address_space(4))) int32_t CInt;
    const CInt *q = (const CInt *)p;
    struct Words {
        int32_t w[sizeof(T) / 4];
    } v;
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); i++) v.w[i] = q[i];
    return __builtin_bit_cast(T, v);
}

__device__ __forceinline__ float lane_float(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float comp(const float4 &p, int d) {
    return d == 0 ? p.x : (d == 1 ? p.y : p.z);
}

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return __builtin_isfinite(x) && __builtin_isfinite(y) && __builtin_isfinite(z);
}

// Split dimension of a tight box: z first, x or y only when strictly wider
// (src/fast3tree.c:736-747).
__host__ __device__ __forceinline__ int split_dim(const float *lo, const float *hi) {
    int dim = 2;
    float w = hi[2] - lo[2];
    for (int k = 0; k < 2; k++) {
        float e = hi[k] - lo[k];
        if (e > w) {
            w = e;
            dim = k;
        }
    }
    return dim;
}

// Midpoint in float, halved exactly (src/fast3tree.c:757).
__host__ __device__ __forceinline__ float split_lim(const float *lo, const float *hi, int dim) {
    // selects instead of hi[dim]: a runtime index would send the box to scratch
    const float h = dim == 0 ? hi[0] : (dim == 1 ? hi[1] : hi[2]);
    const float l = dim == 0 ? lo[0] : (dim == 1 ? lo[1] : lo[2]);
    return 0.5f * (h + l);
}

// Where fast3tree's in-place partition (src/fast3tree.c:762-773) moves a
// point.  The loop swaps every point it finds above the plane with the
// current last slot and examines the swapped-in point next; the removal of
// non-finite points (src/fast3tree.c:875-887) is the same loop, and the
// strip of potential-only records (src/properties.c:539-544) moves the kept
// records the same way.  For a node of n points of which ns stay ("left":
// pos <= lim, finite, kept) the result has a closed form in the stable
// listing st of the node's positions: st(k) for k < ns is the position of
// the k-th left point, st(ns + k) that of the k-th right point (0-based,
// from the left).  For the point at position p, the rank-th of its side:
//   left,  p < ns:  stays at p;
//   left,  p >= ns: with c = ns - 1 - rank left points after it, it fills the
//                   slot of the (c + 1)-th right point, st(ns + c);
//   right: with c = rank right points before it, r_c = n for c = 0, else
//          the position of the c-th left point from the right, st(ns - c);
//          it goes to r_c - 1 when p < r_c, else to p - 1.  There are
//          p - c left points before p, so ns - p + c after it, and p < r_c
//          exactly when p <= ns: a right point beyond ns needs no lookup.
// Only the points on the wrong side of ns look anything up.  Checked
// against the loop on random class patterns: every destination is the
// loop's and each slot is used once.
template <typename ST>
__host__ __device__ __forceinline__ int32_t inplace_dest(int32_t p, bool left, int32_t rank, int32_t n,
                                                         int32_t ns, ST st) {
    if (left) return p < ns ? p : st(ns + (ns - 1 - rank));
    if (p > ns) return p - 1;
    return (rank == 0 ? n : st(ns - rank)) - 1;
}

// Position of the k-th (0-based) set bit of m; k < popcount(m).  A binary
// search over halves, quarters, ... of the word with popcounts.
__host__ __device__ __forceinline__ int select_bit(uint64_t m, int k) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __builtin_popcountll(m & ((1ull << w) - 1ull));
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += w;
        }
    }
    return pos;
}

// Order-preserving map of a float onto an int, for atomic min/max.
__device__ __forceinline__ int float_key(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : (i ^ 0x7fffffff);
}
__device__ __forceinline__ float key_float(int i) {
    return __int_as_float(i >= 0 ? i : (i ^ 0x7fffffff));
}

}  // namespace hp
