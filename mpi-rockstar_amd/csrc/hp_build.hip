// hp_build.hip -- GPU build of the reference's potential tree and its
// multipoles, for every halo of a batch at once.
//
// The tree is MPI-Rockstar's POTENTIAL fast3tree (src/potential.c:57-64,
// src/fast3tree.c:736-908): a binary k-d tree over the finite points of one
// compute_potential call, splitting a node of more than 10 points at the float
// midpoint of the widest side of its tight box (z first, x or y only when
// strictly wider), points with pos[dim] <= lim going left.  A split that puts
// every point on one side makes the node a leaf.  Node membership is a pure
// function of the point set, so a stable partition reproduces the reference's
// nodes exactly; only the order inside a node differs (the reference swaps in
// place).
//
// The build is level-synchronous over the whole batch:
//   level 0   gather: records -> float4 (x, y, z, record index), tight box and
//             non-finite count per halo (tiles of one halo, block reductions,
//             integer atomics on order-preserving float keys);
//             roots: one workgroup per halo writes the root node, moves any
//             non-finite records to the end of the halo's range (stable), and
//             queues the root.
//   level L   small tasks (<= big threshold): one wavefront per node, ballot
//             ranks, two passes (count, then stable scatter + child boxes);
//             big tasks: tiles of 4096 points, count pass -> per-task scan ->
//             scatter pass.
// Children of at most 10 points are leaves: their points are written straight
// into the final tree-order array, with their leaf id.
//
// Multipoles (src/potential.c:71-124): leaf mass and centre from f64 sums,
// internal centres combined in float from the two children (right term
// first) climbing with per-node arrival counters; dmin over the node's whole
// range in f64, one wavefront per node, or tiles for big nodes.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

// ------------------------------------------------------------------ level 0

struct RootTile {
    int32_t halo;
    int32_t off;
};

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// box keys per halo: [lo0 lo1 lo2 hi0 hi1 hi2] as order-preserving ints,
// bad: non-finite count
__global__ void k_root_init(int32_t *box, int32_t *bad, int nhalos) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhalos) return;
    for (int k = 0; k < 3; k++) {
        box[6 * h + k] = float_key(__int_as_float(0x7f800000));       // +inf
        box[6 * h + 3 + k] = float_key(__int_as_float((int)0xff800000));  // -inf
    }
    bad[h] = 0;
}

__global__ __launch_bounds__(kTileThreads) void k_gather(
    const struct potential *__restrict__ po, const int64_t *__restrict__ hfirst,
    const int64_t *__restrict__ hcount, const RootTile *__restrict__ tiles,
    float4 *__restrict__ A, int32_t *__restrict__ box, int32_t *__restrict__ bad) {
    __shared__ float s_lo[3][kTileThreads / kWave], s_hi[3][kTileThreads / kWave];
    __shared__ int s_bad[kTileThreads / kWave];
    const RootTile t = tiles[blockIdx.x];
    const int64_t first = hfirst[t.halo];
    const int64_t end = first + hcount[t.halo];
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = __int_as_float(0x7f800000);
        hi[k] = __int_as_float((int)0xff800000);
    }
    int nbad = 0;
    for (int r = 0; r < kTile / kTileThreads; r++) {
        int64_t i = first + t.off + r * kTileThreads + threadIdx.x;
        if (i >= end) break;
        const float *q = po[i].pos;
        float x = q[0], y = q[1], z = q[2];
        A[i] = make_float4(x, y, z, __int_as_float((int32_t)i));
        if (finite3(x, y, z)) {
            lo[0] = fminf(lo[0], x); hi[0] = fmaxf(hi[0], x);
            lo[1] = fminf(lo[1], y); hi[1] = fmaxf(hi[1], y);
            lo[2] = fminf(lo[2], z); hi[2] = fmaxf(hi[2], z);
        } else {
            nbad++;
        }
    }
    const int w = threadIdx.x / kWave;
    for (int k = 0; k < 3; k++) {
        float a = wave_min(lo[k]), b = wave_max(hi[k]);
        if (lane_id() == 0) {
            s_lo[k][w] = a;
            s_hi[k][w] = b;
        }
    }
    int nb = wave_sum(nbad);
    if (lane_id() == 0) s_bad[w] = nb;
    __syncthreads();
    if (threadIdx.x == 0) {
        float a[3], b[3];
        int tb = 0;
        for (int k = 0; k < 3; k++) {
            a[k] = s_lo[k][0];
            b[k] = s_hi[k][0];
        }
        for (int j = 1; j < kTileThreads / kWave; j++) {
            for (int k = 0; k < 3; k++) {
                a[k] = fminf(a[k], s_lo[k][j]);
                b[k] = fmaxf(b[k], s_hi[k][j]);
            }
        }
        for (int j = 0; j < kTileThreads / kWave; j++) tb += s_bad[j];
        for (int k = 0; k < 3; k++) {
            atomicMin(&box[6 * t.halo + k], float_key(a[k]));
            atomicMax(&box[6 * t.halo + 3 + k], float_key(b[k]));
        }
        if (tb) atomicAdd(&bad[t.halo], tb);
    }
}

struct BuildArgs {
    Node *nodes;
    int32_t *parent;
    int32_t *leaf_of;
    float4 *fin;            // final tree order
    Task *small_next;
    Task *big;              // append-only
    int2 *tiles;            // append-only (task index, offset)
    int32_t *fat;           // leaves of more than kFatLeaf points
    int32_t *counters;
    int32_t big_threshold;
    int32_t cap_small, cap_big, cap_tiles, cap_nodes, cap_fat;
};

__device__ void note_fat_leaf(const BuildArgs &a, int32_t id, int32_t count) {
    if (count <= kFatLeaf) return;
    int32_t k = atomicAdd(&a.counters[kFatCount], 1);
    if (k < a.cap_fat) a.fat[k] = id;
    else atomicOr(&a.counters[kErrFlags], kErrCapacity);
}

// Node ids and small-task slots are allocated together: counters[kNodeCount]
// and counters[kSmallNext] are adjacent 32-bit halves of one 64-bit counter,
// so a split costs one atomic round trip.
__device__ __forceinline__ uint64_t alloc_nodes_and_small(const BuildArgs &a, int32_t nodes,
                                                          int32_t small) {
    static_assert(kNodeCount == 0 && kSmallNext == 1, "packed counter layout");
    return atomicAdd(reinterpret_cast<unsigned long long *>(a.counters),
                     (unsigned long long)nodes | ((unsigned long long)small << 32));
}

__device__ Task make_task(int32_t id, const float *lo, const float *hi, int32_t depth) {
    Task t;
    t.node = id;
    t.dim = split_dim(lo, hi);
    t.lim = split_lim(lo, hi, t.dim);
    t.depth = depth;
    t.tile0 = 0;
    t.ntiles = 0;
    t.pad[0] = t.pad[1] = 0;
    return t;
}

// Queues node `id` (more than 10 points, tight box known) at a small-task
// slot already allocated.
__device__ void push_small_at(const BuildArgs &a, int32_t slot, int32_t id, const float *lo,
                              const float *hi, int32_t depth) {
    if (slot >= a.cap_small) {
        atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    a.small_next[slot] = make_task(id, lo, hi, depth);
}

// Queues a node of more than big_threshold points with its tiles.
__device__ void push_big(const BuildArgs &a, int32_t id, const float *lo, const float *hi,
                         int32_t count, int32_t depth) {
    Task t = make_task(id, lo, hi, depth);
    const int32_t nt = (count + kTile - 1) / kTile;
    const int32_t bi = atomicAdd(&a.counters[kBigNext], 1);
    const int32_t t0 = atomicAdd(&a.counters[kTileNext], nt);
    if (bi >= a.cap_big || t0 + nt > a.cap_tiles) {
        atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    t.tile0 = t0;
    t.ntiles = nt;
    a.big[bi] = t;
    for (int32_t j = 0; j < nt; j++) a.tiles[t0 + j] = make_int2(bi, j * kTile);
}

__device__ void write_node(Node *nd, const float *lo, const float *hi, int32_t first,
                           int32_t count, int32_t skip) {
    Node n;
    for (int k = 0; k < 3; k++) {
        n.lo[k] = lo[k];
        n.hi[k] = hi[k];
        n.com[k] = 0.0f;
    }
    n.mass = 0.0f;
    n.dmin2 = 0.0f;
    n.dmin = 0.0f;
    n.first = first;
    n.count = count;
    n.left = -1;
    n.skip = skip;
    *nd = n;
}

// One workgroup per halo: root node, stable removal of non-finite records to
// the tail (src/fast3tree.c:875-887), root queued or finished as a leaf.
__global__ __launch_bounds__(256) void k_roots(const int64_t *__restrict__ hfirst,
                                               const int64_t *__restrict__ hcount,
                                               const int32_t *__restrict__ box,
                                               const int32_t *__restrict__ bad,
                                               float4 *__restrict__ A, BuildArgs a) {
    __shared__ int s_cnt[2][4];
    const int h = blockIdx.x;
    const int32_t first = (int32_t)hfirst[h];
    const int32_t count = (int32_t)hcount[h];
    const int32_t nbad = bad[h];
    const int32_t nfin = count - nbad;
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = nfin ? key_float(box[6 * h + k]) : 0.0f;
        hi[k] = nfin ? key_float(box[6 * h + 3 + k]) : 0.0f;
    }
    if (nbad) {
        // rare path: stable partition finite | non-finite through the final
        // array, then the finite part back into A for splitting
        int32_t runf = 0, runb = 0;
        const int w = threadIdx.x / kWave;
        for (int32_t base = 0; base < count; base += 256) {
            int32_t i = base + threadIdx.x;
            bool v = i < count;
            float4 p = v ? A[first + i] : make_float4(0, 0, 0, 0);
            bool f = v && finite3(p.x, p.y, p.z);
            bool b = v && !f;
            uint64_t mf = ballot(f), mb = ballot(b);
            if (lane_id() == 0) {
                s_cnt[0][w] = __popcll(mf);
                s_cnt[1][w] = __popcll(mb);
            }
            __syncthreads();
            int of = 0, ob = 0, tf = 0, tb = 0;
            for (int j = 0; j < 4; j++) {
                if (j < w) {
                    of += s_cnt[0][j];
                    ob += s_cnt[1][j];
                }
                tf += s_cnt[0][j];
                tb += s_cnt[1][j];
            }
            if (f) a.fin[first + runf + of + rank_below(mf)] = p;
            if (b) {
                int32_t d = first + nfin + runb + ob + rank_below(mb);
                a.fin[d] = p;
                a.leaf_of[d] = -1;
            }
            runf += tf;
            runb += tb;
            __syncthreads();
        }
        __syncthreads();
        // A keeps the compact finite points of every call (the direct kernel
        // reads them there; the split levels read them when nfin > 10)
        for (int32_t i = threadIdx.x; i < nfin; i += 256) A[first + i] = a.fin[first + i];
    } else if (nfin <= kLeafPoints) {
        for (int32_t i = threadIdx.x; i < nfin; i += 256) a.fin[first + i] = A[first + i];
    }
    if (nfin <= kLeafPoints)
        for (int32_t i = threadIdx.x; i < nfin; i += 256) a.leaf_of[first + i] = h;
    if (threadIdx.x == 0) {
        write_node(a.nodes + h, lo, hi, first, nfin, -1);
        a.parent[h] = -1;
        if (nfin > a.big_threshold) {
            push_big(a, h, lo, hi, nfin, 0);
        } else if (nfin > kLeafPoints) {
            const uint64_t old = alloc_nodes_and_small(a, 0, 1);
            push_small_at(a, (int32_t)(old >> 32), h, lo, hi, 0);
        }
    }
}

// ------------------------------------------------------------ small tasks

// Writes the two children of a split node and queues those that need a
// split: small ones at slots P, P + 1 (allocated with the node ids), big ones
// with their tiles.
__device__ void make_children(const BuildArgs &a, const Task &t, const Node &nd, int32_t L,
                              int32_t P, int32_t nl, const float *loL, const float *hiL,
                              const float *loR, const float *hiR) {
    const int32_t nr = nd.count - nl;
    write_node(a.nodes + L, loL, hiL, nd.first, nl, L + 1);
    write_node(a.nodes + L + 1, loR, hiR, nd.first + nl, nr, nd.skip);
    a.parent[L] = a.parent[L + 1] = t.node;
    a.nodes[t.node].left = L;
    if (nl > a.big_threshold) push_big(a, L, loL, hiL, nl, t.depth + 1);
    else if (nl > kLeafPoints) push_small_at(a, P++, L, loL, hiL, t.depth + 1);
    if (nr > a.big_threshold) push_big(a, L + 1, loR, hiR, nr, t.depth + 1);
    else if (nr > kLeafPoints) push_small_at(a, P, L + 1, loR, hiR, t.depth + 1);
}

__device__ __forceinline__ int32_t small_pushes(int32_t nl, int32_t nr, int32_t thr) {
    return (nl > kLeafPoints && nl <= thr ? 1 : 0) + (nr > kLeafPoints && nr <= thr ? 1 : 0);
}

__device__ __forceinline__ void split_small_task(const float4 *__restrict__ src,
                                                 float4 *__restrict__ dst,
                                                 const Task *__restrict__ tasks, int32_t wid,
                                                 int32_t ntasks, const BuildArgs &a);

// One wavefront per task of at most kSmallMax points, held in registers:
// one read per point, a count pass and a stable scatter pass over ballots.
constexpr int kSplitThreads = 1024;  // 16 wavefronts share one allocation

__global__ __launch_bounds__(kSplitThreads) void k_split_small(const float4 *__restrict__ src,
                                                               float4 *__restrict__ dst,
                                                               const Task *__restrict__ tasks,
                                                               BuildArgs a) {
    const int32_t ntasks = a.counters[kCurSmall];
    constexpr int kWaves = kSplitThreads / kWave;
    // the workgroup walks the task list kWaves tasks at a time; every
    // wavefront takes part in every round (the allocation is per round)
    for (int32_t base = blockIdx.x * kWaves; base < ntasks; base += gridDim.x * kWaves)
        split_small_task(src, dst, tasks, base + (int32_t)(threadIdx.x / kWave), ntasks, a);
}

// One round of k_split_small: wavefront w of the workgroup splits task wid
// (if wid < ntasks).  Node ids and small-task slots are allocated once per
// workgroup: the wavefronts' demands are summed in LDS and thread 0 makes a
// single atomic request.
__device__ __forceinline__ void split_small_task(const float4 *__restrict__ src,
                                                 float4 *__restrict__ dst,
                                                 const Task *__restrict__ tasks, int32_t wid,
                                                 int32_t ntasks, const BuildArgs &a) {
    constexpr int kWaves = kSplitThreads / kWave;
    __shared__ uint32_t s_nodes[kWaves], s_push[kWaves];
    __shared__ unsigned long long s_base;
    const int lane = lane_id();
    const int w = (int)(threadIdx.x / kWave);
    const bool has = wid < ntasks;
    Task t;
    Node nd;
    if (has) {
        t = tasks[wid];
        nd = a.nodes[t.node];
    } else {
        t.node = -1;
        t.dim = 2;
        t.lim = 0.0f;
        t.depth = 0;
        nd.first = 0;
        nd.count = 0;
        nd.skip = -1;
    }
    const int32_t first = nd.first, count = nd.count;
    const int dim = t.dim;
    const float lim = t.lim;
    const int nch = (count + kWave - 1) / kWave;  // <= kSmallChunks

    float4 pts[kSmallChunks];
#pragma unroll
    for (int c = 0; c < kSmallChunks; c++) {
        const int32_t i = c * kWave + lane;
        pts[c] = (c < nch && i < count) ? src[first + i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    // pass 1: left count and both children's boxes
    float lo[2][3], hi[2][3];
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) {
            lo[s][k] = __int_as_float(0x7f800000);
            hi[s][k] = __int_as_float((int)0xff800000);
        }
    int32_t nl = 0;
#pragma unroll
    for (int c = 0; c < kSmallChunks; c++) {
        if (c < nch) {
            const int32_t i = c * kWave + lane;
            const float4 p = pts[c];
            const bool v = i < count;
            const bool f = v && comp(p, dim) <= lim;
            nl += __popcll(ballot(f));
            if (v) {
                const int s = f ? 0 : 1;
                lo[s][0] = fminf(lo[s][0], p.x); hi[s][0] = fmaxf(hi[s][0], p.x);
                lo[s][1] = fminf(lo[s][1], p.y); hi[s][1] = fmaxf(hi[s][1], p.y);
                lo[s][2] = fminf(lo[s][2], p.z); hi[s][2] = fmaxf(hi[s][2], p.z);
            }
        }
    }
    const bool failed = has && (nl == 0 || nl == count);  // src/fast3tree.c:805-809
    const bool splits = has && !failed;
    const int32_t nr = count - nl;
    // workgroup allocation: exclusive offsets of every wavefront's demand
    if (lane == 0) {
        s_nodes[w] = splits ? 2u : 0u;
        s_push[w] = splits ? (uint32_t)small_pushes(nl, nr, a.big_threshold) : 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tn = 0, tp = 0;
        for (int q = 0; q < kWaves; q++) {
            tn += s_nodes[q];
            tp += s_push[q];
        }
        s_base = (tn | tp) ? alloc_nodes_and_small(a, (int32_t)tn, (int32_t)tp) : 0ull;
    }
    __syncthreads();
    uint32_t on = (uint32_t)s_base, op = (uint32_t)(s_base >> 32);
    for (int q = 0; q < w; q++) {
        on += s_nodes[q];
        op += s_push[q];
    }
    __syncthreads();  // s_* are reused by the next round
    if (failed) {  // a leaf
#pragma unroll
        for (int c = 0; c < kSmallChunks; c++) {
            const int32_t i = c * kWave + lane;
            if (c < nch && i < count) {
                a.fin[first + i] = pts[c];
                a.leaf_of[first + i] = t.node;
            }
        }
        if (lane == 0) note_fat_leaf(a, t.node, count);
        return;
    }
    if (!splits) return;
    const int32_t L = (int32_t)on, P = (int32_t)op;
    if (L + 2 > a.cap_nodes) {
        if (lane == 0) atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    const bool leafL = nl <= kLeafPoints, leafR = nr <= kLeafPoints;
    // pass 2: stable scatter
    int32_t bl = 0, br = 0;
#pragma unroll
    for (int c = 0; c < kSmallChunks; c++) {
        if (c >= nch) continue;
        const int32_t i = c * kWave + lane;
        const bool v = i < count;
        const float4 p = pts[c];
        const bool f = v && comp(p, dim) <= lim;
        const bool g = v && !f;
        const uint64_t mL = ballot(f), mR = ballot(g);
        if (v) {
            const int32_t d = f ? first + bl + rank_below(mL) : first + nl + br + rank_below(mR);
            if (f ? leafL : leafR) {
                a.fin[d] = p;
                a.leaf_of[d] = L + (f ? 0 : 1);
            } else {
                dst[d] = p;
            }
        }
        bl += __popcll(mL);
        br += __popcll(mR);
    }
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) {
            lo[s][k] = wave_min(lo[s][k]);
            hi[s][k] = wave_max(hi[s][k]);
        }
    if (lane == 0) make_children(a, t, nd, L, P, nl, lo[0], hi[0], lo[1], hi[1]);
}

// -------------------------------------------------------------- big tasks

__device__ void block_box_reduce(float (*lo)[3], float (*hi)[3], int &nl, TilePart *out) {
    __shared__ float s_lo[2][3][kTileThreads / kWave], s_hi[2][3][kTileThreads / kWave];
    __shared__ int s_n[kTileThreads / kWave];
    const int w = threadIdx.x / kWave;
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) {
            float a = wave_min(lo[s][k]), b = wave_max(hi[s][k]);
            if (lane_id() == 0) {
                s_lo[s][k][w] = a;
                s_hi[s][k][w] = b;
            }
        }
    int n = wave_sum(nl);
    if (lane_id() == 0) s_n[w] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        TilePart tp;
        tp.nleft = 0;
        tp.pad[0] = tp.pad[1] = tp.pad[2] = 0;
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++) {
                float a = s_lo[s][k][0], b = s_hi[s][k][0];
                for (int j = 1; j < kTileThreads / kWave; j++) {
                    a = fminf(a, s_lo[s][k][j]);
                    b = fmaxf(b, s_hi[s][k][j]);
                }
                tp.lo[s][k] = a;
                tp.hi[s][k] = b;
            }
        for (int j = 0; j < kTileThreads / kWave; j++) tp.nleft += s_n[j];
        *out = tp;
    }
    __syncthreads();
}

__device__ void big_count_tile(const float4 *__restrict__ src, const int2 *__restrict__ tiles,
                               const Task *__restrict__ big, const Node *__restrict__ nodes,
                               TilePart *__restrict__ parts, int32_t ti);

// Count pass over one tile of a big task.
__global__ __launch_bounds__(kTileThreads) void k_big_count(const float4 *__restrict__ src,
                                                            const int2 *__restrict__ tiles,
                                                            const Task *__restrict__ big,
                                                            const Node *__restrict__ nodes,
                                                            TilePart *__restrict__ parts,
                                                            const int32_t *__restrict__ counters) {
    const int32_t t1 = counters[kTileEnd];
    for (int32_t ti = counters[kTileBegin] + blockIdx.x; ti < t1; ti += gridDim.x)
        big_count_tile(src, tiles, big, nodes, parts, ti);
}

__device__ void big_count_tile(const float4 *__restrict__ src, const int2 *__restrict__ tiles,
                               const Task *__restrict__ big, const Node *__restrict__ nodes,
                               TilePart *__restrict__ parts, int32_t ti) {
    const int2 tl = tiles[ti];
    const Task t = big[tl.x];
    const Node nd = nodes[t.node];
    float lo[2][3], hi[2][3];
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) {
            lo[s][k] = __int_as_float(0x7f800000);
            hi[s][k] = __int_as_float((int)0xff800000);
        }
    int nl = 0;
    for (int r = 0; r < kTile / kTileThreads; r++) {
        int32_t i = tl.y + r * kTileThreads + threadIdx.x;
        if (i >= nd.count) break;
        float4 p = src[nd.first + i];
        int s = comp(p, t.dim) <= t.lim ? 0 : 1;
        nl += 1 - s;
        lo[s][0] = fminf(lo[s][0], p.x); hi[s][0] = fmaxf(hi[s][0], p.x);
        lo[s][1] = fminf(lo[s][1], p.y); hi[s][1] = fmaxf(hi[s][1], p.y);
        lo[s][2] = fminf(lo[s][2], p.z); hi[s][2] = fmaxf(hi[s][2], p.z);
    }
    block_box_reduce(lo, hi, nl, parts + ti);
}

// One thread per big task of the level: scan of the tile counts, children.
__device__ void big_split_task(const Task *__restrict__ big, TilePart *__restrict__ parts,
                               BigSplit *__restrict__ res, const BuildArgs &a, int32_t bi);

__global__ void k_big_split(const Task *__restrict__ big, TilePart *__restrict__ parts,
                            BigSplit *__restrict__ res, BuildArgs a) {
    const int32_t b1 = a.counters[kBigEnd];
    for (int32_t bi = a.counters[kBigBegin] + blockIdx.x * blockDim.x + threadIdx.x; bi < b1;
         bi += gridDim.x * blockDim.x)
        big_split_task(big, parts, res, a, bi);
}

__device__ void big_split_task(const Task *__restrict__ big, TilePart *__restrict__ parts,
                               BigSplit *__restrict__ res, const BuildArgs &a, int32_t bi) {
    const Task t = big[bi];
    const Node nd = a.nodes[t.node];
    int32_t nl = 0;
    float lo[2][3], hi[2][3];
    for (int s = 0; s < 2; s++)
        for (int d = 0; d < 3; d++) {
            lo[s][d] = __int_as_float(0x7f800000);
            hi[s][d] = __int_as_float((int)0xff800000);
        }
    for (int32_t j = 0; j < t.ntiles; j++) {
        TilePart &p = parts[t.tile0 + j];
        int32_t c = p.nleft;
        p.pad[0] = nl;  // left points in the task's earlier tiles
        nl += c;
        for (int s = 0; s < 2; s++)
            for (int d = 0; d < 3; d++) {
                lo[s][d] = fminf(lo[s][d], p.lo[s][d]);
                hi[s][d] = fmaxf(hi[s][d], p.hi[s][d]);
            }
    }
    BigSplit r;
    r.nleft = nl;
    r.left = -1;
    r.leafL = r.leafR = 0;
    if (nl == 0 || nl == nd.count) {
        note_fat_leaf(a, t.node, nd.count);
    } else {
        const int32_t nr = nd.count - nl;
        const uint64_t al = alloc_nodes_and_small(a, 2, small_pushes(nl, nr, a.big_threshold));
        const int32_t L = (int32_t)al, P = (int32_t)(al >> 32);
        if (L + 2 > a.cap_nodes) {
            atomicOr(&a.counters[kErrFlags], kErrCapacity);
        } else {
            r.left = L;
            r.leafL = nl <= kLeafPoints;
            r.leafR = nr <= kLeafPoints;
            make_children(a, t, nd, L, P, nl, lo[0], hi[0], lo[1], hi[1]);
        }
    }
    res[bi] = r;
}

// Scatter pass over one tile of a big task: stable ranks in rounds of 256.
__device__ void big_scatter_tile(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                 const int2 *__restrict__ tiles, const Task *__restrict__ big,
                                 const TilePart *__restrict__ parts,
                                 const BigSplit *__restrict__ res, const BuildArgs &a, int32_t ti);

__global__ __launch_bounds__(kTileThreads) void k_big_scatter(
    const float4 *__restrict__ src, float4 *__restrict__ dst, const int2 *__restrict__ tiles,
    const Task *__restrict__ big, const TilePart *__restrict__ parts,
    const BigSplit *__restrict__ res, BuildArgs a) {
    const int32_t t1 = a.counters[kTileEnd];
    for (int32_t ti = a.counters[kTileBegin] + blockIdx.x; ti < t1; ti += gridDim.x)
        big_scatter_tile(src, dst, tiles, big, parts, res, a, ti);
}

__device__ void big_scatter_tile(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                 const int2 *__restrict__ tiles, const Task *__restrict__ big,
                                 const TilePart *__restrict__ parts,
                                 const BigSplit *__restrict__ res, const BuildArgs &a, int32_t ti) {
    __shared__ int s_cnt[2][kTileThreads / kWave];
    const int2 tl = tiles[ti];
    const Task t = big[tl.x];
    const Node nd = a.nodes[t.node];
    const BigSplit r = res[tl.x];
    const int w = threadIdx.x / kWave;
    if (r.left < 0) {  // failed split: the node is a leaf
        for (int j = threadIdx.x; j < kTile; j += kTileThreads) {
            int32_t i = tl.y + j;
            if (i >= nd.count) break;
            a.fin[nd.first + i] = src[nd.first + i];
            a.leaf_of[nd.first + i] = t.node;
        }
        return;
    }
    const int32_t lpre = parts[ti].pad[0];
    const int32_t rpre = tl.y - lpre;
    int32_t runL = 0, runR = 0;
    for (int rr = 0; rr < kTile / kTileThreads; rr++) {
        int32_t i = tl.y + rr * kTileThreads + threadIdx.x;
        bool v = i < nd.count;
        float4 p = v ? src[nd.first + i] : make_float4(0, 0, 0, 0);
        bool f = v && comp(p, t.dim) <= t.lim;
        bool g = v && !f;
        uint64_t mL = ballot(f), mR = ballot(g);
        if (lane_id() == 0) {
            s_cnt[0][w] = __popcll(mL);
            s_cnt[1][w] = __popcll(mR);
        }
        __syncthreads();
        int oL = 0, oR = 0, tL = 0, tR = 0;
        for (int j = 0; j < kTileThreads / kWave; j++) {
            if (j < w) {
                oL += s_cnt[0][j];
                oR += s_cnt[1][j];
            }
            tL += s_cnt[0][j];
            tR += s_cnt[1][j];
        }
        if (v) {
            int32_t d;
            bool leaf;
            int32_t id;
            if (f) {
                d = nd.first + lpre + runL + oL + rank_below(mL);
                leaf = r.leafL;
                id = r.left;
            } else {
                d = nd.first + r.nleft + rpre + runR + oR + rank_below(mR);
                leaf = r.leafR;
                id = r.left + 1;
            }
            if (leaf) {
                a.fin[d] = p;
                a.leaf_of[d] = id;
            } else {
                dst[d] = p;
            }
        }
        runL += tL;
        runR += tR;
        __syncthreads();
    }
}

// Moves the level state forward on the device: the pushes of the level just
// run become the next level's work; depth_end[level] records the node count
// (node ids grow with depth).
__global__ void k_level_advance(int32_t *__restrict__ c, int32_t *__restrict__ depth_end) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t ns = c[kSmallNext];
    c[kCurSmall] = ns;
    c[kSmallNext] = 0;
    c[kBigBegin] = c[kBigEnd];
    c[kBigEnd] = c[kBigNext];
    c[kTileBegin] = c[kTileEnd];
    c[kTileEnd] = c[kTileNext];
    const int32_t lvl = c[kLevel];
    if (lvl < kMaxLevels) depth_end[lvl] = c[kNodeCount];
    else c[kErrFlags] |= kErrCapacity;
    c[kLevel] = lvl + 1;
    c[kActive] = (ns > 0 || c[kBigEnd] > c[kBigBegin]) ? 1 : 0;
}

// --------------------------------------------------------------- multipoles

// The multipole kernels take their sizes from the device counters (node
// count, fat-leaf count, per-depth node ranges, big tasks and tiles), so a
// whole build can be replayed from a graph.

// Leaf mass and centre (src/potential.c:100-107): f64 sums of the float
// positions, one thread per leaf of at most kFatLeaf points.
__global__ void k_leaf_com(Node *__restrict__ nodes, const float4 *__restrict__ fin,
                           const int32_t *__restrict__ counters, double particle_mass) {
    const int32_t nnodes = counters[kNodeCount];
    for (int32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < nnodes;
         id += gridDim.x * blockDim.x) {
        const Node nd = nodes[id];
        if (nd.left >= 0 || nd.count <= 0 || nd.count > kFatLeaf) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int32_t i = 0; i < nd.count; i++) {
            const float4 p = fin[nd.first + i];
            s0 += (double)p.x;
            s1 += (double)p.y;
            s2 += (double)p.z;
        }
        const double n = (double)nd.count;
        Node *me = nodes + id;
        me->com[0] = s0 != 0.0 ? (float)(s0 / n) : 0.0f;
        me->com[1] = s1 != 0.0 ? (float)(s1 / n) : 0.0f;
        me->com[2] = s2 != 0.0 ? (float)(s2 / n) : 0.0f;
        me->mass = (float)(particle_mass * n);
    }
}

// The same for leaves of more than kFatLeaf points (failed splits), one
// wavefront each.  Sums of floats in f64 are exact here, so the order of the
// wavefront reduction does not matter.
__global__ __launch_bounds__(256) void k_fat_leaf_com(Node *__restrict__ nodes,
                                                      const int32_t *__restrict__ list,
                                                      const int32_t *__restrict__ counters,
                                                      int32_t cap, const float4 *__restrict__ fin,
                                                      double particle_mass) {
    const int32_t nlist = min(counters[kFatCount], cap);
    const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
    const int lane = lane_id();
    for (int32_t k = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave); k < nlist;
         k += nw) {
        const int32_t id = list[k];
        const Node nd = nodes[id];
        double s[3] = {0.0, 0.0, 0.0};
        for (int32_t i = lane; i < nd.count; i += kWave) {
            const float4 p = fin[nd.first + i];
            s[0] += (double)p.x;
            s[1] += (double)p.y;
            s[2] += (double)p.z;
        }
        for (int k2 = 0; k2 < 3; k2++) s[k2] = wave_sum(s[k2]);
        if (lane == 0) {
            const double n = (double)nd.count;
            Node *me = nodes + id;
            for (int k2 = 0; k2 < 3; k2++) me->com[k2] = s[k2] != 0.0 ? (float)(s[k2] / n) : 0.0f;
            me->mass = (float)(particle_mass * n);
        }
    }
}

// Internal nodes of depth d: float mass and centre from the two children,
// right term first (src/potential.c:115-122).  Node ids grow with depth:
// depth d holds [depth_end[d - 1], depth_end[d]), depth 0 the roots; the
// host runs the depths deepest first.
__global__ void k_internal_com(Node *__restrict__ nodes, const int32_t *__restrict__ depth_end,
                               int32_t d) {
    const int32_t begin = d ? depth_end[d - 1] : 0, end = depth_end[d];
    for (int32_t id = begin + blockIdx.x * blockDim.x + threadIdx.x; id < end;
         id += gridDim.x * blockDim.x) {
        Node *me = nodes + id;
        const int32_t L = me->left;
        if (L < 0) continue;
        const Node *l = nodes + L, *r = nodes + L + 1;
        const float ml = l->mass, mr = r->mass;
        const float m = ml + mr;
        for (int k = 0; k < 3; k++) {
            const float rr = r->com[k] * mr;
            const float ll = l->com[k] * ml;
            me->com[k] = m != 0.0f ? (rr + ll) / m : 0.0f;
        }
        me->mass = m;
    }
}

__device__ __forceinline__ double d2_ref(const float *c, const float4 &p) {
    // src/potential.c:22-29: float differences, double squares, summed in order
    double dx = (double)(c[0] - p.x);
    double dy = (double)(c[1] - p.y);
    double dz = (double)(c[2] - p.z);
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return s;
}

__device__ __forceinline__ void finish_dmin(Node *nd, double big, double sum, int32_t count) {
    // src/potential.c:79-81
    double b = sqrt(big) / 2.0;
    float dmin = (float)(b + sqrt(b * b + sum / ((double)count * 1.0)));
    nd->dmin = dmin;
    nd->dmin2 = dmin * dmin;
}

// One wavefront per node of at most big_threshold points.
__global__ __launch_bounds__(256) void k_dmin_small(Node *__restrict__ nodes,
                                                    const float4 *__restrict__ fin,
                                                    const int32_t *__restrict__ counters,
                                                    int32_t big_threshold) {
    const int32_t nnodes = counters[kNodeCount];
    const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
    const int lane = lane_id();
    for (int32_t id = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave); id < nnodes;
         id += nw) {
        const Node nd = nodes[id];
        if (nd.count <= 0 || nd.count > big_threshold) continue;
        double big = 0.0, sum = 0.0;
        for (int32_t i = lane; i < nd.count; i += kWave) {
            const double d = d2_ref(nd.com, fin[nd.first + i]);
            big = d > big ? d : big;
            sum += d;
        }
        big = wave_max(big);
        sum = wave_sum(sum);
        if (lane == 0) finish_dmin(nodes + id, big, sum, nd.count);
    }
}

// Tile partials of dmin for big nodes (every big task of every level), then
// one thread per big node sums its tiles in order.
__global__ __launch_bounds__(kTileThreads) void k_dmin_tiles(const Node *__restrict__ nodes,
                                                             const float4 *__restrict__ fin,
                                                             const int2 *__restrict__ tiles,
                                                             const Task *__restrict__ big,
                                                             const int32_t *__restrict__ counters,
                                                             double2 *__restrict__ part) {
    __shared__ double s_b[kTileThreads / kWave], s_s[kTileThreads / kWave];
    const int32_t ntiles = counters[kTileNext];
    for (int32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const int2 tl = tiles[ti];
        const Node nd = nodes[big[tl.x].node];
        double bmax = 0.0, sum = 0.0;
        for (int r = 0; r < kTile / kTileThreads; r++) {
            const int32_t i = tl.y + r * kTileThreads + threadIdx.x;
            if (i >= nd.count) break;
            const double d = d2_ref(nd.com, fin[nd.first + i]);
            bmax = d > bmax ? d : bmax;
            sum += d;
        }
        bmax = wave_max(bmax);
        sum = wave_sum(sum);
        const int w = threadIdx.x / kWave;
        if (lane_id() == 0) {
            s_b[w] = bmax;
            s_s[w] = sum;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double b = s_b[0], s = s_s[0];
            for (int j = 1; j < kTileThreads / kWave; j++) {
                b = s_b[j] > b ? s_b[j] : b;
                s += s_s[j];
            }
            part[ti] = make_double2(b, s);
        }
        __syncthreads();
    }
}

__global__ void k_dmin_big(Node *__restrict__ nodes, const Task *__restrict__ big,
                           const int32_t *__restrict__ counters, const double2 *__restrict__ part) {
    const int32_t nbig = counters[kBigNext];
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nbig; k += gridDim.x * blockDim.x) {
        const Task t = big[k];
        double b = 0.0, s = 0.0;
        for (int32_t j = 0; j < t.ntiles; j++) {
            const double2 p = part[t.tile0 + j];
            b = p.x > b ? p.x : b;
            s += p.y;
        }
        finish_dmin(nodes + t.node, b, s, nodes[t.node].count);
    }
}

}  // namespace hp
