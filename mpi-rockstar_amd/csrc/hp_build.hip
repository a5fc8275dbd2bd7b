// hp_build.hip -- GPU build of the reference's potential tree and its
// multipoles, for every halo of a batch at once.
//
// The tree is MPI-Rockstar's POTENTIAL fast3tree (src/potential.c:57-64,
// src/fast3tree.c:736-908): a binary k-d tree over the finite points of one
// compute_potential call, splitting a node of more than 10 points at the float
// midpoint of the widest side of its tight box (z first, x or y only when
// strictly wider), points with pos[dim] <= lim going left.  A split that puts
// every point on one side makes the node a leaf.  Every partition puts the
// points in the order the reference's in-place swap loop leaves them
// (src/fast3tree.c:762-773, closed form in inplace_dest): each split lists
// the node's points by side (the stable order), then moves every point to
// its slot.  So the final tree order, and with it every leaf's summation
// order, is the reference's.
//
// The build runs over the whole batch at once:
//   roots     gather: records -> float4 (x, y, z, record index), tight box and
//             non-finite count per call (tiles of one call, block reductions,
//             integer atomics on order-preserving float keys); roots: one
//             thread per call writes the root node and queues it (calls with
//             non-finite points get a workgroup that moves those records to
//             the end of the call's range, stably).
//   levels    nodes of more than kSmallMax points: tiles of 1024 points,
//             count pass -> one wavefront per node (scan of the tile counts,
//             child boxes) -> scatter pass; host-driven, several levels per
//             counter read.
//   subtrees  every node of at most kSmallMax points is finished in one go:
//             65..kSmallMax points by a workgroup (a kernel each for nodes of
//             at most 256, 512 and 1,024 points), point-parallel per level
//             (block prefix of the left flags, segmented DPP scans for the
//             child boxes); at most 64 points by one wavefront, a point per
//             lane.  Both keep the points and the node table in LDS, compute
//             centres of mass bottom-up and dmin over the final order, and
//             write leaves straight into the final tree-order array.
// Multipoles of the nodes made by the split levels (src/potential.c:71-124):
// leaf mass and centre from f64 sums, internal centres combined in float
// from the two children (right term first) level by level, dmin over the
// node's whole range in f64 by tiles.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

__device__ __forceinline__ double d2_ref(const float *c, const float4 &p) {
    // src/potential.c:22-29: float differences, double squares, summed in
    // order (a float times a float is exact in double, so each fma rounds
    // as the separate product and sum do)
    const double dx = (double)(c[0] - p.x);
    const double dy = (double)(c[1] - p.y);
    const double dz = (double)(c[2] - p.z);
    double s = dx * dx;
    s = fma(dy, dy, s);
    s = fma(dz, dz, s);
    return s;
}

__device__ __forceinline__ void finish_dmin(Node *nd, double big, double sum, int32_t count) {
    // src/potential.c:79-81
    double b = sqrt(big) / 2.0;
    float dmin = (float)(b + sqrt(b * b + sum / ((double)count * 1.0)));
    nd->dmin = dmin;
    nd->dmin2 = dmin * dmin;
}

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
// bad: non-finite count, pot: potential-only records (ke < 0) per call, and
// pot[nhalos] > 0 when some call has any
// (and the build's counters: all zero but the node count, one id per call's
// root; nothing reads them before k_roots)
__global__ void k_root_init(int32_t *box, int32_t *bad, int32_t *pot, int nhalos,
                            int32_t *counters) {
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < kNumCounters) counters[h] = h == kNodeCount ? nhalos : 0;
    if (h == 0) pot[nhalos] = 0;
    if (h >= nhalos) return;
    for (int k = 0; k < 3; k++) {
        box[6 * h + k] = float_key(__int_as_float(0x7f800000));       // +inf
        box[6 * h + 3 + k] = float_key(__int_as_float((int)0xff800000));  // -inf
    }
    bad[h] = 0;
    pot[h] = 0;
}

// The records into float4 (x, y, z, record index), each call's tight box and
// non-finite and potential-only counts; and each record's sort key for the
// bound pass (potential-only flag above the r2 bits), read here where the
// records stream in order.
__global__ __launch_bounds__(kTileThreads) void k_gather(
    const struct potential *__restrict__ po, const int64_t *__restrict__ hfirst,
    const int64_t *__restrict__ hcount, const RootTile *__restrict__ tiles,
    float4 *__restrict__ A, int32_t *__restrict__ box, int32_t *__restrict__ bad,
    int32_t *__restrict__ pot, int32_t nhalos, uint64_t *__restrict__ rkey) {
    __shared__ float s_lo[3][kTileThreads / kWave], s_hi[3][kTileThreads / kWave];
    __shared__ int s_bad[kTileThreads / kWave], s_pot[kTileThreads / kWave];
    const RootTile t = tiles[blockIdx.x];
    const int64_t first = hfirst[t.halo];
    const int64_t end = first + hcount[t.halo];
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = __int_as_float(0x7f800000);
        hi[k] = __int_as_float((int)0xff800000);
    }
    int nbad = 0, npot = 0;
    for (int r = 0; r < kTile / kTileThreads; r++) {
        int64_t i = first + t.off + r * kTileThreads + threadIdx.x;
        if (i >= end) break;
        const float *q = po[i].pos;
        float x = q[0], y = q[1], z = q[2];
        const bool potonly = po[i].ke < 0.0f;
        npot += potonly ? 1 : 0;
        rkey[i] = ((uint64_t)(potonly ? 1 : 0) << 32) | (uint64_t)__float_as_uint(po[i].r2);
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
    int nb = wave_sum(nbad), np_ = wave_sum(npot);
    if (lane_id() == 0) {
        s_bad[w] = nb;
        s_pot[w] = np_;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a[3], b[3];
        int tb = 0, tp = 0;
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
        for (int j = 0; j < kTileThreads / kWave; j++) {
            tb += s_bad[j];
            tp += s_pot[j];
        }
        for (int k = 0; k < 3; k++) {
            atomicMin(&box[6 * t.halo + k], float_key(a[k]));
            atomicMax(&box[6 * t.halo + 3 + k], float_key(b[k]));
        }
        if (tb) atomicAdd(&bad[t.halo], tb);
        if (tp) {
            atomicAdd(&pot[t.halo], tp);
            atomicOr(&pot[nhalos], 1);
        }
    }
}

struct BuildArgs {
    Node *nodes;
    int32_t *parent;
    int32_t *sides;         // per position: a partition's points listed by side
    int32_t *leaf_of;
    float4 *fin;            // final tree order
    Task *small_next;       // wavefront subtree tasks
    Task *group_next;       // workgroup subtree tasks
    Task *big;              // append-only
    int2 *tiles;            // append-only (task index, offset)
    int32_t *fat;           // leaves of more than kFatLeaf points
    int32_t *counters;
    // per call-size bucket (tune_bucket): nodes above bucket_big points are
    // split by the level passes; subtrees of at most bucket_wave points are
    // built by one wavefront (32 or 64)
    int32_t bucket_big[kTuneBuckets];
    int32_t bucket_wave[kTuneBuckets];
    int32_t cap_small, cap_group, cap_big, cap_tiles, cap_nodes, cap_fat;
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

// The launch setting of a call's nodes (its size bucket's).
struct Sizes {
    int32_t big;   // nodes above this many points: split levels
    int32_t wave;  // subtrees of at most this many points: one wavefront
};
__device__ __forceinline__ Sizes sizes_of(const BuildArgs &a, int32_t bucket) {
    bucket = min(max(bucket, 0), kTuneBuckets - 1);
    return Sizes{a.bucket_big[bucket], a.bucket_wave[bucket]};
}

// Subtree tasks: a node of more than 10 and at most s.wave points is built
// by one wavefront, a larger one of at most s.big points by one workgroup.
// Each kind has its own list.
__device__ __forceinline__ bool wave_sized(const Sizes &s, int32_t n) {
    return n > kLeafPoints && n <= s.wave && n <= s.big;
}
__device__ __forceinline__ bool group_sized(const Sizes &s, int32_t n) {
    return n > s.wave && n <= s.big;
}

// Queues node `id` (tight box known) at a slot already allocated in the
// list of its kind; the task carries the node's range, so that the subtree
// kernels can pick their tasks without reading the node.
__device__ void push_task_at(const BuildArgs &a, Task *list, int32_t cap, int32_t slot, int32_t id,
                             const float *lo, const float *hi, int32_t depth, int32_t first,
                             int32_t count) {
    if (slot >= cap) {
        atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    Task t = make_task(id, lo, hi, depth);
    t.pad[0] = first;
    t.pad[1] = count;
    list[slot] = t;
}
__device__ __forceinline__ void push_small_at(const BuildArgs &a, int32_t slot, int32_t id,
                                              const float *lo, const float *hi, int32_t depth,
                                              int32_t first, int32_t count) {
    push_task_at(a, a.small_next, a.cap_small, slot, id, lo, hi, depth, first, count);
}
__device__ __forceinline__ void push_group_at(const BuildArgs &a, int32_t slot, int32_t id,
                                              const float *lo, const float *hi, int32_t depth,
                                              int32_t first, int32_t count) {
    push_task_at(a, a.group_next, a.cap_group, slot, id, lo, hi, depth, first, count);
}

// A big-task slot and its tiles are reserved together: counters[kBigNext]
// and counters[kTileNext] are the two halves of one 64-bit counter.
__device__ __forceinline__ uint64_t alloc_big_and_tiles(const BuildArgs &a, int32_t ntiles) {
    static_assert(kTileNext == kBigNext + 1 && (kBigNext & 1) == 0, "packed counter layout");
    return atomicAdd(reinterpret_cast<unsigned long long *>(a.counters + kBigNext),
                     1ull | ((unsigned long long)ntiles << 32));
}

// Queues a node of more than its bucket's split size with its tiles; the
// task keeps the bucket (pad[0]) for its children.
__device__ void push_big(const BuildArgs &a, int32_t id, const float *lo, const float *hi,
                         int32_t count, int32_t depth, int32_t bucket) {
    Task t = make_task(id, lo, hi, depth);
    t.pad[0] = bucket;
    const int32_t nt = (count + kTile - 1) / kTile;
    const uint64_t al = alloc_big_and_tiles(a, nt);
    const int32_t bi = (int32_t)al, t0 = (int32_t)(al >> 32);
    if (bi >= a.cap_big || t0 + nt > a.cap_tiles) {
        atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    t.tile0 = t0;
    t.ntiles = nt;
    a.big[bi] = t;
    for (int32_t j = 0; j < nt; j++) a.tiles[t0 + j] = make_int2(bi, j * kTile);
}

// A big task queued by a whole wavefront (wavefront-uniform arguments) at a
// task slot and tiles its caller reserved (alloc_big_and_tiles); the lanes
// write the tiles.
__device__ void push_big_wave_at(const BuildArgs &a, int32_t bi, int32_t t0, int32_t id,
                                 const float *lo, const float *hi, int32_t count, int32_t depth,
                                 int32_t bucket) {
    const int lane = lane_id();
    const int32_t nt = (count + kTile - 1) / kTile;
    if (bi >= a.cap_big || t0 + nt > a.cap_tiles) {
        if (lane == 0) atomicOr(&a.counters[kErrFlags], kErrCapacity);
        return;
    }
    if (lane == 0) {
        Task t = make_task(id, lo, hi, depth);
        t.tile0 = t0;
        t.ntiles = nt;
        t.pad[0] = bucket;
        a.big[bi] = t;
    }
    for (int32_t j = lane; j < nt; j += kWave) a.tiles[t0 + j] = make_int2(bi, j * kTile);
}

__device__ __forceinline__ void write_node(Node *nd, const float *lo, const float *hi, int32_t first,
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

// Root nodes, one thread per call.  A call with non-finite points is listed
// for k_roots_bad; every other call gets its root node here, its points in
// tree order when it is a single leaf, and its root queued as a task.  The
// subtree-task slots of a wavefront's calls are taken with one atomic.
__global__ __launch_bounds__(256) void k_roots(const int64_t *__restrict__ hfirst,
                                               const int64_t *__restrict__ hcount,
                                               const int32_t *__restrict__ box,
                                               const int32_t *__restrict__ bad,
                                               const float4 *__restrict__ A,
                                               int32_t *__restrict__ badlist, int nhalos,
                                               BuildArgs a) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = h < nhalos;
    int32_t first = 0, nfin = 0, bucket = 0;
    bool isbad = false;
    if (v) {
        first = (int32_t)hfirst[h];
        const int32_t count = (int32_t)hcount[h];
        const int32_t nbad = bad[h];
        nfin = count - nbad;
        isbad = nbad > 0;
        bucket = tune_bucket(count);
    }
    const bool ok = v && !isbad;
    const Sizes sz = sizes_of(a, bucket);
    const bool wsmall = ok && wave_sized(sz, nfin), gsmall = ok && group_sized(sz, nfin);
    const uint64_t ms = ballot(wsmall), mg = ballot(gsmall);
    int32_t slot0 = 0, gslot0 = 0;
    if (lane_id() == 0) {
        if (ms) slot0 = (int32_t)(alloc_nodes_and_small(a, 0, __popcll(ms)) >> 32);
        if (mg) gslot0 = atomicAdd(&a.counters[kGroupNext], __popcll(mg));
    }
    slot0 = __shfl(slot0, 0, 64);
    gslot0 = __shfl(gslot0, 0, 64);
    if (isbad) {
        const int32_t k = atomicAdd(&a.counters[kBadRoots], 1);
        badlist[k] = h;
    }
    if (!ok) return;
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = nfin ? key_float(box[6 * h + k]) : 0.0f;
        hi[k] = nfin ? key_float(box[6 * h + 3 + k]) : 0.0f;
    }
    if (nfin <= kLeafPoints) {
        for (int32_t i = 0; i < nfin; i++) {
            a.fin[first + i] = A[first + i];
            a.leaf_of[first + i] = h;
        }
    }
    write_node(a.nodes + h, lo, hi, first, nfin, -1);
    a.parent[h] = -1;
    if (nfin > sz.big) push_big(a, h, lo, hi, nfin, 0, bucket);
    else if (wsmall) push_small_at(a, slot0 + rank_below(ms), h, lo, hi, 0, first, nfin);
    else if (gsmall) push_group_at(a, gslot0 + rank_below(mg), h, lo, hi, 0, first, nfin);
}

// The listed calls with non-finite points, one workgroup each: removal of
// the non-finite records to the tail in the order of the reference's swap
// loop (src/fast3tree.c:875-887), root node, root queued or finished as a
// leaf.
__global__ __launch_bounds__(256) void k_roots_bad(const int64_t *__restrict__ hfirst,
                                                   const int64_t *__restrict__ hcount,
                                                   const int32_t *__restrict__ box,
                                                   const int32_t *__restrict__ bad,
                                                   float4 *__restrict__ A,
                                                   const int32_t *__restrict__ badlist,
                                                   BuildArgs a) {
    __shared__ int s_cnt[2][4];
    const int32_t nlist = a.counters[kBadRoots];
    for (int32_t li = blockIdx.x; li < nlist; li += gridDim.x) {
        const int h = badlist[li];
        const int32_t first = (int32_t)hfirst[h];
        const int32_t count = (int32_t)hcount[h];
        const int32_t nbad = bad[h];
        const int32_t nfin = count - nbad;
        float lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = nfin ? key_float(box[6 * h + k]) : 0.0f;
            hi[k] = nfin ? key_float(box[6 * h + 3 + k]) : 0.0f;
        }
        // finite | non-finite in the order of the reference's swap loop
        // (inplace_dest): pass 0 lists the positions by class in the stable
        // order, pass 1 moves every point to its slot in the final array;
        // then the finite part goes back into A for splitting
        const int w = threadIdx.x / kWave;
        const int32_t *sd = a.sides + first;
        for (int pass = 0; pass < 2; pass++) {
            int32_t runf = 0, runb = 0;
            for (int32_t base = 0; base < count; base += 256) {
                int32_t i = base + threadIdx.x;
                bool vv = i < count;
                float4 p = vv ? A[first + i] : make_float4(0, 0, 0, 0);
                bool f = vv && finite3(p.x, p.y, p.z);
                bool bb = vv && !f;
                uint64_t mf = ballot(f), mb = ballot(bb);
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
                if (vv) {
                    const int32_t rank = f ? runf + of + rank_below(mf) : runb + ob + rank_below(mb);
                    if (pass == 0) {
                        a.sides[first + (f ? rank : nfin + rank)] = i;
                    } else {
                        const int32_t d = first + inplace_dest(i, f, rank, count, nfin,
                                                               [&](int32_t k) { return sd[k]; });
                        a.fin[d] = p;
                        if (!f) a.leaf_of[d] = -1;
                    }
                }
                runf += tf;
                runb += tb;
                __syncthreads();
            }
        }
        __syncthreads();
        // A keeps the compact finite points of every call (the direct kernel
        // reads them there; the split levels read them when nfin > 10)
        for (int32_t i = threadIdx.x; i < nfin; i += 256) A[first + i] = a.fin[first + i];
        if (nfin <= kLeafPoints)
            for (int32_t i = threadIdx.x; i < nfin; i += 256) a.leaf_of[first + i] = h;
        if (threadIdx.x == 0) {
            write_node(a.nodes + h, lo, hi, first, nfin, -1);
            a.parent[h] = -1;
            const int32_t bucket = tune_bucket((int32_t)hcount[h]);
            const Sizes sz = sizes_of(a, bucket);
            if (nfin > sz.big) {
                push_big(a, h, lo, hi, nfin, 0, bucket);
            } else if (wave_sized(sz, nfin)) {
                const uint64_t old = alloc_nodes_and_small(a, 0, 1);
                push_small_at(a, (int32_t)(old >> 32), h, lo, hi, 0, first, nfin);
            } else if (group_sized(sz, nfin)) {
                push_group_at(a, atomicAdd(&a.counters[kGroupNext], 1), h, lo, hi, 0, first, nfin);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ small tasks

// Writes the two children of a split node and queues those that need a
// split: small ones at slots P, P + 1 (allocated with the node ids), big ones
// at task slots B, B + 1 with their tiles from T0.
// Called by a whole wavefront with wavefront-uniform arguments.  P, G, B
// and T0 are reserved by the caller in the wavefront and workgroup subtree
// lists and the big-task list.
__device__ void make_children(const BuildArgs &a, const Task &t, const Node &nd, int32_t L,
                              int32_t P, int32_t G, int32_t B, int32_t T0, int32_t nl,
                              const float *loL, const float *hiL, const float *loR,
                              const float *hiR) {
    const int32_t nr = nd.count - nl;
    const Sizes sz = sizes_of(a, t.pad[0]);
    if (lane_id() == 0) {
        write_node(a.nodes + L, loL, hiL, nd.first, nl, L + 1);
        write_node(a.nodes + L + 1, loR, hiR, nd.first + nl, nr, nd.skip);
        a.parent[L] = a.parent[L + 1] = t.node;
        a.nodes[t.node].left = L;
        if (wave_sized(sz, nl)) push_small_at(a, P++, L, loL, hiL, t.depth + 1, nd.first, nl);
        if (group_sized(sz, nl)) push_group_at(a, G++, L, loL, hiL, t.depth + 1, nd.first, nl);
        if (wave_sized(sz, nr)) push_small_at(a, P, L + 1, loR, hiR, t.depth + 1, nd.first + nl, nr);
        if (group_sized(sz, nr)) push_group_at(a, G, L + 1, loR, hiR, t.depth + 1, nd.first + nl, nr);
    }
    // big children at the task slots from B and the tiles from T0 (left first)
    if (nl > sz.big) push_big_wave_at(a, B, T0, L, loL, hiL, nl, t.depth + 1, t.pad[0]);
    if (nr > sz.big)
        push_big_wave_at(a, B + (nl > sz.big ? 1 : 0), T0 + (nl > sz.big ? (nl + kTile - 1) / kTile : 0),
                         L + 1, loR, hiR, nr, t.depth + 1, t.pad[0]);
}

__device__ __forceinline__ int32_t wave_pushes(const Sizes &s, int32_t nl, int32_t nr) {
    return (wave_sized(s, nl) ? 1 : 0) + (wave_sized(s, nr) ? 1 : 0);
}
__device__ __forceinline__ int32_t group_pushes(const Sizes &s, int32_t nl, int32_t nr) {
    return (group_sized(s, nl) ? 1 : 0) + (group_sized(s, nr) ? 1 : 0);
}

// ------------------------------------------------------------- subtrees
//
// A node of 65..kSmallMax points is finished by one workgroup, a node of at
// most 64 points by one wavefront (below).  Both keep the points in LDS and
// number the subtree's nodes locally in creation order: levels in order,
// the two children of a split as a pair.
//
// Workgroup subtrees are point-parallel per level over the whole subtree:
// thread t takes positions t, t + 256, ...; a block-wide prefix of the
// left flags (one 64-bit ballot per chunk of 64 positions) gives every
// node's left count and every point's stable scatter position at once; the
// children's tight boxes come from segmented scans over runs of lanes, one
// LDS atomic per run and coordinate.  Per-node bookkeeping runs a thread per
// node.  Local ids map onto a block of 2n - 2 global ids reserved with one
// atomic when the subtree starts; ids it does not use are never referenced.
// Centres of mass are combined bottom-up in LDS and dmin is computed over
// the final order in LDS.

constexpr int kWaveTask = 64;   // subtrees up to this size: one wavefront
constexpr int kWaveMin = 32;    // the smallest wavefront/workgroup boundary (wave_max) used
constexpr int kRootSkip = -2;   // local skip pointer: the subtree root's own skip
constexpr int kSubThreads = 256;
constexpr int kSubWaves = kSubThreads / kWave;
constexpr uint8_t kNoEntry = 0xff;

// The workgroup subtree kernel comes in three sizes, for nodes of at most
// 256, 512 and 1,024 points: its LDS grows with the size, and the smaller
// sizes fit more workgroups on a CU.  The tasks are split by size into one
// list per kernel (k_group_classes) before the kernels run.
constexpr int kSubClasses = 3;
__host__ __device__ constexpr int sub_class(int32_t n) { return n > 512 ? 2 : (n > 256 ? 1 : 0); }
template <int kCap>
struct SubSize {
    static constexpr int chunks = kCap / kWave;                     // ballot chunks
    static constexpr int list = 2 * (kCap / (kLeafPoints + 1)) + 2; // nodes of one level
    static constexpr int nodes = 2 * kCap;                          // local ids (2n - 1 at most)
    static constexpr int per_thread = nodes / kSubThreads;
    static_assert(list < kNoEntry, "level entries fit a byte");
    static_assert(kCap >= kSubThreads && kCap <= kSmallMax && kCap % kSubThreads == 0,
                  "a whole number of positions per thread");
};

__device__ __forceinline__ uint64_t lanes_below(int i) {
    return i >= 64 ? ~0ull : ((1ull << i) - 1ull);
}

// bits [lf, lf + n) of a wavefront mask (lf + n <= 64)
__device__ __forceinline__ uint64_t lane_run(int lf, int n) {
    return (n >= 64 ? ~0ull : ((1ull << n) - 1ull)) << lf;
}

// Inclusive scan of v over this lane's run [seg, lane] of the wavefront
// (seg <= lane; runs are contiguous), on DPP moves: row_shr 1, 2, 4, 8
// within rows of 16, then row_bcast 15 and 31 across rows.
template <typename Op>
__device__ __forceinline__ float seg_scan(float v, int seg, Op op) {
    const int lane = lane_id(), r = lane & 15;
    float u;
    u = dpp_move<0x111>(v);
    if (r >= 1 && lane - 1 >= seg) v = op(v, u);
    u = dpp_move<0x112>(v);
    if (r >= 2 && lane - 2 >= seg) v = op(v, u);
    u = dpp_move<0x114>(v);
    if (r >= 4 && lane - 4 >= seg) v = op(v, u);
    u = dpp_move<0x118>(v);
    if (r >= 8 && lane - 8 >= seg) v = op(v, u);
    u = dpp_move<0x142>(v);  // row_bcast:15, used by rows 1 and 3
    if (((lane >> 4) & 1) && (lane & ~15) - 1 >= seg) v = op(v, u);
    u = dpp_move<0x143>(v);  // row_bcast:31, used by rows 2 and 3
    if (lane >= 32 && 31 >= seg) v = op(v, u);
    return v;
}

struct SubEntry {       // a node of the level being split, 20 bytes
    float lim;          // split plane, from the node's tight box
    int16_t loc;        // local node id
    int16_t lf, n, nl;  // first point (subtree-relative), points, left count
    int16_t child;      // local id of the left child (splits), else -1
    int16_t skip;       // local skip pointer
    uint8_t kind, dim;  // 1: split; split dimension
};

// Splits the workgroup subtree tasks by size into the lists of the three
// kernel sizes: list c at out + c * cap_group, counted in counters[kClassNext + c].
// Each task also gets its node ids here, 2n - 2 for n points (the most a
// subtree can need; ids it does not use are never referenced), the first
// kept in tile0: one atomic per wavefront instead of one per subtree.
__global__ __launch_bounds__(256) void k_group_classes(const Task *__restrict__ in, BuildArgs a,
                                                       Task *__restrict__ out) {
    const int32_t n = min(a.counters[kGroupNext], a.cap_group);
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k - lane_id() >= n) return;  // whole wavefronts past the list
    Task t;
    int c = -1;
    int32_t need = 0;
    if (k < n) {
        t = in[k];
        c = sub_class(t.pad[1]);
        need = 2 * t.pad[1] - 2;
    }
    const int32_t incl = wave_scan_add(need);
    int32_t ids = 0;
    if (lane_id() == kWave - 1)
        ids = (int32_t)alloc_nodes_and_small(a, incl, 0);
    t.tile0 = read_lane(ids, kWave - 1) + incl - need;
    for (int q = 0; q < kSubClasses; q++) {
        const uint64_t m = ballot(c == q);
        if (!m) continue;
        int32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(&a.counters[kClassNext + q], __popcll(m));
        base = __shfl(base, 0, 64);
        if (c == q) out[(int64_t)q * a.cap_group + base + rank_below(m)] = t;
    }
}

// One workgroup per task of the list `cls` (nodes of at most kCap points).
// The two smaller sizes ask the compiler for 7 wavefronts per SIMD (72
// VGPRs, a few values spilled outside the level loop): their LDS fits more
// workgroups than their registers otherwise allowed, and the latency-bound
// levels gain more from the extra wavefronts than the spills cost.
template <int kCap>
__global__ __launch_bounds__(kSubThreads, kCap <= 512 ? 7 : 1) void k_subtree(const float4 *__restrict__ A,
                                                         const float4 *__restrict__ B,
                                                         const Task *__restrict__ tasks,
                                                         BuildArgs a, double particle_mass,
                                                         int cls) {
    constexpr int kSubChunks = SubSize<kCap>::chunks, kSubList = SubSize<kCap>::list,
                  kSubNodes = SubSize<kCap>::nodes, kSubPerThread = SubSize<kCap>::per_thread;
    // The points stay where they were loaded (P); the order is a 16-bit
    // permutation, PM[buf][i] = the point at position i.  Once the levels are
    // done, the final order is moved to PM[0] and the local nodes' centres of
    // mass (CM) take the room of the level arrays: the LDS fits four
    // workgroups per CU at 1,024 points.
    struct Levels {
        int16_t PM[2][kCap];
        uint8_t Q[2][kCap];               // level entry of each position
        SubEntry E[2][kSubList];
        int32_t BOX[kSubList][6];         // next level's tight boxes, float keys
    };
    constexpr int kCM = (int)((sizeof(Levels) - sizeof(int16_t) * kCap) / sizeof(float4));
    struct Centres {
        int16_t PMF[kCap];                // the final order
        float4 CM[kCM];                   // centres of mass of local nodes j < kCM
    };
    static_assert(sizeof(Centres) <= sizeof(Levels) && kCM >= kCap / 2, "centre room");
    __shared__ float4 P[kCap];
    __shared__ union Pool {
        Levels lv;
        Centres cm;
    } pool;
    auto &PM = pool.lv.PM;
    auto &Q = pool.lv.Q;
    auto &E = pool.lv.E;
    auto &BOX = pool.lv.BOX;
    __shared__ int16_t s_left[kSubNodes];     // left child of each local node, -1: leaf
    __shared__ uint16_t s_lvl[kCap + 2]; // first local id of each level
    __shared__ uint64_t s_mask[kSubChunks + 1];
    __shared__ int32_t s_pre[kSubChunks + 1];
    __shared__ int32_t s_nn[2], s_M, s_nlev, s_base, s_next;
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    const int32_t ntasks = a.counters[kClassNext + cls];
    const float inf = __int_as_float(0x7f800000), ninf = __int_as_float((int)0xff800000);
    // A dynamic queue over the list: thread 0 takes the workgroup's next
    // entry when it starts the current one.  Entries are taken from the end
    // of the list: the largest subtrees (children of the big nodes of the
    // last split levels) are queued last and should start first.
    if (tid == 0) s_next = atomicAdd(&a.counters[kClassCursor + cls], 1);
    __syncthreads();
    for (int32_t k = s_next; k < ntasks; k = s_next) {
        int32_t k_after = 0;
        if (tid == 0) k_after = atomicAdd(&a.counters[kClassCursor + cls], 1);
        const Task t = tasks[ntasks - 1 - k];
        const int32_t first = t.pad[0], n = t.pad[1];
        const int32_t root_skip = a.nodes[t.node].skip;
        const float4 *src = (t.depth & 1) ? B : A;
        for (int i = tid; i < n; i += kSubThreads) {
            P[i] = src[first + i];
            PM[0][i] = (int16_t)i;
            Q[0][i] = 0;
        }
        if (tid <= kSubChunks) s_mask[tid] = 0;  // chunks past n stay empty
        const int rounds = (n + kSubThreads - 1) / kSubThreads;
        if (tid == 0) {
            SubEntry e;
            e.loc = 0;
            e.lf = 0;
            e.n = (int16_t)n;
            e.nl = 0;
            e.child = -1;
            e.skip = kRootSkip;
            e.kind = 0;
            e.dim = (uint8_t)t.dim;
            e.lim = t.lim;
            E[0][0] = e;
            s_nn[0] = 1;
            s_M = 1;
            s_nlev = 0;
            s_lvl[0] = 0;
            s_lvl[1] = 1;
            s_left[0] = -1;
            const int32_t R = 2 * n - 2;
            int32_t base = t.tile0;  // R ids reserved by k_group_classes
            if (base + R > a.cap_nodes) {  // out of node ids: the root stays a leaf
                atomicOr(&a.counters[kErrFlags], kErrCapacity);
                base = -1;
                s_nn[0] = 0;
            }
            s_base = base;
        }
        __syncthreads();
        const int32_t base = s_base;
        auto gid = [&](int j) { return j == 0 ? t.node : base + j - 1; };
        auto gskip = [&](int s) { return s == kRootSkip ? root_skip : gid(s); };
        int cur = 0, buf = 0;
        while (true) {
            const int nn = s_nn[cur];
            if (nn == 0) break;
            const int nxt = cur ^ 1;
            const int c0 = s_M;  // first local id of the next level
            // 1. left flags (src/fast3tree.c:751-774), one ballot per chunk
            for (int r = 0; r < rounds; r++) {
                const int i = r * kSubThreads + tid;
                bool gl = false;
                if (i < n) {
                    const uint8_t e = Q[cur][i];
                    if (e != kNoEntry && E[cur][e].n > kLeafPoints)
                        gl = comp(P[PM[buf][i]], E[cur][e].dim) <= E[cur][e].lim;
                }
                const uint64_t m = ballot(gl);
                if (lane == 0) s_mask[r * kSubWaves + w] = m;
            }
            __syncthreads();
            // 2. wavefront 0: chunk prefix, left counts, splits, child ids
            //    and the next level's entries
            if (w == 0) {
                const uint64_t mk = lane < kSubChunks ? s_mask[lane] : 0ull;
                const int cnt = __popcll(mk);
                int incl = cnt;
                for (int o = 1; o < kSubChunks; o <<= 1) {
                    const int u = __shfl_up(incl, o, 64);
                    if (lane >= o) incl += u;
                }
                const int excl = incl - cnt;
                const int total = __shfl(incl, kSubChunks - 1, 64);
                if (lane < kSubChunks) s_pre[lane] = excl;
                if (lane == kSubChunks) s_pre[kSubChunks] = total;
                // left flags in [0, x)
                auto S = [&](int x) {
                    const int c = x >> 6;
                    const int b = __shfl(excl, c < kSubChunks ? c : 0, 64);
                    return (c < kSubChunks ? b : total) +
                           __popcll(s_mask[c] & lanes_below(x & 63));
                };
                int M = c0;
                for (int k0 = 0; k0 < nn; k0 += kWave) {
                    const int kk = k0 + lane;
                    const bool v = kk < nn;
                    SubEntry en;
                    en.lf = en.n = 0;
                    if (v) en = E[cur][kk];
                    const int s0 = S(en.lf), s1 = S(en.lf + en.n);
                    const int nl = s1 - s0;
                    // src/fast3tree.c:805-809: a split with an empty side is a leaf
                    const bool split = v && en.n > kLeafPoints && nl > 0 && nl < en.n;
                    const uint64_t ms = ballot(split);
                    const int C = M + 2 * rank_below(ms);
                    if (v) {
                        E[cur][kk].nl = (int16_t)nl;
                        E[cur][kk].kind = split ? 1 : 0;
                        E[cur][kk].child = split ? (int16_t)C : (int16_t)-1;
                        s_left[en.loc] = split ? (int16_t)C : (int16_t)-1;
                    }
                    if (split) {
                        const int slot = C - c0;
                        SubEntry L, R;
                        L.loc = (int16_t)C;
                        L.lf = en.lf;
                        L.n = (int16_t)nl;
                        L.skip = (int16_t)(C + 1);
                        R.loc = (int16_t)(C + 1);
                        R.lf = (int16_t)(en.lf + nl);
                        R.n = (int16_t)(en.n - nl);
                        R.skip = en.skip;
                        L.nl = R.nl = 0;
                        L.child = R.child = -1;
                        L.kind = R.kind = 0;
                        L.dim = R.dim = 0;
                        L.lim = R.lim = 0.0f;
                        E[nxt][slot] = L;
                        E[nxt][slot + 1] = R;
                        for (int d = 0; d < 3; d++) {
                            BOX[slot][d] = BOX[slot + 1][d] = float_key(inf);
                            BOX[slot][3 + d] = BOX[slot + 1][3 + d] = float_key(ninf);
                        }
                        a.nodes[gid(en.loc)].left = gid(C);
                        a.parent[gid(C)] = a.parent[gid(C + 1)] = gid(en.loc);
                    }
                    M += 2 * __popcll(ms);
                }
                if (lane == 0) {
                    s_M = M;
                    s_nn[nxt] = M - c0;
                    if (M > c0) {
                        s_nlev = s_nlev + 1;
                        s_lvl[s_nlev + 1] = (uint16_t)M;
                    }
                }
            }
            __syncthreads();
            // 3. scatter: split points to their child's place in the next
            //    buffer, in the order of the reference's in-place partition
            //    (inplace_dest); leaf points are in tree order already.  The
            //    node's k-th left (right) point is the (Sx(lf) + k)-th set
            //    (lf - Sx(lf) + k)-th clear bit of the left flags: a binary
            //    search over the chunk prefixes, then a select in the word.
            auto Sx = [&](int x) {
                const int c = x >> 6;
                return s_pre[c] + __popcll(s_mask[c] & lanes_below(x & 63));
            };
            // r-th set / clear bit overall, searched among chunks [lo, hi]
            auto nth_left = [&](int r, int lo, int hi) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_pre[mid] <= r) lo = mid;
                    else hi = mid - 1;
                }
                return lo * 64 + select_bit(s_mask[lo], r - s_pre[lo]);
            };
            auto nth_right = [&](int r, int lo, int hi) {
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (64 * mid - s_pre[mid] <= r) lo = mid;
                    else hi = mid - 1;
                }
                return lo * 64 + select_bit(~s_mask[lo], r - (64 * lo - s_pre[lo]));
            };
            for (int r = 0; r < rounds; r++) {
                const int i = r * kSubThreads + tid;
                if (i >= n) break;
                const int16_t pi = PM[buf][i];
                const float4 p = P[pi];
                const uint8_t e = Q[cur][i];
                int dst = i;
                uint8_t q = kNoEntry;
                if (e != kNoEntry) {
                    const SubEntry en = E[cur][e];
                    if (en.kind) {
                        const bool gl = comp(p, en.dim) <= en.lim;
                        const int lf = en.lf, s0 = Sx(lf), sl = Sx(i) - s0, nl = en.nl;
                        const int clo = lf >> 6, chi = (lf + en.n - 1) >> 6;  // the node's chunks
                        dst = lf + inplace_dest(i - lf, gl, gl ? sl : i - lf - sl, en.n, nl,
                                                [&](int k) {
                                                    return k < nl
                                                        ? nth_left(s0 + k, clo, chi) - lf
                                                        : nth_right(lf - s0 + k - nl, clo, chi) - lf;
                                                });
                        q = (uint8_t)(en.child - c0 + (gl ? 0 : 1));
                    } else {
                        a.fin[first + i] = p;
                        a.leaf_of[first + i] = gid(en.loc);
                    }
                }
                PM[buf ^ 1][dst] = pi;
                Q[nxt][dst] = q;
            }
            __syncthreads();
            // 4. tight boxes of the next level's nodes
            const int nn_next = s_nn[nxt];
            if (nn_next > 0) {
                for (int r = 0; r < rounds; r++) {
                    const int i = r * kSubThreads + tid;
                    const uint8_t q = i < n ? Q[nxt][i] : kNoEntry;
                    const bool fresh = q != kNoEntry;
                    if (!ballot(fresh)) continue;
                    const float4 p = fresh ? P[PM[buf ^ 1][i]] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const int lf = fresh ? E[nxt][q].lf : i;
                    const int chunk0 = i - lane;
                    const int seg = lf > chunk0 ? lf - chunk0 : 0;
                    const int sg = fresh ? seg : lane;
                    auto mn = [](float x, float y) { return fminf(x, y); };
                    auto mx = [](float x, float y) { return fmaxf(x, y); };
                    const float l0 = seg_scan(p.x, sg, mn), l1 = seg_scan(p.y, sg, mn),
                                l2 = seg_scan(p.z, sg, mn);
                    const float h0 = seg_scan(p.x, sg, mx), h1 = seg_scan(p.y, sg, mx),
                                h2 = seg_scan(p.z, sg, mx);
                    const int qn = __shfl_down((int)q, 1, 64);
                    const bool last = fresh && (lane == kWave - 1 || i + 1 >= n || qn != (int)q);
                    if (last) {
                        atomicMin(&BOX[q][0], float_key(l0));
                        atomicMin(&BOX[q][1], float_key(l1));
                        atomicMin(&BOX[q][2], float_key(l2));
                        atomicMax(&BOX[q][3], float_key(h0));
                        atomicMax(&BOX[q][4], float_key(h1));
                        atomicMax(&BOX[q][5], float_key(h2));
                    }
                }
            }
            __syncthreads();
            // 5. the next level's node records and split planes, a thread each
            for (int c = tid; c < nn_next; c += kSubThreads) {
                SubEntry &en = E[nxt][c];
                float lo[3], hi[3];
                for (int d = 0; d < 3; d++) {
                    lo[d] = key_float(BOX[c][d]);
                    hi[d] = key_float(BOX[c][3 + d]);
                }
                const int dim = split_dim(lo, hi);
                en.dim = (uint8_t)dim;
                en.lim = split_lim(lo, hi, dim);
                write_node(a.nodes + gid(en.loc), lo, hi, first + en.lf, en.n, gskip(en.skip));
            }
            __syncthreads();
            cur = nxt;
            buf ^= 1;
        }
        if (base < 0) {  // the root stays a leaf over the points in input order
            for (int i = tid; i < n; i += kSubThreads) {
                a.fin[first + i] = P[i];
                a.leaf_of[first + i] = t.node;
            }
        }
        // centres of mass and dmin (src/potential.c:71-124): the final order
        // to PM[0] (PMF), the rest of the level arrays' room for com (x, y,
        // z) and mass
        if (buf) {
            for (int i = tid; i < n; i += kSubThreads) PM[0][i] = PM[1][i];
            __syncthreads();
        }
        const int M = s_M, nlev = s_nlev;
        const int16_t *PMF = pool.cm.PMF;
        float4 *CM = pool.cm.CM;
        auto F = [&](int i) { return P[PMF[i]]; };
        auto cm_get = [&](int j) {
            if (j < kCM) return CM[j];
            const Node *np = a.nodes + gid(j);
            return make_float4(np->com[0], np->com[1], np->com[2], np->mass);
        };
        auto cm_put = [&](int j, float4 v) {
            if (j < kCM) CM[j] = v;
            Node *np = a.nodes + gid(j);
            np->com[0] = v.x;
            np->com[1] = v.y;
            np->com[2] = v.z;
            np->mass = v.w;
        };
        int32_t my_lf[kSubPerThread], my_n[kSubPerThread];
#pragma unroll
        for (int q = 0; q < kSubPerThread; q++) {
            const int j = tid + q * kSubThreads;
            my_lf[q] = my_n[q] = 0;
            if (j < M) {
                const Node *np = a.nodes + gid(j);
                my_lf[q] = np->first - first;
                my_n[q] = np->count;
            }
        }
        // leaves: f64 sums in tree order (src/potential.c:100-107)
#pragma unroll
        for (int q = 0; q < kSubPerThread; q++) {
            const int j = tid + q * kSubThreads;
            if (j >= M || s_left[j] >= 0) continue;
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int i = my_lf[q]; i < my_lf[q] + my_n[q]; i++) {
                const float4 p = F(i);
                sx += (double)p.x;
                sy += (double)p.y;
                sz += (double)p.z;
            }
            const double dn = (double)my_n[q];
            cm_put(j, make_float4(sx != 0.0 ? (float)(sx / dn) : 0.0f,
                                  sy != 0.0 ? (float)(sy / dn) : 0.0f,
                                  sz != 0.0 ? (float)(sz / dn) : 0.0f, (float)(particle_mass * dn)));
        }
        __threadfence_block();
        __syncthreads();
        // internal nodes, deepest level first (src/potential.c:115-122:
        // float, right term first).  One wavefront walks all levels: a level
        // reads only values this wavefront wrote at deeper levels (or the
        // leaves, made visible by the barrier above), so the levels need no
        // workgroup barrier, only program order within the wavefront.
        if (w == 0) {
            for (int L = nlev; L >= 0; L--) {
                const int l1 = s_lvl[L + 1];
                for (int j = s_lvl[L] + lane; j < l1; j += kWave) {
                    const int cl = s_left[j];
                    if (cl < 0) continue;
                    const float4 l = cm_get(cl), r = cm_get(cl + 1);
                    const float m = l.w + r.w;
                    float c[3];
                    const float lc[3] = {l.x, l.y, l.z}, rc[3] = {r.x, r.y, r.z};
                    for (int d = 0; d < 3; d++) {
                        const float rr = rc[d] * r.w;
                        const float ll = lc[d] * l.w;
                        c[d] = m != 0.0f ? (rr + ll) / m : 0.0f;
                    }
                    cm_put(j, make_float4(c[0], c[1], c[2], m));
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
        }
        __threadfence_block();
        __syncthreads();
        // dmin: nodes of at most 64 points a thread each, in tree order
#pragma unroll
        for (int q = 0; q < kSubPerThread; q++) {
            const int j = tid + q * kSubThreads;
            if (j >= M || my_n[q] > kWave) continue;
            const float4 c4 = cm_get(j);
            const float com[3] = {c4.x, c4.y, c4.z};
            double big = 0.0, sum = 0.0;
#pragma unroll 4
            for (int i = my_lf[q]; i < my_lf[q] + my_n[q]; i++) {
                const double d = d2_ref(com, F(i));
                big = d > big ? d : big;
                sum += d;
            }
            finish_dmin(a.nodes + gid(j), big, sum, my_n[q]);
        }
        // larger nodes a wavefront each: the wavefront takes the nodes of
        // its lanes' chunk (j = tid + k * kSubThreads) that have more than
        // 64 points, one after another, reading each node's range lane-parallel
#pragma unroll
        for (int k = 0; k < kSubPerThread; k++) {
            const int j0 = w * kWave + k * kSubThreads;
            if (j0 >= M) break;
            uint64_t wide = ballot(tid + k * kSubThreads < M && my_n[k] > kWave);
            while (wide) {
                const int q = __builtin_ctzll(wide);
                wide &= wide - 1;
                const int32_t qn = read_lane(my_n[k], q), ql = read_lane(my_lf[k], q);
                const float4 c4 = cm_get(j0 + q);
                const float com[3] = {c4.x, c4.y, c4.z};
                double big = 0.0, sum = 0.0;
                for (int32_t i = lane; i < qn; i += kWave) {
                    const double d = d2_ref(com, F(ql + i));
                    big = d > big ? d : big;
                    sum += d;
                }
                big = wave_max(big);
                sum = wave_sum(sum);
                if (lane == 0) finish_dmin(a.nodes + gid(j0 + q), big, sum, qn);
            }
        }
        if (tid == 0) s_next = k_after;
        __syncthreads();
    }
}

// ---------------------------------------------------- wavefront subtrees
//
// A subtree of at most kWaveTask points is built by one wavefront, one point
// per lane.  Each level, every lane tests its point against its node's split
// plane; ballots give the left counts and the stable scatter positions (a
// node's points occupy a contiguous run of lanes), and the children's tight
// boxes come from segmented scans across the lanes.  The node table stays in
// LDS with local ids (children in pairs, levels in order) until the subtree
// is complete; then one atomic reserves exactly its node ids.  Centres of
// mass and dmin are computed with a lane per node, summing the node's points
// in tree order, the order of the reference's loops.

constexpr int kWaveNodes = 2 * kWaveTask;  // a tree over n points has at most 2n - 1 nodes

// Node ids of the wavefront subtrees, assigned before they are built: a
// subtree of n points gets 2n - 2 ids (the most it can need; ids it does not
// use are never referenced) from one block above every id in use.  Task k's
// first id is counters[kWaveIdBase] plus the exclusive prefix of the needs
// before it: k_wave_need, the library's exclusive scan, then k_wave_base.
// The subtree kernel then needs no atomic per task.
__host__ __device__ __forceinline__ int32_t wave_ids(int32_t n) { return 2 * n - 2; }
__global__ void k_wave_need(const Task *__restrict__ tasks, BuildArgs a, int32_t *__restrict__ need,
                            int32_t m) {
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x)
        need[k] = k < ntasks ? wave_ids(tasks[k].pad[1]) : 0;
}
// after the scan (off holds cap_small + 1 entries): the block of ids
__global__ void k_wave_base(BuildArgs a, const int32_t *__restrict__ off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    const int32_t base0 = a.counters[kNodeCount];
    a.counters[kWaveIdBase] = base0;
    a.counters[kNodeCount] = base0 + off[ntasks];
}

__global__ __launch_bounds__(kWave) void k_subtree_wave(const float4 *__restrict__ A,
                                                        const float4 *__restrict__ B,
                                                        const Task *__restrict__ tasks,
                                                        BuildArgs a, double particle_mass,
                                                        const int32_t *__restrict__ wave_off) {
    // under 5 KB of LDS, so that 8 wavefronts fit per SIMD: the nodes' boxes
    // go to their global records as they are found (only each node's split
    // plane stays), parents are written from the parent's side
    __shared__ float4 P[kWaveTask];
    __shared__ int16_t Q[kWaveTask];
    __shared__ float s_lim[kWaveNodes];
    __shared__ uint8_t s_dim[kWaveNodes];
    __shared__ float s_com[3][kWaveNodes], s_mass[kWaveNodes];
    __shared__ int16_t s_lf[kWaveNodes], s_cnt[kWaveNodes], s_left[kWaveNodes];
    __shared__ int16_t s_skip[kWaveNodes];
    __shared__ int16_t s_lvl[kWaveTask + 2];  // first local id of each level
    const int lane = lane_id();
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    const int32_t id_base = a.counters[kWaveIdBase];
    // tasks k = blockIdx.x, blockIdx.x + gridDim.x, ...: they are all small
    // and alike, and a fixed stride needs no queue atomic per task
    for (int32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const Task t = tasks[k];
        const int32_t n = t.pad[1];
        if (n > kWaveTask) {  // never queued here; guards the lane-per-point layout
            if (lane == 0) atomicOr(&a.counters[kErrFlags], kErrBadTree);
            continue;
        }
        // the task carries the root's range and split plane, so the points
        // load at once; the root's node record is needed only at the end
        const float4 *src = (t.depth & 1) ? B : A;
        const int32_t first = t.pad[0];
        const bool v = lane < n;
        float4 p = v ? src[first + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const Node root = a.nodes[t.node];
        // the task's node ids (2n - 2 from the scan); a subtree that does not
        // fit keeps its root as a leaf (checked once the levels are done)
        const int32_t base0 = id_base + wave_off[k];
        const bool fits = base0 + wave_ids(n) <= a.cap_nodes;
        if (lane == 0) {
            s_lf[0] = 0;
            s_cnt[0] = (int16_t)n;
            s_left[0] = -1;
            s_skip[0] = kRootSkip;
            s_lvl[0] = 0;
            s_lvl[1] = 1;
        }
        __syncthreads();
        int nid = 0, M = 1, nlev = 0;
        bool pend = v && n > kLeafPoints;
        while (ballot(pend)) {
            int lf = 0, c = 0, dim = 0;
            float lim = 0.0f;
            if (pend) {
                lf = s_lf[nid];
                c = s_cnt[nid];
                if (nid == 0) {  // make_task's plane from the root's box
                    dim = t.dim;
                    lim = t.lim;
                } else {
                    dim = s_dim[nid];
                    lim = s_lim[nid];
                }
            }
            // src/fast3tree.c:751-774 (partition), 805-809 (failed split)
            const bool gl = pend && comp(p, dim) <= lim;
            const uint64_t mL = ballot(gl);
            const uint64_t mN = pend ? lane_run(lf, c) : 0ull;
            const int nl = __popcll(mL & mN);
            const bool split = pend && nl > 0 && nl < c;
            const uint64_t mH = ballot(split && lane == lf);
            const int C = M + 2 * __popcll(mH & lanes_below(lf));
            int dst = lane, nn = nid;
            // the in-place partition's order (inplace_dest): the node's k-th
            // left (right) point is the k-th set bit of its left (right)
            // lanes' mask
            const uint64_t below = mN & lanes_below(lane);
            const int rank = gl ? __popcll(mL & below) : __popcll(~mL & below);
            if (split) {
                const uint64_t mLn = mL & mN, mRn = ~mL & mN;
                dst = lf + inplace_dest(lane - lf, gl, rank, c, nl, [&](int k) {
                    return (k < nl ? select_bit(mLn, k) : select_bit(mRn, k - nl)) - lf;
                });
                nn = gl ? C : C + 1;
                if (lane == lf) {
                    s_lf[C] = (int16_t)lf;
                    s_cnt[C] = (int16_t)nl;
                    s_lf[C + 1] = (int16_t)(lf + nl);
                    s_cnt[C + 1] = (int16_t)(c - nl);
                    s_left[C] = s_left[C + 1] = -1;
                    s_skip[C] = (int16_t)(C + 1);
                    s_skip[C + 1] = s_skip[nid];
                    s_left[nid] = (int16_t)C;
                }
            }
            if (v) {
                P[dst] = p;
                Q[dst] = nn;
            }
            __syncthreads();
            if (v) {
                p = P[lane];
                nid = Q[lane];
            }
            const int M0 = M;
            M += 2 * __popcll(mH);
            // tight boxes of the new children: segmented scans over their runs
            const bool fresh = v && nid >= M0;
            const int seg = fresh ? s_lf[nid] : lane;
            auto mn = [](float x, float y) { return fminf(x, y); };
            auto mx = [](float x, float y) { return fmaxf(x, y); };
            const float l0 = seg_scan(p.x, seg, mn), l1 = seg_scan(p.y, seg, mn),
                        l2 = seg_scan(p.z, seg, mn);
            const float h0 = seg_scan(p.x, seg, mx), h1 = seg_scan(p.y, seg, mx),
                        h2 = seg_scan(p.z, seg, mx);
            const int cnt_new = fresh ? s_cnt[nid] : 0;
            if (fresh && lane == seg + cnt_new - 1) {
                const float lo3[3] = {l0, l1, l2}, hi3[3] = {h0, h1, h2};
                const int dm = split_dim(lo3, hi3);
                s_dim[nid] = (uint8_t)dm;
                s_lim[nid] = split_lim(lo3, hi3, dm);
                if (fits) {
                    Node *g = a.nodes + base0 + nid - 1;
                    for (int d = 0; d < 3; d++) {
                        g->lo[d] = lo3[d];
                        g->hi[d] = hi3[d];
                    }
                }
            }
            if (M > M0) {
                nlev++;
                if (lane == 0) s_lvl[nlev + 1] = (int16_t)M;
            }
            __syncthreads();
            pend = fresh && cnt_new > kLeafPoints;
        }
        // every lane's node is a leaf now; local id j > 0 is node base + j - 1
        // of the ids reserved for the task (its box is there already)
        int32_t base = base0;
        if (M > 1 && !fits) {
            if (lane == 0) atomicOr(&a.counters[kErrFlags], kErrCapacity);
            base = -1;
        }
        if (base < 0) {  // out of node ids: the root stays a leaf
            M = 1;
            nlev = 0;
            nid = 0;
            if (lane == 0) s_left[0] = -1;
        }
        if (v) {
            P[lane] = p;
            a.fin[first + lane] = p;
            a.leaf_of[first + lane] = nid ? base + nid - 1 : t.node;
        }
        __syncthreads();
        // leaf centres: f64 sums in tree order (src/potential.c:100-107)
        for (int j = lane; j < M; j += kWave) {
            if (s_left[j] >= 0) continue;
            const int lf = s_lf[j], c = s_cnt[j];
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int i = lf; i < lf + c; i++) {
                const float4 q = P[i];
                sx += (double)q.x;
                sy += (double)q.y;
                sz += (double)q.z;
            }
            const double dn = (double)c;
            s_com[0][j] = sx != 0.0 ? (float)(sx / dn) : 0.0f;
            s_com[1][j] = sy != 0.0 ? (float)(sy / dn) : 0.0f;
            s_com[2][j] = sz != 0.0 ? (float)(sz / dn) : 0.0f;
            s_mass[j] = (float)(particle_mass * dn);
        }
        __syncthreads();
        // internal centres, deepest level first (src/potential.c:115-122:
        // float, right term first)
        for (int L = nlev; L >= 0; L--) {
            for (int j = s_lvl[L] + lane; j < s_lvl[L + 1]; j += kWave) {
                const int cl = s_left[j];
                if (cl < 0) continue;
                const float ml = s_mass[cl], mr = s_mass[cl + 1];
                const float m = ml + mr;
                for (int d = 0; d < 3; d++) {
                    const float rr = s_com[d][cl + 1] * mr;
                    const float ll = s_com[d][cl] * ml;
                    s_com[d][j] = m != 0.0f ? (rr + ll) / m : 0.0f;
                }
                s_mass[j] = m;
            }
            __syncthreads();
        }
        // dmin over each node's points in tree order (src/potential.c:71-82),
        // then the node records
        for (int j = lane; j < M; j += kWave) {
            const int lf = s_lf[j], c = s_cnt[j];
            const float com[3] = {s_com[0][j], s_com[1][j], s_com[2][j]};
            double big = 0.0, sum = 0.0;
            for (int i = lf; i < lf + c; i++) {
                const double d = d2_ref(com, P[i]);
                big = d > big ? d : big;
                sum += d;
            }
            Node nd;
            for (int d = 0; d < 3; d++) nd.com[d] = com[d];
            nd.mass = s_mass[j];
            nd.first = first + lf;
            nd.count = c;
            const int cl = s_left[j], sk = s_skip[j];
            nd.left = cl >= 0 ? base + cl - 1 : -1;
            nd.skip = sk == kRootSkip ? root.skip : base + sk - 1;
            finish_dmin(&nd, big, sum, c);
            const int32_t id = j ? base + j - 1 : t.node;
            // every field but the box (the root's is the task's, the others
            // were written with their level)
            Node *g = a.nodes + id;
            for (int d = 0; d < 3; d++) g->com[d] = nd.com[d];
            g->mass = nd.mass;
            g->dmin2 = nd.dmin2;
            g->first = nd.first;
            g->count = nd.count;
            g->left = nd.left;
            g->skip = nd.skip;
            g->dmin = nd.dmin;
            if (cl >= 0) a.parent[base + cl - 1] = a.parent[base + cl] = id;
        }
        __syncthreads();
    }
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

// Count pass over the tiles of split level `level`.
__global__ __launch_bounds__(kTileThreads) void k_big_count(const float4 *__restrict__ src,
                                                            const int2 *__restrict__ tiles,
                                                            const Task *__restrict__ big,
                                                            const Node *__restrict__ nodes,
                                                            TilePart *__restrict__ parts,
                                                            const int32_t *__restrict__ lv,
                                                            int32_t level) {
    const int32_t t1 = lvl_tile(lv)[level];
    for (int32_t ti = lvl_tile(lv)[level - 1] + blockIdx.x; ti < t1; ti += gridDim.x)
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

// The split of big task bi (t, its node nd), by one wavefront once its
// tiles are counted: the left count and the children's boxes from the tiles'
// parts (lanes over tiles), then the children.  The node ids and subtree
// slots, the workgroup-subtree slots and the big children's task slots and
// tiles are reserved by three lanes at once (one round trip).
__device__ void split_task(const TilePart *__restrict__ parts, BigSplit *__restrict__ res,
                           const BuildArgs &a, int32_t bi, const Task &t, const Node &nd) {
    const int lane = lane_id();
    const float inf = __int_as_float(0x7f800000), ninf = __int_as_float((int)0xff800000);
    float lo[2][3] = {{inf, inf, inf}, {inf, inf, inf}};
    float hi[2][3] = {{ninf, ninf, ninf}, {ninf, ninf, ninf}};
    int32_t nl = 0;
    for (int32_t j = lane; j < t.ntiles; j += kWave) {
        const TilePart *p = parts + t.tile0 + j;
        nl += p->nleft;
        for (int s = 0; s < 2; s++)
            for (int d = 0; d < 3; d++) {
                lo[s][d] = fminf(lo[s][d], p->lo[s][d]);
                hi[s][d] = fmaxf(hi[s][d], p->hi[s][d]);
            }
    }
    nl = wave_sum(nl);
    for (int s = 0; s < 2; s++)
        for (int d = 0; d < 3; d++) {
            lo[s][d] = wave_min(lo[s][d]);
            hi[s][d] = wave_max(hi[s][d]);
        }
    // nl, nd and the reduced boxes are wavefront-uniform from here on
    BigSplit r;
    r.nleft = nl;
    r.left = -1;
    r.leafL = r.leafR = 0;
    if (nl == 0 || nl == nd.count) {
        if (lane == 0) note_fat_leaf(a, t.node, nd.count);
    } else {
        const int32_t nr = nd.count - nl;
        const Sizes sz = sizes_of(a, t.pad[0]);
        const int32_t ng = group_pushes(sz, nl, nr);
        const int32_t ntl = nl > sz.big ? (nl + kTile - 1) / kTile : 0;
        const int32_t ntr = nr > sz.big ? (nr + kTile - 1) / kTile : 0;
        const int32_t nbig = (ntl ? 1 : 0) + (ntr ? 1 : 0);
        uint64_t al = 0, bal = 0;
        int32_t G = 0;
        if (lane == 0) al = alloc_nodes_and_small(a, 2, wave_pushes(sz, nl, nr));
        if (lane == 1 && ng) G = atomicAdd(&a.counters[kGroupNext], ng);
        if (lane == 2 && nbig)
            bal = atomicAdd(reinterpret_cast<unsigned long long *>(a.counters + kBigNext),
                            (unsigned long long)nbig | ((unsigned long long)(ntl + ntr) << 32));
        al = read_lane(al, 0);
        G = read_lane(G, 1);
        bal = read_lane(bal, 2);
        const int32_t L = (int32_t)al, P = (int32_t)(al >> 32);
        if (L + 2 > a.cap_nodes) {
            if (lane == 0) atomicOr(&a.counters[kErrFlags], kErrCapacity);
        } else {
            r.left = L;
            r.leafL = nl <= kLeafPoints;
            r.leafR = nr <= kLeafPoints;
            make_children(a, t, nd, L, P, G, (int32_t)bal, (int32_t)(bal >> 32), nl, lo[0], hi[0],
                          lo[1], hi[1]);
        }
    }
    if (lane == 0) res[bi] = r;
}

// The two passes over one tile of a big task, with the points' ranks on
// their side of the plane in rounds of 256 (a block prefix of ballots):
//   kSides:   lists the node's points by side, a.sides[first + k] = the
//             position of the k-th left point, then of the right points
//             (the stable order).  Each workgroup takes its tile's left
//             prefix and the task's left total from the count pass's parts
//             (and keeps the prefix in its part for the scatter pass); the
//             workgroup of the task's first tile also splits the task.
//   kScatter: moves each point to the slot the reference's in-place
//             partition gives it (inplace_dest over that list).  Its first
//             workgroup moves the level state forward (level_done).
enum { kSides = 0, kScatter = 1 };

template <int kPass>
__device__ void big_tile_pass(const float4 *__restrict__ src, float4 *__restrict__ dst,
                              const int2 *__restrict__ tiles, const Task *__restrict__ big,
                              TilePart *__restrict__ parts, BigSplit *__restrict__ res,
                              const BuildArgs &a, int32_t ti);
__device__ void level_done(int32_t *c, int32_t *lv, int32_t level);

template <int kPass>
__global__ __launch_bounds__(kTileThreads) void k_big_tiles(
    const float4 *__restrict__ src, float4 *__restrict__ dst, const int2 *__restrict__ tiles,
    const Task *__restrict__ big, TilePart *__restrict__ parts, BigSplit *__restrict__ res,
    BuildArgs a, int32_t *__restrict__ lv, int32_t level) {
    // (the scatter pass makes no pushes: the counters it reads here are final)
    if (kPass == kScatter && blockIdx.x == 0 && threadIdx.x == 0) level_done(a.counters, lv, level);
    const int32_t t1 = lvl_tile(lv)[level];
    for (int32_t ti = lvl_tile(lv)[level - 1] + blockIdx.x; ti < t1; ti += gridDim.x)
        big_tile_pass<kPass>(src, dst, tiles, big, parts, res, a, ti);
}

template <int kPass>
__device__ void big_tile_pass(const float4 *__restrict__ src, float4 *__restrict__ dst,
                              const int2 *__restrict__ tiles, const Task *__restrict__ big,
                              TilePart *__restrict__ parts, BigSplit *__restrict__ res,
                              const BuildArgs &a, int32_t ti) {
    __shared__ int s_cnt[2][kTileThreads / kWave];
    const int2 tl = tiles[ti];
    const Task t = big[tl.x];
    const Node nd = a.nodes[t.node];
    const int w = threadIdx.x / kWave;
    int32_t lpre, nleft;
    BigSplit r;
    if (kPass == kSides) {
        // left points of the task's earlier tiles and of all its tiles (every
        // wavefront sums them: no barrier needed to share them)
        const int32_t me = tl.y / kTile;
        int32_t before = 0, all = 0;
        for (int32_t j = lane_id(); j < t.ntiles; j += kWave) {
            const int32_t c = parts[t.tile0 + j].nleft;
            before += j < me ? c : 0;
            all += c;
        }
        lpre = wave_sum(before);
        nleft = wave_sum(all);
        if (threadIdx.x == 0) parts[ti].pad[0] = lpre;
        if (me == 0 && w == 0) split_task(parts, res, a, tl.x, t, nd);
        if (nleft == 0 || nleft == nd.count) return;  // failed split: no listing
    } else {
        r = res[tl.x];
        nleft = r.nleft;
        if (r.left < 0) {  // failed split: the node is a leaf, its order unchanged
            for (int j = threadIdx.x; j < kTile; j += kTileThreads) {
                int32_t i = tl.y + j;
                if (i >= nd.count) break;
                a.fin[nd.first + i] = src[nd.first + i];
                a.leaf_of[nd.first + i] = t.node;
            }
            return;
        }
        lpre = parts[ti].pad[0];
    }
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
            const int32_t rank = f ? lpre + runL + oL + rank_below(mL)
                                   : rpre + runR + oR + rank_below(mR);
            if (kPass == kSides) {
                a.sides[nd.first + (f ? rank : nleft + rank)] = i;
            } else {
                const int32_t *sd = a.sides + nd.first;
                const int32_t d = nd.first + inplace_dest(i, f, rank, nd.count, nleft,
                                                          [&](int32_t k) { return sd[k]; });
                const bool leaf = f ? r.leafL : r.leafR;  // left points land in [0, nleft)
                const int32_t id = f ? r.left : r.left + 1;
                if (leaf) {
                    a.fin[d] = p;
                    a.leaf_of[d] = id;
                } else {
                    dst[d] = p;
                }
            }
        }
        runL += tL;
        runR += tR;
        __syncthreads();
    }
}

// The level state on the device.  After the roots, level 1's big tasks and
// tiles are the pushes so far; after split level L's scatter pass (whose
// counters are final: the split happened in the sides pass), level L + 1's
// are the pushes of level L.  depth_end[L] records the node count (node
// ids grow with depth).  A dropped push (capacity) ends the levels.
__device__ void level_done(int32_t *c, int32_t *lv, int32_t level) {
    const int32_t nodes = c[kNodeCount];
    int32_t err = c[kErrFlags];
    if (level >= kMaxLevels) {
        atomicOr(&c[kErrFlags], kErrCapacity);
        err |= kErrCapacity;
    } else {
        lv[level] = nodes;
    }
    const bool stop = (err & kErrCapacity) != 0;
    int32_t *big = lvl_big(lv), *tile = lvl_tile(lv);
    big[level + 1] = stop ? big[level] : c[kBigNext];
    tile[level + 1] = stop ? tile[level] : c[kTileNext];
    c[kGlobalNodes] = nodes;
    c[kLevel] = level + 1;
    c[kActive] = big[level + 1] > big[level] ? 1 : 0;
}
__global__ void k_levels_init(int32_t *__restrict__ c, int32_t *__restrict__ lv) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    lvl_big(lv)[0] = 0;
    lvl_tile(lv)[0] = 0;
    level_done(c, lv, 0);
}

// --------------------------------------------------------------- multipoles

// The multipole kernels take their sizes from the device counters (node
// count, fat-leaf count, per-depth node ranges, big tasks and tiles), so a
// whole build can be replayed from a graph.

// Leaves made by the split levels (at most 10 points): mass, centre
// (src/potential.c:100-107, f64 sums) and dmin (src/potential.c:71-82), one
// thread each.  Subtree nodes have ids at or above counters[kGlobalNodes]
// and are finished by their workgroup.
__global__ void k_leaf_com(Node *__restrict__ nodes, const float4 *__restrict__ fin,
                           const int32_t *__restrict__ counters, double particle_mass) {
    const int32_t nnodes = counters[kGlobalNodes];
    for (int32_t id = blockIdx.x * blockDim.x + threadIdx.x; id < nnodes;
         id += gridDim.x * blockDim.x) {
        const Node nd = nodes[id];
        if (nd.left >= 0 || nd.count <= 0 || nd.count > kLeafPoints) continue;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        for (int32_t i = 0; i < nd.count; i++) {
            const float4 p = fin[nd.first + i];
            s0 += (double)p.x;
            s1 += (double)p.y;
            s2 += (double)p.z;
        }
        const double n = (double)nd.count;
        float com[3];
        com[0] = s0 != 0.0 ? (float)(s0 / n) : 0.0f;
        com[1] = s1 != 0.0 ? (float)(s1 / n) : 0.0f;
        com[2] = s2 != 0.0 ? (float)(s2 / n) : 0.0f;
        double big = 0.0, sum = 0.0;
        for (int32_t i = 0; i < nd.count; i++) {
            const double d = d2_ref(com, fin[nd.first + i]);
            big = d > big ? d : big;
            sum += d;
        }
        Node *me = nodes + id;
        me->com[0] = com[0];
        me->com[1] = com[1];
        me->com[2] = com[2];
        me->mass = (float)(particle_mass * n);
        finish_dmin(me, big, sum, nd.count);
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

// The device wall clock at this point of the stream.
__global__ void k_stamp(uint64_t *t, int i) {
    if (threadIdx.x == 0 && blockIdx.x == 0) t[i] = (uint64_t)wall_clock64();
}

// After the last split level of a recorded step: trees still being split
// need more levels than were recorded.
__global__ void k_depth_check(int32_t *__restrict__ c) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && c[kActive]) c[kErrFlags] |= kErrDeep;
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
