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

// ------------------------------------------------------------- subtrees
//
// Every node of at most kSmallMax points is finished by one workgroup: its
// points are staged in LDS (two buffers, ping-pong per level), the subtree
// is split level by level (the wavefronts share the nodes of a level),
// leaves go straight to tree order, and the centres of mass (bottom-up by
// level) and dmin (over the final order) of every subtree node are computed
// before the workgroup moves on.  Node ids come from a reservation made
// once per subtree (refilled if a subtree needs more); unused reserved ids
// are written as empty nodes.

constexpr int kWaveTask = 64;   // subtrees up to this size: one wavefront
constexpr int kSubThreads = 256;
constexpr int kSubWaves = kSubThreads / kWave;
constexpr int kSubList = 2 * (kSmallMax / (kLeafPoints + 1)) + 2;  // entries of one level

struct SubNode {
    int32_t id;
    uint16_t lf;     // first point, relative to the subtree
    uint16_t n;
    uint16_t nl;
    uint8_t kind;    // 0 leaf, 1 split
    uint8_t dim;     // split plane of the node (from its tight box)
    float lim;
    int32_t child;   // first child id (splits)
    int32_t skip;    // the node's skip pointer
};

__device__ __forceinline__ void write_empty_node(const BuildArgs &a, int32_t id) {
    const float z[3] = {0.0f, 0.0f, 0.0f};
    write_node(a.nodes + id, z, z, 0, 0, -1);
    a.parent[id] = -1;
}

__device__ __forceinline__ int32_t even_up(int32_t x) { return (x + 1) & ~1; }

__device__ __forceinline__ SubNode sub_entry(int32_t id, int lf, int n, const float *lo,
                                             const float *hi, int32_t skip) {
    SubNode e;
    e.id = id;
    e.lf = (uint16_t)lf;
    e.n = (uint16_t)n;
    e.nl = 0;
    e.kind = 0;
    const int d = split_dim(lo, hi);
    e.dim = (uint8_t)d;
    e.lim = split_lim(lo, hi, d);
    e.child = -1;
    e.skip = skip;
    return e;
}

__global__ __launch_bounds__(kSubThreads) void k_subtree(const float4 *__restrict__ A,
                                                         const float4 *__restrict__ B,
                                                         const Task *__restrict__ tasks,
                                                         BuildArgs a, double particle_mass) {
    __shared__ float4 P[2][kSmallMax];
    __shared__ SubNode E[2][kSubList];
    __shared__ int32_t made[2 * kSmallMax];
    __shared__ uint16_t lvl_end[kSmallMax + 1];
    __shared__ int32_t s_n[2], s_made, s_levels, s_pool, s_pool_end, s_base;
    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    // from the end of the list: the largest subtrees (children of the big
    // nodes of the last split levels) are queued last and should start first
    for (int32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const Task t = tasks[ntasks - 1 - k];
        const Node root = a.nodes[t.node];
        const float4 *src = (t.depth & 1) ? B : A;
        const int32_t first = root.first, count = root.count;
        if (count <= kWaveTask) continue;  // a wavefront subtree
        for (int32_t i = tid; i < count; i += kSubThreads) P[0][i] = src[first + i];
        if (tid == 0) {
            E[0][0] = sub_entry(t.node, 0, count, root.lo, root.hi, root.skip);
            s_n[0] = 1;
            made[0] = t.node;
            s_made = 1;
            lvl_end[0] = 1;
            s_levels = 1;
            const int32_t R = even_up(count * 2 / 5 + 8);
            const int32_t base = (int32_t)alloc_nodes_and_small(a, R, 0);
            s_pool = base;
            s_pool_end = base + R;
            if (base + R > a.cap_nodes) {
                atomicOr(&a.counters[kErrFlags], kErrCapacity);
                s_n[0] = 0;  // leave this subtree unsplit
                s_pool = s_pool_end = 0;
            }
        }
        __syncthreads();
        int cur = 0, buf = 0;
        while (s_n[cur] > 0) {
            const int nn = s_n[cur];
            // pass A: left count of every node of the level (src/fast3tree.c:751-774)
            for (int k = w; k < nn; k += kSubWaves) {
                const SubNode e = E[cur][k];
                int kind = 0, nl = 0;
                if (e.n > kLeafPoints) {
                    for (int c = 0; c < e.n; c += kWave) {
                        const int i = c + lane;
                        const bool f = i < e.n && comp(P[buf][e.lf + i], e.dim) <= e.lim;
                        nl += __popcll(ballot(f));
                    }
                    kind = (nl == 0 || nl == e.n) ? 0 : 1;  // src/fast3tree.c:805-809
                }
                if (lane == 0) {
                    E[cur][k].kind = (uint8_t)kind;
                    E[cur][k].nl = (uint16_t)nl;
                }
            }
            __syncthreads();
            // wavefront 0: node ids for the splits, in level order
            if (w == 0) {
                int32_t need = 0;
                for (int c = 0; c < nn; c += kWave)
                    need += 2 * __popcll(ballot(c + lane < nn && E[cur][c + lane].kind));
                int32_t base = 0;
                if (lane == 0) {
                    if (s_pool + need > s_pool_end) {
                        for (int32_t j = s_pool; j < s_pool_end; j++) write_empty_node(a, j);
                        const int32_t R = even_up(max(need, count / 4 + 8));
                        const int32_t nb = (int32_t)alloc_nodes_and_small(a, R, 0);
                        s_pool = nb;
                        s_pool_end = nb + R;
                        if (nb + R > a.cap_nodes) {
                            atomicOr(&a.counters[kErrFlags], kErrCapacity);
                            s_pool = s_pool_end = 0;
                            need = -1;
                        }
                    }
                    base = s_pool;
                }
                base = __shfl(base, 0, 64);
                need = __shfl(need, 0, 64);
                int32_t done = 0;
                for (int c = 0; c < nn; c += kWave) {
                    const int k = c + lane;
                    bool split = k < nn && E[cur][k].kind;
                    if (need < 0 && k < nn) {  // out of node ids: finish as leaves
                        E[cur][k].kind = 0;
                        split = false;
                    }
                    const uint64_t m = ballot(split);
                    if (split) E[cur][k].child = base + 2 * (done + rank_below(m));
                    done += __popcll(m);
                }
                if (lane == 0) {
                    s_base = base;
                    if (need >= 0) s_pool = base + 2 * done;
                    s_n[cur ^ 1] = 2 * done;
                }
            }
            __syncthreads();
            const int32_t level_base = s_base;
            // pass B: leaves to tree order, splits to the other buffer
            for (int k = w; k < nn; k += kSubWaves) {
                const SubNode e = E[cur][k];
                if (!e.kind) {
                    for (int i = lane; i < e.n; i += kWave) {
                        a.fin[first + e.lf + i] = P[buf][e.lf + i];
                        a.leaf_of[first + e.lf + i] = e.id;
                    }
                    if (lane == 0) {
                        // src/potential.c:100-107: f64 sums in tree order
                        double sx = 0.0, sy = 0.0, sz = 0.0;
                        for (int i = 0; i < e.n; i++) {
                            const float4 p = P[buf][e.lf + i];
                            sx += (double)p.x;
                            sy += (double)p.y;
                            sz += (double)p.z;
                        }
                        const double n = (double)e.n;
                        Node *me = a.nodes + e.id;
                        me->com[0] = sx != 0.0 ? (float)(sx / n) : 0.0f;
                        me->com[1] = sy != 0.0 ? (float)(sy / n) : 0.0f;
                        me->com[2] = sz != 0.0 ? (float)(sz / n) : 0.0f;
                        me->mass = (float)(particle_mass * n);
                    }
                    continue;
                }
                const int slot = e.child - level_base;
                const float inf = __int_as_float(0x7f800000), ninf = __int_as_float((int)0xff800000);
                float lx0 = inf, ly0 = inf, lz0 = inf, hx0 = ninf, hy0 = ninf, hz0 = ninf;
                float lx1 = inf, ly1 = inf, lz1 = inf, hx1 = ninf, hy1 = ninf, hz1 = ninf;
                int32_t bl = 0, br = 0;
                for (int c = 0; c < e.n; c += kWave) {
                    const int i = c + lane;
                    const bool v = i < e.n;
                    const float4 p = v ? P[buf][e.lf + i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const bool f = v && comp(p, e.dim) <= e.lim;
                    const bool g = v && !f;
                    const uint64_t mL = ballot(f), mR = ballot(g);
                    if (v) {
                        const int d = f ? e.lf + bl + rank_below(mL) : e.lf + e.nl + br + rank_below(mR);
                        P[buf ^ 1][d] = p;
                    }
                    if (f) {
                        lx0 = fminf(lx0, p.x); hx0 = fmaxf(hx0, p.x);
                        ly0 = fminf(ly0, p.y); hy0 = fmaxf(hy0, p.y);
                        lz0 = fminf(lz0, p.z); hz0 = fmaxf(hz0, p.z);
                    }
                    if (g) {
                        lx1 = fminf(lx1, p.x); hx1 = fmaxf(hx1, p.x);
                        ly1 = fminf(ly1, p.y); hy1 = fmaxf(hy1, p.y);
                        lz1 = fminf(lz1, p.z); hz1 = fmaxf(hz1, p.z);
                    }
                    bl += __popcll(mL);
                    br += __popcll(mR);
                }
                const float lo0[3] = {wave_min(lx0), wave_min(ly0), wave_min(lz0)};
                const float hi0[3] = {wave_max(hx0), wave_max(hy0), wave_max(hz0)};
                const float lo1[3] = {wave_min(lx1), wave_min(ly1), wave_min(lz1)};
                const float hi1[3] = {wave_max(hx1), wave_max(hy1), wave_max(hz1)};
                if (lane == 0) {
                    const int32_t C = e.child;
                    write_node(a.nodes + C, lo0, hi0, first + e.lf, e.nl, C + 1);
                    write_node(a.nodes + C + 1, lo1, hi1, first + e.lf + e.nl, e.n - e.nl, e.skip);
                    a.parent[C] = a.parent[C + 1] = e.id;
                    a.nodes[e.id].left = C;
                    E[cur ^ 1][slot] = sub_entry(C, e.lf, e.nl, lo0, hi0, C + 1);
                    E[cur ^ 1][slot + 1] = sub_entry(C + 1, e.lf + e.nl, e.n - e.nl, lo1, hi1, e.skip);
                    made[s_made + slot] = C;
                    made[s_made + slot + 1] = C + 1;
                }
            }
            __syncthreads();
            if (tid == 0) {
                s_made += s_n[cur ^ 1];
                lvl_end[s_levels++] = (uint16_t)s_made;
            }
            cur ^= 1;
            buf ^= 1;
            __syncthreads();
        }
        // unused reserved ids become empty nodes
        for (int32_t j = s_pool + tid; j < s_pool_end; j += kSubThreads) write_empty_node(a, j);
        __threadfence_block();
        __syncthreads();
        // centres of mass of the internal nodes, deepest level first
        // (src/potential.c:115-122: float, right term first)
        for (int lv = s_levels - 1; lv >= 0; lv--) {
            const int32_t b0 = lv ? lvl_end[lv - 1] : 0, b1 = lvl_end[lv];
            for (int32_t j = b0 + tid; j < b1; j += kSubThreads) {
                Node *me = a.nodes + made[j];
                const int32_t L = me->left;
                if (L < 0) continue;
                const Node *l = a.nodes + L, *r = a.nodes + L + 1;
                const float ml = l->mass, mr = r->mass;
                const float m = ml + mr;
                for (int k = 0; k < 3; k++) {
                    const float rr = r->com[k] * mr;
                    const float ll = l->com[k] * ml;
                    me->com[k] = m != 0.0f ? (rr + ll) / m : 0.0f;
                }
                me->mass = m;
            }
            __threadfence_block();
            __syncthreads();
        }
        // dmin over the final order (src/potential.c:71-82)
        for (int32_t i = tid; i < count; i += kSubThreads) P[0][i] = a.fin[first + i];
        __syncthreads();
        // nodes of at most 64 points: a thread each, summing in tree order
        for (int32_t j = tid; j < s_made; j += kSubThreads) {
            Node *np = a.nodes + made[j];
            const int32_t cnt = np->count;
            if (cnt > kWave) continue;
            const int32_t rel = np->first - first;
            const float com[3] = {np->com[0], np->com[1], np->com[2]};
            double big = 0.0, sum = 0.0;
            for (int32_t i = rel; i < rel + cnt; i++) {
                const double d = d2_ref(com, P[0][i]);
                big = d > big ? d : big;
                sum += d;
            }
            finish_dmin(np, big, sum, cnt);
        }
        // larger nodes: a wavefront each; each wavefront takes 64 nodes at a
        // time, their records loaded lane-parallel, then walked one by one
        for (int32_t j0 = w * kWave; j0 < s_made; j0 += kSubWaves * kWave) {
            const int32_t jj = j0 + lane;
            int32_t my_rel = 0, my_cnt = 0;
            float my_c0 = 0.0f, my_c1 = 0.0f, my_c2 = 0.0f;
            if (jj < s_made) {
                const Node *np = a.nodes + made[jj];
                my_rel = np->first - first;
                my_cnt = np->count;
                my_c0 = np->com[0];
                my_c1 = np->com[1];
                my_c2 = np->com[2];
            }
            const int nb = min(kWave, s_made - j0);
            for (int q = 0; q < nb; q++) {
                const int32_t rel = __shfl(my_rel, q, 64), cnt = __shfl(my_cnt, q, 64);
                if (cnt <= kWave) continue;  // done above
                const float com[3] = {__shfl(my_c0, q, 64), __shfl(my_c1, q, 64),
                                      __shfl(my_c2, q, 64)};
                double big = 0.0, sum = 0.0;
                for (int32_t i = lane; i < cnt; i += kWave) {
                    const double d = d2_ref(com, P[0][rel + i]);
                    big = d > big ? d : big;
                    sum += d;
                }
                big = wave_max(big);
                sum = wave_sum(sum);
                if (lane == 0) finish_dmin(a.nodes + made[j0 + q], big, sum, cnt);
            }
        }
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
constexpr int kRootSkip = -2;              // local skip pointer: the subtree root's own skip

__device__ __forceinline__ uint64_t lanes_below(int i) {
    return i >= 64 ? ~0ull : ((1ull << i) - 1ull);
}

// bits [lf, lf + n) of a wavefront mask (lf + n <= 64)
__device__ __forceinline__ uint64_t lane_run(int lf, int n) {
    return (n >= 64 ? ~0ull : ((1ull << n) - 1ull)) << lf;
}

__global__ __launch_bounds__(kWave) void k_subtree_wave(const float4 *__restrict__ A,
                                                        const float4 *__restrict__ B,
                                                        const Task *__restrict__ tasks,
                                                        BuildArgs a, double particle_mass) {
    __shared__ float4 P[kWaveTask];
    __shared__ int32_t Q[kWaveTask];
    __shared__ float s_lo[3][kWaveNodes], s_hi[3][kWaveNodes];
    __shared__ float s_com[3][kWaveNodes], s_mass[kWaveNodes];
    __shared__ int16_t s_lf[kWaveNodes], s_cnt[kWaveNodes], s_left[kWaveNodes];
    __shared__ int16_t s_par[kWaveNodes], s_skip[kWaveNodes];
    __shared__ int16_t s_lvl[kWaveTask + 2];  // first local id of each level
    const int lane = lane_id();
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    for (int32_t k = blockIdx.x; k < ntasks; k += gridDim.x) {
        const Task t = tasks[k];
        const Node root = a.nodes[t.node];
        const int32_t n = root.count;
        if (n > kWaveTask) continue;  // a workgroup subtree
        const float4 *src = (t.depth & 1) ? B : A;
        const int32_t first = root.first;
        const bool v = lane < n;
        float4 p = v ? src[first + lane] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (lane == 0) {
            for (int d = 0; d < 3; d++) {
                s_lo[d][0] = root.lo[d];
                s_hi[d][0] = root.hi[d];
            }
            s_lf[0] = 0;
            s_cnt[0] = (int16_t)n;
            s_left[0] = -1;
            s_par[0] = -1;
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
                const float lo3[3] = {s_lo[0][nid], s_lo[1][nid], s_lo[2][nid]};
                const float hi3[3] = {s_hi[0][nid], s_hi[1][nid], s_hi[2][nid]};
                dim = split_dim(lo3, hi3);
                lim = split_lim(lo3, hi3, dim);
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
            if (split) {
                const uint64_t below = mN & lanes_below(lane);
                dst = gl ? lf + __popcll(mL & below) : lf + nl + __popcll(~mL & below);
                nn = gl ? C : C + 1;
                if (lane == lf) {
                    s_lf[C] = (int16_t)lf;
                    s_cnt[C] = (int16_t)nl;
                    s_lf[C + 1] = (int16_t)(lf + nl);
                    s_cnt[C + 1] = (int16_t)(c - nl);
                    s_left[C] = s_left[C + 1] = -1;
                    s_par[C] = s_par[C + 1] = (int16_t)nid;
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
            float l0 = p.x, l1 = p.y, l2 = p.z, h0 = p.x, h1 = p.y, h2 = p.z;
            for (int o = 1; o < kWave; o <<= 1) {
                const float u0 = __shfl_up(l0, o, 64), u1 = __shfl_up(l1, o, 64),
                            u2 = __shfl_up(l2, o, 64), w0 = __shfl_up(h0, o, 64),
                            w1 = __shfl_up(h1, o, 64), w2 = __shfl_up(h2, o, 64);
                if (lane - o >= seg) {
                    l0 = fminf(l0, u0);
                    l1 = fminf(l1, u1);
                    l2 = fminf(l2, u2);
                    h0 = fmaxf(h0, w0);
                    h1 = fmaxf(h1, w1);
                    h2 = fmaxf(h2, w2);
                }
            }
            const int cnt_new = fresh ? s_cnt[nid] : 0;
            if (fresh && lane == seg + cnt_new - 1) {
                s_lo[0][nid] = l0;
                s_lo[1][nid] = l1;
                s_lo[2][nid] = l2;
                s_hi[0][nid] = h0;
                s_hi[1][nid] = h1;
                s_hi[2][nid] = h2;
            }
            if (M > M0) {
                nlev++;
                if (lane == 0) s_lvl[nlev + 1] = (int16_t)M;
            }
            __syncthreads();
            pend = fresh && cnt_new > kLeafPoints;
        }
        // every lane's node is a leaf now; reserve the subtree's node ids
        int32_t base = 0;
        if (lane == 0 && M > 1) {
            base = (int32_t)alloc_nodes_and_small(a, M - 1, 0);
            if (base + M - 1 > a.cap_nodes) {
                atomicOr(&a.counters[kErrFlags], kErrCapacity);
                base = -1;
            }
        }
        base = __shfl(base, 0, 64);
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
            for (int d = 0; d < 3; d++) {
                nd.lo[d] = s_lo[d][j];
                nd.hi[d] = s_hi[d][j];
                nd.com[d] = com[d];
            }
            nd.mass = s_mass[j];
            nd.first = first + lf;
            nd.count = c;
            const int cl = s_left[j], sk = s_skip[j];
            nd.left = cl >= 0 ? base + cl - 1 : -1;
            nd.skip = sk == kRootSkip ? root.skip : base + sk - 1;
            finish_dmin(&nd, big, sum, c);
            const int32_t id = j ? base + j - 1 : t.node;
            a.nodes[id] = nd;
            if (j) a.parent[id] = s_par[j] ? base + s_par[j] - 1 : t.node;
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
// One wavefront per big task of the level: exclusive scan of the tiles'
// left counts (lanes over tiles), the children's boxes, then the split.
__global__ __launch_bounds__(256) void k_big_split(const Task *__restrict__ big,
                                                   TilePart *__restrict__ parts,
                                                   BigSplit *__restrict__ res, BuildArgs a) {
    const int32_t b1 = a.counters[kBigEnd];
    const int32_t nw = (int32_t)(gridDim.x * blockDim.x / kWave);
    const int lane = lane_id();
    for (int32_t bi = a.counters[kBigBegin] + (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
         bi < b1; bi += nw) {
        const Task t = big[bi];
        const Node nd = a.nodes[t.node];
        const float inf = __int_as_float(0x7f800000), ninf = __int_as_float((int)0xff800000);
        float lo[2][3] = {{inf, inf, inf}, {inf, inf, inf}};
        float hi[2][3] = {{ninf, ninf, ninf}, {ninf, ninf, ninf}};
        int32_t nl = 0;
        for (int32_t j0 = 0; j0 < t.ntiles; j0 += kWave) {
            const int32_t jj = j0 + lane;
            const bool v = jj < t.ntiles;
            TilePart *p = parts + t.tile0 + (v ? jj : 0);
            const int32_t c = v ? p->nleft : 0;
            int32_t incl = c;
            for (int o = 1; o < kWave; o <<= 1) {
                const int32_t u = __shfl_up(incl, o, 64);
                if (lane >= o) incl += u;
            }
            if (v) {
                p->pad[0] = nl + incl - c;  // left points in the task's earlier tiles
                for (int s = 0; s < 2; s++)
                    for (int d = 0; d < 3; d++) {
                        lo[s][d] = fminf(lo[s][d], p->lo[s][d]);
                        hi[s][d] = fmaxf(hi[s][d], p->hi[s][d]);
                    }
            }
            nl += __shfl(incl, kWave - 1, 64);
        }
        for (int s = 0; s < 2; s++)
            for (int d = 0; d < 3; d++) {
                lo[s][d] = wave_min(lo[s][d]);
                hi[s][d] = wave_max(hi[s][d]);
            }
        if (lane != 0) continue;
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
    c[kGlobalNodes] = c[kNodeCount];
    c[kBigBegin] = c[kBigEnd];
    c[kBigEnd] = c[kBigNext];
    c[kTileBegin] = c[kTileEnd];
    c[kTileEnd] = c[kTileNext];
    const int32_t lvl = c[kLevel];
    if (lvl < kMaxLevels) depth_end[lvl] = c[kNodeCount];
    else c[kErrFlags] |= kErrCapacity;
    c[kLevel] = lvl + 1;
    c[kActive] = c[kBigEnd] > c[kBigBegin] ? 1 : 0;
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
