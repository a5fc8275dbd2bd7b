// hp_subfof.hip -- the phase-space sub-FOF step of the halo finder's
// _find_subs for a batch of FOF groups: _find_subfofs_better2
// (src/groupies.c:287-314).  Per group:
//   1. norm_sd (src/groupies.c:788-830): f64 mean, centring in float, the
//      6x6 second moments of float products, calc_deviations
//      (src/jacobi.c:137-175) on each 3x3 block, scaling by sigma_x and
//      sigma_v;
//   2. the 6D GROUPIES fast3tree (40 points per leaf, src/groupies.c:24-28)
//      rebuilt in place: the tree order is the reference's exactly, because
//      the sample and the linking both walk it;
//   3. next-closest distances of the sample (src/fast3tree.c:953-1030);
//   4. find_median_r (src/groupies.c:744-786) with the reference's rand_r
//      pivots: the linking length;
//   5. the linking (_find_subfofs_at_r, src/groupies.c:274-285 with
//      src/fof.c): every pair closer than that length joins one component;
//   6. the output as the reference leaves it: the particles sorted by its
//      union-find roots (collapse_smallfofs, partition_sort_particles) and
//      the sub-FOFs in root order (build_fullfofs, src/fof.c:159-252).
//
// The tree is built level by level over every group of the batch, one
// workgroup per node that still splits.  fast3tree's in-place partition
// (src/fast3tree.c:751-774) swaps each point above the midpoint with the
// current last slot and examines the swapped-in point next.  Its result has
// a closed form that a workgroup evaluates in parallel.  With the points
// of a node at positions 0..n-1, bigs (above the midpoint) and smalls,
// b_k the position of the k-th big from the left and r_k that of the k-th
// small from the right (r_0 = n), K the number of k with b_k < r_k, and S
// the number of points the loop reads from the left (b_{K+1} + 1 when that
// big lies before r_K, else r_K):
//   small at p < S: stays at p          big at p < S, c bigs before it: r_c - 1
//   small at p >= S, c smalls after it: b_{c+1}        big at p >= S: p - 1
// find_median_r's rad_partition is the same loop over [left, right - 1)
// after the pivot moves to the end, so the selection uses the same form.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

constexpr int kSubTile = 2048;        // norm_sd tile (one wavefront)
constexpr int kLeaf6 = 40;            // POINTS_PER_LEAF, src/groupies.c:25
constexpr int kSfThreads = 256;      // split and selection workgroups
constexpr int kSampleMax = 10000;     // MAX_PARTICLES_TO_SAMPLE, src/groupies.c:286

struct SParticle {  // struct particle, src/particle.h:5-8
    int64_t id;
    float pos[6];
};
static_assert(sizeof(SParticle) == 32, "particle layout");

struct SubGroup {
    int64_t first;       // the caller's first particle
    int32_t off, n;      // working slots [off, off + n)
    int32_t top;         // a whole FOF group: sigma_x scaled (src/groupies.c:821-822)
    int32_t nsample;     // distances drawn for find_median_r
    int32_t sample_off;  // into the distance and drawn-position buffers
    int32_t drawn;       // 1: sample positions from the drawn list, else 0..n-1
    int32_t k;           // find_median_r's k
    uint32_t seed;       // rand_r state when find_median_r starts
    int32_t tile0, ntiles;
    int32_t draw_off;    // into the drawn-position lists (drawn groups only)
    int32_t pad;
};

struct SubTile {
    int32_t group, off;
};

struct Node6 {
    float lo[6], hi[6];
    int32_t start, n;    // working slots
    int32_t div;         // split dimension, -1 for a leaf
    int32_t left, right, parent, skip;  // skip: next node after this subtree, depth first
    int32_t where;       // during the build: the buffer holding the node's points (0 or 1)
};
static_assert(sizeof(Node6) == 80, "node layout");

enum SfCounter {
    kSfNodes = 0,
    kSfActive = 1,     // nodes to split next level with kSfThreads
    kSfEmit = 2,
    kSfErr = 3,
    kSfActiveBig = 4,  // nodes above kSfBigNode points: split with 1024 threads
    kSfFused = 5,      // nodes of at most kFuseMax points: finished by k_sub_fused
    kSfEdges = 6,      // 64-bit (6 and 7): pairs the linking walk listed
    kSfRange = 8,      // output-order ranges above kSortMid values to split next level
    kSfRangeSmall = 9, // output-order ranges finished by one wavefront
    kSfRangeMid = 10,  // output-order ranges of at most kSortMid values to split next level
    kSfCounters = 11
};
constexpr int kSfBigNode = 8192;

__device__ __forceinline__ float pcoord(const SParticle &p, int d) {
    return d == 0 ? p.pos[0]
                  : (d == 1 ? p.pos[1]
                            : (d == 2 ? p.pos[2] : (d == 3 ? p.pos[3] : (d == 4 ? p.pos[4] : p.pos[5]))));
}
__device__ __forceinline__ float acoord(const float (&a)[6], int d) {
    return d == 0 ? a[0] : (d == 1 ? a[1] : (d == 2 ? a[2] : (d == 3 ? a[3] : (d == 4 ? a[4] : a[5]))));
}

// glibc's rand_r: the reference draws its sample and its selection pivots
// with it (src/groupies.c:300-304, 742).
__host__ __device__ __forceinline__ int rand_r_glibc(uint32_t *seed) {
    uint32_t next = *seed;
    int32_t result;
    next *= 1103515245u;
    next += 12345u;
    result = (int32_t)((next / 65536u) % 2048u);
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (int32_t)((next / 65536u) % 1024u);
    next *= 1103515245u;
    next += 12345u;
    result <<= 10;
    result ^= (int32_t)((next / 65536u) % 1024u);
    *seed = next;
    return result;
}

// ------------------------------------------------------------------ norm_sd

// copy the group's particles into the working slots, f64 sums per tile,
// non-finite input flagged (the reference's tree build asserts on it)
__global__ __launch_bounds__(kWave) void k_sub_load(const SParticle *__restrict__ parts,
                                                    const SubGroup *__restrict__ groups,
                                                    const SubTile *__restrict__ tiles,
                                                    SParticle *__restrict__ w,
                                                    int32_t *__restrict__ slot_group,
                                                    double *__restrict__ psum,
                                                    int32_t *__restrict__ gerr) {
    const SubTile tl = tiles[blockIdx.x];
    const SubGroup g = groups[tl.group];
    double s[6] = {0, 0, 0, 0, 0, 0};
    bool bad = false;
    const int32_t end = min(tl.off + kSubTile, g.n);
    for (int32_t i = tl.off + lane_id(); i < end; i += kWave) {
        const SParticle p = parts[g.first + i];
        w[g.off + i] = p;
        slot_group[g.off + i] = tl.group;
        for (int j = 0; j < 6; j++) {
            bad |= !__builtin_isfinite(p.pos[j]);
            s[j] += (double)p.pos[j];
        }
    }
    for (int j = 0; j < 6; j++) s[j] = wave_sum(s[j]);
    if (ballot(bad) && lane_id() == 0) gerr[tl.group] = 1;
    if (lane_id() == 0)
        for (int j = 0; j < 6; j++) psum[6 * blockIdx.x + j] = s[j];
}

// src/groupies.c:797-802: the mean, tiles combined in order
__global__ void k_sub_mean(int32_t ng, const SubGroup *__restrict__ groups,
                           const double *__restrict__ psum, double *__restrict__ mean) {
    const int32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const SubGroup g = groups[gi];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t t = g.tile0; t < g.tile0 + g.ntiles; t++)
        for (int j = 0; j < 6; j++) s[j] += psum[6 * t + j];
    for (int j = 0; j < 6; j++) mean[6 * gi + j] = g.n ? s[j] / (double)g.n : 0.0;
}

// src/groupies.c:806-813: centre in float, second moments of float products
__global__ __launch_bounds__(kWave) void k_sub_centre(const SubGroup *__restrict__ groups,
                                                      const SubTile *__restrict__ tiles,
                                                      const double *__restrict__ mean,
                                                      SParticle *__restrict__ w,
                                                      double *__restrict__ pcorr) {
    const SubTile tl = tiles[blockIdx.x];
    const SubGroup g = groups[tl.group];
    double mu[6], c[21];
    for (int j = 0; j < 6; j++) mu[j] = mean[6 * tl.group + j];
    for (int q = 0; q < 21; q++) c[q] = 0.0;
    const int32_t end = min(tl.off + kSubTile, g.n);
    for (int32_t i = tl.off + lane_id(); i < end; i += kWave) {
        SParticle p = w[g.off + i];
        for (int j = 0; j < 6; j++) p.pos[j] = (float)((double)p.pos[j] - mu[j]);
        w[g.off + i] = p;
        int q = 0;
        for (int j = 0; j < 6; j++)
            for (int k = j; k < 6; k++) {
                const float prod = p.pos[j] * p.pos[k];
                c[q++] += (double)prod;
            }
    }
    for (int q = 0; q < 21; q++) c[q] = wave_sum(c[q]);
    if (lane_id() == 0)
        for (int q = 0; q < 21; q++) pcorr[21 * blockIdx.x + q] = c[q];
}

// src/jacobi.c:137-151: sqrt of the smallest eigenvalue, or of the trace
__device__ double set_eig(const double (&in)[3][3]) {
    Jacobi3 jb;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) jb.m[i][j] = in[i][j];
    jb.run();
    double r = jb.eig[0];
    if (jb.eig[1] < r) r = jb.eig[1];
    if (jb.eig[2] < r) r = jb.eig[2];
    if (r <= 0) return sqrt(in[0][0] + in[1][1] + in[2][2]);
    return sqrt(r);
}

// src/groupies.c:815-825 with calc_deviations (src/jacobi.c:154-175)
__global__ void k_sub_sigma(int32_t ng, const SubGroup *__restrict__ groups,
                            const double *__restrict__ pcorr, double scaling,
                            double *__restrict__ sig) {
    const int32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const SubGroup g = groups[gi];
    double c[21];
    for (int q = 0; q < 21; q++) c[q] = 0.0;
    for (int32_t t = g.tile0; t < g.tile0 + g.ntiles; t++)
        for (int q = 0; q < 21; q++) c[q] += pcorr[21 * t + q];
    double m[6][6];
    int q = 0;
    for (int j = 0; j < 6; j++)
        for (int k = j; k < 6; k++) {
            m[j][k] = g.n ? c[q] / (double)g.n : 0.0;
            m[k][j] = m[j][k];
            q++;
        }
    double a[3][3], b[3][3];
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
            a[j][k] = m[j][k];
            b[j][k] = m[j + 3][k + 3];
        }
    double sx = set_eig(a);
    const double sv = set_eig(b);
    if (g.top) sx *= scaling;
    sig[2 * gi] = sx;
    sig[2 * gi + 1] = sv;
}

// src/groupies.c:827-829, then the root box of each group's tree
__global__ __launch_bounds__(kWave) void k_sub_scale(const SubGroup *__restrict__ groups,
                                                     const SubTile *__restrict__ tiles,
                                                     const double *__restrict__ sig,
                                                     SParticle *__restrict__ w,
                                                     float *__restrict__ pbox,
                                                     int32_t *__restrict__ gerr) {
    const SubTile tl = tiles[blockIdx.x];
    const SubGroup g = groups[tl.group];
    const double sx = sig[2 * tl.group], sv = sig[2 * tl.group + 1];
    const bool scale = sx != 0.0 && sv != 0.0;
    float lo[6], hi[6];
    for (int j = 0; j < 6; j++) {
        lo[j] = __builtin_inff();
        hi[j] = -__builtin_inff();
    }
    bool bad = false;
    const int32_t end = min(tl.off + kSubTile, g.n);
    for (int32_t i = tl.off + lane_id(); i < end; i += kWave) {
        SParticle p = w[g.off + i];
        if (scale) {
            for (int j = 0; j < 6; j++)
                p.pos[j] = (float)((double)p.pos[j] / (j < 3 ? sx : sv));
            w[g.off + i] = p;
        }
        for (int j = 0; j < 6; j++) {
            bad |= !__builtin_isfinite(p.pos[j]);
            lo[j] = p.pos[j] < lo[j] ? p.pos[j] : lo[j];
            hi[j] = p.pos[j] > hi[j] ? p.pos[j] : hi[j];
        }
    }
    for (int j = 0; j < 6; j++) {
        lo[j] = wave_min(lo[j]);
        hi[j] = wave_max(hi[j]);
    }
    if (ballot(bad) && lane_id() == 0) gerr[tl.group] = 1;
    if (lane_id() == 0)
        for (int j = 0; j < 6; j++) {
            pbox[12 * blockIdx.x + j] = lo[j];
            pbox[12 * blockIdx.x + 6 + j] = hi[j];
        }
}

// root node of each group (node index = group index); roots that split go
// on the first active list
// the node goes on the list that suits its size: split level by level
// with 1024 or kSfThreads threads, or finished whole by k_sub_fused (the
// fused list follows the two level lists in `big`'s buffer, at fused_at)
constexpr int kFuseMaxNode = 1024;  // kFuseMax, declared with its kernel below
__device__ __forceinline__ void push_active(int32_t id, int32_t n, int32_t *__restrict__ small,
                                            int32_t *__restrict__ big,
                                            int32_t *__restrict__ counters,
                                            int32_t *__restrict__ fused) {
    if (n <= kFuseMaxNode) fused[atomicAdd(&counters[kSfFused], 1)] = id;
    else if (n > kSfBigNode) big[atomicAdd(&counters[kSfActiveBig], 1)] = id;
    else small[atomicAdd(&counters[kSfActive], 1)] = id;
}

__global__ void k_sub_roots(int32_t ng, const SubGroup *__restrict__ groups,
                            const float *__restrict__ pbox, Node6 *__restrict__ nodes,
                            int32_t *__restrict__ active, int32_t *__restrict__ active_big,
                            int32_t *__restrict__ counters, int32_t *__restrict__ fused) {
    const int32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const SubGroup g = groups[gi];
    Node6 nd;
    for (int j = 0; j < 6; j++) {
        nd.lo[j] = __builtin_inff();
        nd.hi[j] = -__builtin_inff();
    }
    for (int32_t t = g.tile0; t < g.tile0 + g.ntiles; t++)
        for (int j = 0; j < 6; j++) {
            const float l = pbox[12 * t + j], h = pbox[12 * t + 6 + j];
            nd.lo[j] = l < nd.lo[j] ? l : nd.lo[j];
            nd.hi[j] = h > nd.hi[j] ? h : nd.hi[j];
        }
    if (!g.n)
        for (int j = 0; j < 6; j++) nd.lo[j] = nd.hi[j] = 0.0f;
    nd.start = g.off;
    nd.n = g.n;
    nd.div = -1;
    nd.left = nd.right = nd.parent = nd.skip = -1;
    nd.where = 0;
    nodes[gi] = nd;
    if (g.n > kLeaf6) push_active(gi, g.n, active, active_big, counters, fused);
}

// ------------------------------------------------------------- tree build

// exclusive rank of b among the workgroup's threads, and the count
template <int NT = kSfThreads>
__device__ __forceinline__ int block_rank(bool b, int *total, int *s_w) {
    const uint64_t m = ballot(b);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) s_w[w] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    for (int j = 0; j < NT / kWave; j++) {
        off += j < w ? s_w[j] : 0;
        tot += s_w[j];
    }
    __syncthreads();
    *total = tot;
    return off + rank_below(m);
}

// K and S of the closed form (file header) for one partition; b and r are
// the positions of the bigs in order and of the smalls in order
__device__ void partition_ks(const int32_t *__restrict__ bpos, const int32_t *__restrict__ spos,
                             int32_t n, int32_t nbig, int32_t nsmall, int32_t *K, int32_t *S) {
    // b_k = bpos[k - 1]; r_k = spos[nsmall - k]; b_k < r_k holds for k <= K
    int32_t lo = 0, hi = min(nbig, nsmall);
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) / 2;
        if (bpos[mid - 1] < spos[nsmall - mid]) lo = mid;
        else hi = mid - 1;
    }
    const int32_t rK = lo == 0 ? n : spos[nsmall - lo];
    *K = lo;
    *S = (lo < nbig && bpos[lo] < rK) ? bpos[lo] + 1 : rK;
}

// destination of the element at p (class big, rank among its class) in the
// closed form
__device__ __forceinline__ int32_t partition_dest(int32_t p, bool big, int32_t rank, int32_t n,
                                                  int32_t nsmall, int32_t S,
                                                  const int32_t *__restrict__ bpos,
                                                  const int32_t *__restrict__ spos) {
    // the clamps never bind for a partition the form describes; they keep
    // every read inside the node's lists whatever the input
    if (p < S) {
        if (!big) return p;
        return (rank == 0 ? n : spos[max(nsmall - rank, 0)]) - 1;
    }
    if (!big) return bpos[min(max(nsmall - 1 - rank, 0), n - 1)];
    return max(p - 1, 0);
}

// one level: every listed node splits (src/fast3tree.c:736-846), one
// workgroup per node.  A node's points are read from the buffer the last
// level wrote them to and its halves go to the other one.
template <int NT>
__global__ __launch_bounds__(NT) void k_sub_split(
    Node6 *__restrict__ nodes, const int32_t *__restrict__ active, SParticle *__restrict__ wa,
    SParticle *__restrict__ wb, int32_t *__restrict__ rk, int32_t *__restrict__ bpos,
    int32_t *__restrict__ spos, int32_t *__restrict__ next, int32_t *__restrict__ next_big,
    int32_t *__restrict__ counters, int32_t node_cap, int32_t *__restrict__ fused) {
    constexpr int NW = NT / kWave;
    __shared__ int s_w[NW];
    __shared__ int32_t s_k[2];
    __shared__ float s_box[NW][24];
    const int32_t me = active[blockIdx.x];
    const Node6 nd = nodes[me];
    const SParticle *__restrict__ w = nd.where ? wb : wa;
    SParticle *__restrict__ w2 = nd.where ? wa : wb;
    // widest dimension, the last one first (src/fast3tree.c:736-747)
    int dim = 5;
    float wd = nd.hi[5] - nd.lo[5];
    for (int i = 0; i < 5; i++) {
        const float e = nd.hi[i] - nd.lo[i];
        if (e > wd) {
            wd = e;
            dim = i;
        }
    }
    const float hd = acoord(nd.hi, dim), ld = acoord(nd.lo, dim);
    if (hd == ld) return;  // no split, no reordering: stays a leaf where it is
    const float lim = 0.5 * (hd + ld);
    const int32_t s0 = nd.start, n = nd.n;
    // pass 1: classes, ranks within the class, positions of each class
    int32_t nbig = 0;
    for (int32_t base = 0; base < n; base += NT) {
        const int32_t p = base + threadIdx.x;
        const bool in = p < n;
        const bool big = in && pcoord(w[s0 + p], dim) > lim;
        int tot;
        const int rb = block_rank<NT>(big, &tot, s_w);
        if (in) {
            const int32_t rank = big ? nbig + rb : p - (nbig + rb);
            rk[s0 + p] = big ? ~rank : rank;
            if (big) bpos[s0 + rank] = p;
            else spos[s0 + rank] = p;
        }
        nbig += tot;
    }
    const int32_t nsmall = n - nbig;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t K, S;
        partition_ks(bpos + s0, spos + s0, n, nbig, nsmall, &K, &S);
        s_k[0] = S;
    }
    __syncthreads();
    const int32_t S = s_k[0];
    // pass 2: scatter to the other buffer, boxes of the two halves
    float lo[2][6], hi[2][6];
    for (int c = 0; c < 2; c++)
        for (int j = 0; j < 6; j++) {
            lo[c][j] = __builtin_inff();
            hi[c][j] = -__builtin_inff();
        }
    for (int32_t p = threadIdx.x; p < n; p += NT) {
        const int32_t r = rk[s0 + p];
        const bool big = r < 0;
        const int32_t d = partition_dest(p, big, big ? ~r : r, n, nsmall, S, bpos + s0, spos + s0);
        const SParticle q = w[s0 + p];
        w2[s0 + d] = q;
        if (d < nsmall) {
            for (int j = 0; j < 6; j++) {
                lo[0][j] = q.pos[j] < lo[0][j] ? q.pos[j] : lo[0][j];
                hi[0][j] = q.pos[j] > hi[0][j] ? q.pos[j] : hi[0][j];
            }
        } else {
            for (int j = 0; j < 6; j++) {
                lo[1][j] = q.pos[j] < lo[1][j] ? q.pos[j] : lo[1][j];
                hi[1][j] = q.pos[j] > hi[1][j] ? q.pos[j] : hi[1][j];
            }
        }
    }
    for (int c = 0; c < 2; c++)
        for (int j = 0; j < 6; j++) {
            lo[c][j] = wave_min(lo[c][j]);
            hi[c][j] = wave_max(hi[c][j]);
        }
    const int wv = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int c = 0; c < 2; c++)
            for (int j = 0; j < 6; j++) {
                s_box[wv][12 * c + j] = lo[c][j];
                s_box[wv][12 * c + 6 + j] = hi[c][j];
            }
    __syncthreads();
    if (threadIdx.x != 0) return;
    Node6 me_n = nd;
    me_n.where = 1 - nd.where;  // its points, reordered, are in the other buffer now
    if (nsmall == n || nsmall == 0) {  // src/fast3tree.c:804-808 (after the swaps)
        me_n.div = -1;
        nodes[me] = me_n;
        return;
    }
    const int32_t l = atomicAdd(&counters[kSfNodes], 2);
    if (l + 2 > node_cap) {
        atomicOr(&counters[kSfErr], 1);
        me_n.div = -1;
        nodes[me] = me_n;
        return;
    }
    Node6 L, R;
    for (int j = 0; j < 6; j++) {
        float a = __builtin_inff(), b = -__builtin_inff(), c = __builtin_inff(), e = -__builtin_inff();
        for (int x = 0; x < NW; x++) {
            a = s_box[x][j] < a ? s_box[x][j] : a;
            b = s_box[x][6 + j] > b ? s_box[x][6 + j] : b;
            c = s_box[x][12 + j] < c ? s_box[x][12 + j] : c;
            e = s_box[x][18 + j] > e ? s_box[x][18 + j] : e;
        }
        L.lo[j] = a;
        L.hi[j] = b;
        R.lo[j] = c;
        R.hi[j] = e;
    }
    L.start = s0;
    L.n = nsmall;
    R.start = s0 + nsmall;
    R.n = n - nsmall;
    L.div = R.div = -1;
    L.left = L.right = R.left = R.right = -1;
    L.parent = R.parent = me;
    L.skip = l + 1;
    R.skip = me_n.skip;
    L.where = R.where = me_n.where;
    nodes[l] = L;
    nodes[l + 1] = R;
    me_n.div = dim;
    me_n.left = l;
    me_n.right = l + 1;
    nodes[me] = me_n;
    if (L.n > kLeaf6) push_active(l, L.n, next, next_big, counters, fused);
    if (R.n > kLeaf6) push_active(l + 1, R.n, next, next_big, counters, fused);
}

// A node of at most kFuseMax points is split down to its leaves by one
// workgroup, in LDS, all nodes of a level at once: the coordinates stay in
// place and a permutation of local positions moves; one block-wide scan
// ranks every node's points, one thread per node finds its K and S, the
// children's boxes come from atomic min/max on order-preserving keys.  The
// same splits as k_sub_split, node for node; the particles are written once,
// to the other buffer, at the end.
constexpr int kFuseMax = kFuseMaxNode;
constexpr int kFuseList = 2 * (kFuseMax / (kLeaf6 + 1)) + 2;  // nodes of one level
constexpr int kFusePer = kFuseMax / kSfThreads;                // positions per thread
constexpr uint8_t kFuseDone = 0xff;
static_assert(kFuseList < kFuseDone, "level entries fit a byte");

struct FuseNode {
    float lo[6], hi[6];
    int32_t start, n;  // local positions
    int32_t gid;       // global node id
    int32_t dim;       // split dimension, -1: no split (a leaf)
    float lim;
    int32_t nsmall, S;
    int32_t child;     // first of the children's entries in the next list, -1: none
};

__global__ __launch_bounds__(kSfThreads) void k_sub_fused(Node6 *__restrict__ nodes,
                                                          const int32_t *__restrict__ roots,
                                                          SParticle *__restrict__ wa,
                                                          SParticle *__restrict__ wb,
                                                          int32_t *__restrict__ counters,
                                                          int32_t node_cap) {
    __shared__ float X[6][kFuseMax];
    __shared__ int16_t I[2][kFuseMax];
    __shared__ uint8_t NO[2][kFuseMax];
    __shared__ int32_t SC[kFuseMax + 1];
    __shared__ int32_t BP[kFuseMax], SP[kFuseMax];
    __shared__ FuseNode E[2][kFuseList];
    __shared__ int32_t BK[kFuseList][2][12];
    __shared__ int s_w[kSfThreads / kWave];
    __shared__ int32_t s_cnt[2], s_g0;
    const int tid = threadIdx.x;
    const int32_t root = roots[blockIdx.x];
    const Node6 rn = nodes[root];
    const int32_t s0 = rn.start, N = rn.n;
    const SParticle *__restrict__ src = rn.where ? wb : wa;
    SParticle *__restrict__ dst = rn.where ? wa : wb;
    const int32_t wnew = 1 - rn.where;  // the buffer the subtree's points end in
    for (int32_t p = tid; p < N; p += kSfThreads) {
        const SParticle q = src[s0 + p];
        for (int k = 0; k < 6; k++) X[k][p] = q.pos[k];
        I[0][p] = (int16_t)p;
        NO[0][p] = 0;
    }
    if (tid == 0) {
        FuseNode &e = E[0][0];
        for (int k = 0; k < 6; k++) {
            e.lo[k] = rn.lo[k];
            e.hi[k] = rn.hi[k];
        }
        e.start = 0;
        e.n = N;
        e.gid = root;
        s_cnt[0] = 1;
    }
    __syncthreads();
    int cur = 0;
    for (int level = 0; level < kFuseMax; level++) {
        const int32_t ne = s_cnt[cur];
        if (ne == 0) break;
        const int nxt = cur ^ 1;
        // 1. each node: dimension and midpoint (src/fast3tree.c:736-760)
        if (tid < ne) {
            FuseNode &e = E[cur][tid];
            int dim = 5;
            float wd = e.hi[5] - e.lo[5];
            for (int k = 0; k < 5; k++) {
                const float w2 = e.hi[k] - e.lo[k];
                if (w2 > wd) {
                    wd = w2;
                    dim = k;
                }
            }
            const float hd = acoord(e.hi, dim), ld = acoord(e.lo, dim);
            e.dim = hd == ld ? -1 : dim;  // a flat extent: a leaf, no reordering
            e.lim = 0.5 * (hd + ld);
            e.child = -1;
        }
        if (tid == 0) s_cnt[nxt] = 0;
        for (int x = tid; x < kFuseList * 24; x += kSfThreads)
            (&BK[0][0][0])[x] = (x % 12) < 6 ? 0x7fffffff : (int32_t)0x80000000;
        __syncthreads();
        // 2. classes and one scan of the bigs over every position
        int32_t cnt_local = 0;
        bool bigv[kFusePer];
        for (int r = 0; r < kFusePer; r++) {
            const int32_t p = tid * kFusePer + r;
            bool b = false;
            if (p < N) {
                const uint8_t en = NO[cur][p];
                if (en != kFuseDone) {
                    const FuseNode &e = E[cur][en];
                    if (e.dim >= 0) b = X[e.dim][I[cur][p]] > e.lim;
                }
            }
            bigv[r] = b;
            cnt_local += b ? 1 : 0;
        }
        // exclusive scan of the per-thread counts over the workgroup
        int32_t incl = wave_scan_add(cnt_local);
        if (lane_id() == kWave - 1) s_w[tid / kWave] = incl;
        __syncthreads();
        int32_t off = 0;
        for (int j = 0; j < tid / kWave; j++) off += s_w[j];
        int32_t run = off + incl - cnt_local;
        for (int r = 0; r < kFusePer; r++) {
            const int32_t p = tid * kFusePer + r;
            if (p < N) SC[p] = run;
            run += bigv[r] ? 1 : 0;
        }
        if (tid == kSfThreads - 1) SC[N] = run;  // the last thread ends with the total
        __syncthreads();
        // 3. ranks within the node and the two position lists
        for (int r = 0; r < kFusePer; r++) {
            const int32_t p = tid * kFusePer + r;
            if (p >= N) continue;
            const uint8_t en = NO[cur][p];
            if (en == kFuseDone) continue;
            const FuseNode &e = E[cur][en];
            if (e.dim < 0) continue;
            const int32_t rel = p - e.start;
            const int32_t rb = SC[p] - SC[e.start];
            if (bigv[r]) BP[e.start + rb] = rel;
            else SP[e.start + (rel - rb)] = rel;
        }
        __syncthreads();
        // 4. each node: K and S, its children's entries and global ids
        if (tid < ne) {
            FuseNode &e = E[cur][tid];
            if (e.dim >= 0) {
                const int32_t nbig = SC[e.start + e.n] - SC[e.start];
                const int32_t nsm = e.n - nbig;
                int32_t K, S;
                partition_ks(BP + e.start, SP + e.start, e.n, nbig, nsm, &K, &S);
                e.nsmall = nsm;
                e.S = S;
                if (nsm > 0 && nsm < e.n) {  // global ids after the level is counted
                    const int32_t c = atomicAdd(&s_cnt[nxt], 2);
                    e.child = c;
                    FuseNode &L = E[nxt][c];
                    FuseNode &R = E[nxt][c + 1];
                    L.start = e.start;
                    L.n = nsm;
                    R.start = e.start + nsm;
                    R.n = e.n - nsm;
                }
            }
        }
        __syncthreads();
        // 5. every position of a splitting node to its place, the children's
        //    boxes; the level's children take their global ids with one atomic
        //    (child slot c is node s_g0 + c)
        if (tid == 0) {
            const int32_t nc = s_cnt[nxt];
            int32_t g0 = nc ? atomicAdd(&counters[kSfNodes], nc) : 0;
            if (g0 + nc > node_cap) {  // no room: the level's nodes stay leaves
                atomicOr(&counters[kSfErr], 1);
                g0 = -1;
            }
            s_g0 = g0;
        }
        for (int r = 0; r < kFusePer; r++) {
            const int32_t p = tid * kFusePer + r;
            if (p >= N) continue;
            const uint8_t en = NO[cur][p];
            if (en == kFuseDone) {
                I[nxt][p] = I[cur][p];
                NO[nxt][p] = kFuseDone;
                continue;
            }
            const FuseNode &e = E[cur][en];
            if (e.dim < 0) {  // no split: a leaf where it is
                I[nxt][p] = I[cur][p];
                NO[nxt][p] = kFuseDone;
                continue;
            }
            const int32_t rel = p - e.start;
            const int32_t rb = SC[p] - SC[e.start];
            const int32_t rank = bigv[r] ? rb : rel - rb;
            const int32_t d = partition_dest(rel, bigv[r], rank, e.n, e.nsmall, e.S, BP + e.start,
                                             SP + e.start);
            const int32_t q = e.start + d;
            const int16_t id = I[cur][p];
            I[nxt][q] = id;
            if (e.child < 0) {  // all on one side: reordered, then a leaf
                NO[nxt][q] = kFuseDone;
                continue;
            }
            const int side = d < e.nsmall ? 0 : 1;
            const int32_t ce = e.child + side;
            const int32_t cn = side ? e.n - e.nsmall : e.nsmall;
            NO[nxt][q] = cn > kLeaf6 ? (uint8_t)ce : kFuseDone;
            for (int k = 0; k < 6; k++) {
                const int32_t key = float_key(X[k][id]);
                atomicMin(&BK[en][side][k], key);
                atomicMax(&BK[en][side][6 + k], key);
            }
        }
        __syncthreads();
        // 6. the node records (children as leaves or as next level's nodes)
        if (tid < ne) {
            const FuseNode &e = E[cur][tid];
            Node6 me_n = nodes[e.gid];
            me_n.where = wnew;
            if (e.child >= 0 && s_g0 >= 0) {
                const int32_t l = s_g0 + e.child;
                E[nxt][e.child].gid = l;
                E[nxt][e.child + 1].gid = l + 1;
                Node6 C[2];
                for (int c = 0; c < 2; c++) {
                    FuseNode &ce = E[nxt][e.child + c];
                    for (int k = 0; k < 6; k++) {
                        ce.lo[k] = key_float(BK[tid][c][k]);
                        ce.hi[k] = key_float(BK[tid][c][6 + k]);
                        C[c].lo[k] = ce.lo[k];
                        C[c].hi[k] = ce.hi[k];
                    }
                    C[c].start = s0 + ce.start;
                    C[c].n = ce.n;
                    C[c].div = -1;
                    C[c].left = C[c].right = -1;
                    C[c].parent = e.gid;
                    C[c].where = wnew;
                }
                C[0].skip = l + 1;
                C[1].skip = me_n.skip;
                nodes[l] = C[0];
                nodes[l + 1] = C[1];
                me_n.div = e.dim;
                me_n.left = l;
                me_n.right = l + 1;
            } else {
                me_n.div = -1;
            }
            nodes[e.gid] = me_n;
        }
        __syncthreads();
        // the next level holds only the children of more than kLeaf6 points:
        // compact the entries, renumbering the positions' entries
        if (tid == 0) {
            int32_t m = 0;
            const int32_t nn2 = s_cnt[nxt];
            for (int32_t c = 0; c < nn2; c++) {
                if (E[nxt][c].n > kLeaf6 && s_g0 >= 0) {
                    if (m != c) E[nxt][m] = E[nxt][c];
                    BP[c] = m;  // old entry -> new entry (BP is free here)
                    m++;
                } else {
                    BP[c] = -1;
                }
            }
            s_cnt[nxt] = m;
        }
        __syncthreads();
        for (int32_t p = tid; p < N; p += kSfThreads) {
            const uint8_t en = NO[nxt][p];
            if (en != kFuseDone) NO[nxt][p] = (uint8_t)BP[en];
        }
        __syncthreads();
        cur = nxt;
    }
    // the particles in their final order, to the other buffer
    for (int32_t k = tid; k < N; k += kSfThreads) dst[s0 + k] = src[s0 + I[cur][k]];
}

// after the build: the leaves whose points the last write left in the
// second buffer copy them to the first, which the searches read
__global__ __launch_bounds__(kWave) void k_sub_collect(const Node6 *__restrict__ nodes,
                                                       SParticle *__restrict__ wa,
                                                       const SParticle *__restrict__ wb) {
    const Node6 nd = nodes[blockIdx.x];
    if (nd.div >= 0 || !nd.where) return;
    for (int32_t p = lane_id(); p < nd.n; p += kWave) wa[nd.start + p] = wb[nd.start + p];
}

// ---------------------------------------------------------- tree searches

// src/fast3tree.c:294-309
__device__ __forceinline__ bool box_apart(const Node6 &nd, const float (&c)[6], float r) {
    float d = 0.0f, e;
    const float r2 = r * r;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if ((e = c[i] - nd.lo[i]) < 0) {
            d += e * e;
            if (d >= r2) return true;
        } else if ((e = c[i] - nd.hi[i]) > 0) {
            d += e * e;
            if (d >= r2) return true;
        }
    }
    return false;
}

__device__ __forceinline__ float dist6(const float (&c)[6], const SParticle &p) {
    float d = 0.0f;
    for (int j = 0; j < 6; j++) {
        const float dx = c[j] - p.pos[j];
        d += dx * dx;
    }
    return d;
}

// the child the reference visits first (src/fast3tree.c:980-984)
__device__ __forceinline__ int32_t near_child(const Node6 &nd, const float (&c)[6]) {
    const int d = nd.div;
    const float mid = acoord(nd.lo, d) + acoord(nd.hi, d);
    return (double)acoord(c, d) > 0.5 * (double)mid ? nd.right : nd.left;
}

// The reference's own-leaf step (src/fast3tree.c:999-1025): the leaf the
// query descends to, climbed while its parent's split dimension is flat
// there (that subtree, `skip`, is left out of the search), and the start
// value: the smallest nonzero squared distance in it, or its squared
// diagonal.
__device__ __forceinline__ float nc_start(const float (&c)[6], int32_t root,
                                          const SParticle *__restrict__ w,
                                          const Node6 *__restrict__ nodes, int32_t *skip) {
    int32_t nd = root;
    while (nodes[nd].div >= 0) {
        const Node6 &x = nodes[nd];
        nd = acoord(c, x.div) <= acoord(nodes[x.left].hi, x.div) ? x.left : x.right;
    }
    while (nd != root) {
        const int pd = nodes[nodes[nd].parent].div;
        if (acoord(nodes[nd].lo, pd) != acoord(nodes[nd].hi, pd)) break;
        nd = nodes[nd].parent;
    }
    *skip = nd;
    float md = 0.0f;
    const Node6 x = nodes[nd];
    for (int k = 0; k < 6; k++) {
        const float e = x.hi[k] - x.lo[k];
        md += e * e;
    }
    for (int32_t p = 0; p < x.n; p++) {
        const float d = dist6(c, w[x.start + p]);
        if (d != 0.0f && d < md) md = d;
    }
    return md;
}

// The search of src/fast3tree.c:956-987 in the reference's visiting order
// (near child first, every entry tests the box against the current r, a
// leaf narrows r), as a walk over parent links.
__device__ float nc_ordered(const float (&c)[6], int32_t root, int32_t skip, float r,
                            const SParticle *__restrict__ w, const Node6 *__restrict__ nodes) {
    int32_t cur = root;
    bool down = true;
    for (;;) {
        if (down) {
            const Node6 x = nodes[cur];
            if (cur == skip || box_apart(x, c, r)) {
                down = false;
            } else if (x.div < 0) {
                float r2 = r * r;
                for (int32_t p = 0; p < x.n; p++) {
                    const float d = dist6(c, w[x.start + p]);
                    if (d < r2) r2 = d;
                }
                r = (float)sqrt((double)r2);
                down = false;
            } else {
                cur = near_child(x, c);
                continue;
            }
        }
        if (cur == root) break;
        const int32_t par = nodes[cur].parent;
        const Node6 px = nodes[par];
        const int32_t nc = near_child(px, c);
        if (cur == nc) {
            cur = nc == px.left ? px.right : px.left;
            down = true;
        } else {
            cur = par;
        }
    }
    return r;
}

// a node whose box is at least thr (squared) from c: every point in it is
__device__ __forceinline__ bool box_far(const Node6 &nd, const float (&c)[6], float thr) {
    float d = 0.0f, e;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        if ((e = c[i] - nd.lo[i]) < 0) d += e * e;
        else if ((e = c[i] - nd.hi[i]) > 0) d += e * e;
    }
    return d >= thr;
}

// fast3tree_find_next_closest_distance (src/fast3tree.c:992-1030) for 64
// sampled particles per wavefront, neighbours in the tree order.
//
// The reference's answer depends on its visiting order only through
// rounding: a leaf that improves nothing returns sqrt(r * r) = r (exact for
// IEEE floats), so with T = fl(r0 * r0) from the start value and D the
// smallest squared distance outside the skipped subtree, the answer is r0
// when D >= T and sqrt(D) otherwise -- unless a value D' with
// D < D' <= D (1 + 2^-21) was taken first, whose round trip
// fl(fl(sqrt(D'))^2) can fall to D or below and reject it.  The wavefront
// walks the tree once for all its lanes (a node is entered when any lane
// may find a value below its bound there), each lane keeps the two smallest
// distinct values, and a lane whose two are that close repeats the search
// in the reference's own order.
__global__ __launch_bounds__(kWave) void k_sub_next_closest(
    const SubGroup *__restrict__ groups, const SubTile *__restrict__ qtiles,
    const int32_t *__restrict__ qpos, const int32_t *__restrict__ qidx,
    const SParticle *__restrict__ w, const Node6 *__restrict__ nodes, float *__restrict__ dist,
    int ordered) {
    const SubTile tl = qtiles[blockIdx.x];
    const SubGroup g = groups[tl.group];
    const int32_t q = tl.off + lane_id();
    const bool valid = q < g.nsample;
    const int32_t root = tl.group;
    float c[6] = {0, 0, 0, 0, 0, 0};
    int32_t skip = -1, s_lo = 0, s_hi = 0;
    float r0 = 0.0f, T = 0.0f;
    if (valid) {
        const SParticle me = w[g.off + (g.drawn ? qpos[g.draw_off + q] : q)];
        for (int k = 0; k < 6; k++) c[k] = me.pos[k];
        r0 = (float)sqrt((double)nc_start(c, root, w, nodes, &skip));
        T = r0 * r0;
        s_lo = nodes[skip].start;
        s_hi = s_lo + nodes[skip].n;
    }
    float m1 = __builtin_inff(), m2 = __builtin_inff();
    float thr = valid ? T : 0.0f;  // nothing at or above T is ever taken
    int32_t cur = ordered ? -1 : root;
    while (cur >= 0) {
        cur = uniform(cur);
        const Node6 x = load_const(nodes + cur);
        const bool inside = x.start >= s_lo && x.start + x.n <= s_hi;  // the skipped subtree
        const bool need = valid && !inside && !box_far(x, c, thr);
        if (!ballot(need)) {
            cur = x.skip;
            continue;
        }
        if (x.div >= 0) {
            cur = x.left;
            continue;
        }
        // the leaf's points at wavefront-uniform addresses: scalar loads
        const int32_t x0 = uniform(x.start), xn = uniform(x.n);
        for (int32_t t = 0; t < xn; t++) {
            const SParticle pt = load_const(w + x0 + t);
            float d = 0.0f;
            for (int k = 0; k < 6; k++) {
                const float dx = c[k] - pt.pos[k];
                d += dx * dx;
            }
            if (need) {
                if (d < m1) {
                    m2 = m1;
                    m1 = d;
                } else if (d > m1 && d < m2) {
                    m2 = d;
                }
            }
        }
        if (need && m1 < thr) thr = fminf(thr, m1 * 1.000002f);
        cur = x.skip;
    }
    if (!valid) return;
    float r;
    if (ordered) {
        r = nc_ordered(c, root, skip, r0, w, nodes);
    } else if (!(m1 < T)) {
        r = r0;
    } else if (m2 > m1 * 1.000002f) {
        r = (float)sqrt((double)m1);
    } else {
        r = nc_ordered(c, root, skip, r0, w, nodes);
    }
    dist[g.sample_off + (g.drawn ? qidx[g.draw_off + q] : q)] = r;
}

// ------------------------------------------------------ linking length

// find_median_r (src/groupies.c:744-786) per group, one workgroup: the
// reference's quickselect step by step, each rad_partition evaluated in
// parallel with the closed form; buf holds the distances (a copy), buf2,
// rk, bpos and spos are scratch of the same extent
__global__ __launch_bounds__(kSfThreads) void k_sub_median(
    const SubGroup *__restrict__ groups, float frac, float *__restrict__ buf,
    float *__restrict__ buf2, int32_t *__restrict__ rk, int32_t *__restrict__ bpos,
    int32_t *__restrict__ spos, float *__restrict__ target) {
    __shared__ int s_w[kSfThreads / kWave];
    __shared__ int32_t s_v[4];
    __shared__ float s_piv;
    const SubGroup g = groups[blockIdx.x];
    const int32_t ns = g.nsample;
    if (ns < 1) {
        if (threadIdx.x == 0) target[blockIdx.x] = 0.0f;
        return;
    }
    float *v = buf + g.sample_off;
    float *v2 = buf2 + g.sample_off;
    int32_t *rko = rk + g.sample_off, *bp = bpos + g.sample_off, *sp = spos + g.sample_off;
    if (ns < 2) {
        if (threadIdx.x == 0) target[blockIdx.x] = v[0];
        return;
    }
    const int32_t k = g.k;
    uint32_t seed = g.seed;  // thread 0's copy is the live one
    int32_t left = 0, right = ns;
    for (;;) {
        if (threadIdx.x == 0) {
            const float u = (float)(rand_r_glibc(&seed) % 2147483647) / (float)2147483647;
            const float at_f = (float)left + u * (float)(right - left);
            // u rounds to 1.0f for the top 63 of 2^31 draws; the reference
            // then reads one past the range, here the last slot is taken
            const int32_t at = min((int32_t)at_f, right - 1);
            const float pv = v[at];
            v[at] = v[right - 1];
            v[right - 1] = pv;
            s_piv = pv;
        }
        __threadfence_block();
        __syncthreads();
        const float pv = s_piv;
        const int32_t m = right - 1 - left;  // the loop's range [left, right - 1)
        int32_t nbig = 0;
        for (int32_t base = 0; base < m; base += kSfThreads) {
            const int32_t p = base + threadIdx.x;
            const bool in = p < m;
            const bool big = in && v[left + p] > pv;
            int tot;
            const int rb = block_rank(big, &tot, s_w);
            if (in) {
                const int32_t rank = big ? nbig + rb : p - (nbig + rb);
                rko[left + p] = big ? ~rank : rank;
                if (big) bp[left + rank] = p;
                else sp[left + rank] = p;
            }
            nbig += tot;
        }
        const int32_t nsmall = m - nbig;
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t K, S;
            partition_ks(bp + left, sp + left, m, nbig, nsmall, &K, &S);
            s_v[0] = S;
        }
        __syncthreads();
        const int32_t S = s_v[0];
        for (int32_t p = threadIdx.x; p < m; p += kSfThreads) {
            const int32_t r = rko[left + p];
            const bool big = r < 0;
            v2[left + partition_dest(p, big, big ? ~r : r, m, nsmall, S, bp + left, sp + left)] =
                v[left + p];
        }
        __threadfence_block();
        __syncthreads();
        for (int32_t p = threadIdx.x; p < m; p += kSfThreads) v[left + p] = v2[left + p];
        __threadfence_block();
        __syncthreads();
        if (threadIdx.x == 0) {
            // rad_partition's end: the pivot moves to the first slot above it
            const int32_t si = left + nsmall;
            const float t = v[right - 1];
            v[right - 1] = v[si];
            v[si] = t;
            int32_t done = 0;
            if (k == si || v[left] == v[right - 1]) done = 1;
            else if (k < si) right = si;
            else left = si + 1;
            s_v[1] = done;
            s_v[2] = left;
            s_v[3] = right;
        }
        __threadfence_block();
        __syncthreads();
        if (s_v[1]) break;
        left = s_v[2];
        right = s_v[3];
    }
    if (threadIdx.x == 0) target[blockIdx.x] = v[k];
}

// --------------------------------------------------------------- linking
//
// _find_subfofs_at_r (src/groupies.c:274-285) visits the particles in tree
// order.  Particle i's sphere search (fast3tree_find_sphere_skip,
// src/fast3tree.c:419-465) lists i itself, then every later particle closer
// than the length, in array order, and link_particle_to_fof
// (src/fof.c:113-137) gives the listed particles a small FOF or merges
// theirs.  The output is sorted by the final union-find roots
// (build_fullfofs, src/fof.c:223-252), so the roots' numbers fix it.  They
// are found here without replaying the loop particle by particle:
//   * particle x gets its small FOF at the first step that lists it: kmin(x),
//     its earliest earlier neighbour, or its own step when it has none;
//   * step i works with i's small FOF when i has an earlier neighbour, else
//     with that of its first listed particle j numbered before (kmin(j) < i),
//     else it creates the next number.  These pointers lead to earlier
//     steps and end at the creating one: the numbers are a pointer chase and
//     a count of the creating steps;
//   * a merge (merge_smallfofs, src/fof.c:72-89) hangs the root of step i's
//     small FOF under the root of the listed particle's.  Only pairs (i, x)
//     with x numbered before step i under another number can merge.  One
//     workgroup per group replays them in step order, a chunk of pairs at a
//     time, keeping only the pairs whose roots still differ when the chunk
//     starts (a joined pair stays joined).

constexpr int32_t kNoEarlier = 0x7fffffff;

struct EdgeRec {
    int32_t i, j, rank;  // step i lists j as its rank-th later particle
};
constexpr int kLinkStage = 256;  // pairs a wavefront stages before one append

__global__ void k_sub_init(int32_t m, int32_t *__restrict__ kmin, int32_t *__restrict__ deg,
                           int32_t *__restrict__ csize) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;  // m + 1 entries: the scans' totals
    kmin[i] = kNoEarlier;
    deg[i] = 0;
    csize[i] = 0;
}

// Each particle's sphere search over the later slots: the later particles
// closer than the group's length, appended to `edges` in the order the
// search lists them.  64 consecutive slots per wavefront walk the tree
// together: a node is entered when some lane's sphere reaches it and it
// holds slots after that lane's, and each leaf's points are loaded once and
// passed lane to lane.  Every pair is tested with the reference's float
// distance.  The append counter keeps counting past cap_edges (nothing is
// written there), so the host can size the list and run again.
__global__ __launch_bounds__(kWave) void k_sub_link(const SubGroup *__restrict__ groups,
                                                    const SubTile *__restrict__ ltiles,
                                                    const float *__restrict__ target,
                                                    const SParticle *__restrict__ w,
                                                    const Node6 *__restrict__ nodes,
                                                    EdgeRec *__restrict__ edges, int64_t cap_edges,
                                                    int32_t *__restrict__ deg,
                                                    int32_t *__restrict__ counters) {
    __shared__ float4 s_pts[kWave][2];
    __shared__ EdgeRec s_e[kLinkStage];
    const SubTile tl = ltiles[blockIdx.x];
    const SubGroup g = groups[tl.group];
    const int32_t i = g.off + tl.off + lane_id();
    const bool valid = tl.off + lane_id() < g.n;
    const float r = target[tl.group];
    const float r2 = r * r;
    float c[6] = {0, 0, 0, 0, 0, 0};
    if (valid) {
        const SParticle q = w[i];
        for (int k = 0; k < 6; k++) c[k] = q.pos[k];
    }
    unsigned long long *ecount = reinterpret_cast<unsigned long long *>(counters + kSfEdges);
    int32_t listed = 0;  // this lane's later particles so far
    int32_t staged = 0;  // pairs in s_e (wavefront-uniform)
    // the staged pairs to the list with one append
    auto flush = [&]() {
        unsigned long long at = 0;
        if (lane_id() == 0) at = atomicAdd(ecount, (unsigned long long)staged);
        const int64_t pos = (int64_t)read_lane(at, 0);
        __syncthreads();
        for (int32_t k = lane_id(); k < staged; k += kWave)
            if (pos + k < cap_edges) edges[pos + k] = s_e[k];
        __syncthreads();
        staged = 0;
    };
    int32_t cur = tl.group;
    while (cur >= 0) {
        cur = uniform(cur);
        const Node6 x = load_const(nodes + cur);
        const bool need = valid && x.start + x.n > i + 1 && !box_apart(x, c, r);
        if (!ballot(need)) {
            cur = x.skip;
            continue;
        }
        if (x.div >= 0) {
            cur = x.left;
            continue;
        }
        // the leaf's points, one per lane, staged in LDS and read back by
        // every lane at the same address (a broadcast; a leaf whose split
        // failed can hold more than a wavefront of points)
        for (int32_t base = 0; base < x.n; base += kWave) {
            const int32_t cnt = min(x.n - base, kWave);
            if (lane_id() < cnt) {
                const SParticle q = w[x.start + base + lane_id()];
                s_pts[lane_id()][0] = make_float4(q.pos[0], q.pos[1], q.pos[2], q.pos[3]);
                s_pts[lane_id()][1] = make_float4(q.pos[4], q.pos[5], 0.0f, 0.0f);
            }
            __syncthreads();
            uint64_t hit = 0;
            for (int32_t t = 0; t < cnt; t++) {
                const float4 a = s_pts[t][0], b = s_pts[t][1];
                const float pp[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
                float d = 0.0f;
                for (int k = 0; k < 6; k++) {
                    const float dx = c[k] - pp[k];
                    d += dx * dx;
                }
                const int32_t j = x.start + base + t;
                if (need && j > i && d < r2) hit |= 1ull << t;
            }
            __syncthreads();
            // the chunk's pairs: lanes in order, each lane's in order, staged
            // in LDS (appended directly when they alone overfill it)
            const int32_t nh = __popcll(hit);
            const int32_t incl = wave_scan_add(nh);
            const int32_t tot = __builtin_amdgcn_readlane(incl, kWave - 1);
            if (tot) {
                if (staged + tot > kLinkStage) flush();
                int32_t rank = listed;
                if (tot > kLinkStage) {
                    unsigned long long at = 0;
                    if (lane_id() == 0) at = atomicAdd(ecount, (unsigned long long)tot);
                    int64_t pos = (int64_t)read_lane(at, 0) + (incl - nh);
                    while (hit) {
                        const int t = __builtin_ctzll(hit);
                        hit &= hit - 1;
                        if (pos < cap_edges) edges[pos] = EdgeRec{i, x.start + base + t, rank};
                        pos++;
                        rank++;
                    }
                } else {
                    int32_t at = staged + incl - nh;
                    while (hit) {
                        const int t = __builtin_ctzll(hit);
                        hit &= hit - 1;
                        s_e[at++] = EdgeRec{i, x.start + base + t, rank++};
                    }
                    staged += tot;
                }
                listed += nh;
            }
        }
        cur = x.skip;
    }
    if (staged) flush();
    if (valid) deg[i] = listed;
}

// the lists in place (off: exclusive scan of the counts) and each
// particle's earliest earlier neighbour
__global__ void k_sub_adjacent(int64_t ne, const EdgeRec *__restrict__ edges,
                               const int32_t *__restrict__ off, int32_t *__restrict__ adj,
                               int32_t *__restrict__ kmin) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const EdgeRec r = edges[e];
    adj[off[r.i] + r.rank] = r.j;
    atomicMin(kmin + r.j, r.i);
}

// the step whose small FOF each particle takes (-1: never listed; a creating
// step points at itself), and created[i] = 1 for a creating step
__global__ void k_sub_parent(int32_t m, const int32_t *__restrict__ off,
                             const int32_t *__restrict__ adj, const int32_t *__restrict__ kmin,
                             int32_t *__restrict__ par, int32_t *__restrict__ created) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    if (i == m) {
        created[m] = 0;
        return;
    }
    const int32_t k = kmin[i];
    int32_t p = -1;
    if (k != kNoEarlier) {
        p = k;
    } else if (off[i + 1] > off[i]) {
        p = i;
        for (int32_t e = off[i]; e < off[i + 1]; e++) {
            const int32_t kj = kmin[adj[e]];
            if (kj < i) {
                p = kj;
                break;
            }
        }
    }
    par[i] = p;
    created[i] = p == i;
}

// each particle's small FOF number: that of the creating step its pointers
// lead to (first_label: exclusive count of creating steps, over the batch).
// Path halving: every pointer only ever moves to an earlier step of the same
// chain, so racing updates keep each chain intact.
__global__ void k_sub_label(int32_t m, int32_t *__restrict__ par,
                            const int32_t *__restrict__ first_label, int32_t *__restrict__ lab) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    int32_t x = __hip_atomic_load(par + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (x < 0) {
        lab[i] = -1;
        return;
    }
    x = i;
    for (;;) {
        const int32_t y = __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (y == x) break;
        const int32_t z = __hip_atomic_load(par + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (z != y) __hip_atomic_store(par + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = z;
    }
    lab[i] = first_label[x];
}

constexpr int32_t kEventDedup = 64;

// The pairs of step i that can merge, (i's number, x's number) for each
// listed x numbered before step i under another number, in list order.  A
// number met earlier in the same step is dropped: the first merge with it
// joined the two.  Written at the step's own list positions (stage), counted
// in ecnt.
__global__ void k_sub_events(int32_t m, const int32_t *__restrict__ off,
                             const int32_t *__restrict__ adj, const int32_t *__restrict__ kmin,
                             const int32_t *__restrict__ lab, int2 *__restrict__ stage,
                             int32_t *__restrict__ ecnt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > m) return;
    if (i == m) {
        ecnt[m] = 0;
        return;
    }
    const int32_t e0 = off[i], e1 = off[i + 1];
    int32_t k = 0;
    if (e1 > e0) {
        const int32_t f = lab[i];
        for (int32_t e = e0; e < e1; e++) {
            const int32_t x = adj[e];
            if (kmin[x] >= i) continue;
            const int32_t lx = lab[x];
            if (lx == f) continue;
            // checked against the step's first kEventDedup pairs only: a
            // repeat kept past them is one more pair that cannot merge
            bool seen = false;
            for (int32_t q = 0; q < min(k, kEventDedup) && !seen; q++) seen = stage[e0 + q].y == lx;
            if (!seen) stage[e0 + k++] = make_int2(f, lx);
        }
    }
    ecnt[i] = k;
}

// the pairs of every step in one list, in step order (eoff: exclusive scan
// of ecnt)
__global__ void k_sub_compact(int32_t m, const int32_t *__restrict__ off,
                              const int2 *__restrict__ stage, const int32_t *__restrict__ eoff,
                              int2 *__restrict__ ev) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int32_t k = eoff[i + 1] - eoff[i];
    for (int32_t q = 0; q < k; q++) ev[eoff[i] + q] = stage[off[i] + q];
}

constexpr int kTokThreads = 1024;
constexpr int kTokLds = 12288;  // a group's numbers held in LDS up to this many
constexpr int kTokHash = 1024;  // one chunk's root pairs, deduplicated

// root of x with path halving, in the group's table u (LDS, or global when
// the group has more than kTokLds numbers; one workgroup only)
template <bool kLds>
__device__ __forceinline__ int32_t tok_find(int32_t *u, int32_t x) {
    if constexpr (kLds) {
        volatile int32_t *v = u;
        for (;;) {
            const int32_t y = v[x];
            if (y == x) return x;
            const int32_t z = v[y];
            if (z != y) v[x] = z;
            x = z;
        }
    } else {
        for (;;) {
            const int32_t y = __hip_atomic_load(u + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (y == x) return x;
            const int32_t z = __hip_atomic_load(u + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (z != y) __hip_atomic_store(u + x, z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            x = z;
        }
    }
}

template <bool kLds>
__device__ __forceinline__ void tok_set(int32_t *u, int32_t x, int32_t v) {
    if constexpr (kLds) *(volatile int32_t *)(u + x) = v;
    else __hip_atomic_store(u + x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct TokShared {
    int2 pair[kTokThreads];
    unsigned long long key[kTokHash];
    int32_t first[kTokHash];
    int w[kTokThreads / kWave];
};

// One group's merges, a chunk of kTokThreads pairs at a time.  Every thread
// finds its pair's two roots as the chunk starts; a pair whose roots are
// equal then, or equal as a set to an earlier pair's of the chunk, cannot
// merge.  One thread replays the others in order: the root of step i's small
// FOF under the root of the listed particle's.
template <bool kLds>
__device__ void tok_group(int32_t *u, int32_t nl, int32_t q0, int32_t q1, int32_t L0,
                          const int2 *__restrict__ ev, int32_t *__restrict__ root,
                          TokShared &S) {
    for (int32_t k = threadIdx.x; k < nl; k += kTokThreads) tok_set<kLds>(u, k, k);
    __syncthreads();
    for (int32_t base = q0; base < q1; base += kTokThreads) {
        const int32_t q = base + threadIdx.x;
        int32_t ra = 0, rb = 0;
        if (q < q1) {
            const int2 p = ev[q];
            ra = tok_find<kLds>(u, p.x - L0);
            rb = tok_find<kLds>(u, p.y - L0);
        }
        for (int k = threadIdx.x; k < kTokHash; k += kTokThreads) {
            S.key[k] = 0ull;
            S.first[k] = kTokThreads;
        }
        __syncthreads();
        // the earliest pair of each root set claims its slot (open
        // addressing; at most kTokThreads of kTokHash slots are taken)
        int h = 0;
        if (ra != rb) {
            const uint32_t lo = (uint32_t)min(ra, rb), hi = (uint32_t)max(ra, rb);
            const unsigned long long key = ((unsigned long long)lo << 32) | hi;
            h = (int)((lo * 0x9E3779B1u) ^ (hi * 0x85EBCA77u)) & (kTokHash - 1);
            for (;;) {
                const unsigned long long was = atomicCAS(&S.key[h], 0ull, key);
                if (was == 0ull || was == key) break;
                h = (h + 1) & (kTokHash - 1);
            }
            atomicMin(&S.first[h], (int)threadIdx.x);
        }
        __syncthreads();
        const bool keep = ra != rb && S.first[h] == (int)threadIdx.x;
        int tot;
        const int at = block_rank<kTokThreads>(keep, &tot, S.w);
        if (keep) S.pair[at] = make_int2(ra, rb);
        __syncthreads();
        if (threadIdx.x == 0)
            for (int k = 0; k < tot; k++) {
                const int32_t a = tok_find<kLds>(u, S.pair[k].x), b = tok_find<kLds>(u, S.pair[k].y);
                if (a != b) tok_set<kLds>(u, a, b);
            }
        __syncthreads();
    }
    for (int32_t k = threadIdx.x; k < nl; k += kTokThreads) {
        const int32_t rt = tok_find<kLds>(u, k);
        __hip_atomic_store(root + L0 + k, rt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// The merges of each group in step order: root[L0 + k] = the root of number
// k (group-local numbers).  A group with more than kTokLds numbers keeps its
// table in root itself.
__global__ __launch_bounds__(kTokThreads) void k_sub_tokens(
    const SubGroup *__restrict__ groups, const int32_t *__restrict__ eoff,
    const int2 *__restrict__ ev, const int32_t *__restrict__ first_label,
    int32_t *__restrict__ root) {
    __shared__ int32_t s_u[kTokLds];
    __shared__ TokShared S;
    const SubGroup g = groups[blockIdx.x];
    if (g.n < 2) return;
    const int32_t L0 = first_label[g.off];
    const int32_t nl = first_label[g.off + g.n] - L0;
    if (nl <= 0) return;
    const int32_t q0 = eoff[g.off], q1 = eoff[g.off + g.n];
    if (nl <= kTokLds) tok_group<true>(s_u, nl, q0, q1, L0, ev, root, S);
    else tok_group<false>(root + L0, nl, q0, q1, L0, ev, root, S);
}

// each particle's sort value (its root's number in the group, -1 without a
// neighbour; collapse_smallfofs, src/fof.c:159-170), its slot, and the size
// of each root's set
__global__ void k_sub_values(int32_t m, const int32_t *__restrict__ slot_group,
                             const SubGroup *__restrict__ groups, const int32_t *__restrict__ lab,
                             const int32_t *__restrict__ first_label,
                             const int32_t *__restrict__ root, int32_t *__restrict__ val,
                             int32_t *__restrict__ idx, int32_t *__restrict__ csize) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = q < m;
    int32_t v = -1, key = 0;
    if (in) {
        const int32_t l = lab[q];
        if (l >= 0) {
            v = root[l];
            key = first_label[groups[slot_group[q]].off] + v;
        }
        val[q] = v;
        idx[q] = q;
    }
    // one atomic per distinct root in the wavefront
    const bool linked = v >= 0;
    uint64_t todo = ballot(linked);
    while (todo) {
        const int lead = __builtin_ctzll(todo);
        const int32_t kl = __builtin_amdgcn_readlane(key, lead);
        const uint64_t same = ballot(linked && key == kl);
        if (lane_id() == lead) atomicAdd(csize + kl, (int)__popcll(same));
        todo &= ~same;
    }
}

// --------------------------------------------------------- the output order
//
// partition_sort_particles (src/fof.c:184-221) on the sort values: each range
// splits about the middle of its values with the loop of the tree build
// (the closed form in the file header, `big` = value above the middle),
// level by level over every group.  A range of at most kSortSmall runs the
// loop itself in one wavefront.  Ranges move between two buffers of values
// and slots; every finished range ends in the first.

constexpr int kSortSmall = 128;
constexpr int kSortMid = 4096;  // ranges up to this size split with 256 threads
constexpr int kSortThreads1 = 1024;
constexpr int kSortThreads2 = 256;

struct SortRange {
    int32_t start, n, where;
    int32_t lo, hi;  // the range's least and greatest values (lo > hi: not known yet)
    int32_t pad;
};

// a range goes on the list that suits its size; next holds the large ones
// from 0 and the mid-sized ones from cap_next down
__device__ __forceinline__ void push_range(SortRange r, SortRange *__restrict__ next,
                                           SortRange *__restrict__ small,
                                           int32_t *__restrict__ counters, int32_t cap_next,
                                           int32_t cap_small) {
    if (r.n > kSortMid) {
        const int32_t k = atomicAdd(&counters[kSfRange], 1);
        if (k < cap_next) next[k] = r;
        else atomicOr(&counters[kSfErr], 4);
    } else if (r.n > kSortSmall) {
        const int32_t k = atomicAdd(&counters[kSfRangeMid], 1);
        if (k < cap_next) next[2 * cap_next - 1 - k] = r;
        else atomicOr(&counters[kSfErr], 4);
    } else {
        const int32_t k = atomicAdd(&counters[kSfRangeSmall], 1);
        if (k < cap_small) small[k] = r;
        else atomicOr(&counters[kSfErr], 4);
    }
}

__global__ void k_sort_roots(int32_t ng, const SubGroup *__restrict__ groups,
                             SortRange *__restrict__ next, SortRange *__restrict__ small,
                             int32_t *__restrict__ counters, int32_t cap_next, int32_t cap_small) {
    const int32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= ng) return;
    const SubGroup g = groups[gi];
    if (g.n >= 2)
        push_range(SortRange{g.off, g.n, 0, 1, 0, 0}, next, small, counters, cap_next, cap_small);
}

// exclusive prefix sum over the workgroup's threads, and the total
template <int NT>
__device__ __forceinline__ int block_excl_scan(int x, int *total, int *s_w) {
    constexpr int NW = NT / kWave;
    const int incl = wave_scan_add(x);
    const int wv = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) s_w[wv] = incl;
    __syncthreads();
    int off = 0, tot = 0;
    for (int j = 0; j < NW; j++) {
        off += j < wv ? s_w[j] : 0;
        tot += s_w[j];
    }
    __syncthreads();
    *total = tot;
    return off + incl - x;
}

template <int NT>
__device__ __forceinline__ void block_minmax(int32_t &lo, int32_t &hi, int32_t *s_w) {
    constexpr int NW = NT / kWave;
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int wv = threadIdx.x / kWave;
    if (lane_id() == 0) {
        s_w[wv] = lo;
        s_w[NW + wv] = hi;
    }
    __syncthreads();
    for (int k = 0; k < NW; k++) {
        lo = min(lo, s_w[k]);
        hi = max(hi, s_w[NW + k]);
    }
    __syncthreads();
}

// one level: every listed range splits, one workgroup per range: 1024
// threads (four values each per round) for the large ranges, 256 (one value
// each) for the mid-sized ones, whose list runs down from `ranges`
template <int NT>
__global__ __launch_bounds__(NT) void k_sort_split(
    const SortRange *__restrict__ ranges, int32_t *__restrict__ v0, int32_t *__restrict__ v1,
    int32_t *__restrict__ x0, int32_t *__restrict__ x1, int32_t *__restrict__ rk,
    int32_t *__restrict__ bpos, int32_t *__restrict__ spos, SortRange *__restrict__ next,
    SortRange *__restrict__ small, int32_t *__restrict__ counters, int32_t cap_next,
    int32_t cap_small) {
    static_assert(NT == 1024 || NT == 256, "launch shapes");
    __shared__ int s_w[4 * (NT / kWave)];
    __shared__ int32_t s_k[1];
    const SortRange sr = NT == 1024 ? ranges[blockIdx.x] : *(ranges - blockIdx.x);
    const int32_t s0 = sr.start, n = sr.n;
    const int32_t *__restrict__ vs = sr.where ? v1 : v0;
    const int32_t *__restrict__ xs = sr.where ? x1 : x0;
    int32_t *__restrict__ vd = sr.where ? v0 : v1;
    int32_t *__restrict__ xd = sr.where ? x0 : x1;
    int32_t lo = sr.lo, hi = sr.hi;
    if (lo > hi) {  // a group's whole range: its values' span first
        lo = 0x7fffffff;
        hi = -0x7fffffff - 1;
        for (int32_t p = threadIdx.x; p < n; p += NT) {
            const int32_t v = vs[s0 + p];
            lo = min(lo, v);
            hi = max(hi, v);
        }
        block_minmax<NT>(lo, hi, s_w);
    }
    if (lo == hi) {  // one value: in order already, to the first buffer
        if (sr.where)
            for (int32_t p = threadIdx.x; p < n; p += NT) {
                v0[s0 + p] = v1[s0 + p];
                x0[s0 + p] = x1[s0 + p];
            }
        return;
    }
    const int32_t pivot = (int32_t)((int64_t)lo + ((int64_t)hi - (int64_t)lo) / 2);
    // classes, ranks within the class, positions of each class
    int32_t nbig = 0;
    for (int32_t base = 0; base < n; base += 4 * NT) {
        const int32_t p0 = base + 4 * (int32_t)threadIdx.x;
        int nb = 0;
        bool big[4];
        for (int k = 0; k < 4; k++) {
            big[k] = p0 + k < n && vs[s0 + p0 + k] > pivot;
            nb += big[k];
        }
        int tot;
        int32_t run = nbig + block_excl_scan<NT>(nb, &tot, s_w);  // bigs before p0
        for (int k = 0; k < 4; k++) {
            const int32_t p = p0 + k;
            if (p >= n) break;
            if (big[k]) {
                rk[s0 + p] = ~run;
                bpos[s0 + run] = p;
                run++;
            } else {
                rk[s0 + p] = p - run;
                spos[s0 + p - run] = p;
            }
        }
        nbig += tot;
    }
    const int32_t nsmall = n - nbig;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t K, S;
        partition_ks(bpos + s0, spos + s0, n, nbig, nsmall, &K, &S);
        s_k[0] = S;
    }
    __syncthreads();
    const int32_t S = s_k[0];
    // the scatter, and the value spans of the two halves
    int32_t llo = 0x7fffffff, lhi = -0x7fffffff - 1, rlo = llo, rhi = lhi;
    for (int32_t base = 0; base < n; base += 4 * NT) {
        // four independent positions per thread, their loads in flight together
        int32_t r[4], v[4], x[4], d[4];
        for (int k = 0; k < 4; k++) {
            const int32_t p = base + k * NT + threadIdx.x;
            r[k] = p < n ? rk[s0 + p] : 0;
            v[k] = p < n ? vs[s0 + p] : 0;
            x[k] = p < n ? xs[s0 + p] : 0;
        }
        for (int k = 0; k < 4; k++) {
            const int32_t p = base + k * NT + threadIdx.x;
            const bool big = r[k] < 0;
            d[k] = p < n ? partition_dest(p, big, big ? ~r[k] : r[k], n, nsmall, S, bpos + s0,
                                          spos + s0)
                         : 0;
        }
        for (int k = 0; k < 4; k++) {
            const int32_t p = base + k * NT + threadIdx.x;
            if (p >= n) break;
            vd[s0 + d[k]] = v[k];
            xd[s0 + d[k]] = x[k];
            if (r[k] < 0) {
                rlo = min(rlo, v[k]);
                rhi = max(rhi, v[k]);
            } else {
                llo = min(llo, v[k]);
                lhi = max(lhi, v[k]);
            }
        }
    }
    block_minmax<NT>(llo, lhi, s_w);
    block_minmax<NT>(rlo, rhi, s_w + 2 * (NT / kWave));
    if (threadIdx.x != 0) return;
    const int32_t wh = 1 - sr.where;
    const SortRange L{s0, nsmall, wh, llo, lhi, 0}, R{s0 + nsmall, n - nsmall, wh, rlo, rhi, 0};
    for (const SortRange &c : {L, R}) {
        if (c.n >= 2) {
            push_range(c, next, small, counters, cap_next, cap_small);
        } else if (c.n == 1 && wh) {
            v0[c.start] = v1[c.start];
            x0[c.start] = x1[c.start];
        }
    }
}

// a range of at most kSortSmall: the reference's loop itself, by one lane,
// on a copy in LDS; the result to the first buffer
__global__ __launch_bounds__(kWave) void k_sort_small(const SortRange *__restrict__ ranges,
                                                      int32_t *__restrict__ v0,
                                                      int32_t *__restrict__ v1,
                                                      int32_t *__restrict__ x0,
                                                      int32_t *__restrict__ x1) {
    __shared__ int32_t s_v[kSortSmall], s_x[kSortSmall];
    __shared__ int2 s_stack[kSortSmall];
    const SortRange sr = ranges[blockIdx.x];
    const int32_t s0 = sr.start, n = min(sr.n, kSortSmall);
    const int32_t *__restrict__ vs = sr.where ? v1 : v0;
    const int32_t *__restrict__ xs = sr.where ? x1 : x0;
    for (int32_t p = lane_id(); p < n; p += kWave) {
        s_v[p] = vs[s0 + p];
        s_x[p] = xs[s0 + p];
    }
    __syncthreads();
    if (lane_id() == 0) {
        // ranges still to sort; each pop pushes at most two ranges of at
        // least two values each, disjoint, so the stack never holds more
        // than n / 2 of them
        int32_t sp = 0;
        s_stack[sp++] = make_int2(0, n);
        while (sp) {
            const int2 rg = s_stack[--sp];
            const int32_t a = rg.x, b = rg.y;
            int32_t lo = s_v[a], hi = s_v[a];
            for (int32_t p = a + 1; p < b; p++) {
                lo = min(lo, s_v[p]);
                hi = max(hi, s_v[p]);
            }
            if (lo == hi) continue;
            const int32_t pivot = (int32_t)((int64_t)lo + ((int64_t)hi - (int64_t)lo) / 2);
            int32_t p = a, last = b - 1;
            while (p < last) {
                if (s_v[p] > pivot) {  // swap with the last unexamined slot, look again
                    const int32_t tv = s_v[p], tx = s_x[p];
                    s_v[p] = s_v[last];
                    s_x[p] = s_x[last];
                    s_v[last] = tv;
                    s_x[last] = tx;
                    last--;
                } else {
                    p++;
                }
            }
            if (s_v[last] <= pivot) last++;
            if (last - a >= 2) s_stack[sp++] = make_int2(a, last);
            if (b - last >= 2) s_stack[sp++] = make_int2(last, b);
        }
    }
    __syncthreads();
    for (int32_t p = lane_id(); p < n; p += kWave) {
        v0[s0 + p] = s_v[p];
        x0[s0 + p] = s_x[p];
    }
}

// the particles to the caller's array in the sorted order; one sub-FOF per
// run of a root of at least min_p particles (build_fullfofs)
__global__ void k_sub_emit(int32_t m, const int32_t *__restrict__ slot_group,
                           const SubGroup *__restrict__ groups, const int32_t *__restrict__ val,
                           const int32_t *__restrict__ idx, const int32_t *__restrict__ first_label,
                           const int32_t *__restrict__ csize, const SParticle *__restrict__ w,
                           SParticle *__restrict__ parts, int64_t min_p, int4 *__restrict__ subs,
                           int32_t cap, int32_t *__restrict__ counters) {
    const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int32_t gi = slot_group[q];
    const SubGroup g = groups[gi];
    parts[g.first + (q - g.off)] = w[idx[q]];
    const int32_t v = val[q];
    if (v < 0) return;  // no neighbour
    if (q > g.off && val[q - 1] == v) return;
    const int32_t size = csize[first_label[g.off] + v];
    if (size < min_p) return;
    const int32_t e = atomicAdd(&counters[kSfEmit], 1);
    if (e < cap) subs[e] = make_int4(gi, q - g.off, size, 0);
    else atomicOr(&counters[kSfErr], 2);
}

}  // namespace hp
