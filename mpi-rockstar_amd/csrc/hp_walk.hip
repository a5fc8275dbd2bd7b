// hp_walk.hip -- Barnes-Hut walk and direct-sum kernels.
//
// The reference walks the tree once per sink leaf (src/potential.c:133-168,
// 178-180): for a source node n2 it
//   * computes the self-leaf pair sum when n2 is the sink (skipped above 20
//     points),
//   * opens n2 when the sink's box lies inside n2's box or the monopole is
//     not acceptable (|com_sink - com_n2|^2 > dmin^2 in float fails), doing
//     leaf-leaf pair sums at leaves,
//   * otherwise adds the unsoftened monopole m_n2 / r.
// Two facts make a wavefront-wide walk exact:
//   1. "the sink's box lies inside n2's box" holds exactly when n2 is an
//      ancestor of the sink leaf: any other node is separated from the sink by
//      the split plane of their lowest common ancestor.  Node ranges nest the
//      same way, so the test is `first <= p < first + count`.
//   2. Lanes are sink particles (64 consecutive points in tree order, a few
//      leaves).  Every lane takes its own leaf's decisions and the wavefront
//      visits the union of the lanes' nodes in depth-first order.  Node data
//      are wavefront-uniform (scalar loads).  The default walk decides both
//      children of an opened node in one step, with a short stack of pending
//      right children in registers; the node-by-node walk, on skip pointers
//      and with no stack, is the fallback for trees deeper than the stack.
// The acceptance test reproduces the reference's rounding exactly (float
// differences, double squares summed in order, rounded to float, against
// the float square of dmin); FP64 runs at the FP32 rate on this GPU, so the
// exact test costs no more than a float estimate with a fallback would.
// Lane sets (active, inside, accepting, opening) are wavefront masks handled
// by the scalar unit.
//
// Pair arithmetic is FP32 (rsqrt, clamp at FORCE_RES for particle pairs,
// src/potential.c:31-36) with compensated accumulation: terms are summed in
// short FP32 blocks (<= 16 terms) that are folded into a (hi, lo) pair by
// two-sum.  HP_ACC_NAIVE keeps one FP32 sum; HP_ACC_F64 evaluates every term
// in FP64 the way the reference does.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <int ACC>
struct Acc;

template <>
struct Acc<HP_ACC_COMPENSATED> {
    float hi = 0.0f, lo = 0.0f, blk = 0.0f;
    __device__ __forceinline__ void add(float t) { blk += t; }
    __device__ __forceinline__ void fold() {
        float s, e;
        two_sum(hi, blk, s, e);
        hi = s;
        lo += e;
        blk = 0.0f;
    }
    __device__ __forceinline__ void add_block(float t) {
        float s, e;
        two_sum(hi, t, s, e);
        hi = s;
        lo += e;
    }
    __device__ __forceinline__ double value() const {
        return ((double)hi + (double)lo) + (double)blk;
    }
};

template <>
struct Acc<HP_ACC_NAIVE> {
    float s = 0.0f;
    __device__ __forceinline__ void add(float t) { s += t; }
    __device__ __forceinline__ void fold() {}
    __device__ __forceinline__ void add_block(float t) { s += t; }
    __device__ __forceinline__ double value() const { return (double)s; }
};

template <>
struct Acc<HP_ACC_F64> {
    double s = 0.0;
    __device__ __forceinline__ void add(double t) { s += t; }
    __device__ __forceinline__ void fold() {}
    __device__ __forceinline__ void add_block(double t) { s += t; }
    __device__ __forceinline__ double value() const { return s; }
};

// src/potential.c:126-131, bit for bit: float differences, their squares
// summed in double in the reference's order, the sum rounded to float and
// compared with the float square of dmin.  A float times a float is exact in
// double, so fma(d, d, s) rounds s + d * d exactly as the separate product
// and sum do.
__device__ __forceinline__ bool accept_exact(float a0, float a1, float a2, const float *c,
                                             float dmin2) {
    const double dx = (double)(a0 - c[0]);
    const double dy = (double)(a1 - c[1]);
    const double dz = (double)(a2 - c[2]);
    double s = dx * dx;
    s = fma(dy, dy, s);
    s = fma(dz, dz, s);
    return (float)s > dmin2;
}

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// v[l] = x for the lane l (wavefront-uniform); no lane for l outside [0, 64).
__device__ __forceinline__ void write_lane(int32_t &v, int32_t x, int32_t l) {
    v = lane_id() == l ? x : v;
}

struct WalkArgs {
    const Node *nodes;
    const float4 *fin;
    const int32_t *leaf_of;
    const Group *groups;
    int32_t ngroups;
    const int32_t *counters;  // node count (valid ids are [0, counters[0]))
    int32_t nrec;     // records of the batch
    int32_t cap_nodes;
    int32_t pairs;    // 1: sibling-pair walk, 0: node-by-node walk
    float pm;         // PARTICLE_MASS (float)
    float inv_fr;     // 1 / FORCE_RES (float)
    double pm_d;
    double fr_d;
    struct potential *po;
    int32_t *err;
    unsigned long long *work;  // node visits, p2p pairs, p2m terms
};

// One wavefront's walk state: lane = sink particle.
template <int ACC, bool COUNT>
struct Walker {
    float4 me;                  // the sink (x, y, z, record index)
    float lc0, lc1, lc2;        // its leaf's centre of mass
    int32_t p;                  // its tree position
    Acc<ACC> acc;
    int since = 0;
    unsigned long long n_vis = 0, n_p2p = 0, n_p2m = 0;
    float pm, inv_fr;
    double pm_d, fr_d;
    const float4 *__restrict__ fin;

    __device__ __forceinline__ void reset() {
        acc = Acc<ACC>();
        since = 0;
        n_vis = n_p2p = n_p2m = 0;
    }

    // Node nd for the lanes of `act` (those whose walk reaches it): the
    // monopole where it is accepted, the leaf's particle pairs where a leaf
    // is opened.  Returns the lanes that open nd (none for a leaf).  Lane
    // sets are wavefront masks (the compares write them directly), combined
    // on the scalar unit.
    __device__ __forceinline__ uint64_t visit(const NodeHead &nd, uint64_t act) {
        const uint64_t inr = ballot((uint32_t)(p - nd.first) < (uint32_t)nd.count);
        const uint64_t cand = act & ~inr;
        // acceptance, src/potential.c:126-131 (not evaluated where every
        // active lane lies inside nd and so opens it)
        const uint64_t ok = cand ? ballot(accept_exact(lc0, lc1, lc2, nd.com, nd.dmin2)) : 0;
        const uint64_t p2m = cand & ok;
        if constexpr (COUNT) {
            n_vis += 1;
            n_p2m += __popcll(p2m);
        }
        // monopole, unsoftened (src/potential.c:162-166)
        if (p2m) {
            const bool take = __builtin_amdgcn_inverse_ballot_w64(p2m);
            if constexpr (ACC == HP_ACC_F64) {
                double ex = (double)me.x - (double)nd.com[0];
                double ey = (double)me.y - (double)nd.com[1];
                double ez = (double)me.z - (double)nd.com[2];
                double t = (double)nd.mass / sqrt(ex * ex + ey * ey + ez * ez);
                acc.add(take ? t : 0.0);
            } else {
                const float ex = me.x - nd.com[0], ey = me.y - nd.com[1], ez = me.z - nd.com[2];
                const float r2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                const float t = nd.mass * rsq(r2);
                acc.add(take ? t : 0.0f);
                if constexpr (ACC == HP_ACC_COMPENSATED) {
                    if (++since == 16) {
                        acc.fold();
                        since = 0;
                    }
                }
            }
        }
        // an internal node is opened by a lane whose leaf lies inside it or
        // that does not accept it
        if (nd.left >= 0) return act & (inr | ~ok);
        // a leaf: particle pairs, softened (src/potential.c:38-55); the self
        // leaf only up to 20 points (src/potential.c:144-146)
        const uint64_t p2p = act & ((nd.count <= kDegenerate ? inr : 0ull) | (~inr & ~ok));
        if (p2p) {
            if constexpr (COUNT) {
                n_p2p += (unsigned long long)__popcll(p2p) * (unsigned long long)nd.count -
                         (unsigned long long)__popcll(p2p & inr);
            }
            const bool mine = __builtin_amdgcn_inverse_ballot_w64(p2p);
            // the self pair is left out only where a lane's own leaf is summed
            if (p2p & inr) leaf_pairs<true>(nd, mine);
            else leaf_pairs<false>(nd, mine);
        }
        return 0;
    }

    // The pairs of this lane's sink with the leaf's points, added when p2p.
    // The points are read with scalar loads, kSrcBatch at a time (the point
    // buffer is padded for the last batch), and used as scalar operands.
    template <bool SELF>
    __device__ __forceinline__ void leaf_pairs(const NodeHead &nd, bool p2p) {
        struct Batch {
            float4 q[kSrcBatch];
        };
        for (int32_t j0 = 0; j0 < nd.count; j0 += kSrcBatch) {
            const int32_t cnt = min(kSrcBatch, nd.count - j0);
            const Batch b = load_const(reinterpret_cast<const Batch *>(fin + nd.first + j0));
            // all of the batch's loads are issued before the first use (the
            // compiler would otherwise load each point where it is used and
            // wait for each load on its own)
            static_assert(kSrcBatch == 8, "the batch is pinned in eight steps");
            asm volatile("" ::"s"(b.q[0].x), "s"(b.q[0].y), "s"(b.q[0].z), "s"(b.q[1].x),
                         "s"(b.q[1].y), "s"(b.q[1].z), "s"(b.q[2].x), "s"(b.q[2].y),
                         "s"(b.q[2].z), "s"(b.q[3].x), "s"(b.q[3].y), "s"(b.q[3].z));
            asm volatile("" ::"s"(b.q[4].x), "s"(b.q[4].y), "s"(b.q[4].z), "s"(b.q[5].x),
                         "s"(b.q[5].y), "s"(b.q[5].z), "s"(b.q[6].x), "s"(b.q[6].y),
                         "s"(b.q[6].z), "s"(b.q[7].x), "s"(b.q[7].y), "s"(b.q[7].z));
            const int32_t self = p - (nd.first + j0);  // this lane's own slot, if any
            if constexpr (ACC == HP_ACC_F64) {
                double s = 0.0;
#pragma unroll
                for (int j = 0; j < kSrcBatch; j++) {
                    if (j < cnt) {
                        double ex = (double)(me.x - b.q[j].x), ey = (double)(me.y - b.q[j].y),
                               ez = (double)(me.z - b.q[j].z);
                        double r = sqrt((ex * ex + ey * ey) + ez * ez);
                        r = r < fr_d ? fr_d : r;
                        double t = pm_d * (1.0 / r);
                        s += (!SELF || j != self) ? t : 0.0;
                    }
                }
                acc.add_block(p2p ? s : 0.0);
            } else {
                float s = 0.0f;
#pragma unroll
                for (int j = 0; j < kSrcBatch; j++) {
                    if (j < cnt) {
                        const float ex = me.x - b.q[j].x, ey = me.y - b.q[j].y, ez = me.z - b.q[j].z;
                        const float r2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                        const float t = fminf(rsq(r2), inv_fr);
                        s += (!SELF || j != self) ? t : 0.0f;
                    }
                }
                acc.add_block(p2p ? pm * s : 0.0f);
            }
        }
    }

    // Sibling-pair walk.  Both children of an opened node are decided in one
    // step (their two heads share a 128-byte line), which halves the chain of
    // dependent node loads.  The lanes that open the right child are kept on
    // a stack of pending pairs, entry k in lane k of three registers (first
    // child, lane mask); the walk descends into the left child first, so the
    // visiting order is depth-first as in the node-by-node walk and every
    // lane takes the same decisions.  Returns false when more than 64 pairs
    // would be pending (a tree deeper than 64 levels below the call's root);
    // the caller then walks node by node.
    __device__ bool walk_pairs(const Node *__restrict__ nodes, const Node &rn, uint64_t live,
                               int32_t nnodes, int32_t cap, int32_t *err) {
        const int lane = lane_id();
        if (rn.left < 0) {  // a one-leaf tree: the self-leaf pairs
            visit(*reinterpret_cast<const NodeHead *>(&rn), live);
            return true;
        }
        int32_t st_c = 0, st_lo = 0, st_hi = 0;
        int32_t sp = 0;
        uint64_t m = live;
        int32_t c = rn.left;
        int32_t guard = cap, fail = 0;
        while (c >= 0) {
            if (sp > kWave) break;  // a pending pair was lost: the caller walks node by node
            --guard;
            fail = (uint32_t)c >= (uint32_t)(nnodes - 1) ? kErrBadTree : (guard < 0 ? kErrWalkCap : 0);
            if (fail != 0) break;
            const NodeHead n0 = load_const(reinterpret_cast<const NodeHead *>(nodes + c));
            const NodeHead n1 = load_const(reinterpret_cast<const NodeHead *>(nodes + c + 1));
            // both loads are issued before either node is used
            asm volatile("" ::"s"(n1.com[0]), "s"(n1.dmin2), "s"(n1.left), "s"(n0.com[0]));
            const uint64_t m0 = visit(n0, m);
            const uint64_t m1 = visit(n1, m);
            if (m1) {
                // entry sp of the stack: lane sp of each register (a lane
                // past 63 selects none, and the check above ends the walk)
                write_lane(st_c, n1.left, sp);
                write_lane(st_lo, (int32_t)(uint32_t)m1, sp);
                write_lane(st_hi, (int32_t)(uint32_t)(m1 >> 32), sp);
                sp++;
            }
            if (m0) {
                c = n0.left;
                m = m0;
            } else if (sp > 0 && sp <= kWave) {
                sp--;
                c = __builtin_amdgcn_readlane(st_c, sp);
                m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(st_hi, sp) << 32) |
                    (uint32_t)__builtin_amdgcn_readlane(st_lo, sp);
            } else {
                c = -1;
            }
        }
        if (fail && lane == 0) atomicOr(err, fail);
        return sp <= kWave;
    }

    // Node-by-node walk.  The wavefront visits the union of the lanes'
    // nodes in depth-first order, moving to the left child if any lane
    // opened the node and to its skip pointer otherwise; a lane that did not
    // open node a resumes exactly at a.skip, which the wavefront always
    // lands on when it leaves a's subtree.
    __device__ void walk_nodes(const Node *__restrict__ nodes, int32_t root, uint64_t live,
                               int32_t nnodes, int32_t cap, int32_t *err) {
        const int lane = lane_id();
        int32_t resume = __builtin_amdgcn_inverse_ballot_w64(live) ? root : -2;
        int32_t c = root;
        int32_t guard = 0;
        while (c >= 0) {
            if (++guard > cap) {
                if (lane == 0) atomicOr(err, kErrWalkCap);
                return;
            }
            c = uniform(c);
            if ((uint32_t)c >= (uint32_t)nnodes) {
                if (lane == 0) atomicOr(err, kErrBadTree);
                return;
            }
            const Node nd = load_const(nodes + c);
            const uint64_t act = ballot(resume == c);
            const uint64_t open = visit(*reinterpret_cast<const NodeHead *>(&nd), act);
            if (__builtin_amdgcn_inverse_ballot_w64(act))
                resume = __builtin_amdgcn_inverse_ballot_w64(open) ? nd.left : nd.skip;
            c = open ? nd.left : nd.skip;
        }
    }
};

template <int ACC, bool COUNT>
__global__ __launch_bounds__(64) void k_walk(WalkArgs a) {
    const int32_t gid = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (gid >= a.ngroups || (*a.err & kErrDeep)) return;
    const int lane = lane_id();
    const Group g = load_const(a.groups + gid);
    if (g.root < 0) return;  // an empty entry of the XCD layout
    const int32_t nnodes = min(load_const(a.counters + kNodeCount), a.cap_nodes);
    const int32_t root = uniform(g.root);
    const Node rn = load_const(a.nodes + root);
    Walker<ACC, COUNT> w;
    w.p = g.p0 + lane;
    const bool inseg = w.p < g.pend;
    const bool live = (uint32_t)(w.p - rn.first) < (uint32_t)rn.count;
    w.me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (inseg) w.me = a.fin[w.p];
    w.lc0 = w.lc1 = w.lc2 = 0.0f;
    bool bad_leaf = false;
    if (live) {
        const int32_t lf = a.leaf_of[w.p];
        bad_leaf = (uint32_t)lf >= (uint32_t)nnodes;
        const Node *ln = a.nodes + (bad_leaf ? 0 : lf);
        w.lc0 = ln->com[0];
        w.lc1 = ln->com[1];
        w.lc2 = ln->com[2];
    }
    if (ballot(bad_leaf)) {
        if (lane == 0) atomicOr(a.err, kErrBadTree);
        return;
    }
    // the arguments used inside the loops are made resident here, so that
    // the loops' waits are for their own node loads only
    w.pm = a.pm;
    w.inv_fr = __builtin_canonicalizef(a.inv_fr);  // fminf needs no per-pair canonicalisation
    w.pm_d = a.pm_d;
    w.fr_d = a.fr_d;
    w.fin = a.fin;
    const Node *__restrict__ nodes = a.nodes;
    asm volatile("" ::"s"(w.pm), "s"(w.pm_d), "s"(w.fr_d), "s"(w.fin), "s"(nodes));
    const int32_t cap = 2 * rn.count + 4;
    bool done = false;
    const uint64_t live_m = ballot(live);
    if (a.pairs) done = w.walk_pairs(nodes, rn, live_m, nnodes, cap, a.err);
    if (!done) {
        w.reset();
        w.walk_nodes(nodes, root, live_m, nnodes, cap, a.err);
    }
    if constexpr (ACC == HP_ACC_COMPENSATED) w.acc.fold();
    if (inseg) {
        const int32_t idx = __float_as_int(w.me.w);
        if ((uint32_t)idx < (uint32_t)a.nrec) a.po[idx].pe = live ? w.acc.value() : 0.0;
        else atomicOr(a.err, kErrBadTree);
    }
    if constexpr (COUNT) {
        if (lane == 0) {
            atomicAdd(&a.work[0], w.n_vis);
            atomicAdd(&a.work[1], w.n_p2p);
            atomicAdd(&a.work[2], w.n_p2m);
        }
    }
}

// Exact pairwise sum over a halo's finite points, source tiles staged in LDS.
// One workgroup of 256 sinks per block entry (halo root, first sink).  For a
// halo of at most 10 finite points this is the reference's answer: its tree
// is a single leaf and the walk reduces to the self-leaf pair sum.
struct DirectArgs {
    const Node *nodes;
    const float4 *fin;
    const int2 *blocks;   // (root node, first sink offset)
    int32_t nblocks;
    const int32_t *flags;  // error flags (nothing is written under kErrDeep)
    float pm;
    float inv_fr;
    double pm_d;
    double fr_d;
    struct potential *po;
};

template <int ACC>
__global__ __launch_bounds__(256) void k_direct(DirectArgs a) {
    __shared__ float4 s_src[256];
    if (*a.flags & kErrDeep) return;
    const int2 b = a.blocks[blockIdx.x];
    const Node rn = a.nodes[b.x];
    const int32_t i = b.y + threadIdx.x;
    const bool live = i < rn.count;
    float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live) me = a.fin[rn.first + i];
    Acc<ACC> acc;
    for (int32_t t0 = 0; t0 < rn.count; t0 += 256) {
        __syncthreads();
        if (t0 + (int32_t)threadIdx.x < rn.count) s_src[threadIdx.x] = a.fin[rn.first + t0 + threadIdx.x];
        __syncthreads();
        const int32_t nt = min(256, rn.count - t0);
        for (int32_t j0 = 0; j0 < nt; j0 += 16) {
            const int32_t j1 = min(j0 + 16, nt);
            if constexpr (ACC == HP_ACC_F64) {
                double s = 0.0;
                for (int32_t j = j0; j < j1; j++) {
                    const float4 q = s_src[j];
                    double ex = (double)(me.x - q.x), ey = (double)(me.y - q.y),
                           ez = (double)(me.z - q.z);
                    double r = sqrt((ex * ex + ey * ey) + ez * ez);
                    r = r < a.fr_d ? a.fr_d : r;
                    s += (t0 + j != i) ? a.pm_d * (1.0 / r) : 0.0;
                }
                acc.add_block(s);
            } else {
                float s = 0.0f;
                for (int32_t j = j0; j < j1; j++) {
                    const float4 q = s_src[j];
                    const float ex = me.x - q.x, ey = me.y - q.y, ez = me.z - q.z;
                    const float r2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                    const float t = fminf(rsq(r2), a.inv_fr);
                    s += (t0 + j != i) ? t : 0.0f;
                }
                acc.add_block(a.pm * s);
            }
        }
    }
    if (live) a.po[__float_as_int(me.w)].pe = acc.value();
}

// The records in tree order, the order the reference's compute_potential
// leaves them in (its tree build permutes po in place, src/fast3tree.c:
// 762-773, 875-887): out[j] = po[record at tree position j].  Each
// workgroup's 256 records go out through LDS as 16-byte words, so that the
// stores of a wavefront are contiguous (the loads are a gather).
__global__ __launch_bounds__(256) void k_tree_records(const struct potential *__restrict__ po,
                                                      const float4 *__restrict__ fin, int64_t n,
                                                      struct potential *__restrict__ out) {
    __shared__ float4 s_rec[3 * 256];
    const int64_t j0 = (int64_t)blockIdx.x * 256;
    const int64_t j = j0 + threadIdx.x;
    const int m = (int)min((int64_t)256, n - j0);
    if (j < n) {
        const float4 *src = reinterpret_cast<const float4 *>(po + __float_as_int(fin[j].w));
        for (int q = 0; q < 3; q++) s_rec[3 * threadIdx.x + q] = src[q];
    }
    __syncthreads();
    float4 *dst = reinterpret_cast<float4 *>(out + j0);
    for (int q = threadIdx.x; q < 3 * m; q += 256) dst[q] = s_rec[q];
}

// pe = 0 before the direct kernel, which writes finite records only
__global__ void k_zero_pe(struct potential *po, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) po[i].pe = 0.0;
}
__global__ void k_zero_pe_calls(struct potential *po, const int64_t *hfirst,
                                const int64_t *hcount, const int2 *blocks, const int32_t *flags) {
    if (*flags & kErrDeep) return;
    const int2 b = blocks[blockIdx.x];
    const int64_t i = b.y + threadIdx.x;
    if (i < hcount[b.x]) po[hfirst[b.x] + i].pe = 0.0;
}

template __global__ void k_walk<HP_ACC_COMPENSATED, false>(WalkArgs);
template __global__ void k_walk<HP_ACC_COMPENSATED, true>(WalkArgs);
template __global__ void k_walk<HP_ACC_NAIVE, false>(WalkArgs);
template __global__ void k_walk<HP_ACC_NAIVE, true>(WalkArgs);
template __global__ void k_walk<HP_ACC_F64, false>(WalkArgs);
template __global__ void k_walk<HP_ACC_F64, true>(WalkArgs);
template __global__ void k_direct<HP_ACC_COMPENSATED>(DirectArgs);
template __global__ void k_direct<HP_ACC_NAIVE>(DirectArgs);
template __global__ void k_direct<HP_ACC_F64>(DirectArgs);

}  // namespace hp
