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
//      leaves).  Every lane takes its own leaf's decisions; the wavefront
//      visits the union of the lanes' nodes in depth-first order, moving to
//      the left child if any lane opened the node and to the node's skip
//      pointer otherwise.  A lane that did not open node a resumes exactly at
//      a.skip, which the wavefront always lands on when it leaves a's
//      subtree.  No stack is needed and node data are wavefront-uniform
//      (scalar loads).
// The acceptance test reproduces the reference's rounding: float differences,
// double squares summed in order, rounded to float, against the float square
// of dmin.  A float estimate settles all but a band of 2^-20 around the
// threshold; lanes inside the band take the exact double path.
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

// src/potential.c:126-131, bit for bit
__device__ __forceinline__ bool accept_exact(float a0, float a1, float a2, const float *c,
                                             float dmin2) {
    double dx = (double)(a0 - c[0]);
    double dy = (double)(a1 - c[1]);
    double dz = (double)(a2 - c[2]);
    double s = dx * dx;
    s = s + dy * dy;
    s = s + dz * dz;
    return (float)s > dmin2;
}

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

struct WalkArgs {
    const Node *nodes;
    const float4 *fin;
    const int32_t *leaf_of;
    const Group *groups;
    int32_t ngroups;
    const int32_t *counters;  // node count (valid ids are [0, counters[0]))
    int32_t nrec;     // records of the batch
    int32_t cap_nodes;
    float pm;         // PARTICLE_MASS (float)
    float inv_fr;     // 1 / FORCE_RES (float)
    double pm_d;
    double fr_d;
    struct potential *po;
    int32_t *err;
    unsigned long long *work;  // visits, p2p pairs, p2m terms
};

template <int ACC, bool COUNT>
__global__ __launch_bounds__(256) void k_walk(WalkArgs a) {
    const int32_t gid = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (gid >= a.ngroups) return;
    const int lane = lane_id();
    const Group g = load_const(a.groups + gid);
    if (g.root < 0) return;  // an empty entry of the XCD layout
    const int32_t nnodes = min(load_const(a.counters + kNodeCount), a.cap_nodes);
    const int32_t root = uniform(g.root);
    const Node rn = load_const(a.nodes + root);
    const int32_t p = g.p0 + lane;
    const bool inseg = p < g.pend;
    const bool live = (uint32_t)(p - rn.first) < (uint32_t)rn.count;
    float4 me = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (inseg) me = a.fin[p];
    float lc0 = 0.0f, lc1 = 0.0f, lc2 = 0.0f;
    bool bad_leaf = false;
    if (live) {
        const int32_t lf = a.leaf_of[p];
        bad_leaf = (uint32_t)lf >= (uint32_t)nnodes;
        const Node *ln = a.nodes + (bad_leaf ? 0 : lf);
        lc0 = ln->com[0];
        lc1 = ln->com[1];
        lc2 = ln->com[2];
    }
    if (ballot(bad_leaf)) {
        if (lane == 0) atomicOr(a.err, kErrBadTree);
        return;
    }
    int32_t resume = live ? root : -2;
    Acc<ACC> acc;
    int32_t c = root;
    int32_t guard = 0;
    const int32_t cap = 2 * rn.count + 4;
    int since = 0;
    unsigned long long n_vis = 0, n_p2p = 0, n_p2m = 0;

    // the arguments used inside the loop are made resident here, so that the
    // loop's waits are for its own node loads only
    const float pm = a.pm, inv_fr = a.inv_fr;
    const double pm_d = a.pm_d, fr_d = a.fr_d;
    const float4 *__restrict__ fin = a.fin;
    const Node *__restrict__ nodes = a.nodes;
    asm volatile("" ::"s"(pm), "s"(inv_fr), "s"(pm_d), "s"(fr_d), "s"(fin), "s"(nodes));
    while (c >= 0) {
        if (++guard > cap) {
            if (lane == 0) atomicOr(a.err, kErrWalkCap);
            break;
        }
        c = uniform(c);
        if ((uint32_t)c >= (uint32_t)nnodes) {
            if (lane == 0) atomicOr(a.err, kErrBadTree);
            break;
        }
        const Node nd = load_const(nodes + c);
        const bool act = resume == c;
        const bool inr = (uint32_t)(p - nd.first) < (uint32_t)nd.count;
        const bool cand = act & !inr;
        // acceptance, src/potential.c:126-131
        const float dx = lc0 - nd.com[0], dy = lc1 - nd.com[1], dz = lc2 - nd.com[2];
        const float f = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        bool acc_ok = f > nd.dmin2;
        const bool amb = cand & (fabsf(f - nd.dmin2) <= nd.dmin2 * 0x1p-20f);
        if (__builtin_expect(ballot(amb) != 0, 0)) {
            const bool ex = accept_exact(lc0, lc1, lc2, nd.com, nd.dmin2);
            acc_ok = amb ? ex : acc_ok;
        }
        const bool p2m = cand & acc_ok;

        if constexpr (COUNT) {
            n_vis += 1;
            n_p2m += __popcll(ballot(p2m));
        }
        // monopole, unsoftened (src/potential.c:162-166)
        {
            if constexpr (ACC == HP_ACC_F64) {
                double ex = (double)me.x - (double)nd.com[0];
                double ey = (double)me.y - (double)nd.com[1];
                double ez = (double)me.z - (double)nd.com[2];
                double t = (double)nd.mass / sqrt(ex * ex + ey * ey + ez * ez);
                acc.add(p2m ? t : 0.0);
            } else {
                const float ex = me.x - nd.com[0], ey = me.y - nd.com[1], ez = me.z - nd.com[2];
                const float r2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                const float t = nd.mass * rsq(r2);
                acc.add(p2m ? t : 0.0f);
                if constexpr (ACC == HP_ACC_COMPENSATED) {
                    if (++since == 16) {
                        acc.fold();
                        since = 0;
                    }
                }
            }
        }
        if (nd.left >= 0) {
            // an internal node: opened by a lane whose leaf lies inside it
            // or that does not accept it
            const bool open = act & (inr | !acc_ok);
            if (act) resume = open ? nd.left : nd.skip;
            c = ballot(open) != 0 ? nd.left : nd.skip;
            continue;
        }
        // a leaf: particle pairs, softened (src/potential.c:38-55); the self
        // leaf only up to 20 points (src/potential.c:144-146).  The leaf's
        // sources are loaded once, one per lane, and broadcast lane by lane.
        const bool small = nd.count <= kDegenerate;
        const bool p2p = act & ((inr & small) | (!inr & !acc_ok));
        if (ballot(p2p)) {
            if constexpr (COUNT) n_p2p += (unsigned long long)wave_sum(p2p ? (int)nd.count - (inr ? 1 : 0) : 0);
            for (int32_t j0 = 0; j0 < nd.count; j0 += kWave) {
                const int32_t cnt = min(kWave, nd.count - j0);
                const float4 src = lane < cnt ? fin[nd.first + j0 + lane]
                                              : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                const int32_t self = p - (nd.first + j0);  // this lane's own slot, if any
                for (int32_t jb = 0; jb < cnt; jb += 16) {
                    const int32_t je = min(jb + 16, cnt);
                    if constexpr (ACC == HP_ACC_F64) {
                        double s = 0.0;
                        for (int32_t j = jb; j < je; j++) {
                            const float qx = lane_float(src.x, j), qy = lane_float(src.y, j),
                                        qz = lane_float(src.z, j);
                            double ex = (double)(me.x - qx), ey = (double)(me.y - qy),
                                   ez = (double)(me.z - qz);
                            double r = sqrt((ex * ex + ey * ey) + ez * ez);
                            r = r < fr_d ? fr_d : r;
                            double t = pm_d * (1.0 / r);
                            s += (p2p && j != self) ? t : 0.0;
                        }
                        acc.add_block(s);
                    } else {
                        float s = 0.0f;
                        for (int32_t j = jb; j < je; j++) {
                            const float qx = lane_float(src.x, j), qy = lane_float(src.y, j),
                                        qz = lane_float(src.z, j);
                            const float ex = me.x - qx, ey = me.y - qy, ez = me.z - qz;
                            const float r2 = fmaf(ez, ez, fmaf(ey, ey, ex * ex));
                            const float t = fminf(rsq(r2), inv_fr);
                            s += (p2p && j != self) ? t : 0.0f;
                        }
                        acc.add_block(pm * s);
                    }
                }
            }
        }
        if (act) resume = nd.skip;
        c = nd.skip;
    }
    if constexpr (ACC == HP_ACC_COMPENSATED) acc.fold();
    if (inseg) {
        const int32_t idx = __float_as_int(me.w);
        if ((uint32_t)idx < (uint32_t)a.nrec) a.po[idx].pe = live ? acc.value() : 0.0;
        else atomicOr(a.err, kErrBadTree);
    }
    if constexpr (COUNT) {
        if (lane == 0) {
            atomicAdd(&a.work[0], n_vis);
            atomicAdd(&a.work[1], n_p2p);
            atomicAdd(&a.work[2], n_p2m);
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
    float pm;
    float inv_fr;
    double pm_d;
    double fr_d;
    struct potential *po;
};

template <int ACC>
__global__ __launch_bounds__(256) void k_direct(DirectArgs a) {
    __shared__ float4 s_src[256];
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
// 762-773, 875-887): out[j] = po[record at tree position j].
__global__ void k_tree_records(const struct potential *__restrict__ po,
                               const float4 *__restrict__ fin, int64_t n,
                               struct potential *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = po[__float_as_int(fin[j].w)];
}

// pe = 0 before the direct kernel, which writes finite records only
__global__ void k_zero_pe(struct potential *po, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) po[i].pe = 0.0;
}
__global__ void k_zero_pe_calls(struct potential *po, const int64_t *hfirst,
                                const int64_t *hcount, const int2 *blocks) {
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
