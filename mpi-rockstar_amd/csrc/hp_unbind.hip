// hp_unbind.hip -- the bound pass after the potential
// (src/properties.c:539-550 and the bound test of src/properties.c:403),
// for every call of a batch at once:
//   1. keys: (call, potential-only flag, r2 bits).  A stable LSD radix sort
//      (8-bit digits) orders every call's records by r2 with the potential-
//      only records (ke < 0) last; ties keep their input order, which is what
//      glibc's merge-sort qsort does with the reference's dist_compare
//      (src/groupies.c:202-210).  r2 >= 0, so its bits sort as unsigned.
//   2. gather the records into sorted order.
//   3. core / bulk velocity (src/properties.c:165-214).  rvir_max is the
//      largest j with j^2 > r2^3 (4pi/3 rho_vir)^2 and core_max the largest j
//      with 100 r2_j < r2[rvir_max] (at least 100).  The reference's running
//      Welford mean and variance over the sorted velocities are restated with
//      f64 prefix sums (exact for the first moment), the variance of a prefix
//      being S2 - S1^2 / n per axis.  The reference keeps the last prefix i
//      (i < 10 always, then whenever var_i < best (i - 3) i) before core_max
//      for corevel and before rvir_max for bulkvel: that is the first index
//      of the minimum of var_i / ((i - 3) i) over i in [9, limit), or
//      limit - 1 when limit <= 10.  Found with a segmented argmin.
//   4. kinetic energy about corevel (sub-halos) or bulkvel (hosts), as
//      compute_kinetic_energy does (float ke, double terms), and the bound
//      flag !(pe < ke).
// The per-call tables are tiles of 2048 sorted records that never straddle
// two calls.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;              // per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kScanTile = 2048;

// ---------------------------------------------------------------- radix sort

__global__ void k_make_keys(const struct potential *__restrict__ po,
                            const int32_t *__restrict__ rec_call, int32_t nrec,
                            uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    const float r2 = po[i].r2;
    const uint64_t pot = po[i].ke < 0.0f ? 1ull : 0ull;
    keys[i] = ((uint64_t)(uint32_t)rec_call[i] << 33) | (pot << 32) |
              (uint64_t)__float_as_uint(r2);
    vals[i] = i;
}

// digit histogram of one tile: hist[digit * ntiles + tile]
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                             int32_t n, int shift,
                                                             int32_t ntiles,
                                                             int32_t *__restrict__ hist) {
    __shared__ int32_t s_h[256];
    s_h[threadIdx.x] = 0;
    __syncthreads();
    const int32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        int32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * ntiles + blockIdx.x] = s_h[threadIdx.x];
}

// exclusive scan of hist (length m) in place; one workgroup, chunked
__global__ __launch_bounds__(1024) void k_scan_i32(int32_t *__restrict__ v, int32_t m) {
    __shared__ int32_t s_w[16];
    __shared__ int32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int32_t base = 0; base < m; base += 1024) {
        int32_t i = base + threadIdx.x;
        int32_t x = i < m ? v[i] : 0;
        // inclusive wave scan
        int32_t y = x;
        for (int o = 1; o < 64; o <<= 1) {
            int32_t t = __shfl_up(y, o, 64);
            if (lane_id() >= o) y += t;
        }
        const int w = threadIdx.x / kWave;
        if (lane_id() == 63) s_w[w] = y;
        __syncthreads();
        int32_t off = s_carry;
        for (int j = 0; j < w; j++) off += s_w[j];
        if (i < m) v[i] = off + y - x;
        __syncthreads();
        if (threadIdx.x == 1023) s_carry = off + y;
        __syncthreads();
    }
}

// stable scatter of one tile: rounds of 256 consecutive keys, wave-level
// multisplit by digit, running per-digit counters across rounds
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin,
    uint64_t *__restrict__ kout, int32_t *__restrict__ vout, int32_t n, int shift,
    int32_t ntiles, const int32_t *__restrict__ offs) {
    __shared__ int32_t s_base[256];
    __shared__ int32_t s_wc[kSortThreads / kWave][256];
    const int w = threadIdx.x / kWave;
    s_base[threadIdx.x] = offs[threadIdx.x * ntiles + blockIdx.x];
    const int32_t tile0 = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        for (int j = 0; j < kSortThreads / kWave; j++) s_wc[j][threadIdx.x] = 0;
        __syncthreads();
        const int32_t i = tile0 + r * kSortThreads + threadIdx.x;
        const bool v = i < n;
        uint64_t key = v ? kin[i] : 0;
        const int32_t val = v ? vin[i] : 0;
        const int dg = (int)((key >> shift) & 255);
        // lanes of this wave with the same digit
        uint64_t peers = ballot(v);
        for (int b = 0; b < 8; b++) {
            uint64_t m = ballot(((dg >> b) & 1) != 0);
            peers &= ((dg >> b) & 1) ? m : ~m;
        }
        const int rk = rank_below(peers);
        const int cnt = __popcll(peers);
        // the lowest peer publishes the wave's count of this digit
        if (v && rk == 0) s_wc[w][dg] = cnt;
        __syncthreads();
        if (v) {
            int32_t off = s_base[dg];
            for (int j = 0; j < w; j++) off += s_wc[j][dg];
            kout[off + rk] = key;
            vout[off + rk] = val;
        }
        __syncthreads();
        // advance the running base by this round's totals
        int32_t tot = 0;
        for (int j = 0; j < kSortThreads / kWave; j++) tot += s_wc[j][threadIdx.x];
        s_base[threadIdx.x] += tot;
        __syncthreads();
    }
}

// ----------------------------------------------------------------- gather

__global__ void k_gather_sorted(const struct potential *__restrict__ po,
                                const int32_t *__restrict__ order, int32_t nrec,
                                struct potential *__restrict__ out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nrec) out[i] = po[order[i]];
}

// ------------------------------------------------------------ core velocity

struct CallInfo {
    int32_t first;      // sorted records of the call: [first, first + kept)
    int32_t kept;
    int32_t rvir_max;   // -1: no virial prefix
    int32_t core_max;
};

// kept counts: number of records with ke >= 0 per call (sorted order puts
// them first); rvir_max by a per-call max reduction
__global__ void k_call_init(const hp_halo *__restrict__ halos, int32_t nh,
                            CallInfo *__restrict__ ci) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    CallInfo c;
    c.first = (int32_t)halos[h].first;
    c.kept = 0;
    c.rvir_max = -1;
    c.core_max = 100;
    ci[h] = c;
}

__global__ void k_kept_rvir(const struct potential *__restrict__ srt,
                            const int32_t *__restrict__ rec_call, int32_t nrec,
                            double rvir_thresh, CallInfo *__restrict__ ci) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    const struct potential &r = srt[i];
    if (r.ke < 0.0f) return;
    const int32_t h = rec_call[i];
    const int32_t j = i - ci[h].first;
    atomicAdd(&ci[h].kept, 1);
    // src/properties.c:173-176: float r2^3, then double products
    const float r6 = (r.r2 * r.r2) * r.r2;
    if ((double)((int64_t)j * (int64_t)j) > (double)r6 * (rvir_thresh * rvir_thresh))
        atomicMax(&ci[h].rvir_max, j);
}

// core_max: records with 100 r2 < r2[rvir_max] form a prefix of the sorted
// call; binary search (src/properties.c:179-184)
__global__ void k_core_max(const struct potential *__restrict__ srt, int32_t nh,
                           CallInfo *__restrict__ ci) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    CallInfo c = ci[h];
    if (c.rvir_max < 1) return;
    const float ref = srt[c.first + c.rvir_max].r2;
    int32_t lo = 0, hi = c.kept;  // first index with !(r2 * 100 < ref)
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if ((double)srt[c.first + mid].r2 * 100.0 < (double)ref) lo = mid + 1;
        else hi = mid;
    }
    int32_t cm = lo - 1;
    ci[h].core_max = cm < 100 ? 100 : cm;
}

// Scan tiles: tile t covers sorted records [first + off, first + min(off +
// 2048, limit)) of call h, limit = rvir_max.
struct ScanTile {
    int32_t call;
    int32_t off;
};

struct Moments {
    double s1[3];
    double s2[3];
};

// per-tile sums of v and v^2
__global__ __launch_bounds__(256) void k_vel_sums(const struct potential *__restrict__ srt,
                                                  const ScanTile *__restrict__ tiles,
                                                  int32_t ntiles, const CallInfo *__restrict__ ci,
                                                  Moments *__restrict__ tsum) {
    __shared__ double s_m[6][4];
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const CallInfo c = ci[st.call];
    const int32_t lim = c.rvir_max;
    double m[6] = {0, 0, 0, 0, 0, 0};
    for (int r = 0; r < kScanTile / 256; r++) {
        int32_t j = st.off + r * 256 + threadIdx.x;
        if (j >= lim) break;
        const float *v = srt[c.first + j].pos + 3;
        for (int k = 0; k < 3; k++) {
            double x = (double)v[k];
            m[k] += x;
            m[3 + k] += x * x;
        }
    }
    const int w = threadIdx.x / kWave;
    for (int k = 0; k < 6; k++) {
        double s = wave_sum(m[k]);
        if (lane_id() == 0) s_m[k][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = s_m[threadIdx.x][0] + s_m[threadIdx.x][1];
        s = s + s_m[threadIdx.x][2];
        s = s + s_m[threadIdx.x][3];
        if (threadIdx.x < 3) tsum[t].s1[threadIdx.x] = s;
        else tsum[t].s2[threadIdx.x - 3] = s;
    }
}

// per call: exclusive prefix of the tile sums, in tile order
__global__ void k_vel_prefix(const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles, int32_t nh,
                             Moments *__restrict__ tsum) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double a[6] = {0, 0, 0, 0, 0, 0};
    const int32_t t0 = call_tile0[h], nt = call_ntiles[h];
    for (int32_t t = t0; t < t0 + nt; t++) {
        Moments m = tsum[t];
        double b[6] = {m.s1[0], m.s1[1], m.s1[2], m.s2[0], m.s2[1], m.s2[2]};
        for (int k = 0; k < 3; k++) {
            tsum[t].s1[k] = a[k];
            tsum[t].s2[k] = a[3 + k];
        }
        for (int k = 0; k < 6; k++) a[k] += b[k];
    }
}

// Candidate of the running-minimum selection: value and index.
struct Cand {
    double val;
    int32_t idx;
    int32_t pad;
};

__device__ __forceinline__ Cand better(Cand a, Cand b) {
    // smaller value wins; ties go to the smaller index (the reference keeps a
    // prefix only when it is strictly better)
    if (b.val < a.val || (b.val == a.val && b.idx < a.idx)) return b;
    return a;
}

__device__ __forceinline__ Cand wave_better(Cand c) {
    for (int o = 32; o > 0; o >>= 1) {
        Cand d;
        d.val = __shfl_xor(c.val, o, 64);
        d.idx = __shfl_xor(c.idx, o, 64);
        d.pad = 0;
        c = better(c, d);
    }
    return c;
}

// per tile: inclusive prefix moments per record -> var_i / ((i - 3) i), and
// the tile's best candidates for the core window and the bulk window
__global__ __launch_bounds__(256) void k_vel_select(const struct potential *__restrict__ srt,
                                                    const ScanTile *__restrict__ tiles,
                                                    int32_t ntiles,
                                                    const CallInfo *__restrict__ ci,
                                                    const Moments *__restrict__ tsum,
                                                    Cand *__restrict__ tcand) {
    __shared__ double s_run[6];
    __shared__ double s_w[6][4];
    __shared__ Cand s_c[2][4];
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const CallInfo c = ci[st.call];
    const int32_t lim = c.rvir_max;
    const int32_t core_lim = min(c.core_max, lim);
    if (threadIdx.x < 6)
        s_run[threadIdx.x] = threadIdx.x < 3 ? tsum[t].s1[threadIdx.x] : tsum[t].s2[threadIdx.x - 3];
    Cand bc, bb;
    bc.val = bb.val = __longlong_as_double(0x7ff0000000000000ll);
    bc.idx = bb.idx = 0x7fffffff;
    bc.pad = bb.pad = 0;
    const int w = threadIdx.x / kWave;
    __syncthreads();
    for (int r = 0; r < kScanTile / 256; r++) {
        const int32_t j = st.off + r * 256 + threadIdx.x;
        const bool v = j < lim;
        double x[6] = {0, 0, 0, 0, 0, 0};
        if (v) {
            const float *vel = srt[c.first + j].pos + 3;
            for (int k = 0; k < 3; k++) {
                x[k] = (double)vel[k];
                x[3 + k] = x[k] * x[k];
            }
        }
        // inclusive block scan of the six moments
        double y[6];
        for (int k = 0; k < 6; k++) {
            double s = x[k];
            for (int o = 1; o < 64; o <<= 1) {
                double u = __shfl_up(s, o, 64);
                if (lane_id() >= o) s += u;
            }
            y[k] = s;
            if (lane_id() == 63) s_w[k][w] = s;
        }
        __syncthreads();
        double tot[6];
        for (int k = 0; k < 6; k++) {
            double off = s_run[k];
            for (int q = 0; q < w; q++) off += s_w[k][q];
            y[k] += off;
            tot[k] = s_run[k] + s_w[k][0] + s_w[k][1] + s_w[k][2] + s_w[k][3];
        }
        if (v && j >= 9) {
            const double n = (double)(j + 1);
            double var = 0.0;
            for (int k = 0; k < 3; k++) var += y[3 + k] - y[k] * y[k] / n;
            Cand me;
            me.val = j > 3 ? var / (double)((int64_t)(j - 3) * j) : 0.0;
            me.idx = j;
            me.pad = 0;
            if (j < core_lim) bc = better(bc, me);
            bb = better(bb, me);
        }
        __syncthreads();
        if (threadIdx.x < 6) s_run[threadIdx.x] = tot[threadIdx.x];
        __syncthreads();
    }
    bc = wave_better(bc);
    bb = wave_better(bb);
    if (lane_id() == 0) {
        s_c[0][w] = bc;
        s_c[1][w] = bb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Cand a = s_c[0][0], b = s_c[1][0];
        for (int q = 1; q < 4; q++) {
            a = better(a, s_c[0][q]);
            b = better(b, s_c[1][q]);
        }
        tcand[2 * t] = a;
        tcand[2 * t + 1] = b;
    }
}

// prefix mean of velocities at index j (sum over [0, j])
__device__ void prefix_mean(const struct potential *__restrict__ srt, int32_t first, int32_t j,
                            const Moments *__restrict__ tsum, int32_t tile0, float *out) {
    const int32_t t = tile0 + j / kScanTile;
    double s[3] = {tsum[t].s1[0], tsum[t].s1[1], tsum[t].s1[2]};
    for (int32_t q = (j / kScanTile) * kScanTile; q <= j; q++) {
        const float *v = srt[first + q].pos + 3;
        for (int k = 0; k < 3; k++) s[k] += (double)v[k];
    }
    for (int k = 0; k < 3; k++) out[k] = (float)(s[k] / (double)(j + 1));
}

__device__ double prefix_val(const struct potential *__restrict__ srt, int32_t first, int32_t j,
                             const Moments *__restrict__ tsum, int32_t tile0) {
    if (j <= 3) return 0.0;
    const int32_t t = tile0 + j / kScanTile;
    double s[6] = {tsum[t].s1[0], tsum[t].s1[1], tsum[t].s1[2],
                   tsum[t].s2[0], tsum[t].s2[1], tsum[t].s2[2]};
    for (int32_t q = (j / kScanTile) * kScanTile; q <= j; q++) {
        const float *v = srt[first + q].pos + 3;
        for (int k = 0; k < 3; k++) {
            double x = (double)v[k];
            s[k] += x;
            s[3 + k] += x * x;
        }
    }
    const double n = (double)(j + 1);
    double var = 0.0;
    for (int k = 0; k < 3; k++) var += s[3 + k] - s[k] * s[k] / n;
    return var / (double)((int64_t)(j - 3) * j);
}

// one thread per call: final selection, halo fields
__global__ void k_vel_finish(const struct potential *__restrict__ srt,
                             const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles, int32_t nh,
                             const CallInfo *__restrict__ ci, const Moments *__restrict__ tsum,
                             const Cand *__restrict__ tcand, hp_halo *__restrict__ halos) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const CallInfo c = ci[h];
    hp_halo hh = halos[h];
    hh.kept = c.kept;
    hh.status = 0;
    if (c.rvir_max >= 1) {
        const int32_t lim = c.rvir_max;
        const int32_t core_lim = min(c.core_max, lim);
        const int32_t t0 = call_tile0[h], nt = call_ntiles[h];
        Cand a, b;
        a.val = b.val = __longlong_as_double(0x7ff0000000000000ll);
        a.idx = b.idx = 0x7fffffff;
        a.pad = b.pad = 0;
        for (int32_t t = t0; t < t0 + nt; t++) {
            a = better(a, tcand[2 * t]);
            b = better(b, tcand[2 * t + 1]);
        }
        // windows shorter than 10: the last forced update wins
        const int32_t ic = core_lim <= 10 ? core_lim - 1 : a.idx;
        const int32_t ib = lim <= 10 ? lim - 1 : b.idx;
        float vc[3], vb[3];
        prefix_mean(srt, c.first, ic, tsum, t0, vc);
        prefix_mean(srt, c.first, ib, tsum, t0, vb);
        hh.n_core = ic;
        hh.min_vel_err = (float)prefix_val(srt, c.first, ic, tsum, t0);
        hh.min_bulkvel_err = (float)prefix_val(srt, c.first, ib, tsum, t0);
        for (int k = 0; k < 3; k++) {
            hh.corevel[k] = vc[k];
            hh.bulkvel[k] = vb[k];
        }
        for (int k = 0; k < 3; k++) hh.pos[3 + k] = hh.corevel[k];  // src/properties.c:213
    }
    halos[h] = hh;
}

// ------------------------------------------------------- kinetic energy

// src/potential.c:194-211 on the sorted kept records of each call; bound flag
// !(pe < ke) (src/properties.c:403).  Potential-only records keep ke < 0.
__global__ void k_kinetic_bound(struct potential *__restrict__ srt,
                                const int32_t *__restrict__ rec_call, int32_t nrec,
                                const hp_halo *__restrict__ halos, double hub_a, double conv,
                                uint8_t *__restrict__ bound, int32_t *__restrict__ nbound) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec) return;
    struct potential r = srt[i];
    const int32_t h = rec_call[i];
    if (r.ke < 0.0f) {
        if (bound) bound[i] = 0;
        return;
    }
    const hp_halo &hh = halos[h];
    const float *vc = hh.is_sub ? hh.corevel : hh.bulkvel;
    float ke = 0.0f;
    for (int k = 0; k < 3; k++) {
        const float dvp = r.pos[k + 3] - vc[k];
        const float dxp = r.pos[k] - hh.pos[k];
        const double dv = (double)dvp + hub_a * (double)dxp;
        ke = (float)((double)ke + dv * dv);
    }
    ke = (float)((double)ke * conv);
    r.ke = ke;
    srt[i].ke = ke;
    const bool b = !(r.pe < (double)ke);
    if (bound) bound[i] = b ? 1 : 0;
    if (b) atomicAdd(&nbound[h], 1);
}

// the plain reference entry point: every record with ke >= 0, one centre
__global__ void k_kinetic_plain(struct potential *__restrict__ po, int64_t n, float vc0, float vc1,
                                float vc2, float pc0, float pc1, float pc2, double hub_a,
                                double conv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (po[i].ke < 0.0f) return;
    const float vc[3] = {vc0, vc1, vc2}, pc[3] = {pc0, pc1, pc2};
    float ke = 0.0f;
    for (int k = 0; k < 3; k++) {
        const float dvp = po[i].pos[k + 3] - vc[k];
        const float dxp = po[i].pos[k] - pc[k];
        const double dv = (double)dvp + hub_a * (double)dxp;
        ke = (float)((double)ke + dv * dv);
    }
    po[i].ke = (float)((double)ke * conv);
}

__global__ void k_write_nbound(const int32_t *__restrict__ nbound, int32_t nh,
                               hp_halo *__restrict__ halos) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h < nh) halos[h].n_bound = nbound[h];
}

}  // namespace hp
