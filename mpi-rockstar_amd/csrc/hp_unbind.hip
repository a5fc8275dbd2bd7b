// hp_unbind.hip -- the bound pass after the potential
// (src/properties.c:539-550 and the bound test of src/properties.c:403),
// for every call of a batch at once:
//   1. keys: (potential-only flag, r2 bits), sorted inside each call (the
//      calls tile the record array in order, so each call's sorted records
//      stay in its own range): every call's records by r2 with the
//      potential-only records (ke < 0) last; ties keep the order the
//      reference's compute_potential and potential-only strip leave (see
//      "sort input" below), which is what glibc's merge-sort qsort does with
//      the reference's dist_compare (src/groupies.c:202-210).  r2 >= 0, so
//      its bits sort as unsigned.  Calls of at most 64 records are ranked by one wavefront,
//      calls of at most 1,024 sorted by one workgroup (bitonic, in LDS),
//      larger ones by a stable LSD
//      radix sort (3 passes of 11-bit digits) over tiles that never straddle
//      two calls.
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
constexpr int kRadixBits = 10;              // digit width: 5 passes cover 50 key bits
constexpr int kRadix = 1 << kRadixBits;
constexpr int kRadixPer = kRadix / kSortThreads;  // buckets per thread
constexpr int kScanTile = 2048;

// kScanTile consecutive records of one call, from record `off` of the call
struct ScanTile {
    int32_t call;
    int32_t off;
};

// Each record's call, filled once per plan (per-record kernels read it).
__global__ void k_rec_calls(const int64_t *__restrict__ hfirst, const int64_t *__restrict__ hcount,
                            int32_t nh, int32_t *__restrict__ rec_call) {
    const int32_t h = blockIdx.x;
    if (h >= nh) return;
    for (int64_t j = threadIdx.x; j < hcount[h]; j += blockDim.x) rec_call[hfirst[h] + j] = h;
}

// ---------------------------------------------------------------- radix sort


// digit histogram of one tile: hist[digit * ntiles + tile]
__global__ __launch_bounds__(kSortThreads) void k_radix_hist(const uint64_t *__restrict__ keys,
                                                             int32_t n, int shift,
                                                             int32_t ntiles,
                                                             int32_t *__restrict__ hist) {
    __shared__ int32_t s_h[kRadix];
    for (int q = 0; q < kRadixPer; q++) s_h[q * kSortThreads + threadIdx.x] = 0;
    __syncthreads();
    const int32_t base = blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        int32_t i = base + r * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & (kRadix - 1)], 1);
    }
    __syncthreads();
    for (int q = 0; q < kRadixPer; q++) {
        const int d = q * kSortThreads + threadIdx.x;
        hist[d * ntiles + blockIdx.x] = s_h[d];
    }
}

// Exclusive scan of hist (length m) in place, in three launches: per block
// of 4096 (1024 threads x 4), the block totals, then the offsets.
constexpr int kScanBlock = 4096;

__device__ __forceinline__ int32_t block_excl_scan_1024(int32_t x, int32_t *total) {
    __shared__ int32_t s_w[16];
    int32_t y = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t t = __shfl_up(y, o, 64);
        if (lane_id() >= o) y += t;
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 63) s_w[w] = y;
    __syncthreads();
    int32_t off = 0, tot = 0;
    for (int j = 0; j < 16; j++) {
        if (j < w) off += s_w[j];
        tot += s_w[j];
    }
    __syncthreads();
    *total = tot;
    return off + y - x;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(int32_t *__restrict__ v, int32_t m,
                                                      int32_t *__restrict__ bsum) {
    const int32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    int32_t x[4], s = 0;
    for (int k = 0; k < 4; k++) {
        x[k] = base + k < m ? v[base + k] : 0;
        s += x[k];
    }
    int32_t tot;
    int32_t off = block_excl_scan_1024(s, &tot);
    for (int k = 0; k < 4; k++) {
        if (base + k < m) v[base + k] = off;
        off += x[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_top(int32_t *__restrict__ bsum, int32_t nb) {
    __shared__ int32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int32_t base = 0; base < nb; base += 1024) {
        const int32_t i = base + threadIdx.x;
        const int32_t x = i < nb ? bsum[i] : 0;
        int32_t tot;
        const int32_t ex = block_excl_scan_1024(x, &tot);
        if (i < nb) bsum[i] = s_carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_scan_add(int32_t *__restrict__ v, int32_t m,
                                                   const int32_t *__restrict__ bsum) {
    const int32_t add = bsum[blockIdx.x];
    const int32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4;
    for (int k = 0; k < 4; k++)
        if (base + k < m) v[base + k] += add;
}

// The wavefront subtrees' node ids (k_wave_need, the scan and k_wave_base)
// in one launch, for a list of at most kScanBlock entries (a small batch or
// one call): off[k] = the ids of the tasks before k.
__global__ __launch_bounds__(1024) void k_wave_ids_one(const Task *__restrict__ tasks, BuildArgs a,
                                                       int32_t *__restrict__ off, int32_t m) {
    const int32_t ntasks = min(a.counters[kSmallNext], a.cap_small);
    const int32_t base = threadIdx.x * 4;
    int32_t x[4], s = 0;
    for (int k = 0; k < 4; k++) {
        const int32_t i = base + k;
        x[k] = i < m && i < ntasks ? wave_ids(tasks[i].pad[1]) : 0;
        s += x[k];
    }
    int32_t tot;
    int32_t o = block_excl_scan_1024(s, &tot);
    for (int k = 0; k < 4; k++) {
        if (base + k < m) off[base + k] = o;
        o += x[k];
    }
    if (threadIdx.x == 0) {
        const int32_t base0 = a.counters[kNodeCount];
        a.counters[kWaveIdBase] = base0;
        a.counters[kNodeCount] = base0 + tot;
    }
}

// stable scatter of one tile: its keys are loaded up front, then scattered
// in rounds of 256 consecutive keys, wave-level multisplit by digit, running
// per-digit bases across rounds
__global__ __launch_bounds__(kSortThreads) void k_radix_scatter(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin,
    uint64_t *__restrict__ kout, int32_t *__restrict__ vout, int32_t n, int shift,
    int32_t ntiles, const int32_t *__restrict__ offs) {
    __shared__ int32_t s_base[kRadix];
    __shared__ int32_t s_wc[kSortThreads / kWave][kRadix];
    const int w = threadIdx.x / kWave;
    const int32_t tile0 = blockIdx.x * kSortTile;
    uint64_t key[kSortItems];
    int32_t val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int32_t i = tile0 + r * kSortThreads + threadIdx.x;
        key[r] = i < n ? kin[i] : 0ull;
        val[r] = i < n ? vin[i] : 0;
    }
    for (int q = 0; q < kRadixPer; q++) {
        const int d = q * kSortThreads + threadIdx.x;
        s_base[d] = offs[d * ntiles + blockIdx.x];
        for (int j = 0; j < kSortThreads / kWave; j++) s_wc[j][d] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int32_t i = tile0 + r * kSortThreads + threadIdx.x;
        if (tile0 + r * kSortThreads >= n) break;  // block-uniform
        const bool v = i < n;
        const int dg = (int)((key[r] >> shift) & (kRadix - 1));
        // lanes of this wave with the same digit
        uint64_t peers = ballot(v);
        for (int b = 0; b < kRadixBits; b++) {
            const uint64_t m = ballot(((dg >> b) & 1) != 0);
            peers &= ((dg >> b) & 1) ? m : ~m;
        }
        const int rk = rank_below(peers);
        const bool lead = v && rk == 0;  // publishes the wave's count of its digit
        if (lead) s_wc[w][dg] = __popcll(peers);
        __syncthreads();
        if (v) {
            int32_t off = s_base[dg];
            for (int j = 0; j < w; j++) off += s_wc[j][dg];
            kout[off + rk] = key[r];
            vout[off + rk] = val[r];
        }
        __syncthreads();
        // the next round starts after this round's keys of each digit
        if (lead) {
            atomicAdd(&s_base[dg], s_wc[w][dg]);
            s_wc[w][dg] = 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------ sort inside calls

// A call of at most 64 records, one wavefront: record i's place is the
// number of records with a smaller (key, position); the value carried is
// the record index at that position (vin).
__global__ __launch_bounds__(256) void k_seg_wave(const uint64_t *__restrict__ kin,
                                                  const int32_t *__restrict__ vin,
                                                  uint64_t *__restrict__ kout,
                                                  int32_t *__restrict__ vout,
                                                  const int32_t *__restrict__ calls,
                                                  int32_t ncalls, const int64_t *__restrict__ hfirst,
                                                  const int64_t *__restrict__ hcount) {
    const int32_t wv = (int32_t)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    if (wv >= ncalls) return;  // whole wavefronts
    const int lane = lane_id();
    const int32_t h = calls[wv];
    const int32_t first = (int32_t)hfirst[h], n = (int32_t)hcount[h];
    const bool v = lane < n;
    const uint64_t key = v ? kin[first + lane] : 0ull;
    const uint64_t me = (key << 6) | (uint64_t)lane;  // 33-bit key, unique with the position
    int32_t rank = 0;
    for (int j = 0; j < n; j++) rank += ((read_lane(key, j) << 6) | (uint64_t)j) < me ? 1 : 0;
    if (v) {
        kout[first + rank] = key;
        vout[first + rank] = vin[first + lane];
    }
}

// A call of 65..1024 records, one workgroup: a bitonic sort in LDS of the
// (key, position) pairs, unique, so the order is the stable one; padded to
// a power of two with keys above every record's.
constexpr int kSegBlockMax = 1024;
__global__ __launch_bounds__(kSortThreads) void k_seg_block(const uint64_t *__restrict__ kin,
                                                            const int32_t *__restrict__ vin,
                                                            uint64_t *__restrict__ kout,
                                                            int32_t *__restrict__ vout,
                                                            const int32_t *__restrict__ calls,
                                                            const int64_t *__restrict__ hfirst,
                                                            const int64_t *__restrict__ hcount) {
    __shared__ uint64_t s_k[kSegBlockMax];
    const int32_t h = calls[blockIdx.x];
    const int32_t first = (int32_t)hfirst[h], n = (int32_t)hcount[h];
    int32_t np2 = kWave;
    while (np2 < n) np2 <<= 1;  // n <= kSegBlockMax, so np2 <= kSegBlockMax
    for (int32_t i = threadIdx.x; i < np2; i += kSortThreads)
        s_k[i] = i < n ? (kin[first + i] << 10) | (uint64_t)i : ~0ull;
    __syncthreads();
    for (int32_t k = 2; k <= np2; k <<= 1) {
        for (int32_t j = k >> 1; j > 0; j >>= 1) {
            for (int32_t i = threadIdx.x; i < np2; i += kSortThreads) {
                const int32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = s_k[i], b = s_k[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        s_k[i] = b;
                        s_k[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int32_t i = threadIdx.x; i < n; i += kSortThreads) {
        const uint64_t kk = s_k[i];
        kout[first + i] = kk >> 10;
        vout[first + i] = vin[first + (int32_t)(kk & 1023)];
    }
}

// Calls of more than 1,024 records: tiles of up to kSortTile records that
// never straddle two calls.  The digit counts of call c's tiles are stored
// digit-major in their own block of the count array, entry (d, t) at
// hpos(t) + d * ntiles(c), so one exclusive scan over the whole array gives
// each (tile, digit) its place among the large calls' records; delta moves
// that place into the call's own range.
constexpr int kSegBits = 11;  // 3 passes cover the 33 key bits
constexpr int kSegRadix = 1 << kSegBits;
constexpr int kSegRadixPer = kSegRadix / kSortThreads;
struct SegTile {
    int32_t off;    // first record of the tile
    int32_t n;      // records in the tile
    int32_t hpos;   // count-array entry of digit 0
    int32_t ntc;    // tiles of the call
    int32_t delta;  // call's first record minus the large-call records before it
    int32_t pad[3];
};

__global__ __launch_bounds__(kSortThreads) void k_seg_hist(const uint64_t *__restrict__ keys,
                                                           const SegTile *__restrict__ tiles,
                                                           int shift, int32_t *__restrict__ hist) {
    __shared__ int32_t s_h[kSegRadix];
    const SegTile st = tiles[blockIdx.x];
    for (int q = 0; q < kSegRadixPer; q++) s_h[q * kSortThreads + threadIdx.x] = 0;
    __syncthreads();
    for (int r = 0; r < kSortItems; r++) {
        const int32_t i = r * kSortThreads + threadIdx.x;
        if (i < st.n) atomicAdd(&s_h[(keys[st.off + i] >> shift) & (kSegRadix - 1)], 1);
    }
    __syncthreads();
    for (int q = 0; q < kSegRadixPer; q++) {
        const int d = q * kSortThreads + threadIdx.x;
        hist[st.hpos + d * st.ntc] = s_h[d];
    }
}

// Stable scatter of one tile into its call's range.  Wavefront w takes the
// tile's w-th quarter (rounds of 64 consecutive positions); a record's place
// is its digit's base for the tile (the scanned counts), plus the records of
// that digit in the earlier quarters, plus those before it in its own
// quarter.  The ranks come from ballots (the lanes of a round with the same
// digit), counted per wavefront in LDS, so the tile needs two workgroup
// barriers instead of three per round.
constexpr int kSortWaves = kSortThreads / kWave;
__global__ __launch_bounds__(kSortThreads) void k_seg_scatter(
    const uint64_t *__restrict__ kin, const int32_t *__restrict__ vin, uint64_t *__restrict__ kout,
    int32_t *__restrict__ vout, const SegTile *__restrict__ tiles, int shift,
    const int32_t *__restrict__ offs) {
    __shared__ int32_t s_cnt[kSortWaves][kSegRadix];
    const int w = threadIdx.x / kWave, lane = lane_id();
    const SegTile st = tiles[blockIdx.x];
    for (int q = threadIdx.x; q < kSortWaves * kSegRadix; q += kSortThreads) (&s_cnt[0][0])[q] = 0;
    __syncthreads();
    const int32_t q0 = w * kSortItems * kWave;  // this wavefront's quarter
    uint64_t key[kSortItems];
    int32_t val[kSortItems], pos[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        const int32_t i = q0 + r * kWave + lane;
        const bool v = i < st.n;
        key[r] = v ? kin[st.off + i] : 0ull;
        val[r] = v ? vin[st.off + i] : 0;
        const int dg = (int)((key[r] >> shift) & (kSegRadix - 1));
        uint64_t peers = ballot(v);
        for (int b = 0; b < kSegBits; b++) {
            const uint64_t m = ballot(((dg >> b) & 1) != 0);
            peers &= ((dg >> b) & 1) ? m : ~m;
        }
        const int rk = rank_below(peers);
        // the count so far is read by every lane of the digit before its
        // first lane adds the round's records (program order in the wavefront)
        pos[r] = v ? s_cnt[w][dg] + rk : -1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (v && rk == 0) s_cnt[w][dg] += __popcll(peers);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __syncthreads();
    // each digit's place for each quarter: the tile's base, then the quarters in order
    for (int d = threadIdx.x; d < kSegRadix; d += kSortThreads) {
        int32_t run = offs[st.hpos + d * st.ntc] + st.delta;
        for (int q = 0; q < kSortWaves; q++) {
            const int32_t c = s_cnt[q][d];
            s_cnt[q][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; r++) {
        if (pos[r] < 0) continue;
        const int dg = (int)((key[r] >> shift) & (kSegRadix - 1));
        const int32_t dst = s_cnt[w][dg] + pos[r];
        kout[dst] = key[r];
        vout[dst] = val[r];
    }
}

// ------------------------------------------------------------ sort input
//
// The reference sorts each call's records (glibc's qsort: a stable merge
// sort) in the order compute_potential and the potential-only strip leave
// them: the call's tree order (the build reproduces the reference's
// in-place partitions), then the strip of src/properties.c:539-544, which
// overwrites each potential-only record with the current last one.  The
// strip moves the kept records as the in-place partition moves its left
// side (inplace_dest with ns = the kept count):
//   kept record at tree position t < nk: stays at t;
//   kept record at t >= nk, with c kept records after it: goes to the tree
//   position of the (c + 1)-th potential-only record.
// The keys are written at each record's place in that order, so that the
// stable sort breaks ties as the reference does.  Calls without potential-
// only records keep the tree order (k_keys_tree); the others go through
// k_strip_count, k_strip_prefix, k_strip_list and k_strip_keys, which skip
// every other call.  Potential-only records follow the kept ones.
constexpr int kKeyThreads = 256;
constexpr int kKeyPer = kScanTile / kKeyThreads;  // consecutive positions per thread

struct KeyArgs {
    const uint64_t *rkey;       // each record's sort key, in input order
    const float4 *fin;          // tree order (record index in w); null: the input order
    const ScanTile *tiles;
    const int64_t *hfirst, *hcount;
    const int32_t *pot;         // potential-only records of each call; pot[ncalls] > 0 if any
    int32_t ncalls;
    const int32_t *call_tile0, *call_ntiles;
    int32_t *tcnt;              // per tile: its potential-only records, then their prefix
    int32_t *bpos;              // per position: tree positions of the potential-only records
    uint64_t *keys;
    int32_t *vals;
};

__device__ __forceinline__ int32_t record_at(const KeyArgs &a, int64_t pos) {
    return a.fin ? __float_as_int(a.fin[pos].w) : (int32_t)pos;
}

// calls without potential-only records: keys in tree order, a thread per
// position
__global__ void k_keys_tree(KeyArgs a, const int32_t *__restrict__ rec_call, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || a.pot[rec_call[i]] != 0) return;
    const int32_t idx = record_at(a, i);
    a.keys[i] = a.rkey[idx];
    a.vals[i] = idx;
}

// Potential-only flags of this thread's kKeyPer consecutive positions of a
// tile, their exclusive rank among the tile's and the tile's total.
struct StripRun {
    int32_t idx[kKeyPer];
    uint32_t potm;   // bit j: position j of the run is potential-only
    int32_t before;  // potential-only positions of the tile before the run
    int32_t total;
};

__device__ StripRun strip_run(const KeyArgs &a, const ScanTile &tl, int64_t first, int64_t n) {
    __shared__ int32_t s_w[kKeyThreads / kWave];
    StripRun r;
    r.potm = 0;
    const int64_t t0 = tl.off + (int64_t)threadIdx.x * kKeyPer;
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) {
        r.idx[j] = -1;
        if (t0 + j < n) {
            r.idx[j] = record_at(a, first + t0 + j);
            if (a.rkey[r.idx[j]] >> 32) r.potm |= 1u << j;
        }
    }
    const int c = __popc(r.potm);
    const int incl = wave_scan_add(c);
    const int w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) s_w[w] = incl;
    __syncthreads();
    int wb = 0, tot = 0;
    for (int k = 0; k < kKeyThreads / kWave; k++) {
        wb += k < w ? s_w[k] : 0;
        tot += s_w[k];
    }
    __syncthreads();
    r.before = wb + incl - c;
    r.total = tot;
    return r;
}

// (the strip kernels take a tile per workgroup, the grid striding over the
// tiles; they return at once when no call has potential-only records)
__global__ __launch_bounds__(kKeyThreads) void k_strip_count(KeyArgs a, int32_t ntiles) {
    if (a.pot[a.ncalls] == 0) return;  // no call has potential-only records
    for (int32_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const ScanTile tl = a.tiles[k];
        if (a.pot[tl.call] == 0) continue;
        const StripRun r = strip_run(a, tl, a.hfirst[tl.call], a.hcount[tl.call]);
        if (threadIdx.x == 0) a.tcnt[k] = r.total;
    }
}

// per call with potential-only records: exclusive prefix of its tiles' counts
__global__ void k_strip_prefix(KeyArgs a) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (a.pot[a.ncalls] == 0 || h >= a.ncalls || a.pot[h] == 0) return;
    int32_t run = 0;
    for (int32_t k = a.call_tile0[h]; k < a.call_tile0[h] + a.call_ntiles[h]; k++) {
        const int32_t c = a.tcnt[k];
        a.tcnt[k] = run;
        run += c;
    }
}

// the tree positions of each call's potential-only records, in order
__global__ __launch_bounds__(kKeyThreads) void k_strip_list(KeyArgs a, int32_t ntiles) {
    if (a.pot[a.ncalls] == 0) return;  // no call has potential-only records
    for (int32_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const ScanTile tl = a.tiles[k];
        if (a.pot[tl.call] == 0) continue;
        const int64_t first = a.hfirst[tl.call];
        const StripRun r = strip_run(a, tl, first, a.hcount[tl.call]);
        int32_t b = a.tcnt[k] + r.before;
        const int64_t t0 = tl.off + (int64_t)threadIdx.x * kKeyPer;
#pragma unroll
        for (int j = 0; j < kKeyPer; j++)
            if (r.potm >> j & 1u) a.bpos[first + b++] = (int32_t)(t0 + j);
    }
}

// keys at each record's place after the strip
__global__ __launch_bounds__(kKeyThreads) void k_strip_keys(KeyArgs a, int32_t ntiles) {
    if (a.pot[a.ncalls] == 0) return;  // no call has potential-only records
    for (int32_t k = blockIdx.x; k < ntiles; k += gridDim.x) {
        const ScanTile tl = a.tiles[k];
        const int32_t npot = a.pot[tl.call];
        if (npot == 0) continue;
        const int64_t first = a.hfirst[tl.call], n = a.hcount[tl.call];
        const int64_t nk = n - npot;
        const StripRun r = strip_run(a, tl, first, n);
        int32_t b = a.tcnt[k] + r.before;  // potential-only positions before t
        const int64_t t0 = tl.off + (int64_t)threadIdx.x * kKeyPer;
#pragma unroll
        for (int j = 0; j < kKeyPer; j++) {
            const int64_t t = t0 + j;
            if (t >= n) break;
            int64_t q;
            if (r.potm >> j & 1u) {
                q = nk + b++;
            } else {
                const int64_t kept_rank = t - b;
                q = t < nk ? t : (int64_t)a.bpos[first + (nk - 1 - kept_rank)];
            }
            a.keys[first + q] = a.rkey[r.idx[j]];
            a.vals[first + q] = r.idx[j];
        }
    }
}

// ----------------------------------------------------------------- gather

// A record moved as three 16-byte words (three loads and three stores
// instead of one per field group).
struct RecWords {
    float4 w[3];
};
static_assert(sizeof(RecWords) == sizeof(struct potential), "48-byte records");
__device__ __forceinline__ struct potential load_rec(const struct potential *p) {
    const float4 *q = reinterpret_cast<const float4 *>(p);
    RecWords u;
    u.w[0] = q[0];
    u.w[1] = q[1];
    u.w[2] = q[2];
    return __builtin_bit_cast(struct potential, u);
}
__device__ __forceinline__ void store_rec(struct potential *p, const struct potential &r) {
    const RecWords u = __builtin_bit_cast(RecWords, r);
    float4 *q = reinterpret_cast<float4 *>(p);
    q[0] = u.w[0];
    q[1] = u.w[1];
    q[2] = u.w[2];
}

// the records in sorted order, and a compact copy of what the core-velocity
// passes read (r2 and the velocity, 16 B instead of the 48-B record)
__global__ __launch_bounds__(256) void k_gather_sorted(const struct potential *__restrict__ po,
                                                       const int32_t *__restrict__ order,
                                                       int32_t nrec,
                                                       struct potential *__restrict__ out,
                                                       float4 *__restrict__ rv) {
    // each workgroup's 256 records go out through LDS as 16-byte words, so
    // that the stores of a wavefront are contiguous (the loads are a gather)
    __shared__ float4 s_rec[3 * 256];
    const int32_t i0 = blockIdx.x * 256;
    const int32_t i = i0 + threadIdx.x;
    const int32_t m = min(256, nrec - i0);
    if (i < nrec) {
        const struct potential rec = load_rec(po + order[i]);
        const RecWords u = __builtin_bit_cast(RecWords, rec);
        for (int q = 0; q < 3; q++) s_rec[3 * threadIdx.x + q] = u.w[q];
        rv[i] = make_float4(rec.r2, rec.pos[3], rec.pos[4], rec.pos[5]);
    }
    __syncthreads();
    float4 *dst = reinterpret_cast<float4 *>(out + i0);
    for (int q = threadIdx.x; q < 3 * m; q += 256) dst[q] = s_rec[q];
}

// ------------------------------------------------------------ core velocity

struct CallInfo {
    int32_t first;      // sorted records of the call: [first, first + kept)
    int32_t kept;
    int32_t count;      // all records of the call (the potential-only ones follow)
    int32_t rvir_max;   // -1: no virial prefix
    int32_t core_max;
};

// kept = records with ke >= 0: they sort before the potential-only ones, so
// it is the first index whose key has the potential-only bit (binary search)
__global__ void k_call_kept(const uint64_t *__restrict__ skeys, const int64_t *__restrict__ hfirst,
                            const int64_t *__restrict__ hcount, int32_t nh,
                            CallInfo *__restrict__ ci) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const int32_t first = (int32_t)hfirst[h];
    int32_t lo = 0, hi = (int32_t)hcount[h];
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if ((skeys[first + mid] >> 32) & 1ull) hi = mid;
        else lo = mid + 1;
    }
    CallInfo c;
    c.first = first;
    c.kept = lo;
    c.count = (int32_t)hcount[h];
    c.rvir_max = -1;
    c.core_max = 100;
    ci[h] = c;
}

// Scan tiles: tile t covers sorted records [first + off, first + off + 2048)
// of one call; every kernel below clips it to the call's live prefix.

__device__ __forceinline__ int32_t block_max_i32(int32_t v) {
    __shared__ int32_t s_m[4];
    v = wave_max(v);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) s_m[w] = v;
    __syncthreads();
    int32_t m = s_m[0];
    for (int q = 1; q < 4; q++) m = s_m[q] > m ? s_m[q] : m;
    __syncthreads();
    return m;
}

__device__ __forceinline__ int32_t block_sum_i32(int32_t v) {
    __shared__ int32_t s_s[4];
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) s_s[w] = v;
    __syncthreads();
    int32_t m = s_s[0] + s_s[1] + s_s[2] + s_s[3];
    __syncthreads();
    return m;
}

// rvir_max (src/properties.c:173-177): the largest j < kept with
// j^2 > r2_j^3 (4pi/3 rho_vir)^2, r2^3 in float, the rest in double.
// Per tile (k_rvir_vel_tiles, below), then per call.
__global__ void k_rvir_calls(const float4 *__restrict__ rv,
                             const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles,
                             const int32_t *__restrict__ tile_rvir, int32_t nh,
                             CallInfo *__restrict__ ci) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    CallInfo c = ci[h];
    int32_t rvm = -1;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++)
        rvm = tile_rvir[t] > rvm ? tile_rvir[t] : rvm;
    c.rvir_max = rvm;
    if (rvm >= 1) {
        const float ref = rv[c.first + rvm].x;
        int32_t lo = 0, hi = c.kept;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if ((double)rv[c.first + mid].x * 100.0 < (double)ref) lo = mid + 1;
            else hi = mid;
        }
        const int32_t cm = lo - 1;
        c.core_max = cm < 100 ? 100 : cm;
    }
    ci[h] = c;
}

struct Moments {
    double s1[3];
    double s2[3];
};

// One pass over each tile of the kept records (one wavefront per tile): the
// last record inside the virial density threshold (rvir_max above) and the
// tile's sums of v and v^2.  The sums cover the whole tile; only the tiles
// wholly below rvir_max feed the core-velocity prefixes, so the records
// past rvir_max in its own tile do not matter.
__global__ __launch_bounds__(kWave) void k_rvir_vel_tiles(const float4 *__restrict__ rv,
                                                          const ScanTile *__restrict__ tiles,
                                                          int32_t ntiles,
                                                          const CallInfo *__restrict__ ci,
                                                          double rvir_thresh,
                                                          int32_t *__restrict__ tile_rvir,
                                                          Moments *__restrict__ tsum) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const CallInfo c = ci[st.call];
    const int32_t end = min(st.off + kScanTile, c.kept);
    int32_t best = -1;
    double m[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t j = st.off + lane_id(); j < end; j += kWave) {
        const float4 q = rv[c.first + j];
        const float r6 = (q.x * q.x) * q.x;
        if ((double)((int64_t)j * (int64_t)j) > (double)r6 * (rvir_thresh * rvir_thresh))
            best = j > best ? j : best;
        const float v[3] = {q.y, q.z, q.w};
        for (int k = 0; k < 3; k++) {
            const double x = (double)v[k];
            m[k] += x;
            m[3 + k] += x * x;
        }
    }
    best = wave_max(best);
    for (int k = 0; k < 6; k++) m[k] = wave_sum(m[k]);
    if (lane_id() == 0) {
        tile_rvir[t] = best;
        for (int k = 0; k < 3; k++) {
            tsum[t].s1[k] = m[k];
            tsum[t].s2[k] = m[3 + k];
        }
    }
}

__global__ void k_vel_prefix(const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles, int32_t nh,
                             Moments *__restrict__ tsum) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    double a[6] = {0, 0, 0, 0, 0, 0};
    const int32_t t0 = call_tile0[h], nt = call_ntiles[h];
    for (int32_t t = t0; t < t0 + nt; t++) {
        const Moments m = tsum[t];
        const double b[6] = {m.s1[0], m.s1[1], m.s1[2], m.s2[0], m.s2[1], m.s2[2]};
        for (int k = 0; k < 3; k++) {
            tsum[t].s1[k] = a[k];
            tsum[t].s2[k] = a[3 + k];
        }
        for (int k = 0; k < 6; k++) a[k] += b[k];
    }
}

// Candidate prefix of the running-minimum selection: its normalised
// variance, index, and mean velocity.
struct Cand {
    double val;
    int32_t idx;
    double s1[3];  // the prefix's velocity sums: the mean is s1 / (idx + 1), taken at the end
};

__device__ __forceinline__ Cand no_cand() {
    Cand c;
    c.val = __longlong_as_double(0x7ff0000000000000ll);
    c.idx = 0x7fffffff;
    c.s1[0] = c.s1[1] = c.s1[2] = 0.0;
    return c;
}

__device__ __forceinline__ Cand better(const Cand &a, const Cand &b) {
    // smaller value wins; ties go to the smaller index (the reference keeps a
    // prefix only when it is strictly better)
    return (b.val < a.val || (b.val == a.val && b.idx < a.idx)) ? b : a;
}

__device__ __forceinline__ Cand wave_better(Cand c) {
    for (int o = 32; o > 0; o >>= 1) {
        Cand d;
        d.val = __shfl_xor(c.val, o, 64);
        d.idx = __shfl_xor(c.idx, o, 64);
        d.s1[0] = __shfl_xor(c.s1[0], o, 64);
        d.s1[1] = __shfl_xor(c.s1[1], o, 64);
        d.s1[2] = __shfl_xor(c.s1[2], o, 64);
        c = better(c, d);
    }
    return c;
}

// per tile: inclusive prefix moments per record -> var_i / ((i - 3) i), and
// the tile's best candidates for the core window and the bulk window.  One
// wavefront per tile, rounds of 64 records carried lane 63 to the next.
// (6 wavefronts per SIMD asked of the compiler, 80 VGPRs: the kernel is a
// latency chain of f64 scans and divisions per round)
__global__ __launch_bounds__(kWave, 6) void k_vel_select(const float4 *__restrict__ rv,
                                                      const ScanTile *__restrict__ tiles,
                                                      int32_t ntiles,
                                                      const CallInfo *__restrict__ ci,
                                                      const Moments *__restrict__ tsum,
                                                      Cand *__restrict__ tcand) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const CallInfo c = ci[st.call];
    const int32_t lim = c.rvir_max;
    const int32_t core_lim = min(c.core_max, lim);
    const int32_t end = min(st.off + kScanTile, lim);
    const int lane = lane_id();
    double run[6];
    for (int k = 0; k < 3; k++) {
        run[k] = tsum[t].s1[k];
        run[3 + k] = tsum[t].s2[k];
    }
    Cand bc = no_cand(), bb = no_cand();
    for (int32_t base = st.off; base < end; base += kWave) {
        const int32_t j = base + lane;
        const bool v = j < end;
        double x[6] = {0, 0, 0, 0, 0, 0};
        if (v) {
            const float4 q = rv[c.first + j];
            const float vel[3] = {q.y, q.z, q.w};
            for (int k = 0; k < 3; k++) {
                x[k] = (double)vel[k];
                x[3 + k] = x[k] * x[k];
            }
        }
        double y[6];
        for (int k = 0; k < 6; k++) {
            y[k] = wave_scan_add(x[k]) + run[k];
            run[k] = read_lane(y[k], kWave - 1);
        }
        if (v && j >= 9) {
            const double n = (double)(j + 1);
            double var = 0.0;
            for (int k = 0; k < 3; k++) var += y[3 + k] - y[k] * y[k] / n;
            Cand me;
            me.val = j > 3 ? var / (double)((int64_t)(j - 3) * j) : 0.0;
            me.idx = j;
            for (int k = 0; k < 3; k++) me.s1[k] = y[k];
            if (j < core_lim) bc = better(bc, me);
            bb = better(bb, me);
        }
    }
    bc = wave_better(bc);
    bb = wave_better(bb);
    if (lane == 0) {
        tcand[2 * t] = bc;
        tcand[2 * t + 1] = bb;
    }
}

// sum and normalised variance of the first j + 1 records, j < 10, directly
__device__ __forceinline__ Cand small_prefix(const float4 *__restrict__ rv, int32_t first, int32_t j) {
    double s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (int32_t i = 0; i <= j; i++) {
        const float4 w4 = rv[first + i];
        const float v[3] = {w4.y, w4.z, w4.w};
        for (int k = 0; k < 3; k++) {
            s[k] += (double)v[k];
            q[k] += (double)v[k] * (double)v[k];
        }
    }
    const double n = (double)(j + 1);
    double var = 0.0;
    Cand c;
    for (int k = 0; k < 3; k++) {
        c.s1[k] = s[k];
        var += q[k] - s[k] * s[k] / n;
    }
    c.val = j > 3 ? var / (double)((int64_t)(j - 3) * j) : 0.0;
    c.idx = j;
    return c;
}

// one thread per call: final selection, halo fields
__global__ void k_vel_finish(const float4 *__restrict__ rv,
                             const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles, int32_t nh,
                             const CallInfo *__restrict__ ci, const Cand *__restrict__ tcand,
                             hp_halo *__restrict__ halos, const int32_t *__restrict__ flags) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh || (*flags & kErrDeep)) return;
    const CallInfo c = ci[h];
    // the fields are written one by one (a copy of the whole record would
    // sit in scratch)
    hp_halo *hh = halos + h;
    hh->kept = c.kept;
    hh->status = 0;
    if (c.rvir_max >= 1) {
        const int32_t lim = c.rvir_max;
        const int32_t core_lim = min(c.core_max, lim);
        // the best of the tiles' candidates, kept in scalars (a select of
        // whole Cand records would go through scratch memory)
        Cand a = no_cand(), b = no_cand();
        for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) {
            const Cand ca = tcand[2 * t], cb = tcand[2 * t + 1];
            const bool ta = ca.val < a.val || (ca.val == a.val && ca.idx < a.idx);
            const bool tb = cb.val < b.val || (cb.val == b.val && cb.idx < b.idx);
            a.val = ta ? ca.val : a.val;
            a.idx = ta ? ca.idx : a.idx;
            b.val = tb ? cb.val : b.val;
            b.idx = tb ? cb.idx : b.idx;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                a.s1[k] = ta ? ca.s1[k] : a.s1[k];
                b.s1[k] = tb ? cb.s1[k] : b.s1[k];
            }
        }
        // windows of at most 10 records: the last forced update wins
        if (core_lim <= 10) a = small_prefix(rv, c.first, core_lim - 1);
        if (lim <= 10) b = small_prefix(rv, c.first, lim - 1);
        hh->n_core = a.idx;
        hh->min_vel_err = (float)a.val;
        hh->min_bulkvel_err = (float)b.val;
        const double na = (double)a.idx + 1.0, nb = (double)b.idx + 1.0;
        for (int k = 0; k < 3; k++) {
            const float cv = (float)(a.s1[k] / na);
            hh->corevel[k] = cv;
            hh->bulkvel[k] = (float)(b.s1[k] / nb);
            hh->pos[3 + k] = cv;  // src/properties.c:213
        }
    }
}

// ------------------------------------------------------- kinetic energy

// src/potential.c:194-211 on the sorted kept records of each call, about
// corevel (sub-halos) or bulkvel (hosts); bound flag !(pe < ke)
// (src/properties.c:403).  One wavefront per scan tile; bound counts per
// tile.  The records are written to their final place here, the potential-
// only ones unchanged.
// one tile of kScanTile records per workgroup of kKinThreads (several
// wavefronts keep more loads in flight than one)
constexpr int kKinThreads = 256;
__global__ __launch_bounds__(kKinThreads) void k_kinetic_tiles(const struct potential *__restrict__ srt,
                                                         struct potential *__restrict__ out,
                                                         const ScanTile *__restrict__ tiles,
                                                         int32_t ntiles,
                                                         const CallInfo *__restrict__ ci,
                                                         const hp_halo *__restrict__ halos,
                                                         double hub_a, double conv,
                                                         uint8_t *__restrict__ bound,
                                                         int32_t *__restrict__ tile_nb,
                                                         const int32_t *__restrict__ flags) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles || (*flags & kErrDeep)) return;
    const ScanTile st = tiles[t];
    const CallInfo c = ci[st.call];
    const hp_halo &hh = halos[st.call];
    const float *vc = hh.is_sub ? hh.corevel : hh.bulkvel;
    const float cen_v[3] = {vc[0], vc[1], vc[2]};
    const float cen_x[3] = {hh.pos[0], hh.pos[1], hh.pos[2]};
    const int32_t end = min(st.off + kScanTile, c.count);
    int32_t nb = 0;
    // chunks of kKinThreads records moved through LDS as 16-byte words, so
    // that the loads and stores of a wavefront are contiguous
    __shared__ float4 s_rec[3 * kKinThreads];
    for (int32_t j0 = st.off; j0 < end; j0 += kKinThreads) {
        const int32_t m = min(kKinThreads, end - j0);
        const float4 *src = reinterpret_cast<const float4 *>(srt + c.first + j0);
        for (int q = threadIdx.x; q < 3 * m; q += kKinThreads) s_rec[q] = src[q];
        __syncthreads();
        const int32_t j = j0 + (int32_t)threadIdx.x;
        if ((int32_t)threadIdx.x < m && j < c.kept) {
            RecWords u;
            for (int q = 0; q < 3; q++) u.w[q] = s_rec[3 * threadIdx.x + q];
            struct potential rec = __builtin_bit_cast(struct potential, u);
            float ke = 0.0f;
            for (int k = 0; k < 3; k++) {
                const float dvp = rec.pos[k + 3] - cen_v[k];
                const float dxp = rec.pos[k] - cen_x[k];
                const double dv = (double)dvp + hub_a * (double)dxp;
                ke = (float)((double)ke + dv * dv);
            }
            ke = (float)((double)ke * conv);
            rec.ke = ke;
            const bool b = !(rec.pe < (double)ke);
            if (bound) bound[c.first + j] = b ? 1 : 0;
            nb += b ? 1 : 0;
            u = __builtin_bit_cast(RecWords, rec);
            s_rec[3 * threadIdx.x + 2] = u.w[2];  // ke and flags: the third word
        }
        __syncthreads();
        float4 *dst = reinterpret_cast<float4 *>(out + c.first + j0);
        for (int q = threadIdx.x; q < 3 * m; q += kKinThreads) dst[q] = s_rec[q];
        __syncthreads();
    }
    __shared__ int32_t s_nb[kKinThreads / kWave];
    nb = wave_sum(nb);
    if (lane_id() == 0) s_nb[threadIdx.x / kWave] = nb;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t tot = 0;
        for (int w = 0; w < kKinThreads / kWave; w++) tot += s_nb[w];
        tile_nb[t] = tot;
    }
}

// potential-only records: bound = 0
__global__ void k_bound_tail(const int64_t *__restrict__ hfirst, const int64_t *__restrict__ hcount,
                             const CallInfo *__restrict__ ci, int32_t nh, uint8_t *__restrict__ bound,
                             const int32_t *__restrict__ flags) {
    const int32_t h = blockIdx.x;
    if (h >= nh || (*flags & kErrDeep)) return;
    const CallInfo c = ci[h];
    for (int64_t j = c.kept + threadIdx.x; j < hcount[h]; j += blockDim.x) bound[hfirst[h] + j] = 0;
}

__global__ void k_nbound_calls(const int32_t *__restrict__ call_tile0,
                               const int32_t *__restrict__ call_ntiles,
                               const int32_t *__restrict__ tile_nb, int32_t nh,
                               hp_halo *__restrict__ halos, const int32_t *__restrict__ flags) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh || (*flags & kErrDeep)) return;
    int64_t n = 0;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) n += tile_nb[t];
    halos[h].n_bound = n;
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

}  // namespace hp
