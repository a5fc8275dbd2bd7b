// hp_lib.hip -- libhalopot: host engine and C ABI.
//
// One translation unit: the kernel files are included here so that every
// launch sees its kernel's definition.
//
// Engine: one HIP stream (default priority) plus device scratch grown on
// demand.  Each host thread that calls the reference entry points gets its
// own engine (the reference keeps its tree thread-private,
// src/potential.c:66-69); a plan owns one too.
//
// Pipeline of one batch (every call of the batch together):
//   gather + roots -> split levels until no node has more than 10 points ->
//   centres of mass -> dmin -> walk (or the direct kernel) -> [bound pass:
//   each call's records sorted by (potential-only, r2) -> core/bulk velocity -> kinetic
//   energy + bound flags -> records written back in sorted order].
// The split levels are host-driven: after each level the engine reads the
// device counters (four integers) to size the next level's launches.

#include "hp_build.hip"
#include "hp_unbind.hip"
#include "hp_props.hip"
#include "hp_gather.hip"
#include "hp_subfof.hip"
#include "hp_walk.hip"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

// ----------------------------------------------------------- C globals

extern "C" {
// Weak definitions with the reference's defaults (src/config.template.h:7,
// 16, 20, 22-25); the halo finder's own definitions take precedence.
__attribute__((weak, visibility("default"))) double FORCE_RES = 0.003;
__attribute__((weak, visibility("default"))) double PARTICLE_MASS = 0.0;
__attribute__((weak, visibility("default"))) double Om = 0.27;
__attribute__((weak, visibility("default"))) double Ol = 0.73;
__attribute__((weak, visibility("default"))) double W0 = -1.0;
__attribute__((weak, visibility("default"))) double WA = 0.0;
__attribute__((weak, visibility("default"))) double SCALE_NOW = 1.0;
// src/error.h:15, present when the program links src/error.cpp
void rockstar_abort_impl(int code, const char *file, int line) __attribute__((weak));
}

namespace hp {

static thread_local std::string g_last_error;
static std::atomic<hp_error_handler> g_handler{nullptr};

static int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HP_TRY(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_error(HP_EHIP, "%s failed at %s:%d: %s", #call, __FILE__,       \
                             __LINE__, hipGetErrorString(e_));                         \
    } while (0)

#define HP_LAUNCHED()                                                                  \
    do {                                                                               \
        hipError_t e_ = hipGetLastError();                                             \
        if (e_ != hipSuccess)                                                          \
            return set_error(HP_EHIP, "launch failed at %s:%d: %s", __FILE__, __LINE__,\
                             hipGetErrorString(e_));                                   \
    } while (0)

// Fatal error inside a void reference entry point: the reference's exit()
// path (src/error.h:17-20) when linked, else the user handler, else abort.
static void fatal(int code, const char *file, int line) {
    if (rockstar_abort_impl) {
        rockstar_abort_impl(code, file, line);
        return;
    }
    hp_error_handler h = g_handler.load();
    if (h) {
        h(code, file, line);
        return;
    }
    fprintf(stderr, "[halopot error] %s (code %d at %s:%d)\n", g_last_error.c_str(), code,
            file, line);
    abort();
}

// --------------------------------------------------------------- scalars

// src/hubble.c:8-19
static double hubble_scaling(double z) {
    double z1 = 1.0 + z;
    double a = 1.0 / z1;
    double weff = (a != 1.0) ? W0 + WA - WA * (a - 1.0) / log(a) : W0;
    return sqrt(Om * (z1 * z1 * z1) + Ol * pow(a, -3.0 * (1.0 + weff)));
}

// src/groupies.c:97-100
static double vir_density(double a) {
    double x = (Om / pow(a, 3)) / pow(hubble_scaling(1.0 / a - 1.0), 2.0) - 1.0;
    return ((18 * M_PI * M_PI + 82.0 * x - 39 * x * x) / (1.0 + x));
}

// particle_rvir_dens as src/groupies.c:106-135 leaves it for "vir": float
// background density, float threshold
static double particle_rvir_dens() {
    float scale_now = (float)SCALE_NOW;
    float cons = (float)(Om * 2.77519737e11 / PARTICLE_MASS);
    float thresh = (float)(vir_density(scale_now) * cons);
    return (double)thresh;
}

// _calc_mass_definition (src/groupies.c:102-122): a threshold in particles
// per (Mpc/h)^3 for "vir", "<x>b" (x times the background density) or
// "<x>c" (x times the critical density), optionally prefixed "m"; any other
// string is "vir".  Float scale factor, matter fraction, background density
// and result, as there.
static double mass_definition(const char *md) {
    const float scale_now = (float)SCALE_NOW;
    const size_t length = strlen(md);
    const char last = length ? md[length - 1] : 0;
    const float matter_fraction =
        (float)((Om / pow(scale_now, 3)) / pow(hubble_scaling(1.0 / scale_now - 1.0), 2.0));
    const float cons = (float)(Om * 2.77519737e11 / PARTICLE_MASS);  // background density
    const char *mass = md;
    if (mass[0] == 'm' || mass[0] == 'M') mass++;
    float thresh;
    if (last == 'b' || last == 'B') thresh = (float)(atof(mass) * cons);
    else if (last == 'c' || last == 'C') thresh = (float)(atof(mass) * cons / matter_fraction);
    else thresh = (float)(vir_density(scale_now) * cons);
    return (double)thresh;
}

// ----------------------------------------------------------- device buffer

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool carved = false;  // a slice of a region the buffer does not own (never freed)
    int ensure(size_t n) {
        if (n <= bytes) return HP_OK;
        if (carved) return set_error(HP_EINVAL, "carved buffer of %zu bytes asked for %zu", bytes, n);
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = n + n / 8 + 256;
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(HP_ENOMEM, "hipMalloc of %zu bytes failed", want);
        }
        bytes = want;
        return HP_OK;
    }
    void release() {
        if (p && !carved) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        carved = false;
    }
    template <typename T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

#define HP_ENSURE(buf, n)                       \
    do {                                        \
        int r_ = (buf).ensure(n);               \
        if (r_) return r_;                      \
    } while (0)

// Inside the engine: from the engine's carve region when it has one, else
// hipMalloc'd (DevBuf::ensure).
#define HP_GRAB(buf, n)                         \
    do {                                        \
        int r_ = grab((buf), (n));              \
        if (r_) return r_;                      \
    } while (0)

// Device scratch of the sub-FOF step (hp_find_subfofs).
struct SubBufs {
    DevBuf groups, tiles, qtiles, ltiles, drawn, qidx, psum, mean, pcorr, sig, pbox, gerr, w, w2, rk, bpos, spos,
        nodes, act0, act1, counters, slot_group, dist, sel, sel2, target, csize, emit, fused;
    // the linking and the output order: each particle's later neighbours
    // (appended, then listed in place), earliest earlier neighbour, pointer,
    // number, the roots of the numbers, the sort values and slots (two
    // buffers), the ranges still to sort, scan block sums, the merge pairs
    // (counts, then the list)
    DevBuf edges, deg, adj, kmin, par, first_label, lab, root, v0, v1, x0, x1, rng0, rng1, rsmall,
        ssum, ecnt, evt;
    void release() {
        for (DevBuf *b : {&groups, &tiles, &qtiles, &ltiles, &drawn, &qidx, &psum, &mean, &pcorr, &sig, &pbox,
                          &gerr, &w, &w2, &rk, &bpos, &spos, &nodes, &act0, &act1, &counters,
                          &slot_group, &dist, &sel, &sel2, &target, &csize, &emit, &fused,
                          &edges, &deg, &adj, &kmin, &par, &first_label, &lab, &root, &v0, &v1,
                          &x0, &x1, &rng0, &rng1, &rsmall, &ssum, &ecnt, &evt})
            b->release();
    }
};

// Host-side tables of one batch (depend only on the calls' first/count).
struct Tables {
    int64_t nrec = 0, nhalos = 0;
    std::vector<int64_t> first, count;
    std::vector<RootTile> root_tiles;
    std::vector<Group> groups;          // walk wavefronts, 64 sink points each
    std::vector<int2> direct_blocks;    // direct-kernel workgroups
    std::vector<ScanTile> scan_tiles;
    std::vector<int32_t> call_tile0, call_ntiles;
    std::vector<int32_t> seg_wave, seg_block;  // calls sorted by one wavefront / workgroup
    std::vector<SegTile> seg_tiles;            // tiles of the calls sorted by passes
    int method = HP_METHOD_TREE;
    int32_t direct_max = kLeafPoints;   // AUTO: direct at or below this count
};

// Node ids, task slots and record positions are 32-bit: the node capacity
// 5N + H + 1024 (Engine::reserve) must stay below INT32_MAX.
constexpr int64_t kMaxCalls = 1 << 22;
constexpr int64_t kMaxRecords = (INT32_MAX - kMaxCalls - 1024) / 5;
static_assert(5 * kMaxRecords + kMaxCalls + 1024 <= INT32_MAX, "node ids fit 32 bits");

// Workgroups of a launch are dispatched to the XCDs in turn (workgroup b to
// XCD b % xcds), and each XCD has its own L2.  The walk reads a call's tree
// over and over, so the wavefronts of one call should run on one XCD: the
// calls, in the list's order, go to the XCD with the fewest wavefronts so
// far, and the list is rewritten so that entry b is the next of XCD b %
// xcds's wavefronts; XCDs that run out get empty entries (root -1).  A call
// with more than 1/(8 xcds) of all the wavefronts is cut into pieces of that
// size, placed one by one, so that one large call still runs on every XCD.
static void place_on_xcds(std::vector<Group> &g, int xcds) {
    if (xcds <= 1 || g.empty()) return;
    const size_t piece = std::max<size_t>(1, g.size() / (8 * (size_t)xcds));
    std::vector<std::vector<Group>> per(xcds);
    for (size_t i = 0; i < g.size();) {
        size_t j = i;
        while (j < g.size() && g[j].root == g[i].root && j - i < piece) j++;
        int best = 0;
        for (int x = 1; x < xcds; x++)
            if (per[x].size() < per[best].size()) best = x;
        per[best].insert(per[best].end(), g.begin() + i, g.begin() + j);
        i = j;
    }
    size_t rows = 0;
    for (auto &v : per) rows = std::max(rows, v.size());
    g.assign(rows * xcds, Group{-1, 0, 0, 0});
    for (int x = 0; x < xcds; x++)
        for (size_t k = 0; k < per[x].size(); k++) g[k * xcds + x] = per[x][k];
}

static int make_tables(Tables &T, const int64_t *first, const int64_t *count, int64_t nh,
                       int method, bool want_unbind, int32_t direct_max = kLeafPoints) {
    T = Tables();
    T.method = method;
    T.direct_max = direct_max;
    T.nhalos = nh;
    T.first.assign(first, first + nh);
    T.count.assign(count, count + nh);
    int64_t pos = 0;
    for (int64_t h = 0; h < nh; h++) {
        if (count[h] < 0 || first[h] != pos)
            return set_error(HP_EINVAL, "call %lld: first=%lld count=%lld; calls must tile the "
                             "record array in order", (long long)h, (long long)first[h],
                             (long long)count[h]);
        pos += count[h];
    }
    if (nh > kMaxCalls) return set_error(HP_EINVAL, "too many calls in one batch");
    if (pos > kMaxRecords)
        return set_error(HP_EINVAL, "batch of %lld records is too large (at most %lld)",
                         (long long)pos, (long long)kMaxRecords);
    T.nrec = pos;
    for (int64_t h = 0; h < nh; h++) {
        for (int64_t off = 0; off < count[h]; off += kTile)
            T.root_tiles.push_back(RootTile{(int32_t)h, (int32_t)off});
        const bool direct = method == HP_METHOD_DIRECT ||
                            (method == HP_METHOD_AUTO && count[h] <= T.direct_max);
        if (direct) {
            for (int64_t off = 0; off < count[h]; off += 256)
                T.direct_blocks.push_back(make_int2((int32_t)h, (int32_t)off));
        } else {
            for (int64_t off = 0; off < count[h]; off += kGroup)
                T.groups.push_back(Group{(int32_t)h, (int32_t)(first[h] + off),
                                         (int32_t)(first[h] + count[h]), 0});
        }
    }
    // the walk's wavefronts of the largest calls first: they walk the
    // deepest trees, and starting them early shortens the launch's tail
    auto larger = [&](const Group &x, const Group &y) { return count[x.root] > count[y.root]; };
    std::stable_sort(T.groups.begin(), T.groups.end(), larger);
    if (want_unbind) {
        T.call_tile0.resize(nh);
        T.call_ntiles.resize(nh);
        for (int64_t h = 0; h < nh; h++) {
            T.call_tile0[h] = (int32_t)T.scan_tiles.size();
            for (int64_t off = 0; off < count[h]; off += kScanTile)
                T.scan_tiles.push_back(ScanTile{(int32_t)h, (int32_t)off});
            T.call_ntiles[h] = (int32_t)(T.scan_tiles.size() - T.call_tile0[h]);
        }
        int64_t big_before = 0;  // records of the pass-sorted calls before this one
        int32_t tiles_before = 0;
        for (int64_t h = 0; h < nh; h++) {
            const int64_t n = count[h];
            if (n <= 0) continue;
            if (n <= kWave) {
                T.seg_wave.push_back((int32_t)h);
            } else if (n <= kSegBlockMax) {
                T.seg_block.push_back((int32_t)h);
            } else {
                const int32_t ntc = (int32_t)((n + kSortTile - 1) / kSortTile);
                for (int32_t t = 0; t < ntc; t++) {
                    SegTile st{};
                    st.off = (int32_t)(first[h] + (int64_t)t * kSortTile);
                    st.n = (int32_t)std::min<int64_t>(kSortTile, n - (int64_t)t * kSortTile);
                    st.hpos = tiles_before * kSegRadix + t;
                    st.ntc = ntc;
                    st.delta = (int32_t)(first[h] - big_before);
                    T.seg_tiles.push_back(st);
                }
                tiles_before += ntc;
                big_before += n;
            }
        }
    }
    return HP_OK;
}

struct Engine {
    int device = -1;
    hipStream_t stream = nullptr;
    // streams (default priority) on which a recorded step runs the subtree
    // kernels of the other sizes beside the largest, and their events
    hipStream_t side[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr};
    // a recorded step's bound-pass sort branch (forked after the subtrees)
    hipEvent_t sort_fork_ev = nullptr, sort_join_ev = nullptr;
    bool sort_forked = false;
    // and its walk branch for the calls without split levels
    hipEvent_t walk_fork_ev = nullptr, walk_join_ev = nullptr;
    bool walk_forked = false;
    int32_t *h_counters = nullptr;  // pinned mirror
    unsigned long long *h_work = nullptr;
    uint64_t *h_stamps = nullptr;   // pinned: wall-clock stamps between a graph step's phases
    double wall_khz = 1.0e5;        // the device wall clock's rate
    hipEvent_t ev[8] = {};
    // launch setting per call-size bucket (tune_bucket)
    int32_t bucket_big[kTuneBuckets] = {kSmallMax, kSmallMax, kSmallMax, kSmallMax};
    int32_t bucket_wave[kTuneBuckets] = {kWaveTask, kWaveTask, kWaveTask, kWaveTask};
    int32_t big_max() const { return *std::max_element(bucket_big, bucket_big + kTuneBuckets); }
    int32_t big_min() const { return *std::min_element(bucket_big, bucket_big + kTuneBuckets); }
    int xcds = 1;             // XCDs of the device (each with its own L2)
    int32_t walk_groups = 0;  // entries of the loaded walk table (XCD layout)
    // The walk table holds the wavefronts of the calls with split levels
    // first (walk_groups_lv entries), then the others, each part laid out on
    // the XCDs by itself: a recorded step walks the second part as soon as
    // those calls' trees are done.  walk_layout_big: the thresholds the split
    // was made with.
    int32_t walk_groups_lv = 0;
    int32_t walk_layout_big[kTuneBuckets] = {-1, -1, -1, -1};
    int sub_grid[kSubClasses] = {}, wave_grid = 0;  // resident workgroups of the subtree kernels

    // batch tables on the device
    DevBuf hfirst, hcount, root_tiles, groups, direct_blocks, scan_tiles, call_tile0,
        call_ntiles;
    // build
    DevBuf box, bad, pot, badlist, A, B, F, leaf_of, sides, rkey, nodes, parent, fat, small0, group0, groupc, wave_off, wave_hsum, big, tiles, parts,
        splitres, dparts, counters, work, depth_buf;
    // unbind
    DevBuf keys0, keys1, vals0, vals1, hist, hsum, srt, callinfo, tsum, tcand, tile_i32, rec_call,
        strip_cnt;
    DevBuf stage;  // records in tree order before they go back (compute_potential's permutation)
    DevBuf seg_wave_l, seg_block_l, seg_tiles_l, seg_hist, seg_hsum, srt_rv;
    // halo properties
    DevBuf pcnt, pscan, pcall, psums, shape_st, shape_part, shape_cd, scale_calls, r_pe, pe_cnt,
        fit_list, fit_n;
    // record assembly
    DevBuf gsegs, gtiles, gcen;
    SubBufs sub;
    int64_t cap_nodes = 0, cap_small = 0, cap_group = 0, cap_big = 0, cap_tiles = 0, cap_fat = 0;
    std::vector<int32_t> depth_end;  // node count after each depth's level

    // last run
    hp_stats st{};

    int init() {
        if (stream) return HP_OK;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            (void)hipGetLastError();
            return set_error(HP_ENODEV, "no HIP device available (libhalopot has no CPU path)");
        }
        HP_TRY(hipGetDevice(&device));
        hipDeviceProp_t prop;
        HP_TRY(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return set_error(HP_ENODEV, "device %d is %s; libhalopot is built for gfx950", device,
                             prop.gcnArchName);
        // A blocking stream: it waits for work the caller left on the null
        // stream (a device-to-device hipMemcpy, for one, can return before
        // the copy is done), so the records are complete when read.
        HP_TRY(hipStreamCreateWithFlags(&stream, hipStreamDefault));
        if (hipDeviceGetAttribute(&xcds, hipDeviceAttributeNumberOfXccs, device) != hipSuccess ||
            xcds < 1) {
            (void)hipGetLastError();
            xcds = 1;
        }
        if (const char *s = getenv("HALOPOT_XCDS")) xcds = std::max(1, atoi(s));
        HP_TRY(hipHostMalloc((void **)&h_counters, sizeof(int32_t) * kNumCounters));
        HP_TRY(hipHostMalloc((void **)&h_work, sizeof(unsigned long long) * 4));
        HP_TRY(hipHostMalloc((void **)&h_stamps, sizeof(uint64_t) * 8));
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
            khz > 0)
            wall_khz = (double)khz;
        else
            (void)hipGetLastError();
        for (auto &e : ev) HP_TRY(hipEventCreate(&e));
        for (int b = 0; b < 3; b++) {
            HP_TRY(hipStreamCreateWithFlags(&side[b], hipStreamDefault));
            HP_TRY(hipEventCreateWithFlags(&join_ev[b], hipEventDisableTiming));
        }
        HP_TRY(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        HP_TRY(hipEventCreateWithFlags(&sort_fork_ev, hipEventDisableTiming));
        HP_TRY(hipEventCreateWithFlags(&sort_join_ev, hipEventDisableTiming));
        HP_TRY(hipEventCreateWithFlags(&walk_fork_ev, hipEventDisableTiming));
        HP_TRY(hipEventCreateWithFlags(&walk_join_ev, hipEventDisableTiming));
        // The workgroup subtree kernels take tasks from dynamic queues, the
        // wavefront one at a fixed stride, until the lists are empty; their
        // grids are the workgroups that fit on the device at once.
        int per_cu = 0;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<256>, kSubThreads, 0));
        sub_grid[0] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<512>, kSubThreads, 0));
        sub_grid[1] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<1024>, kSubThreads, 0));
        sub_grid[2] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree_wave, kWave, 0));
        wave_grid = std::max(1, per_cu) * prop.multiProcessorCount;
        if (const char *s = getenv("HALOPOT_BIG_THRESHOLD")) {
            const int32_t bt = std::min(std::max(atoi(s), kLeafPoints), kSmallMax);
            for (int b = 0; b < kTuneBuckets; b++) bucket_big[b] = bt;
        }
        return HP_OK;
    }

    // The engine of hp_compute_potential_noalloc: every device buffer is a
    // slice of a caller-given region (carve), the stream is the caller's, and
    // the level counters are read into host memory of the engine itself, so
    // that a run allocates nothing.
    uint8_t *carve = nullptr;
    size_t carve_cap = 0, carve_off = 0;
    bool own_stream = true;
    int32_t host_counters[kNumCounters] = {};
    unsigned long long host_work[4] = {};

    int grab(DevBuf &b, size_t n) {
        if (!carve) return b.ensure(n);
        if (b.carved && b.bytes >= n) return HP_OK;
        const size_t a = (n + 255) & ~(size_t)255;
        if (carve_off + a > carve_cap)
            return set_error(HP_EINVAL, "scratch of %zu bytes is too small (%zu more needed)",
                             carve_cap, carve_off + a - carve_cap);
        b.p = carve + carve_off;
        b.bytes = a;
        b.carved = true;
        carve_off += a;
        return HP_OK;
    }

    template <class Fn>
    void for_each_buf(Fn &&f) {
        DevBuf *all[] = {&hfirst, &hcount, &root_tiles, &groups, &direct_blocks,
                         &scan_tiles, &call_tile0, &call_ntiles, &box, &bad, &pot, &rkey, &badlist, &A, &B, &F,
                         &leaf_of, &sides, &nodes, &parent, &fat, &small0, &group0, &groupc, &wave_off, &wave_hsum, &big, &tiles,
                         &parts, &splitres, &dparts, &counters, &work, &depth_buf, &pcnt, &pscan, &pcall, &psums, &shape_st, &shape_part, &shape_cd, &scale_calls, &r_pe, &pe_cnt, &fit_list, &fit_n, &gsegs, &gtiles, &gcen, &keys0, &keys1, &vals0,
                         &vals1, &hist, &hsum, &srt, &callinfo, &tsum, &tcand, &tile_i32, &rec_call, &strip_cnt, &stage,
                         &seg_wave_l, &seg_block_l, &seg_tiles_l, &seg_hist, &seg_hsum, &srt_rv};
        for (DevBuf *b : all) f(*b);
    }

    // Every buffer is carved again from the start of the region.
    void reset_carve() {
        carve_off = 0;
        for_each_buf([](DevBuf &b) { b.release(); });
    }

    // Set-up of the engine of hp_compute_potential_noalloc on the current
    // device: the checks and launch sizes of init(), no stream, no pinned
    // memory, no events.
    int init_static(void *region, size_t bytes) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            (void)hipGetLastError();
            return set_error(HP_ENODEV, "no HIP device available (libhalopot has no CPU path)");
        }
        HP_TRY(hipGetDevice(&device));
        hipDeviceProp_t prop;
        HP_TRY(hipGetDeviceProperties(&prop, device));
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            return set_error(HP_ENODEV, "device %d is %s; libhalopot is built for gfx950", device,
                             prop.gcnArchName);
        if (hipDeviceGetAttribute(&xcds, hipDeviceAttributeNumberOfXccs, device) != hipSuccess ||
            xcds < 1) {
            (void)hipGetLastError();
            xcds = 1;
        }
        int per_cu = 0;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<256>, kSubThreads, 0));
        sub_grid[0] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<512>, kSubThreads, 0));
        sub_grid[1] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree<1024>, kSubThreads, 0));
        sub_grid[2] = std::max(1, per_cu) * prop.multiProcessorCount;
        HP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_subtree_wave, kWave, 0));
        wave_grid = std::max(1, per_cu) * prop.multiProcessorCount;
        carve = static_cast<uint8_t *>(region);
        carve_cap = bytes;
        own_stream = false;
        h_counters = host_counters;
        h_work = host_work;
        return HP_OK;
    }

    ~Engine() {
        if (stream && own_stream) {
            (void)hipStreamSynchronize(stream);
            drop_graphs();
            for_each_buf([](DevBuf &b) { b.release(); });
            sub.release();
            for (auto &e : ev)
                if (e) (void)hipEventDestroy(e);
            if (h_counters) (void)hipHostFree(h_counters);
            if (h_work) (void)hipHostFree(h_work);
            if (h_stamps) (void)hipHostFree(h_stamps);
            if (fork_ev) (void)hipEventDestroy(fork_ev);
            if (sort_fork_ev) (void)hipEventDestroy(sort_fork_ev);
            if (sort_join_ev) (void)hipEventDestroy(sort_join_ev);
            if (walk_fork_ev) (void)hipEventDestroy(walk_fork_ev);
            if (walk_join_ev) (void)hipEventDestroy(walk_join_ev);
            for (int b = 0; b < 3; b++) {
                if (join_ev[b]) (void)hipEventDestroy(join_ev[b]);
                if (side[b]) {
                    (void)hipStreamSynchronize(side[b]);
                    (void)hipStreamDestroy(side[b]);
                }
            }
            (void)hipStreamDestroy(stream);
        }
    }

    template <typename T>
    int upload(DevBuf &b, const std::vector<T> &v) {
        HP_GRAB(b, sizeof(T) * (v.size() ? v.size() : 1));
        if (!v.empty())
            HP_TRY(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, stream));
        return HP_OK;
    }

    int reserve(const Tables &T, bool unbind) {
        const int64_t N = T.nrec, H = T.nhalos;
        HP_TRY(hipSetDevice(device));
        HP_GRAB(box, sizeof(int32_t) * 6 * (H + 1));
        HP_GRAB(bad, sizeof(int32_t) * (H + 1));
        HP_GRAB(pot, sizeof(int32_t) * (H + 1));
        HP_GRAB(badlist, sizeof(int32_t) * (H + 1));
        HP_GRAB(A, sizeof(float4) * (N + 1));
        HP_GRAB(B, sizeof(float4) * (N + 1));
        HP_GRAB(F, sizeof(float4) * (N + kSrcBatch));  // the walk reads leaf points in batches
        HP_GRAB(leaf_of, sizeof(int32_t) * (N + 1));
        HP_GRAB(sides, sizeof(int32_t) * (N + 1));
        HP_GRAB(rkey, sizeof(uint64_t) * (N + 1));
        // split levels use at most about 0.1N + H ids; a subtree of n points
        // (workgroup or wavefront) reserves 2n - 2
        cap_nodes = 5 * N + H + 1024;
        HP_GRAB(nodes, sizeof(Node) * cap_nodes);
        HP_GRAB(parent, sizeof(int32_t) * cap_nodes);
        cap_fat = N / (kFatLeaf + 1) + 16;
        HP_GRAB(fat, sizeof(int32_t) * cap_fat);
        cap_small = N / (kLeafPoints + 1) + H + 64;
        HP_GRAB(small0, sizeof(Task) * cap_small);
        HP_GRAB(wave_off, sizeof(int32_t) * (cap_small + 2));
        HP_GRAB(wave_hsum, sizeof(int32_t) * ((cap_small + 1) / kScanBlock + 2));
        cap_group = N / (kWaveMin + 1) + H + 64;  // tasks of more than wave_max >= kWaveMin points
        HP_GRAB(group0, sizeof(Task) * cap_group);
        HP_GRAB(groupc, sizeof(Task) * kSubClasses * cap_group);
        const int64_t bt = std::max(big_min(), 1);
        cap_big = 128 * (N / bt + 1) + H + 64;
        cap_tiles = 128 * (N / kTile + 2) + cap_big;
        HP_GRAB(big, sizeof(Task) * cap_big);
        HP_GRAB(splitres, sizeof(BigSplit) * cap_big);
        HP_GRAB(tiles, sizeof(int2) * cap_tiles);
        HP_GRAB(parts, sizeof(TilePart) * cap_tiles);
        HP_GRAB(dparts, sizeof(double2) * cap_tiles);
        HP_GRAB(counters, sizeof(int32_t) * kNumCounters);
        HP_GRAB(depth_buf, sizeof(int32_t) * kLevelWords);
        HP_GRAB(work, sizeof(unsigned long long) * 4);
        if (unbind) {
            HP_GRAB(keys0, sizeof(uint64_t) * (N + 1));
            HP_GRAB(keys1, sizeof(uint64_t) * (N + 1));
            HP_GRAB(vals0, sizeof(int32_t) * (N + 1));
            HP_GRAB(vals1, sizeof(int32_t) * (N + 1));
            const int64_t nt = (N + kSortTile - 1) / kSortTile + 1;
            HP_GRAB(hist, sizeof(int32_t) * kRadix * nt);
            HP_GRAB(hsum, sizeof(int32_t) * (kRadix * nt / kScanBlock + 2));
            HP_GRAB(srt, sizeof(struct potential) * (N + 1));
            HP_GRAB(srt_rv, sizeof(float4) * (N + 1));
            HP_GRAB(callinfo, sizeof(CallInfo) * (H + 1));
            HP_GRAB(tsum, sizeof(Moments) * (T.scan_tiles.size() + 1));
            HP_GRAB(tcand, sizeof(Cand) * 2 * (T.scan_tiles.size() + 1));
            HP_GRAB(tile_i32, sizeof(int32_t) * (T.scan_tiles.size() + 1));
            HP_GRAB(rec_call, sizeof(int32_t) * (N + 1));
            HP_GRAB(strip_cnt, sizeof(int32_t) * (T.scan_tiles.size() + 1));
            const int64_t m = (int64_t)T.seg_tiles.size() * kSegRadix;
            HP_GRAB(seg_hist, sizeof(int32_t) * (m + 1));
            HP_GRAB(seg_hsum, sizeof(int32_t) * (m / kScanBlock + 2));
        }
        return HP_OK;
    }

    // A call has split levels when it holds more points than its bucket's
    // threshold (the device splits on its finite points, never more).
    int layout_walk(const Tables &T) {
        std::vector<Group> lv, rest;
        for (const Group &x : T.groups) {
            const int64_t n = T.count[(size_t)x.root];
            (n > bucket_big[tune_bucket(n)] ? lv : rest).push_back(x);
        }
        place_on_xcds(lv, xcds);
        place_on_xcds(rest, xcds);
        walk_groups_lv = (int32_t)lv.size();
        lv.insert(lv.end(), rest.begin(), rest.end());
        walk_groups = (int32_t)lv.size();
        std::copy(bucket_big, bucket_big + kTuneBuckets, walk_layout_big);
        return upload(groups, lv);
    }

    int64_t max_count = 0;  // the largest call of the loaded tables
    int load_tables(const Tables &T) {
        int r;
        max_count = T.count.empty() ? 0 : *std::max_element(T.count.begin(), T.count.end());
        if ((r = upload(hfirst, T.first))) return r;
        if ((r = upload(hcount, T.count))) return r;
        if ((r = upload(root_tiles, T.root_tiles))) return r;
        if ((r = layout_walk(T))) return r;
        if ((r = upload(direct_blocks, T.direct_blocks))) return r;
        if ((r = upload(scan_tiles, T.scan_tiles))) return r;
        if ((r = upload(call_tile0, T.call_tile0))) return r;
        if ((r = upload(call_ntiles, T.call_ntiles))) return r;
        if ((r = upload(seg_wave_l, T.seg_wave))) return r;
        if ((r = upload(seg_block_l, T.seg_block))) return r;
        if ((r = upload(seg_tiles_l, T.seg_tiles))) return r;
        if (!T.scan_tiles.empty() && rec_call.p && T.nhalos > 0) {  // each record's call
            // sized by these tables: a per-call plan reserves for its largest
            // call but loads the whole batch's tables too
            HP_GRAB(rec_call, sizeof(int32_t) * (T.nrec + 1));
            k_rec_calls<<<(int)T.nhalos, 256, 0, stream>>>(hfirst.as<int64_t>(), hcount.as<int64_t>(),
                                                          (int32_t)T.nhalos, rec_call.as<int32_t>());
            HP_LAUNCHED();
        }
        return HP_OK;
    }

    // Orders the engine's work after what the caller queued on `caller` (a
    // hipStream_t; NULL: nothing beyond the null stream, which the blocking
    // engine stream waits for by itself).
    int join(void *caller) {
        if (!caller) return HP_OK;
        HP_TRY(hipSetDevice(device));
        HP_TRY(hipEventRecord(ev[7], (hipStream_t)caller));
        HP_TRY(hipStreamWaitEvent(stream, ev[7], 0));
        return HP_OK;
    }
    // ... and the caller's later work on `caller` after the engine's.
    int release(void *caller) {
        if (!caller) return HP_OK;
        HP_TRY(hipEventRecord(ev[7], stream));
        HP_TRY(hipStreamWaitEvent((hipStream_t)caller, ev[7], 0));
        return HP_OK;
    }

    int read_counters() {
        HP_TRY(hipMemcpyAsync(h_counters, counters.p, sizeof(int32_t) * kNumCounters,
                              hipMemcpyDeviceToHost, stream));
        HP_TRY(hipStreamSynchronize(stream));
        return HP_OK;
    }

    BuildArgs build_args(Task *small_next) {
        BuildArgs a;
        a.nodes = nodes.as<Node>();
        a.parent = parent.as<int32_t>();
        a.sides = sides.as<int32_t>();
        a.leaf_of = leaf_of.as<int32_t>();
        a.fin = F.as<float4>();
        a.small_next = small_next;
        a.group_next = group0.as<Task>();
        a.big = big.as<Task>();
        a.tiles = tiles.as<int2>();
        a.fat = fat.as<int32_t>();
        a.counters = counters.as<int32_t>();
        for (int b = 0; b < kTuneBuckets; b++) {
            a.bucket_big[b] = bucket_big[b];
            a.bucket_wave[b] = bucket_wave[b];
        }
        a.cap_small = (int32_t)cap_small;
        a.cap_group = (int32_t)cap_group;
        a.cap_big = (int32_t)cap_big;
        a.cap_tiles = (int32_t)cap_tiles;
        a.cap_nodes = (int32_t)cap_nodes;
        a.cap_fat = (int32_t)cap_fat;
        return a;
    }

    static int blocks(int64_t n, int per) { return (int)((n + per - 1) / per); }

    // ------------------------------------------------------ enqueue stages
    // Nothing below reads device memory back: every stage can be recorded
    // into a graph.  The eager driver (potential) checks the level state
    // between batches of levels; the graph driver replays a planned number
    // of levels and checks once.

    int enqueue_roots(const Tables &T, const struct potential *d_po) {
        const int32_t H = (int32_t)T.nhalos;
        hipStream_t s = stream;
        // the per-call state and the build's counters
        k_root_init<<<blocks(std::max<int32_t>(H, kNumCounters), 256), 256, 0, s>>>(
            box.as<int32_t>(), bad.as<int32_t>(), pot.as<int32_t>(), H, counters.as<int32_t>());
        HP_LAUNCHED();
        if (!T.root_tiles.empty()) {
            k_gather<<<(int)T.root_tiles.size(), kTileThreads, 0, s>>>(
                d_po, hfirst.as<int64_t>(), hcount.as<int64_t>(), root_tiles.as<RootTile>(),
                A.as<float4>(), box.as<int32_t>(), bad.as<int32_t>(), pot.as<int32_t>(), H,
                rkey.as<uint64_t>());
            HP_LAUNCHED();
        }
        if (H > 0) {
            k_roots<<<blocks(H, 256), 256, 0, s>>>(hfirst.as<int64_t>(), hcount.as<int64_t>(),
                                                   box.as<int32_t>(), bad.as<int32_t>(),
                                                   A.as<float4>(), badlist.as<int32_t>(), H,
                                                   build_args(small0.as<Task>()));
            HP_LAUNCHED();
            k_roots_bad<<<64, 256, 0, s>>>(hfirst.as<int64_t>(), hcount.as<int64_t>(),
                                           box.as<int32_t>(), bad.as<int32_t>(), A.as<float4>(),
                                           badlist.as<int32_t>(), build_args(small0.as<Task>()));
            HP_LAUNCHED();
        }
        k_levels_init<<<1, 1, 0, s>>>(counters.as<int32_t>(), depth_buf.as<int32_t>());
        HP_LAUNCHED();
        return HP_OK;
    }

    // split levels [first, first + n) of the nodes of more than big_threshold
    // points; level L reads the points from buffer A (odd L) or B (even L)
    // and queues children of at most big_threshold points as subtree tasks
    // workgroups of the level kernels (grid-stride over the level's tiles):
    // the first levels hold about 4,600 tiles of 1,024 points on the
    // one-GPU catalogue; 4,096 measured faster than 2,048 and 8,192
    static constexpr int kLevelGrid = 4096;
    int enqueue_levels(int first, int n) {
        hipStream_t s = stream;
        if (first < 1 || first + n - 1 >= kMaxLevels)
            return set_error(HP_EINVAL, "split levels %d..%d out of range", first, first + n - 1);
        int32_t *lv = depth_buf.as<int32_t>();
        for (int level = first; level < first + n; level++) {
            float4 *src = (level & 1) ? A.as<float4>() : B.as<float4>();
            float4 *dst = (level & 1) ? B.as<float4>() : A.as<float4>();
            const BuildArgs ba = build_args(small0.as<Task>());
            k_big_count<<<kLevelGrid, kTileThreads, 0, s>>>(src, tiles.as<int2>(), big.as<Task>(),
                                                      nodes.as<Node>(), parts.as<TilePart>(), lv,
                                                      level);
            HP_LAUNCHED();
            k_big_tiles<kSides><<<kLevelGrid, kTileThreads, 0, s>>>(
                src, dst, tiles.as<int2>(), big.as<Task>(), parts.as<TilePart>(),
                splitres.as<BigSplit>(), ba, lv, level);
            HP_LAUNCHED();
            k_big_tiles<kScatter><<<kLevelGrid, kTileThreads, 0, s>>>(
                src, dst, tiles.as<int2>(), big.as<Task>(), parts.as<TilePart>(),
                splitres.as<BigSplit>(), ba, lv, level);
            HP_LAUNCHED();
        }
        return HP_OK;
    }

    // subtrees, then the multipoles of the nodes of `levels` split levels
    // fork: a recorded step's branches (subtree sizes; with sort_T the bound
    // pass's sort; with walk_po and walk_o the walk of the calls without
    // split levels)
    int enqueue_multipole(int levels, bool fork = false, const Tables *sort_T = nullptr,
                          struct potential *walk_po = nullptr,
                          const hp_options *walk_o = nullptr) {
        hipStream_t s = stream;
        fork = fork && side[0];
        const int32_t *cnt = counters.as<int32_t>();
        const BuildArgs ba = build_args(small0.as<Task>());
        // A workgroup subtree kernel (and the sorting of its tasks by size) is
        // launched only if a task of its size can exist: a workgroup task
        // holds more than the wavefront size and at most both the split size
        // and the largest call.  Largest first: the big subtrees hold most
        // points and the fewest workgroups per CU.
        const int64_t gmax = std::min<int64_t>(max_count, big_max());
        if (gmax > kWaveMin) {
            k_group_classes<<<blocks(cap_group, 256), 256, 0, s>>>(group0.as<Task>(), ba,
                                                                   groupc.as<Task>());
            HP_LAUNCHED();
        }
        {  // the wavefront subtrees' node ids: exclusive scan of 2n - 2 over the list
            const int32_t m = (int32_t)cap_small + 1;
            const int32_t nbk = (m + kScanBlock - 1) / kScanBlock;
            int32_t *off = wave_off.as<int32_t>();
            if (nbk == 1) {
                k_wave_ids_one<<<1, 1024, 0, s>>>(small0.as<Task>(), ba, off, m);
                HP_LAUNCHED();
            } else {
                k_wave_need<<<std::min(blocks(m, 256), 4096), 256, 0, s>>>(small0.as<Task>(), ba,
                                                                           off, m);
                HP_LAUNCHED();
                k_scan_blocks<<<nbk, 1024, 0, s>>>(off, m, wave_hsum.as<int32_t>());
                HP_LAUNCHED();
                k_scan_top<<<1, 1024, 0, s>>>(wave_hsum.as<int32_t>(), nbk);
                HP_LAUNCHED();
                k_scan_add<<<nbk, 1024, 0, s>>>(off, m, wave_hsum.as<int32_t>());
                HP_LAUNCHED();
                k_wave_base<<<1, 64, 0, s>>>(ba, off);
                HP_LAUNCHED();
            }
        }
        // recorded steps: the other sizes and the wavefront subtrees on their
        // own streams, beside the largest (their task lists and node ids are
        // independent once the ids are reserved), so that each kernel's last
        // tasks overlap the others' instead of leaving the device idle
        hipStream_t s1 = s, s2 = s, s3 = s;
        if (fork) {
            HP_TRY(hipEventRecord(fork_ev, s));
            for (int b = 0; b < 3; b++) HP_TRY(hipStreamWaitEvent(side[b], fork_ev, 0));
            s1 = side[0];
            s2 = side[1];
            s3 = side[2];
        }
        if (gmax > 512) {
            k_subtree<1024><<<sub_grid[2], kSubThreads, 0, s>>>(
                A.as<float4>(), B.as<float4>(), groupc.as<Task>() + 2 * cap_group, ba, PARTICLE_MASS, 2);
            HP_LAUNCHED();
        }
        if (gmax > 256) {
            k_subtree<512><<<sub_grid[1], kSubThreads, 0, s1>>>(
                A.as<float4>(), B.as<float4>(), groupc.as<Task>() + cap_group, ba, PARTICLE_MASS, 1);
            HP_LAUNCHED();
        }
        if (gmax > kWaveMin) {
            k_subtree<256><<<sub_grid[0], kSubThreads, 0, s2>>>(A.as<float4>(), B.as<float4>(),
                                                               groupc.as<Task>(), ba, PARTICLE_MASS, 0);
            HP_LAUNCHED();
        }
        k_subtree_wave<<<wave_grid, kWave, 0, s3>>>(A.as<float4>(), B.as<float4>(), small0.as<Task>(),
                                              build_args(small0.as<Task>()), PARTICLE_MASS,
                                              wave_off.as<int32_t>());
        HP_LAUNCHED();
        if (fork) {
            for (int b = 0; b < 3; b++) {
                HP_TRY(hipEventRecord(join_ev[b], side[b]));
                HP_TRY(hipStreamWaitEvent(s, join_ev[b], 0));
            }
            // the tree order is final: the bound pass's sort on a branch
            // that the bound pass joins
            if (sort_T) {
                HP_TRY(hipEventRecord(sort_fork_ev, s));
                HP_TRY(hipStreamWaitEvent(side[0], sort_fork_ev, 0));
                int r;
                if ((r = enqueue_bound_sort(*sort_T, side[0]))) return r;
                HP_TRY(hipEventRecord(sort_join_ev, side[0]));
                sort_forked = true;
            }
        }
        // the roots that stay leaves (at most 10 points, or as many finite)
        k_leaf_com<<<2048, 256, 0, s>>>(nodes.as<Node>(), F.as<float4>(), cnt, PARTICLE_MASS);
        HP_LAUNCHED();
        // the calls without split levels have their whole trees now: their
        // walk on a branch beside the split levels' nodes (joined after the
        // walk of the others)
        if (fork && sort_T && walk_o) {
            HP_TRY(hipEventRecord(walk_fork_ev, s));
            HP_TRY(hipStreamWaitEvent(side[1], walk_fork_ev, 0));
            int r;
            if ((r = enqueue_walk(*sort_T, walk_po, *walk_o, 2, side[1]))) return r;
            HP_TRY(hipEventRecord(walk_join_ev, side[1]));
            walk_forked = true;
        }
        // Without split levels every other node is a subtree's, and the
        // subtree kernels finished it (its root included): the fat leaves,
        // the internal nodes and dmin of the levels' nodes have no work.
        if (levels == 0) return HP_OK;
        k_fat_leaf_com<<<256, 256, 0, s>>>(nodes.as<Node>(), fat.as<int32_t>(), cnt,
                                           (int32_t)cap_fat, F.as<float4>(), PARTICLE_MASS);
        HP_LAUNCHED();
        for (int d = levels; d >= 0; d--) {
            k_internal_com<<<512, 256, 0, s>>>(nodes.as<Node>(), depth_buf.as<int32_t>(), d);
            HP_LAUNCHED();
        }
        k_dmin_tiles<<<1024, kTileThreads, 0, s>>>(nodes.as<Node>(), F.as<float4>(),
                                                   tiles.as<int2>(), big.as<Task>(), cnt,
                                                   dparts.as<double2>());
        HP_LAUNCHED();
        k_dmin_big<<<64, 256, 0, s>>>(nodes.as<Node>(), big.as<Task>(), cnt, dparts.as<double2>());
        HP_LAUNCHED();
        return HP_OK;
    }

    // part 0: every call; 1: the calls with split levels; 2: the others
    // and the direct-sum calls (on stream st)
    int enqueue_walk(const Tables &T, struct potential *d_po, const hp_options &o, int part = 0,
                     hipStream_t st = nullptr) {
        hipStream_t s = st ? st : stream;
        const int32_t g0 = part == 2 ? walk_groups_lv : 0;
        const int32_t ng = part == 0 ? walk_groups
                                     : (part == 1 ? walk_groups_lv : walk_groups - walk_groups_lv);
        if (!T.groups.empty() && ng > 0) {
            WalkArgs a;
            a.nodes = nodes.as<Node>();
            a.fin = F.as<float4>();
            a.leaf_of = leaf_of.as<int32_t>();
            a.groups = groups.as<Group>() + g0;
            a.ngroups = ng;
            a.counters = counters.as<int32_t>();
            a.nrec = (int32_t)T.nrec;
            a.cap_nodes = (int32_t)cap_nodes;
            a.pairs = o.walk != HP_WALK_NODES;
            a.pm = (float)PARTICLE_MASS;
            a.inv_fr = (float)(1.0 / FORCE_RES);
            a.pm_d = PARTICLE_MASS;
            a.fr_d = FORCE_RES;
            a.po = d_po;
            a.err = counters.as<int32_t>() + kErrFlags;
            a.work = work.as<unsigned long long>();
            const int nb = blocks((int64_t)ng * kWave, 64);
            const bool cnt = o.count_work != 0;
#define HP_WALK(ACC)                                                                   \
    do {                                                                               \
        if (cnt) k_walk<ACC, true><<<nb, 64, 0, s>>>(a);                              \
        else k_walk<ACC, false><<<nb, 64, 0, s>>>(a);                                 \
    } while (0)
            if (o.accumulate == HP_ACC_NAIVE) HP_WALK(HP_ACC_NAIVE);
            else if (o.accumulate == HP_ACC_F64) HP_WALK(HP_ACC_F64);
            else HP_WALK(HP_ACC_COMPENSATED);
#undef HP_WALK
            HP_LAUNCHED();
        }
        if (!T.direct_blocks.empty() && part != 1) {
            DirectArgs d;
            d.nodes = nodes.as<Node>();
            d.fin = T.method == HP_METHOD_DIRECT ? A.as<float4>() : F.as<float4>();
            d.blocks = direct_blocks.as<int2>();
            d.nblocks = (int32_t)T.direct_blocks.size();
            d.flags = counters.as<int32_t>() + kErrFlags;
            d.pm = (float)PARTICLE_MASS;
            d.inv_fr = (float)(1.0 / FORCE_RES);
            d.pm_d = PARTICLE_MASS;
            d.fr_d = FORCE_RES;
            d.po = d_po;
            if (T.method == HP_METHOD_DIRECT) {
                k_zero_pe<<<blocks(T.nrec, 256), 256, 0, s>>>(d_po, T.nrec);
            } else {
                k_zero_pe_calls<<<(int)T.direct_blocks.size(), 256, 0, s>>>(
                    d_po, hfirst.as<int64_t>(), hcount.as<int64_t>(), direct_blocks.as<int2>(),
                    counters.as<int32_t>() + kErrFlags);
            }
            HP_LAUNCHED();
            const int nbk = (int)T.direct_blocks.size();
            if (o.accumulate == HP_ACC_NAIVE) k_direct<HP_ACC_NAIVE><<<nbk, 256, 0, s>>>(d);
            else if (o.accumulate == HP_ACC_F64) k_direct<HP_ACC_F64><<<nbk, 256, 0, s>>>(d);
            else k_direct<HP_ACC_COMPENSATED><<<nbk, 256, 0, s>>>(d);
            HP_LAUNCHED();
        }
        return HP_OK;
    }

    int read_work() {
        HP_TRY(hipMemcpyAsync(h_work, work.p, sizeof(unsigned long long) * 4,
                              hipMemcpyDeviceToHost, stream));
        HP_TRY(hipStreamSynchronize(stream));
        st.visits = (int64_t)h_work[0];
        st.p2p = (int64_t)h_work[1];
        st.p2m = (int64_t)h_work[2];
        return HP_OK;
    }

    // Runs split levels from `level` until the level state is inactive,
    // checking every kLevelBatch levels.  Returns the next level index.
    int finish_levels(int level, int *next) {
        constexpr int kLevelBatch = 8;
        int r;
        for (;;) {
            if ((r = enqueue_levels(level, kLevelBatch))) return r;
            level += kLevelBatch;
            if ((r = read_counters())) return r;
            if (!h_counters[kActive] || h_counters[kErrFlags]) break;
            if (level > kMaxLevels - kLevelBatch)
                return set_error(HP_EINVAL, "tree deeper than %d levels", kMaxLevels);
        }
        *next = level;
        return HP_OK;
    }

    int check_build() {
        if (h_counters[kErrFlags])
            return set_error(HP_EINVAL, "tree build failed (flags %d)", h_counters[kErrFlags]);
        const int32_t nn = h_counters[kNodeCount];
        st.nodes = nn;
        st.leaves = (nn + st.halos) / 2;  // a full binary tree of n nodes has (n + 1) / 2 leaves
        st.levels = h_counters[kLevel] - 1;
        return HP_OK;
    }

    // whether some call of the tables has more points than its size bucket's
    // split size (the root gather counts only finite points, never more)
    bool may_split(const Tables &T) const {
        for (int64_t h = 0; h < T.nhalos; h++)
            if (T.count[h] > bucket_big[tune_bucket(T.count[h])]) return true;
        return false;
    }

    // the checks of a deferred pass, once h_counters holds its counters
    int finish_checks() {
        int r;
        if ((r = check_build()) || (r = check_walk(false))) return r;
        return HP_OK;
    }

    int check_walk(bool read = true) {
        int r;
        if (read && (r = read_counters())) return r;
        if (h_counters[kErrFlags])
            return set_error(HP_EINVAL, "walk stopped early (flags %d): corrupt tree",
                             h_counters[kErrFlags]);
        st.nodes = h_counters[kNodeCount];  // subtree ids included (some unused)
        return HP_OK;
    }

    // The eager driver: learns the depth of the tree as it builds it.  With
    // `permute`, the records end in tree order, as the reference's
    // compute_potential leaves them (the bound pass reads the tree order
    // itself and runs without).
    // With `staged`, the permuted records are left in the engine's stage
    // buffer (*staged points there) instead of being copied back to d_po.
    // With `deferred`, a pass without split levels reads no counters back:
    // *deferred is set, and the caller reads them with its own copy back
    // and calls finish_checks().
    int potential(const Tables &T, struct potential *d_po, const hp_options &o,
                  bool permute = false, const struct potential **staged = nullptr,
                  bool *deferred = nullptr) {
        if (deferred) *deferred = false;
        st = hp_stats();
        st.records = T.nrec;
        st.halos = T.nhalos;
        hipStream_t s = stream;
        const bool timing = o.timing != 0;
        int r;
        if (timing) HP_TRY(hipEventRecord(ev[0], s));
        if ((r = enqueue_roots(T, d_po))) return r;
        bool build_checked = true;
        if (T.method != HP_METHOD_DIRECT) {
            if (may_split(T)) {
                int next = 1;
                if ((r = finish_levels(1, &next))) return r;
                if ((r = check_build())) return r;
            } else {
                // no call above its split size: no split level has work, so
                // none is run, and the build's state is read after the walk
                st.levels = 0;
                build_checked = false;
            }
            if (timing) HP_TRY(hipEventRecord(ev[1], s));
            if ((r = enqueue_multipole((int)st.levels))) return r;
        } else if (timing) {
            HP_TRY(hipEventRecord(ev[1], s));
        }
        if (timing) HP_TRY(hipEventRecord(ev[2], s));
        if (o.count_work) HP_TRY(hipMemsetAsync(work.p, 0, sizeof(unsigned long long) * 4, s));
        if ((r = enqueue_walk(T, d_po, o))) return r;
        if (timing) HP_TRY(hipEventRecord(ev[3], s));
        if (permute && T.method != HP_METHOD_DIRECT && T.nrec > 0) {
            if ((r = grab(stage, sizeof(struct potential) * T.nrec))) return r;
            k_tree_records<<<blocks(T.nrec, 256), 256, 0, s>>>(d_po, F.as<float4>(), T.nrec,
                                                               stage.as<struct potential>());
            HP_LAUNCHED();
            if (staged)
                *staged = stage.as<struct potential>();
            else
                HP_TRY(hipMemcpyAsync(d_po, stage.p, sizeof(struct potential) * T.nrec,
                                      hipMemcpyDeviceToDevice, s));
        }
        // one readback for the build's and the walk's state when the build
        // was not checked before the walk
        if (!build_checked && deferred && !timing && !o.count_work) {
            *deferred = true;
            st.walk_groups = (int64_t)T.groups.size();
            st.walk_launches = T.groups.empty() ? 0 : 1;
            return HP_OK;
        }
        if (!build_checked && ((r = read_counters()) || (r = check_build()))) return r;
        if ((r = check_walk(build_checked))) return r;
        st.walk_groups = (int64_t)T.groups.size();
        st.walk_launches = T.groups.empty() ? 0 : 1;
        if (o.count_work && (r = read_work())) return r;
        if (timing) {
            float t;
            HP_TRY(hipEventElapsedTime(&t, ev[0], ev[1]));
            st.ms_build = t;
            HP_TRY(hipEventElapsedTime(&t, ev[1], ev[2]));
            st.ms_multipole = t;
            HP_TRY(hipEventElapsedTime(&t, ev[2], ev[3]));
            st.ms_walk = t;
            HP_TRY(hipEventElapsedTime(&t, ev[0], ev[3]));
            st.ms_total = t;
        }
        return HP_OK;
    }

    // ------------------------------------------------------------ graphs

    // The configuration globals the recorded kernels take as arguments
    // (PARTICLE_MASS, FORCE_RES: multipoles and walk; SCALE_NOW and the
    // cosmology: rvir threshold, Hubble term and KE factor of the bound pass).
    struct Config {
        double particle_mass, force_res, scale_now, om, ol, w0, wa;
        static Config now() { return Config{PARTICLE_MASS, FORCE_RES, SCALE_NOW, Om, Ol, W0, WA}; }
        bool operator==(const Config &o) const {
            return particle_mass == o.particle_mass && force_res == o.force_res &&
                   scale_now == o.scale_now && om == o.om && ol == o.ol && w0 == o.w0 &&
                   wa == o.wa;
        }
    };

    struct Graphs {
        hipGraphExec_t step = nullptr;  // the whole hp_plan_run, one launch
        int levels = 0;  // split levels recorded in `step`
        const void *po = nullptr, *halos = nullptr, *bound = nullptr;
        Config cfg{};    // the globals the graphs were recorded with
        bool ok = false;
    } g;

    void drop_graphs() {
        if (g.step) {
            (void)hipGraphExecDestroy(g.step);
            g.step = nullptr;
        }
        g.ok = false;
    }

    template <class F>
    int capture(hipGraphExec_t *out, F &&enqueue) {
        hipGraph_t graph = nullptr;
        HP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const int r = enqueue();
        sort_forked = walk_forked = false;
        const hipError_t e = hipStreamEndCapture(stream, &graph);
        if (r) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (e != hipSuccess)
            return set_error(HP_EHIP, "graph capture failed: %s", hipGetErrorString(e));
        const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess)
            return set_error(HP_EHIP, "graph instantiation failed: %s", hipGetErrorString(ie));
        return HP_OK;
    }

    // Records the whole step as one graph: the roots, `levels` split levels
    // (the depth learned eagerly plus kLevelMargin; levels with nothing left
    // to split cost their empty launches only), the depth check, the
    // multipoles, the walk and the bound pass, with a wall-clock stamp
    // before, between and after the phases (events recorded inside a graph
    // are not usable for elapsed times once the graph is recorded again).
    static constexpr int kLevelMargin = 2;
    int capture_all(const Tables &T, struct potential *d_po, hp_halo *d_halos, uint8_t *d_bound,
                    const hp_options &o, int levels) {
        drop_graphs();
        int r;
        const bool tree = T.method != HP_METHOD_DIRECT;
        if (tree && !std::equal(bucket_big, bucket_big + kTuneBuckets, walk_layout_big) &&
            (r = layout_walk(T)))
            return r;
        if ((r = capture(&g.step, [&]() -> int {
                 int q;
                 k_stamp<<<1, 1, 0, stream>>>(h_stamps, 0);
                 HP_LAUNCHED();
                 if ((q = enqueue_roots(T, d_po))) return q;
                 if (tree) {
                     if ((q = enqueue_levels(1, levels))) return q;
                     k_depth_check<<<1, 1, 0, stream>>>(counters.as<int32_t>());
                     HP_LAUNCHED();
                 }
                 k_stamp<<<1, 1, 0, stream>>>(h_stamps, 1);
                 HP_LAUNCHED();
                 sort_forked = walk_forked = false;
                 if (tree && (q = enqueue_multipole(levels, true, &T, d_po, &o))) return q;
                 k_stamp<<<1, 1, 0, stream>>>(h_stamps, 2);
                 HP_LAUNCHED();
                 if ((q = enqueue_walk(T, d_po, o, walk_forked ? 1 : 0))) return q;
                 if (walk_forked) {
                     walk_forked = false;
                     HP_TRY(hipStreamWaitEvent(stream, walk_join_ev, 0));
                 }
                 k_stamp<<<1, 1, 0, stream>>>(h_stamps, 3);
                 HP_LAUNCHED();
                 if ((q = enqueue_unbind(T, d_po, d_halos, d_bound))) return q;
                 k_stamp<<<1, 1, 0, stream>>>(h_stamps, 4);
                 HP_LAUNCHED();
                 return HP_OK;
             })))
            return r;
        g.levels = levels;
        g.cfg = Config::now();
        g.po = d_po;
        g.halos = d_halos;
        g.bound = d_bound;
        g.ok = true;
        return HP_OK;
    }

    // Graph driver of hp_plan_run: one graph launch per step, with no host
    // read between its phases.  The split levels are driven on the device;
    // a tree deeper than the recorded levels raises kErrDeep, under which
    // nothing of the caller's is written, and the step is then run eagerly
    // (the graph is recorded again on the next call).  The graph is recorded
    // again whenever the buffers or any configuration global it was
    // recorded with changed.
    int run_graphs(const Tables &T, struct potential *d_po, hp_halo *d_halos, uint8_t *d_bound,
                   const hp_options &o) {
        hipStream_t s = stream;
        int r;
        if (!g.ok || g.po != d_po || g.halos != d_halos || g.bound != d_bound ||
            !(g.cfg == Config::now())) {
            // learn the depth eagerly, then record
            if ((r = potential(T, d_po, o))) return r;
            if ((r = enqueue_unbind(T, d_po, d_halos, d_bound))) return r;
            HP_TRY(hipStreamSynchronize(s));
            return capture_all(T, d_po, d_halos, d_bound, o, (int)st.levels + 1 + kLevelMargin);
        }
        st = hp_stats();
        st.records = T.nrec;
        st.halos = T.nhalos;
        HP_TRY(hipGraphLaunch(g.step, s));
        if ((r = read_counters())) return r;
        if (h_counters[kErrFlags] & kErrDeep) {
            g.ok = false;
            if ((r = potential(T, d_po, o))) return r;
            if ((r = enqueue_unbind(T, d_po, d_halos, d_bound))) return r;
            HP_TRY(hipStreamSynchronize(s));
            return HP_OK;
        }
        if ((r = check_build())) return r;
        if ((r = check_walk())) return r;
        st.walk_groups = (int64_t)T.groups.size();
        st.walk_launches = T.groups.empty() ? 0 : 1;
        auto ms = [&](int i, int j) {
            return (float)((double)(int64_t)(h_stamps[j] - h_stamps[i]) / wall_khz);
        };
        st.ms_build = ms(0, 1);
        st.ms_multipole = ms(1, 2);
        st.ms_walk = ms(2, 3);
        st.ms_unbind = ms(3, 4);
        st.ms_total = ms(0, 4);
        return HP_OK;
    }

    // ------------------------------------------------------------ bound pass

    int enqueue_sort(int64_t n, int bits) {
        hipStream_t s = stream;
        const int32_t nt = (int32_t)((n + kSortTile - 1) / kSortTile);
        uint64_t *k0 = keys0.as<uint64_t>(), *k1 = keys1.as<uint64_t>();
        int32_t *v0 = vals0.as<int32_t>(), *v1 = vals1.as<int32_t>();
        for (int shift = 0; shift < bits; shift += kRadixBits) {
            k_radix_hist<<<nt, kSortThreads, 0, s>>>(k0, (int32_t)n, shift, nt, hist.as<int32_t>());
            HP_LAUNCHED();
            const int32_t m = kRadix * nt;
            const int32_t nbk = (m + kScanBlock - 1) / kScanBlock;
            k_scan_blocks<<<nbk, 1024, 0, s>>>(hist.as<int32_t>(), m, hsum.as<int32_t>());
            HP_LAUNCHED();
            k_scan_top<<<1, 1024, 0, s>>>(hsum.as<int32_t>(), nbk);
            HP_LAUNCHED();
            k_scan_add<<<nbk, 1024, 0, s>>>(hist.as<int32_t>(), m, hsum.as<int32_t>());
            HP_LAUNCHED();
            k_radix_scatter<<<nt, kSortThreads, 0, s>>>(k0, v0, k1, v1, (int32_t)n, shift, nt,
                                                        hist.as<int32_t>());
            HP_LAUNCHED();
            std::swap(k0, k1);
            std::swap(v0, v1);
        }
        sorted_vals = v0;
        sorted_keys = k0;
        return HP_OK;
    }
    int32_t *sorted_vals = nullptr;
    uint64_t *sorted_keys = nullptr;

    // the halo property pass over the bound pass's results (sorted records,
    // halos with kept and centre); T's tables must be the loaded ones
    int props(const Tables &T, const struct potential *d_po, const hp_halo *d_halos,
              const hp_mass_defs &md, hp_halo_props *d_props) {
        const int32_t H = (int32_t)T.nhalos, nst = (int32_t)T.scan_tiles.size();
        if (H == 0) return HP_OK;
        HP_ENSURE(pcnt, sizeof(PropCount) * (nst + 1));
        HP_ENSURE(pscan, sizeof(PropScan) * (nst + 1));
        HP_ENSURE(psums, sizeof(PropSums) * (nst + 1));
        HP_ENSURE(pcall, sizeof(PropCall) * (H + 1));
        HP_ENSURE(shape_st, sizeof(ShapeState) * 2 * (H + 1));
        HP_ENSURE(shape_part, sizeof(ShapePart) * 2 * (nst + 1));
        HP_ENSURE(shape_cd, sizeof(float4) * (T.nrec + 1));
        HP_ENSURE(scale_calls, sizeof(ScaleCall) * (H + 1));
        HP_ENSURE(r_pe, sizeof(double) * (H + 1));
        HP_ENSURE(pe_cnt, sizeof(int32_t) * (nst + 1));
        HP_ENSURE(fit_list, sizeof(int32_t) * (H + 1));
        HP_ENSURE(fit_n, sizeof(int32_t) * 2);
        PropArgs pa;
        const double four_thirds_pi = 4.0 * M_PI / 3.0;  // src/properties.c:393-398
        pa.dens_thresh = md.thresh_dens[0] * four_thirds_pi;
        for (int k = 0; k < 4; k++) pa.d[k] = md.thresh_dens[k + 1] * four_thirds_pi;
        pa.rvir_thresh = md.rvir_dens * four_thirds_pi;
        pa.thresh0 = md.thresh_dens[0];
        pa.thresh3 = md.thresh_dens[3];
        pa.rvir_dens = md.rvir_dens;
        pa.rvir_dens_z0 = md.rvir_dens_z0;
        pa.force_res = FORCE_RES;
        pa.particle_mass = PARTICLE_MASS;
        pa.scale_now = (float)SCALE_NOW;
        hipStream_t s = stream;
        const ScanTile *tl = scan_tiles.as<ScanTile>();
        const int32_t *t0 = call_tile0.as<int32_t>(), *tn = call_ntiles.as<int32_t>();
        const int64_t *hf = hfirst.as<int64_t>(), *hc = hcount.as<int64_t>();
        if (nst > 0) {
            k_props_counts<<<nst, kWave, 0, s>>>(d_po, tl, nst, d_halos, hf, hc,
                                                 pcnt.as<PropCount>());
            HP_LAUNCHED();
        }
        k_props_prefix<<<blocks(H, 256), 256, 0, s>>>(t0, tn, H, pcnt.as<PropCount>());
        HP_LAUNCHED();
        if (nst > 0) {
            k_props_scan<<<nst, kWave, 0, s>>>(d_po, tl, nst, d_halos, hf, hc,
                                               pcnt.as<PropCount>(), pa, pscan.as<PropScan>());
            HP_LAUNCHED();
        }
        k_props_calls<<<blocks(H, 256), 256, 0, s>>>(t0, tn, H, d_halos, hc,
                                                     pscan.as<PropScan>(), pcall.as<PropCall>());
        HP_LAUNCHED();
        if (nst > 0) {
            k_props_sums<<<nst, kWave, 0, s>>>(d_po, tl, nst, d_halos, hf, pcnt.as<PropCount>(),
                                               pcall.as<PropCall>(), pa, psums.as<PropSums>());
            HP_LAUNCHED();
        }
        k_props_final<<<blocks(H, 256), 256, 0, s>>>(d_po, t0, tn, H, d_halos, hf,
                                                     pcall.as<PropCall>(), psums.as<PropSums>(),
                                                     pa, d_props);
        HP_LAUNCHED();
        // calc_shape in both windows, SHAPE_ITERATIONS rounds
        ShapeState *ss = shape_st.as<ShapeState>();
        k_shape_init<<<blocks(H, 256), 256, 0, s>>>(H, d_halos, hc, pcall.as<PropCall>(), d_props,
                                                    pa, ss);
        HP_LAUNCHED();
        if (nst > 0) {
            k_shape_prep<<<nst, kWave, 0, s>>>(d_po, tl, nst, d_halos, hf, ss,
                                               shape_cd.as<float4>());
            HP_LAUNCHED();
        }
        for (int it = 0; it < kShapeIterations; it++) {
            if (nst > 0) {
                k_shape_tiles<<<std::min(nst, 8192), kWave, 0, s>>>(
                    tl, nst, hf, ss, shape_cd.as<float4>(), shape_part.as<ShapePart>());
                HP_LAUNCHED();
            }
            k_shape_calls<<<blocks(2 * (int64_t)H, 128), 128, 0, s>>>(
                H, it, t0, tn, shape_part.as<ShapePart>(), ss);
            HP_LAUNCHED();
        }
        // calc_scale_radius and the pseudo-evolution masses
        k_fill_i32<<<1, 64, 0, s>>>(fit_n.as<int32_t>(), 0, 1);
        HP_LAUNCHED();
        k_scale_init<<<blocks(H, 256), 256, 0, s>>>(H, d_po, hf, t0, tn, pcall.as<PropCall>(),
                                                    psums.as<PropSums>(),
                                                    scale_calls.as<ScaleCall>(),
                                                    fit_list.as<int32_t>(), fit_n.as<int32_t>());
        HP_LAUNCHED();
        if (nst > 0) {
            k_scale_bins<<<nst, kWave, 0, s>>>(d_po, tl, nst, hf, pcnt.as<PropCount>(),
                                               scale_calls.as<ScaleCall>());
            HP_LAUNCHED();
        }
        k_scale_fit<<<blocks(H, 64), 64, 0, s>>>(fit_list.as<int32_t>(), fit_n.as<int32_t>(),
                                                 pcall.as<PropCall>(), d_props, pa,
                                                 scale_calls.as<ScaleCall>());
        HP_LAUNCHED();
        k_props_late<<<blocks(H, 64), 64, 0, s>>>(H, pcall.as<PropCall>(), ss,
                                                  scale_calls.as<ScaleCall>(), pa,
                                                  r_pe.as<double>(), d_props);
        HP_LAUNCHED();
        if (nst > 0) {
            k_pe_tiles<<<nst, kWave, 0, s>>>(d_po, tl, nst, d_halos, hf, hc, r_pe.as<double>(),
                                             pe_cnt.as<int32_t>());
            HP_LAUNCHED();
        }
        k_pe_calls<<<blocks(H, 256), 256, 0, s>>>(H, t0, tn, pe_cnt.as<int32_t>(), pa, d_props);
        HP_LAUNCHED();
        return HP_OK;
    }

    // The bound pass's sort: each call's records by r2 inside the call.  It
    // needs the tree order only, not the potentials, so a recorded step runs
    // it on a branch of its own beside the multipoles and the walk.
    int enqueue_bound_sort(const Tables &T, hipStream_t s) {
        const int64_t N = T.nrec;
        const int32_t H = (int32_t)T.nhalos;
        if (N == 0 || H == 0) return HP_OK;
        // sort input: each call's records in the order the reference's
        // compute_potential and potential-only strip leave them
        const int32_t nst_k = (int32_t)T.scan_tiles.size();
        KeyArgs ka;
        ka.rkey = rkey.as<uint64_t>();
        ka.ncalls = H;
        ka.fin = T.method == HP_METHOD_DIRECT ? nullptr : F.as<float4>();
        ka.tiles = scan_tiles.as<ScanTile>();
        ka.hfirst = hfirst.as<int64_t>();
        ka.hcount = hcount.as<int64_t>();
        ka.pot = pot.as<int32_t>();
        ka.call_tile0 = call_tile0.as<int32_t>();
        ka.call_ntiles = call_ntiles.as<int32_t>();
        ka.tcnt = strip_cnt.as<int32_t>();
        ka.bpos = sides.as<int32_t>();
        ka.keys = keys0.as<uint64_t>();
        ka.vals = vals0.as<int32_t>();
        if (nst_k > 0) {
            const int sg = std::min(nst_k, 2048);  // strip kernels: a tile per workgroup, grid-stride
            k_keys_tree<<<blocks(N, 256), 256, 0, s>>>(ka, rec_call.as<int32_t>(), N);
            HP_LAUNCHED();
            k_strip_count<<<sg, kKeyThreads, 0, s>>>(ka, nst_k);
            HP_LAUNCHED();
            k_strip_prefix<<<blocks(H, 256), 256, 0, s>>>(ka);
            HP_LAUNCHED();
            k_strip_list<<<sg, kKeyThreads, 0, s>>>(ka, nst_k);
            HP_LAUNCHED();
            k_strip_keys<<<sg, kKeyThreads, 0, s>>>(ka, nst_k);
            HP_LAUNCHED();
        }
        // each call's records sorted inside its own range, into keys1 / vals1
        uint64_t *k0 = keys0.as<uint64_t>(), *k1 = keys1.as<uint64_t>();
        int32_t *v0 = vals0.as<int32_t>(), *v1 = vals1.as<int32_t>();
        const int32_t nw = (int32_t)T.seg_wave.size(), nb = (int32_t)T.seg_block.size();
        const int32_t ntb = (int32_t)T.seg_tiles.size();
        if (nw) {
            k_seg_wave<<<blocks((int64_t)nw * kWave, 256), 256, 0, s>>>(
                k0, v0, k1, v1, seg_wave_l.as<int32_t>(), nw, hfirst.as<int64_t>(),
                hcount.as<int64_t>());
            HP_LAUNCHED();
        }
        if (nb) {
            k_seg_block<<<nb, kSortThreads, 0, s>>>(k0, v0, k1, v1, seg_block_l.as<int32_t>(),
                                                    hfirst.as<int64_t>(), hcount.as<int64_t>());
            HP_LAUNCHED();
        }
        if (ntb) {
            static_assert((33 + kSegBits - 1) / kSegBits == 3, "an odd number of passes ends in keys1");
            const int32_t m = ntb * kSegRadix;
            const int32_t nbk = (m + kScanBlock - 1) / kScanBlock;
            uint64_t *ki = k0, *ko = k1;
            int32_t *vi = v0, *vo = v1;
            const SegTile *st = seg_tiles_l.as<SegTile>();
            for (int shift = 0; shift < 33; shift += kSegBits) {
                k_seg_hist<<<ntb, kSortThreads, 0, s>>>(ki, st, shift, seg_hist.as<int32_t>());
                HP_LAUNCHED();
                k_scan_blocks<<<nbk, 1024, 0, s>>>(seg_hist.as<int32_t>(), m, seg_hsum.as<int32_t>());
                HP_LAUNCHED();
                k_scan_top<<<1, 1024, 0, s>>>(seg_hsum.as<int32_t>(), nbk);
                HP_LAUNCHED();
                k_scan_add<<<nbk, 1024, 0, s>>>(seg_hist.as<int32_t>(), m, seg_hsum.as<int32_t>());
                HP_LAUNCHED();
                k_seg_scatter<<<ntb, kSortThreads, 0, s>>>(ki, vi, ko, vo, st, shift,
                                                           seg_hist.as<int32_t>());
                HP_LAUNCHED();
                std::swap(ki, ko);
                std::swap(vi, vo);
            }
        }
        sorted_keys = k1;
        sorted_vals = v1;
        return HP_OK;
    }

    int enqueue_unbind(const Tables &T, struct potential *d_po, hp_halo *d_halos,
                       uint8_t *d_bound) {
        hipStream_t s = stream;
        const int64_t N = T.nrec;
        const int32_t H = (int32_t)T.nhalos;
        if (N == 0 || H == 0) return HP_OK;
        int r;
        if (sort_forked) {  // the sort ran on its branch: join it
            sort_forked = false;
            HP_TRY(hipStreamWaitEvent(s, sort_join_ev, 0));
        } else if ((r = enqueue_bound_sort(T, s))) {
            return r;
        }
        struct potential *S = srt.as<struct potential>();
        float4 *RV = srt_rv.as<float4>();
        k_gather_sorted<<<blocks(N, 256), 256, 0, s>>>(d_po, sorted_vals, (int32_t)N, S, RV);
        HP_LAUNCHED();
        CallInfo *ci = callinfo.as<CallInfo>();
        const int32_t nst = (int32_t)T.scan_tiles.size();
        int32_t *ti32 = tile_i32.as<int32_t>();
        k_call_kept<<<blocks(H, 256), 256, 0, s>>>(sorted_keys, hfirst.as<int64_t>(),
                                                   hcount.as<int64_t>(), H, ci);
        HP_LAUNCHED();
        const double rvir_thresh = particle_rvir_dens() * (4.0 * M_PI / 3.0);
        if (nst > 0) {
            k_rvir_vel_tiles<<<nst, kWave, 0, s>>>(RV, scan_tiles.as<ScanTile>(), nst, ci,
                                                   rvir_thresh, ti32, tsum.as<Moments>());
            HP_LAUNCHED();
        }
        k_rvir_calls<<<blocks(H, 256), 256, 0, s>>>(RV, call_tile0.as<int32_t>(),
                                                    call_ntiles.as<int32_t>(), ti32, H, ci);
        HP_LAUNCHED();
        if (nst > 0) {
            k_vel_prefix<<<blocks(H, 256), 256, 0, s>>>(call_tile0.as<int32_t>(),
                                                        call_ntiles.as<int32_t>(), H,
                                                        tsum.as<Moments>());
            HP_LAUNCHED();
            k_vel_select<<<nst, kWave, 0, s>>>(RV, scan_tiles.as<ScanTile>(), nst, ci,
                                             tsum.as<Moments>(), tcand.as<Cand>());
            HP_LAUNCHED();
        }
        const int32_t *flags = counters.as<int32_t>() + kErrFlags;
        k_vel_finish<<<blocks(H, 256), 256, 0, s>>>(RV, call_tile0.as<int32_t>(),
                                                    call_ntiles.as<int32_t>(), H, ci,
                                                    tcand.as<Cand>(), d_halos, flags);
        HP_LAUNCHED();
        // kinetic energy (src/potential.c:194-211 with scale_now = SCALE_NOW)
        const float a = (float)SCALE_NOW;
        const double conv = 0.5 * a / 4.30117902e-9;
        const double z1 = 1.0 / a;
        const double hub = 100.0 * hubble_scaling(z1 - 1.0);
        const double hub_a = hub * a;
        if (nst > 0) {
            k_kinetic_tiles<<<nst, kKinThreads, 0, s>>>(S, d_po, scan_tiles.as<ScanTile>(), nst, ci,
                                                d_halos, hub_a, conv, d_bound, ti32, flags);
            HP_LAUNCHED();
        }
        if (d_bound) {
            k_bound_tail<<<H, 256, 0, s>>>(hfirst.as<int64_t>(), hcount.as<int64_t>(), ci, H, d_bound,
                                           flags);
            HP_LAUNCHED();
        }
        k_nbound_calls<<<blocks(H, 256), 256, 0, s>>>(call_tile0.as<int32_t>(),
                                                      call_ntiles.as<int32_t>(), ti32, H, d_halos,
                                                      flags);
        HP_LAUNCHED();
        return HP_OK;
    }
};

// One engine per host thread for the reference entry points.
static Engine *thread_engine(int *err) {
    static thread_local Engine *e = nullptr;
    if (!e) {
        e = new Engine();
        int r = e->init();
        if (r) {
            delete e;
            e = nullptr;
            *err = r;
            return nullptr;
        }
    }
    *err = HP_OK;
    return e;
}

// calc_particle_radii (src/groupies.c:235-268) as a segment list: for each
// base halo, the halos its recursion visits in order, each as a run of
// particles [p_start, p_start + num_p) that becomes the next run of records.
// The recursion is kept on an explicit stack; a halo already on the descent
// path is skipped (src/groupies.c:240-246) and a sibling list that returns
// to the first child is an error (src/groupies.c:254).
static int plan_gather(const hp_tree_halo *tree, int64_t nh, const int64_t *base, int64_t nbase,
                       int64_t nparticles, double ratio, std::vector<GatherSeg> &segs,
                       std::vector<float> &cen, hp_halo *calls, int64_t *total) {
    struct Frame {
        int64_t h, child;
        int64_t steps;      // sibling steps taken (a list longer than the tree is a cycle)
        int32_t potential_only;
        int32_t done;       // children visited (and the parent rule applied)
    };
    std::vector<Frame> st;
    std::vector<int64_t> path;
    int64_t pos = 0;
    segs.clear();
    cen.clear();
    for (int64_t b = 0; b < nbase; b++) {
        const int64_t bh = base[b];
        if (bh < 0 || bh >= nh)
            return set_error(HP_EINVAL, "hp_gather_calls: base[%lld] = %lld out of range",
                             (long long)b, (long long)bh);
        const int64_t first = pos;
        const double small = (double)tree[bh].num_p * 0.03;
        int r = HP_OK;
        // src/groupies.c:212-233, 240-250: enter halo h unless it is on the path
        auto enter = [&](int64_t h, int32_t potential_only) -> int {
            if (h < 0 || h >= nh)
                return set_error(HP_EINVAL, "hp_gather_calls: halo link %lld out of range",
                                 (long long)h);
            for (int64_t q : path)
                if (q == h) return HP_OK;
            const hp_tree_halo &t = tree[h];
            if (t.num_p < 0 || t.p_start < 0 || t.p_start + t.num_p > nparticles)
                return set_error(HP_EINVAL, "hp_gather_calls: halo %lld particles [%lld, +%lld) "
                                 "outside the %lld copies", (long long)h, (long long)t.p_start,
                                 (long long)t.num_p, (long long)nparticles);
            int32_t flags = h == bh ? 1 : 0;
            if (!potential_only && (double)t.num_p < small) flags = 2;
            if (t.num_p > 0)
                segs.push_back(GatherSeg{t.p_start, pos, (int32_t)t.num_p, (int32_t)b, flags,
                                         potential_only});
            pos += t.num_p;
            if (pos > INT32_MAX / 2)
                return set_error(HP_EINVAL, "hp_gather_calls: more than %d records", INT32_MAX / 2);
            path.push_back(h);
            st.push_back(Frame{h, t.child, 0, potential_only, 0});
            return HP_OK;
        };
        if ((r = enter(bh, 0))) return r;
        while (!st.empty()) {
            Frame &f = st.back();
            if (!f.done && f.child > -1) {
                const int64_t c = f.child;
                if (c >= nh)
                    return set_error(HP_EINVAL, "hp_gather_calls: child link %lld out of range",
                                     (long long)c);
                f.child = tree[c].next_cochild;
                if (f.child > -1 && f.child == tree[f.h].child)
                    return set_error(HP_EINVAL, "hp_gather_calls: sibling list of halo %lld "
                                     "returns to its first child", (long long)f.h);
                if (++f.steps > nh)
                    return set_error(HP_EINVAL, "hp_gather_calls: sibling list of halo %lld "
                                     "does not end", (long long)f.h);
                if ((r = enter(c, f.potential_only))) return r;
                continue;
            }
            if (!f.done) {
                f.done = 1;
                // src/groupies.c:257-265: the host's subtree as potential-only
                // records when this sub-halo holds a large enough share
                const int64_t par = tree[f.h].sub_of;
                if (f.h == bh && par > -1) {
                    if (par >= nh)
                        return set_error(HP_EINVAL, "hp_gather_calls: parent link %lld out of "
                                         "range", (long long)par);
                    if ((double)tree[par].num_child_particles * ratio <
                        (double)tree[bh].num_child_particles) {
                        if ((r = enter(par, 1))) return r;
                        continue;
                    }
                }
            }
            path.pop_back();
            st.pop_back();
        }
        for (int k = 0; k < 3; k++) cen.push_back(tree[bh].pos[k]);
        hp_halo &hc = calls[b];
        memset(&hc, 0, sizeof hc);
        hc.first = first;
        hc.count = pos - first;
        for (int k = 0; k < 6; k++) hc.pos[k] = tree[bh].pos[k];
        hc.is_sub = tree[bh].sub_of > -1 ? 1 : 0;
    }
    *total = pos;
    return HP_OK;
}

static hp_options default_options() {
    hp_options o;
    memset(&o, 0, sizeof o);
    o.accumulate = HP_ACC_COMPENSATED;
    o.method = HP_METHOD_TREE;
    return o;
}

}  // namespace hp

namespace hp {

// ----------------------------------------------------------- no allocation
//
// hp_compute_potential_noalloc works in one file-scope device array per GPU:
// every buffer its run needs (points, tree, tasks, counters, the tree-order
// staging of the records) is carved from it, sized for at most
// kNoAllocMaxRecords records.  Host code never passes the array's name to a
// kernel: its address on each device comes from hipGetSymbolAddress, taken
// once per device with that device current, and kept.
constexpr int64_t kNoAllocMaxRecords = 32768;
constexpr size_t kNoAllocScratchBytes = (size_t)20 << 20;  // needs about 16 MiB at the maximum
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kNoAllocDevices = 64;
static std::mutex g_noalloc_mu;
static Engine *g_noalloc[kNoAllocDevices];  // per device; kept until the process ends

// The engine of the current device, set up on the first call there.
static Engine *noalloc_engine(int *err) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) {
        (void)hipGetLastError();
        *err = set_error(HP_ENODEV, "hp_compute_potential_noalloc: no usable current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            *err = set_error(HP_EHIP, "hipGetSymbolAddress of the scratch array failed on "
                             "device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        Engine *en = new Engine();
        const int r = en->init_static(ptr, kNoAllocScratchBytes);
        if (r) {
            delete en;
            *err = r;
            return nullptr;
        }
        g_noalloc[dev] = en;
    }
    *err = HP_OK;
    return g_noalloc[dev];
}

}  // namespace hp

// ------------------------------------------------------------------ C ABI

struct hp_plan {
    hp::Engine eng;
    hp::Tables T;
    std::vector<int64_t> first, count;  // the calls, for re-planning
    std::vector<hp::Tables> per;  // per_call: one single-call table per call
    hp_options opt;
};

using namespace hp;

extern "C" {

__attribute__((visibility("default"))) const char *hp_last_error(void) {
    return g_last_error.c_str();
}

__attribute__((visibility("default"))) void hp_set_error_handler(hp_error_handler h) {
    g_handler.store(h);
}

__attribute__((visibility("default"))) void compute_potential(struct potential *po,
                                                               int64_t num_po) {
    if (num_po <= 0) return;
    int r;
    Tables T;
    int64_t first = 0, count = num_po;
    if ((r = make_tables(T, &first, &count, 1, HP_METHOD_TREE, false)))
        return fatal(r, __FILE__, __LINE__);
    Engine *e = thread_engine(&r);
    if (!e) return fatal(r, __FILE__, __LINE__);
    hp_options o = default_options();
    if ((r = e->reserve(T, false)) || (r = e->load_tables(T))) return fatal(r, __FILE__, __LINE__);
    DevBuf &dpo = e->srt;  // record staging
    if ((r = dpo.ensure(sizeof(struct potential) * num_po))) return fatal(r, __FILE__, __LINE__);
    if (hipMemcpyAsync(dpo.p, po, sizeof(struct potential) * num_po, hipMemcpyHostToDevice,
                       e->stream) != hipSuccess) {
        set_error(HP_EHIP, "copy of %lld records to the device failed", (long long)num_po);
        return fatal(HP_EHIP, __FILE__, __LINE__);
    }
    const struct potential *res = dpo.as<struct potential>();
    bool deferred = false;
    if ((r = e->potential(T, dpo.as<struct potential>(), o, true, &res, &deferred)))
        return fatal(r, __FILE__, __LINE__);
    // the records and, for a deferred pass, the counters, with one wait
    if (hipMemcpyAsync(po, res, sizeof(struct potential) * num_po, hipMemcpyDeviceToHost,
                       e->stream) != hipSuccess ||
        (deferred && hipMemcpyAsync(e->h_counters, e->counters.p, sizeof(int32_t) * kNumCounters,
                                    hipMemcpyDeviceToHost, e->stream) != hipSuccess) ||
        hipStreamSynchronize(e->stream) != hipSuccess) {
        set_error(HP_EHIP, "copy of %lld records from the device failed", (long long)num_po);
        return fatal(HP_EHIP, __FILE__, __LINE__);
    }
    if (deferred && (r = e->finish_checks())) return fatal(r, __FILE__, __LINE__);
}

__attribute__((visibility("default"))) int hp_compute_potential_noalloc(struct potential *d_po,
                                                                        int64_t num_po,
                                                                        void *stream) {
    if (num_po < 0 || (num_po > 0 && !d_po))
        return set_error(HP_EINVAL, "hp_compute_potential_noalloc: bad arguments");
    if (num_po > kNoAllocMaxRecords)
        return set_error(HP_EINVAL, "hp_compute_potential_noalloc: %lld records; at most %lld "
                         "without allocating (use compute_potential or a plan)",
                         (long long)num_po, (long long)kNoAllocMaxRecords);
    if (num_po == 0) return HP_OK;
    int r;
    Tables T;
    int64_t first = 0, count = num_po;
    if ((r = make_tables(T, &first, &count, 1, HP_METHOD_TREE, false))) return r;
    Engine *e = noalloc_engine(&r);
    if (!e) return r;
    e->stream = static_cast<hipStream_t>(stream);
    e->reset_carve();
    if ((r = e->reserve(T, false)) || (r = e->load_tables(T))) return r;
    return e->potential(T, d_po, default_options(), true);
}

__attribute__((visibility("default"))) void compute_kinetic_energy(struct potential *po,
                                                                    int64_t num_po,
                                                                    float *vel_cen,
                                                                    float *pos_cen,
                                                                    float scale_now) {
    if (num_po <= 0) return;
    int r;
    if (num_po > kMaxRecords || !vel_cen || !pos_cen) {
        set_error(HP_EINVAL, "compute_kinetic_energy: bad arguments (%lld records)",
                  (long long)num_po);
        return fatal(HP_EINVAL, __FILE__, __LINE__);
    }
    Engine *e = thread_engine(&r);
    if (!e) return fatal(r, __FILE__, __LINE__);
    DevBuf &dpo = e->srt;
    if ((r = dpo.ensure(sizeof(struct potential) * num_po))) return fatal(r, __FILE__, __LINE__);
    const double conv = 0.5 * scale_now / 4.30117902e-9;
    const double z1 = 1.0 / scale_now;
    const double hub = 100.0 * hubble_scaling(z1 - 1.0);
    const double hub_a = hub * scale_now;
    if (hipMemcpyAsync(dpo.p, po, sizeof(struct potential) * num_po, hipMemcpyHostToDevice,
                       e->stream) != hipSuccess) {
        set_error(HP_EHIP, "copy to the device failed");
        return fatal(HP_EHIP, __FILE__, __LINE__);
    }
    k_kinetic_plain<<<Engine::blocks(num_po, 256), 256, 0, e->stream>>>(
        dpo.as<struct potential>(), num_po, vel_cen[0], vel_cen[1], vel_cen[2], pos_cen[0],
        pos_cen[1], pos_cen[2], hub_a, conv);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(po, dpo.p, sizeof(struct potential) * num_po, hipMemcpyDeviceToHost,
                       e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess) {
        set_error(HP_EHIP, "kinetic energy kernel failed");
        return fatal(HP_EHIP, __FILE__, __LINE__);
    }
}

__attribute__((visibility("default"))) int hp_plan_create(hp_plan **out, const hp_halo *halos,
                                                          int64_t nhalos,
                                                          const hp_options *opt) {
    if (!out || (nhalos > 0 && !halos) || nhalos < 0)
        return set_error(HP_EINVAL, "hp_plan_create: bad arguments");
    if (opt && ((uint32_t)opt->accumulate > HP_ACC_F64 || (uint32_t)opt->method > HP_METHOD_AUTO ||
                (uint32_t)opt->walk > HP_WALK_NODES))
        return set_error(HP_EINVAL, "hp_plan_create: bad options");
    hp_plan *p = new hp_plan();
    p->opt = opt ? *opt : default_options();
    // the calls are checked before any device work
    std::vector<int64_t> f(nhalos), c(nhalos);
    for (int64_t h = 0; h < nhalos; h++) {
        f[h] = halos[h].first;
        c[h] = halos[h].count;
    }
    int r = make_tables(p->T, f.data(), c.data(), nhalos, p->opt.method, true);
    if (!r) r = p->eng.init();
    if (!r) {
        p->first = f;
        p->count = c;
        if (p->opt.per_call) {
            p->per.resize(nhalos);
            int64_t big = 0, bigh = 0;
            for (int64_t h = 0; h < nhalos && !r; h++) {
                int64_t f0 = 0, c0 = c[h];
                r = make_tables(p->per[h], &f0, &c0, 1, p->opt.method, true);
                if (c0 > big) {
                    big = c0;
                    bigh = h;
                }
            }
            if (!r && nhalos) r = p->eng.reserve(p->per[bigh], true);
        }
    }
    if (!r && !p->opt.per_call) r = p->eng.reserve(p->T, true);
    if (!r) r = p->eng.load_tables(p->T);
    if (!r && hipStreamSynchronize(p->eng.stream) != hipSuccess)
        r = set_error(HP_EHIP, "plan upload failed");
    if (r) {
        delete p;
        return r;
    }
    *out = p;
    return HP_OK;
}

// Makes `t` the plan's launch setting (graphs are recorded again on the next
// run).
static int apply_tuning(hp_plan *p, const hp_tuning &t) {
    Engine &e = p->eng;
    int r;
    for (int b = 0; b < kTuneBuckets; b++) {
        e.bucket_big[b] = t.big_threshold[b];
        e.bucket_wave[b] = t.wave_max[b];
    }
    e.drop_graphs();
    if (p->opt.method == HP_METHOD_AUTO && t.direct_max != p->T.direct_max) {
        if ((r = make_tables(p->T, p->first.data(), p->count.data(), (int64_t)p->first.size(),
                             p->opt.method, true, t.direct_max)) ||
            (r = e.load_tables(p->T)))
            return r;
    }
    if ((r = e.reserve(p->T, true))) return r;
    HP_TRY(hipStreamSynchronize(e.stream));
    return HP_OK;
}

// Times the plan's potential pass on the caller's records for settings that
// change no tree, decision or sum, and keeps the fastest: per call-size
// bucket, the size above which nodes are split by the level passes and the
// size up to which a subtree is built by one wavefront; for HP_METHOD_AUTO
// plans, whether calls of at most 10 points use the direct kernel.  First
// one setting for every bucket, then each bucket that holds calls in turn,
// the others fixed.
__attribute__((visibility("default"))) int hp_plan_tune(hp_plan *p, struct potential *d_po,
                                                        int reps, hp_tuning *out) {
    if (!p || !d_po || reps < 1) return set_error(HP_EINVAL, "hp_plan_tune: bad arguments");
    if (p->opt.per_call) return set_error(HP_EINVAL, "hp_plan_tune: per-call plans are not tuned");
    Engine &e = p->eng;
    hp_options o = p->opt;
    o.timing = 0;
    o.count_work = 0;
    const int32_t bts[] = {128, 256, 512, 1024};
    const int32_t wms[] = {kWaveMin, kWaveTask};
    const int32_t dms[] = {kLeafPoints, 0};
    const int ndm = p->opt.method == HP_METHOD_AUTO ? 2 : 1;
    int r;
    hipEvent_t e0 = e.ev[5], e1 = e.ev[6];
    // the potential pass under setting t (its direct crossover already loaded)
    auto time_setting = [&](const hp_tuning &t, float *ms) -> int {
        for (int b = 0; b < kTuneBuckets; b++) {
            e.bucket_big[b] = t.big_threshold[b];
            e.bucket_wave[b] = t.wave_max[b];
        }
        e.drop_graphs();
        int rr;
        if ((rr = e.reserve(p->T, true)) || (rr = e.potential(p->T, d_po, o))) return rr;
        HP_TRY(hipEventRecord(e0, e.stream));
        for (int k = 0; k < reps; k++)
            if ((rr = e.potential(p->T, d_po, o))) return rr;
        HP_TRY(hipEventRecord(e1, e.stream));
        HP_TRY(hipEventSynchronize(e1));
        HP_TRY(hipEventElapsedTime(ms, e0, e1));
        *ms /= (float)reps;
        return HP_OK;
    };
    bool used[kTuneBuckets] = {};
    for (int64_t c : p->T.count) used[tune_bucket(c)] = true;
    hp_tuning best{};
    float best_ms = -1.0f;
    for (int di = 0; di < ndm; di++) {
        if (ndm > 1) {
            if ((r = make_tables(p->T, p->first.data(), p->count.data(), (int64_t)p->first.size(),
                                 p->opt.method, true, dms[di])) ||
                (r = e.load_tables(p->T)))
                return r;
        }
        for (int32_t bt : bts)
            for (int32_t wm : wms) {
                hp_tuning t{};
                for (int b = 0; b < kTuneBuckets; b++) {
                    t.big_threshold[b] = bt;
                    t.wave_max[b] = wm;
                }
                t.direct_max = ndm > 1 ? dms[di] : p->T.direct_max;
                float ms = 0.0f;
                if ((r = time_setting(t, &ms))) return r;
                if (best_ms < 0.0f || ms < best_ms) {
                    best_ms = ms;
                    best = t;
                    best.ms_potential = ms;
                }
            }
    }
    if (ndm > 1 && best.direct_max != p->T.direct_max) {
        if ((r = make_tables(p->T, p->first.data(), p->count.data(), (int64_t)p->first.size(),
                             p->opt.method, true, best.direct_max)) ||
            (r = e.load_tables(p->T)))
            return r;
    }
    for (int b = 0; b < kTuneBuckets; b++) {
        if (!used[b]) continue;
        for (int32_t bt : bts)
            for (int32_t wm : wms) {
                if (bt == best.big_threshold[b] && wm == best.wave_max[b]) continue;
                hp_tuning t = best;
                t.big_threshold[b] = bt;
                t.wave_max[b] = wm;
                float ms = 0.0f;
                if ((r = time_setting(t, &ms))) return r;
                if (ms < best_ms) {
                    best_ms = ms;
                    best = t;
                    best.ms_potential = ms;
                }
            }
    }
    if ((r = apply_tuning(p, best))) return r;
    if (out) *out = best;
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_plan_set_tuning(hp_plan *p, const hp_tuning *t) {
    if (!p || !t) return set_error(HP_EINVAL, "hp_plan_set_tuning: bad arguments");
    if (p->opt.per_call) return set_error(HP_EINVAL, "hp_plan_set_tuning: per-call plans are not tuned");
    for (int b = 0; b < kTuneBuckets; b++) {
        const int32_t bt = t->big_threshold[b], wm = t->wave_max[b];
        if ((bt != 128 && bt != 256 && bt != 512 && bt != 1024) || (wm != kWaveMin && wm != kWaveTask))
            return set_error(HP_EINVAL, "hp_plan_set_tuning: bucket %d: big_threshold %d / wave_max "
                             "%d is not a setting of the tuner", b, bt, wm);
    }
    const int32_t dm = t->direct_max;
    if (p->opt.method == HP_METHOD_AUTO ? (dm != 0 && dm != kLeafPoints) : dm != p->T.direct_max)
        return set_error(HP_EINVAL, "hp_plan_set_tuning: direct_max %d not allowed for this plan",
                         dm);
    return apply_tuning(p, *t);
}

// The plain per-halo path: every call with its own launches, one after
// another on the plan's stream.
static int run_per_call(hp_plan *p, struct potential *d_po, hp_halo *d_halos, uint8_t *d_bound,
                        bool unbind) {
    int r;
    for (size_t h = 0; h < p->per.size(); h++) {
        const Tables &t = p->per[h];
        if (t.nrec == 0) continue;
        struct potential *po_h = d_po + p->T.first[h];
        if ((r = p->eng.load_tables(t))) return r;
        if ((r = p->eng.potential(t, po_h, p->opt, !unbind))) return r;
        if (unbind && (r = p->eng.enqueue_unbind(t, po_h, d_halos + h,
                                                 d_bound ? d_bound + p->T.first[h] : nullptr)))
            return r;
    }
    HP_TRY(hipStreamSynchronize(p->eng.stream));
    return HP_OK;
}

static int plan_potential(hp_plan *p, struct potential *d_po) {
    if (p->opt.per_call) return run_per_call(p, d_po, nullptr, nullptr, false);
    return p->eng.potential(p->T, d_po, p->opt, true);
}

// The entry points below run after the work queued on the caller's stream
// and before the work the caller queues on it afterwards.
__attribute__((visibility("default"))) int hp_plan_potential(hp_plan *p, struct potential *d_po,
                                                             void *stream) {
    if (!p || (!d_po && p->T.nrec)) return set_error(HP_EINVAL, "hp_plan_potential: bad arguments");
    int r = p->eng.join(stream);
    if (!r) r = plan_potential(p, d_po);
    if (!r) r = p->eng.release(stream);
    return r;
}

static int plan_run(hp_plan *p, struct potential *d_po, hp_halo *d_halos, uint8_t *d_bound) {
    if (p->opt.per_call) return run_per_call(p, d_po, d_halos, d_bound, true);
    int r;
    if (p->opt.use_graph && p->opt.method != HP_METHOD_DIRECT && !p->opt.count_work) {
        if ((r = p->eng.run_graphs(p->T, d_po, d_halos, d_bound, p->opt))) return r;
        HP_TRY(hipStreamSynchronize(p->eng.stream));
        return HP_OK;
    }
    if ((r = p->eng.potential(p->T, d_po, p->opt))) return r;
    hipEvent_t e0 = p->eng.ev[5], e1 = p->eng.ev[6];
    if (p->opt.timing) HP_TRY(hipEventRecord(e0, p->eng.stream));
    if ((r = p->eng.enqueue_unbind(p->T, d_po, d_halos, d_bound))) return r;
    if (p->opt.timing) {
        HP_TRY(hipEventRecord(e1, p->eng.stream));
        HP_TRY(hipEventSynchronize(e1));
        float t;
        HP_TRY(hipEventElapsedTime(&t, e0, e1));
        p->eng.st.ms_unbind = t;
        HP_TRY(hipEventElapsedTime(&t, p->eng.ev[0], e1));
        p->eng.st.ms_total = t;
    }
    HP_TRY(hipStreamSynchronize(p->eng.stream));
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_plan_run(hp_plan *p, struct potential *d_po,
                                                       hp_halo *d_halos, uint8_t *d_bound,
                                                       void *stream) {
    if (!p || ((!d_po || !d_halos) && p->T.nrec))
        return set_error(HP_EINVAL, "hp_plan_run: bad arguments");
    int r = p->eng.join(stream);
    if (!r) r = plan_run(p, d_po, d_halos, d_bound);
    if (!r) r = p->eng.release(stream);
    return r;
}

__attribute__((visibility("default"))) int hp_gather_calls(
    const hp_tree_halo *tree, int64_t nhalos, const int64_t *base, int64_t nbase,
    const struct particle *d_particles, int64_t nparticles, double include_host_ratio,
    struct potential *d_po, int64_t cap_records, hp_halo *calls, int64_t *nrecords) {
    if (nhalos < 0 || nbase < 0 || nparticles < 0 || cap_records < 0 || !nrecords ||
        (nbase > 0 && (!tree || !base || !calls)) || (nparticles > 0 && !d_particles))
        return set_error(HP_EINVAL, "hp_gather_calls: bad arguments");
    std::vector<GatherSeg> segs;
    std::vector<float> cen;
    int64_t total = 0;
    int r = plan_gather(tree, nhalos, base, nbase, nparticles, include_host_ratio, segs, cen,
                        calls, &total);
    if (r) return r;
    if (total > cap_records)
        return set_error(HP_EINVAL, "hp_gather_calls: %lld records do not fit the %lld given",
                         (long long)total, (long long)cap_records);
    if (total > 0 && !d_po) return set_error(HP_EINVAL, "hp_gather_calls: bad arguments");
    *nrecords = total;
    if (segs.empty()) return HP_OK;
    std::vector<GatherTile> tiles;
    for (size_t i = 0; i < segs.size(); i++)
        for (int32_t off = 0; off < segs[i].n; off += kGatherTile)
            tiles.push_back(GatherTile{(int32_t)i, off});
    Engine *e = thread_engine(&r);
    if (!e) return r;
    HP_TRY(hipSetDevice(e->device));
    if ((r = e->upload(e->gsegs, segs)) || (r = e->upload(e->gtiles, tiles)) ||
        (r = e->upload(e->gcen, cen)))
        return r;
    const int32_t nt = (int32_t)tiles.size();
    k_gather_records<<<std::min(nt, 8192), 256, 0, e->stream>>>(
        reinterpret_cast<const GParticle *>(d_particles), e->gsegs.as<GatherSeg>(),
        e->gtiles.as<GatherTile>(), nt, e->gcen.as<float>(), d_po);
    HP_TRY(hipGetLastError());
    HP_TRY(hipStreamSynchronize(e->stream));
    return HP_OK;
}

static_assert(sizeof(hp_fof_group) == 40 && offsetof(hp_fof_group, target_r) == 20 &&
                  offsetof(hp_fof_group, first_sub) == 24,
              "hp_fof_group layout");
static_assert(sizeof(hp_subfof) == 16 && sizeof(hp_subfof_config) == 32, "sub-FOF layouts");

__attribute__((visibility("default"))) int hp_find_subfofs(struct particle *d_particles,
                                                           int64_t nparticles, hp_fof_group *groups,
                                                           int64_t ngroups,
                                                           const hp_subfof_config *config,
                                                           hp_subfof *subs, int64_t cap_subs,
                                                           int64_t *nsubs) {
    if (nparticles < 0 || ngroups < 0 || cap_subs < 0 || !nsubs || (ngroups > 0 && !groups) ||
        (nparticles > 0 && !d_particles) || (cap_subs > 0 && !subs))
        return set_error(HP_EINVAL, "hp_find_subfofs: bad arguments");
    const hp_subfof_config dflt = {0.7, 1.0, 10, 0, 0};  // the reference's defaults
    const hp_subfof_config cfg = config ? *config : dflt;
    *nsubs = 0;
    if (!(cfg.fof_fraction >= 0.0 && cfg.fof_fraction < 1.0) || cfg.min_halo_particles < 0)
        return set_error(HP_EINVAL, "hp_find_subfofs: FOF_FRACTION must be in [0, 1)");
    if (ngroups == 0) return HP_OK;
    // host tables: working slots, norm_sd tiles, the sample of each group
    const float thresh = (float)cfg.fof_fraction;
    std::vector<SubGroup> G((size_t)ngroups);
    std::vector<SubTile> tiles, qtiles, ltiles;
    std::vector<int32_t> drawn, qidx;
    std::vector<int64_t> drawn_groups;
    int64_t ndrawn = 0;
    std::vector<std::pair<int64_t, int64_t>> ranges;
    int64_t M = 0, S = 0;
    for (int64_t g = 0; g < ngroups; g++) {
        const int64_t n = groups[g].num_p, f = groups[g].first;
        if (n < 0 || f < 0 || f + n > nparticles)
            return set_error(HP_EINVAL, "hp_find_subfofs: group %lld particles [%lld, +%lld) "
                             "outside the %lld given", (long long)g, (long long)f, (long long)n,
                             (long long)nparticles);
        if (M + n > (1 << 29)) return set_error(HP_EINVAL, "hp_find_subfofs: batch too large");
        ranges.emplace_back(f, n);
        SubGroup &sg = G[(size_t)g];
        sg.first = f;
        sg.off = (int32_t)M;
        sg.n = (int32_t)n;
        sg.top = groups[g].top ? 1 : 0;
        // src/groupies.c:290-311: every particle when there are at most
        // MAX_PARTICLES_TO_SAMPLE (or EXACT_LL_CALC), else rand_r draws
        int64_t ns = cfg.exact_ll_calc ? n : kSampleMax;
        if (ns > n) ns = n;
        uint32_t seed = (uint32_t)n;
        sg.nsample = (int32_t)ns;
        sg.sample_off = (int32_t)S;
        sg.drawn = ns != n;
        // rand_r draws only for the groups that draw (the others search
        // 0..n-1); made below, several groups at a time
        sg.draw_off = (int32_t)ndrawn;
        if (sg.drawn) {
            drawn_groups.push_back(g);
            ndrawn += ns;
        }
        sg.seed = seed;  // advanced past the draws below for the drawn groups
        for (int64_t off = 0; off < n; off += kWave)
            ltiles.push_back(SubTile{(int32_t)g, (int32_t)off});
        sg.k = (int32_t)(int64_t)((float)ns * thresh);  // src/groupies.c:771
        if (ns > 0 && sg.k >= ns)
            return set_error(HP_EINVAL, "hp_find_subfofs: FOF_FRACTION selects past the sample");
        sg.tile0 = (int32_t)tiles.size();
        for (int64_t off = 0; off < n; off += kSubTile)
            tiles.push_back(SubTile{(int32_t)g, (int32_t)off});
        sg.ntiles = (int32_t)(tiles.size() - sg.tile0);
        for (int64_t off = 0; off < ns; off += kWave)
            qtiles.push_back(SubTile{(int32_t)g, (int32_t)off});
        M += n;
        S += ns;
    }
    // src/groupies.c:300-304 for each drawn group: the sample positions in
    // draw order (seed = num_p), then sorted by position (counting sort,
    // stable in draw order) so that a wavefront's queries are neighbours in
    // the tree; qidx keeps each draw's place in the sample, the order
    // find_median_r sees.  Groups are independent: several host threads.
    drawn.resize((size_t)ndrawn);
    qidx.resize((size_t)ndrawn);
    auto draw_group = [&](int64_t g) {
        SubGroup &sg = G[(size_t)g];
        const int64_t n = sg.n, ns = sg.nsample;
        uint32_t seed = (uint32_t)n;
        std::vector<int32_t> js((size_t)ns), start((size_t)n + 1, 0);
        for (int64_t i = 0; i < ns; i++) {
            int64_t j = rand_r_glibc(&seed);
            j <<= 31;
            j += rand_r_glibc(&seed);
            j %= n;
            js[(size_t)i] = (int32_t)j;
            start[(size_t)j + 1]++;
        }
        for (int64_t j = 0; j < n; j++) start[(size_t)j + 1] += start[(size_t)j];
        const size_t base = (size_t)sg.draw_off;
        for (int64_t i = 0; i < ns; i++) {
            const int32_t at = start[(size_t)js[(size_t)i]]++;
            drawn[base + (size_t)at] = js[(size_t)i];
            qidx[base + (size_t)at] = (int32_t)i;
        }
        sg.seed = seed;
    };
    {
        const size_t nd = drawn_groups.size();
        const size_t nthr = std::min<size_t>(nd, std::min(16u, std::max(1u, std::thread::hardware_concurrency())));
        if (nthr <= 1) {
            for (int64_t g : drawn_groups) draw_group(g);
        } else {
            std::atomic<size_t> next{0};
            std::vector<std::thread> pool;
            for (size_t w = 0; w < nthr; w++)
                pool.emplace_back([&] {
                    for (size_t k; (k = next.fetch_add(1)) < nd;) draw_group(drawn_groups[k]);
                });
            for (auto &th : pool) th.join();
        }
    }
    std::sort(ranges.begin(), ranges.end());
    for (size_t i = 1; i < ranges.size(); i++)
        if (ranges[i].second && ranges[i - 1].first + ranges[i - 1].second > ranges[i].first)
            return set_error(HP_EINVAL, "hp_find_subfofs: groups overlap at particle %lld",
                             (long long)ranges[i].first);
    int r;
    Engine *e = thread_engine(&r);
    if (!e) return r;
    HP_TRY(hipSetDevice(e->device));
    SubBufs &B = e->sub;
    const int64_t nt = (int64_t)tiles.size(), ng = ngroups;
    const int64_t node_cap = 2 * M + ng + 2;
    const int64_t cap_emit = M / 2 + ng + 1;
    HP_ENSURE(B.psum, sizeof(double) * 6 * (nt + 1));
    HP_ENSURE(B.mean, sizeof(double) * 6 * ng);
    HP_ENSURE(B.pcorr, sizeof(double) * 21 * (nt + 1));
    HP_ENSURE(B.sig, sizeof(double) * 2 * ng);
    HP_ENSURE(B.pbox, sizeof(float) * 12 * (nt + 1));
    HP_ENSURE(B.gerr, sizeof(int32_t) * ng);
    HP_ENSURE(B.w, sizeof(SParticle) * (M + 1));
    HP_ENSURE(B.w2, sizeof(SParticle) * (M + 1));
    HP_ENSURE(B.rk, sizeof(int32_t) * (M + 1));
    HP_ENSURE(B.bpos, sizeof(int32_t) * (M + 1));
    HP_ENSURE(B.spos, sizeof(int32_t) * (M + 1));
    HP_ENSURE(B.nodes, sizeof(Node6) * node_cap);
    HP_ENSURE(B.act0, 2 * sizeof(int32_t) * (M / (kLeaf6 + 1) + ng + 1));
    HP_ENSURE(B.act1, 2 * sizeof(int32_t) * (M / (kLeaf6 + 1) + ng + 1));
    HP_ENSURE(B.fused, sizeof(int32_t) * (M / (kLeaf6 + 1) + ng + 1));
    HP_ENSURE(B.counters, sizeof(int32_t) * kSfCounters);
    HP_ENSURE(B.slot_group, sizeof(int32_t) * (M + 1));
    HP_ENSURE(B.dist, sizeof(float) * (S + 1));
    HP_ENSURE(B.sel, sizeof(float) * (S + 1));
    HP_ENSURE(B.sel2, sizeof(float) * (S + 1));
    HP_ENSURE(B.target, sizeof(float) * ng);
    HP_ENSURE(B.csize, sizeof(int32_t) * (M + 1));
    HP_ENSURE(B.emit, sizeof(int4) * cap_emit);
    // the linking's pair list starts at 8 per particle and grows to what the
    // walk reports
    int64_t cap_edges = std::max<int64_t>((int64_t)(B.edges.bytes / sizeof(EdgeRec)), 8 * M + 1024);
    HP_ENSURE(B.edges, sizeof(EdgeRec) * cap_edges);
    for (DevBuf *b : {&B.deg, &B.kmin, &B.par, &B.first_label, &B.lab, &B.root, &B.v0, &B.v1, &B.x0,
                      &B.x1, &B.ecnt})
        HP_ENSURE(*b, sizeof(int32_t) * (M + 2));
    // one level's ranges: the large ones from the front, the mid-sized ones
    // from the back
    const int64_t cap_rng = M / (kSortSmall + 1) + ng + 1, cap_rsmall = M / 2 + ng + 1;
    HP_ENSURE(B.rng0, 2 * sizeof(SortRange) * cap_rng);
    HP_ENSURE(B.rng1, 2 * sizeof(SortRange) * cap_rng);
    HP_ENSURE(B.rsmall, sizeof(SortRange) * cap_rsmall);
    HP_ENSURE(B.ssum, sizeof(int32_t) * ((M + 2) / kScanBlock + 2));
    if ((r = e->upload(B.groups, G)) || (r = e->upload(B.tiles, tiles)) ||
        (r = e->upload(B.qtiles, qtiles)) || (r = e->upload(B.drawn, drawn)) ||
        (r = e->upload(B.qidx, qidx)) || (r = e->upload(B.ltiles, ltiles)))
        return r;
    hipStream_t s = e->stream;
    const SubGroup *dg = B.groups.as<SubGroup>();
    const SubTile *dt = B.tiles.as<SubTile>();
    int32_t *cnt = B.counters.as<int32_t>();
    SParticle *parts = reinterpret_cast<SParticle *>(d_particles);
    k_fill_i32<<<1, 64, 0, s>>>(cnt, 0, kSfCounters);
    HP_LAUNCHED();
    k_fill_i32<<<Engine::blocks(ng, 256), 256, 0, s>>>(B.gerr.as<int32_t>(), 0, ng);
    HP_LAUNCHED();
    if (nt) {
        // norm_sd
        k_sub_load<<<(uint32_t)nt, kWave, 0, s>>>(parts, dg, dt, B.w.as<SParticle>(),
                                                  B.slot_group.as<int32_t>(), B.psum.as<double>(),
                                                  B.gerr.as<int32_t>());
        HP_LAUNCHED();
        k_sub_mean<<<Engine::blocks(ng, 64), 64, 0, s>>>((int32_t)ng, dg, B.psum.as<double>(),
                                                 B.mean.as<double>());
        HP_LAUNCHED();
        k_sub_centre<<<(uint32_t)nt, kWave, 0, s>>>(dg, dt, B.mean.as<double>(),
                                                    B.w.as<SParticle>(), B.pcorr.as<double>());
        HP_LAUNCHED();
        k_sub_sigma<<<Engine::blocks(ng, 64), 64, 0, s>>>((int32_t)ng, dg, B.pcorr.as<double>(),
                                                  cfg.initial_metric_scaling, B.sig.as<double>());
        HP_LAUNCHED();
        k_sub_scale<<<(uint32_t)nt, kWave, 0, s>>>(dg, dt, B.sig.as<double>(), B.w.as<SParticle>(),
                                                   B.pbox.as<float>(), B.gerr.as<int32_t>());
        HP_LAUNCHED();
    }
    // the trees, level by level
    k_fill_i32<<<1, 1, 0, s>>>(cnt + kSfNodes, (int32_t)ng, 1);
    HP_LAUNCHED();
    // two lists per level: nodes above kSfBigNode points split with 1024
    // threads, the others with kSfThreads
    const int64_t act_cap = M / (kLeaf6 + 1) + ng + 1;
    int32_t *act = B.act0.as<int32_t>(), *nxt = B.act1.as<int32_t>();
    int32_t *act_big = act + act_cap, *nxt_big = nxt + act_cap;
    int32_t *fused = B.fused.as<int32_t>();
    k_sub_roots<<<Engine::blocks(ng, 64), 64, 0, s>>>((int32_t)ng, dg, B.pbox.as<float>(),
                                              B.nodes.as<Node6>(), act, act_big, cnt, fused);
    HP_LAUNCHED();
    for (int level = 0;; level++) {
        int32_t na[kSfCounters];
        HP_TRY(hipMemcpyAsync(na, cnt, sizeof na, hipMemcpyDeviceToHost, s));
        HP_TRY(hipStreamSynchronize(s));
        const int32_t ns_ = na[kSfActive], nb_ = na[kSfActiveBig];
        if (!ns_ && !nb_) break;
        if (level > 4 * 64) return set_error(HP_EHIP, "hp_find_subfofs: tree too deep");
        k_fill_i32<<<1, 1, 0, s>>>(cnt + kSfActive, 0, 1);
        HP_LAUNCHED();
        k_fill_i32<<<1, 1, 0, s>>>(cnt + kSfActiveBig, 0, 1);
        HP_LAUNCHED();
        if (nb_) {
            k_sub_split<1024><<<(uint32_t)nb_, 1024, 0, s>>>(
                B.nodes.as<Node6>(), act_big, B.w.as<SParticle>(), B.w2.as<SParticle>(),
                B.rk.as<int32_t>(), B.bpos.as<int32_t>(), B.spos.as<int32_t>(), nxt, nxt_big,
                cnt, (int32_t)node_cap, fused);
            HP_LAUNCHED();
        }
        if (ns_) {
            k_sub_split<kSfThreads><<<(uint32_t)ns_, kSfThreads, 0, s>>>(
                B.nodes.as<Node6>(), act, B.w.as<SParticle>(), B.w2.as<SParticle>(),
                B.rk.as<int32_t>(), B.bpos.as<int32_t>(), B.spos.as<int32_t>(), nxt, nxt_big,
                cnt, (int32_t)node_cap, fused);
            HP_LAUNCHED();
        }
        std::swap(act, nxt);
        std::swap(act_big, nxt_big);
    }
    {
        // the nodes of at most kFuseMax points, each finished by one workgroup
        int32_t nf = 0;
        HP_TRY(hipMemcpyAsync(&nf, cnt + kSfFused, sizeof nf, hipMemcpyDeviceToHost, s));
        HP_TRY(hipStreamSynchronize(s));
        if (nf > 0) {
            k_sub_fused<<<(uint32_t)nf, kSfThreads, 0, s>>>(B.nodes.as<Node6>(), fused,
                                                           B.w.as<SParticle>(),
                                                           B.w2.as<SParticle>(), cnt,
                                                           (int32_t)node_cap);
            HP_LAUNCHED();
        }
        int32_t nn = 0;
        HP_TRY(hipMemcpyAsync(&nn, cnt + kSfNodes, sizeof nn, hipMemcpyDeviceToHost, s));
        HP_TRY(hipStreamSynchronize(s));
        nn = (int32_t)std::min<int64_t>(nn, node_cap);
        if (nn > 0) {
            k_sub_collect<<<(uint32_t)nn, kWave, 0, s>>>(B.nodes.as<Node6>(), B.w.as<SParticle>(),
                                                         B.w2.as<SParticle>());
            HP_LAUNCHED();
        }
    }
    // the linking length of each group
    if (!qtiles.empty()) {
        k_sub_next_closest<<<(uint32_t)qtiles.size(), kWave, 0, s>>>(
            dg, B.qtiles.as<SubTile>(), B.drawn.as<int32_t>(), B.qidx.as<int32_t>(),
            B.w.as<SParticle>(), B.nodes.as<Node6>(), B.dist.as<float>(), cfg.ordered_search);
        HP_LAUNCHED();
        HP_TRY(hipMemcpyAsync(B.sel.p, B.dist.p, sizeof(float) * S, hipMemcpyDeviceToDevice, s));
    }
    k_sub_median<<<(uint32_t)ng, kSfThreads, 0, s>>>(dg, thresh, B.sel.as<float>(),
                                                      B.sel2.as<float>(), B.rk.as<int32_t>(),
                                                      B.bpos.as<int32_t>(), B.spos.as<int32_t>(),
                                                      B.target.as<float>());
    HP_LAUNCHED();
    // the linking, the reference's numbers and roots, then its output order
    if (M) {
        const int32_t m = (int32_t)M;
        int32_t *deg = B.deg.as<int32_t>(), *kmin = B.kmin.as<int32_t>();
        int32_t *flab = B.first_label.as<int32_t>();
        auto scan = [&](int32_t *v, int32_t n) {  // exclusive, in place
            const int32_t nbk = (n + kScanBlock - 1) / kScanBlock;
            k_scan_blocks<<<nbk, 1024, 0, s>>>(v, n, B.ssum.as<int32_t>());
            k_scan_top<<<1, 1024, 0, s>>>(B.ssum.as<int32_t>(), nbk);
            k_scan_add<<<nbk, 1024, 0, s>>>(v, n, B.ssum.as<int32_t>());
        };
        k_sub_init<<<Engine::blocks(M + 1, 256), 256, 0, s>>>(m, kmin, deg, B.csize.as<int32_t>());
        HP_LAUNCHED();
        unsigned long long ne = 0;
        for (int pass = 0; pass < 2; pass++) {
            k_fill_i32<<<1, 64, 0, s>>>(cnt + kSfEdges, 0, 2);
            HP_LAUNCHED();
            k_sub_link<<<(uint32_t)ltiles.size(), kWave, 0, s>>>(
                dg, B.ltiles.as<SubTile>(), B.target.as<float>(), B.w.as<SParticle>(),
                B.nodes.as<Node6>(), B.edges.as<EdgeRec>(), cap_edges, deg, cnt);
            HP_LAUNCHED();
            HP_TRY(hipMemcpyAsync(&ne, cnt + kSfEdges, sizeof ne, hipMemcpyDeviceToHost, s));
            HP_TRY(hipStreamSynchronize(s));
            if ((int64_t)ne <= cap_edges) break;
            if (pass || ne >= (unsigned long long)INT32_MAX)
                return set_error(HP_EINVAL, "hp_find_subfofs: %llu linked pairs (at most %d)", ne,
                                 INT32_MAX - 1);
            cap_edges = (int64_t)ne;  // the walk again, with room for every pair
            HP_ENSURE(B.edges, sizeof(EdgeRec) * cap_edges);
        }
        HP_ENSURE(B.adj, sizeof(int32_t) * (ne + 1));
        scan(deg, m + 1);
        HP_LAUNCHED();
        if (ne) {
            k_sub_adjacent<<<(uint32_t)((ne + 255) / 256), 256, 0, s>>>(
                (int64_t)ne, B.edges.as<EdgeRec>(), deg, B.adj.as<int32_t>(), kmin);
            HP_LAUNCHED();
        }
        k_sub_parent<<<Engine::blocks(M + 1, 256), 256, 0, s>>>(m, deg, B.adj.as<int32_t>(), kmin,
                                                        B.par.as<int32_t>(), flab);
        HP_LAUNCHED();
        scan(flab, m + 1);
        HP_LAUNCHED();
        k_sub_label<<<Engine::blocks(M, 256), 256, 0, s>>>(m, B.par.as<int32_t>(), flab,
                                                   B.lab.as<int32_t>());
        HP_LAUNCHED();
        // the pairs that can merge, staged in the pair list's place (free now)
        HP_ENSURE(B.evt, sizeof(int2) * (ne + 1));
        int32_t *ecnt = B.ecnt.as<int32_t>();
        k_sub_events<<<Engine::blocks(M + 1, 256), 256, 0, s>>>(m, deg, B.adj.as<int32_t>(), kmin,
                                                        B.lab.as<int32_t>(),
                                                        B.edges.as<int2>(), ecnt);
        HP_LAUNCHED();
        scan(ecnt, m + 1);
        HP_LAUNCHED();
        k_sub_compact<<<Engine::blocks(M, 256), 256, 0, s>>>(m, deg, B.edges.as<int2>(), ecnt,
                                                     B.evt.as<int2>());
        HP_LAUNCHED();
        k_sub_tokens<<<(uint32_t)ng, kTokThreads, 0, s>>>(dg, ecnt, B.evt.as<int2>(), flab,
                                                          B.root.as<int32_t>());
        HP_LAUNCHED();
        int32_t *v0 = B.v0.as<int32_t>(), *v1 = B.v1.as<int32_t>();
        int32_t *x0 = B.x0.as<int32_t>(), *x1 = B.x1.as<int32_t>();
        k_sub_values<<<Engine::blocks(M, 256), 256, 0, s>>>(m, B.slot_group.as<int32_t>(), dg,
                                                    B.lab.as<int32_t>(), flab,
                                                    B.root.as<int32_t>(), v0, x0,
                                                    B.csize.as<int32_t>());
        HP_LAUNCHED();
        // partition_sort_particles, level by level
        SortRange *ra = B.rng0.as<SortRange>(), *rb = B.rng1.as<SortRange>();
        SortRange *rs = B.rsmall.as<SortRange>();
        k_fill_i32<<<1, 64, 0, s>>>(cnt + kSfRange, 0, 3);
        HP_LAUNCHED();
        k_sort_roots<<<Engine::blocks(ng, 64), 64, 0, s>>>((int32_t)ng, dg, ra, rs, cnt,
                                                   (int32_t)cap_rng, (int32_t)cap_rsmall);
        HP_LAUNCHED();
        int32_t nsmall = 0;
        for (int level = 0;; level++) {
            int32_t na[kSfCounters];
            HP_TRY(hipMemcpyAsync(na, cnt, sizeof na, hipMemcpyDeviceToHost, s));
            HP_TRY(hipStreamSynchronize(s));
            if (na[kSfErr])
                return set_error(HP_EHIP, "hp_find_subfofs: scratch overflow (%d)", na[kSfErr]);
            const int32_t nr = na[kSfRange], nm = na[kSfRangeMid];
            nsmall = na[kSfRangeSmall];
            if (!nr && !nm) break;
            // each level halves the value span of every range it splits
            if (level > 40) return set_error(HP_EHIP, "hp_find_subfofs: output sort too deep");
            k_fill_i32<<<1, 1, 0, s>>>(cnt + kSfRange, 0, 1);
            HP_LAUNCHED();
            k_fill_i32<<<1, 1, 0, s>>>(cnt + kSfRangeMid, 0, 1);
            HP_LAUNCHED();
            if (nr) {
                k_sort_split<kSortThreads1><<<(uint32_t)nr, kSortThreads1, 0, s>>>(
                    ra, v0, v1, x0, x1, B.rk.as<int32_t>(), B.bpos.as<int32_t>(),
                    B.spos.as<int32_t>(), rb, rs, cnt, (int32_t)cap_rng, (int32_t)cap_rsmall);
                HP_LAUNCHED();
            }
            if (nm) {
                k_sort_split<kSortThreads2><<<(uint32_t)nm, kSortThreads2, 0, s>>>(
                    ra + 2 * cap_rng - 1, v0, v1, x0, x1, B.rk.as<int32_t>(),
                    B.bpos.as<int32_t>(), B.spos.as<int32_t>(), rb, rs, cnt, (int32_t)cap_rng,
                    (int32_t)cap_rsmall);
                HP_LAUNCHED();
            }
            std::swap(ra, rb);
        }
        if (nsmall > 0) {
            k_sort_small<<<(uint32_t)nsmall, kWave, 0, s>>>(rs, v0, v1, x0, x1);
            HP_LAUNCHED();
        }
        k_sub_emit<<<Engine::blocks(M, 256), 256, 0, s>>>(
            m, B.slot_group.as<int32_t>(), dg, v0, x0, flab, B.csize.as<int32_t>(),
            B.w.as<SParticle>(), parts, cfg.min_halo_particles, B.emit.as<int4>(),
            (int32_t)cap_emit, cnt);
        HP_LAUNCHED();
    }
    int32_t hc[kSfCounters];
    std::vector<int32_t> gerr((size_t)ng);
    std::vector<float> target((size_t)ng);
    HP_TRY(hipMemcpyAsync(hc, cnt, sizeof hc, hipMemcpyDeviceToHost, s));
    HP_TRY(hipMemcpyAsync(gerr.data(), B.gerr.p, sizeof(int32_t) * ng, hipMemcpyDeviceToHost, s));
    HP_TRY(hipMemcpyAsync(target.data(), B.target.p, sizeof(float) * ng, hipMemcpyDeviceToHost, s));
    HP_TRY(hipStreamSynchronize(s));
    for (int64_t g = 0; g < ng; g++)
        if (gerr[(size_t)g])
            return set_error(HP_EINVAL, "hp_find_subfofs: group %lld has a non-finite coordinate "
                             "(before or after norm_sd)", (long long)g);
    if (hc[kSfErr])
        return set_error(HP_EHIP, "hp_find_subfofs: scratch overflow (%d)", hc[kSfErr]);
    const int32_t ne = hc[kSfEmit];
    std::vector<int4> em((size_t)ne);
    if (ne) {
        HP_TRY(hipMemcpyAsync(em.data(), B.emit.p, sizeof(int4) * ne, hipMemcpyDeviceToHost, s));
        HP_TRY(hipStreamSynchronize(s));
    }
    std::sort(em.begin(), em.end(), [](const int4 &a, const int4 &b) {
        return a.x != b.x ? a.x < b.x : a.y < b.y;
    });
    *nsubs = ne;
    if (ne > cap_subs)
        return set_error(HP_EINVAL, "hp_find_subfofs: %d sub-FOFs do not fit the %lld given", ne,
                         (long long)cap_subs);
    for (int64_t g = 0; g < ng; g++) {
        groups[g].target_r = target[(size_t)g];
        groups[g].first_sub = 0;
        groups[g].num_sub = 0;
    }
    for (int32_t i = 0; i < ne; i++) {
        hp_fof_group &gg = groups[em[(size_t)i].x];
        if (!gg.num_sub) gg.first_sub = i;
        gg.num_sub++;
        subs[i].first = gg.first + em[(size_t)i].y;
        subs[i].num_p = em[(size_t)i].z;
    }
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_mass_definitions(const char *const defs[5],
                                                               hp_mass_defs *out) {
    if (!out) return set_error(HP_EINVAL, "hp_mass_definitions: bad arguments");
    static const char *const dflt[5] = {"vir", "200b", "200c", "500c", "2500c"};
    if (!(PARTICLE_MASS > 0.0) || !(SCALE_NOW > 0.0))
        return set_error(HP_EINVAL, "hp_mass_definitions: PARTICLE_MASS and SCALE_NOW must be > 0");
    for (int k = 0; k < 5; k++)
        out->thresh_dens[k] = mass_definition(defs && defs[k] ? defs[k] : dflt[k]);
    out->rvir_dens = mass_definition("vir");
    // src/groupies.c:123: "vir" at a = 1 with the same float background
    // density, kept in double there
    const float cons = (float)(Om * 2.77519737e11 / PARTICLE_MASS);
    out->rvir_dens_z0 = vir_density(1.0) * cons;
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_plan_props(hp_plan *p, const struct potential *d_po,
                                                         const hp_halo *d_halos,
                                                         const hp_mass_defs *defs,
                                                         hp_halo_props *d_props, void *stream) {
    if (!p || !defs || ((!d_po || !d_halos || !d_props) && p->T.nhalos))
        return set_error(HP_EINVAL, "hp_plan_props: bad arguments");
    int r;
    if ((r = p->eng.join(stream))) return r;
    // a per-call plan's runs load single-call tables; the pass needs the batch's
    if (p->opt.per_call && (r = p->eng.load_tables(p->T))) return r;
    if ((r = p->eng.props(p->T, d_po, d_halos, *defs, d_props))) return r;
    if ((r = p->eng.release(stream))) return r;
    HP_TRY(hipStreamSynchronize(p->eng.stream));
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_plan_stats(const hp_plan *p, hp_stats *out) {
    if (!p || !out) return set_error(HP_EINVAL, "hp_plan_stats: bad arguments");
    *out = p->eng.st;
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_plan_destroy(hp_plan *p) {
    delete p;
    return HP_OK;
}

__attribute__((visibility("default"))) int hp_bound_pass_host(struct potential *po,
                                                              int64_t nrecords, hp_halo *halos,
                                                              int64_t nhalos, uint8_t *bound,
                                                              const hp_options *opt) {
    hp_plan *p = nullptr;
    int r = hp_plan_create(&p, halos, nhalos, opt);
    if (r) return r;
    if (p->T.nrec != nrecords) {
        hp_plan_destroy(p);
        return set_error(HP_EINVAL, "halos cover %lld records, not %lld", (long long)p->T.nrec,
                         (long long)nrecords);
    }
    DevBuf dpo, dh, db;
    r = dpo.ensure(sizeof(struct potential) * (nrecords + 1));
    if (!r) r = dh.ensure(sizeof(hp_halo) * (nhalos + 1));
    if (!r) r = db.ensure(nrecords + 1);
    hipStream_t s = p->eng.stream;
    if (!r && (hipMemcpyAsync(dpo.p, po, sizeof(struct potential) * nrecords,
                              hipMemcpyHostToDevice, s) != hipSuccess ||
               hipMemcpyAsync(dh.p, halos, sizeof(hp_halo) * nhalos, hipMemcpyHostToDevice, s) !=
                   hipSuccess))
        r = set_error(HP_EHIP, "copy to the device failed");
    if (!r)
        r = hp_plan_run(p, dpo.as<struct potential>(), dh.as<hp_halo>(),
                        bound ? db.as<uint8_t>() : nullptr, nullptr);
    if (!r && (hipMemcpyAsync(po, dpo.p, sizeof(struct potential) * nrecords,
                              hipMemcpyDeviceToHost, s) != hipSuccess ||
               hipMemcpyAsync(halos, dh.p, sizeof(hp_halo) * nhalos, hipMemcpyDeviceToHost, s) !=
                   hipSuccess ||
               (bound && hipMemcpyAsync(bound, db.p, nrecords, hipMemcpyDeviceToHost, s) !=
                             hipSuccess) ||
               hipStreamSynchronize(s) != hipSuccess))
        r = set_error(HP_EHIP, "copy from the device failed");
    dpo.release();
    dh.release();
    db.release();
    hp_plan_destroy(p);
    return r;
}

__attribute__((visibility("default"))) int64_t hp_debug_tree(const struct potential *po,
                                                             int64_t n, int32_t *perm,
                                                             hp_node_out *out, int64_t cap) {
    if (n < 0 || (n > 0 && !po)) return set_error(HP_EINVAL, "hp_debug_tree: bad arguments");
    int r;
    Engine *e = thread_engine(&r);
    if (!e) return r;
    Tables T;
    int64_t first = 0, count = n;
    if ((r = make_tables(T, &first, &count, 1, HP_METHOD_TREE, false))) return r;
    if ((r = e->reserve(T, false)) || (r = e->load_tables(T))) return r;
    DevBuf &dpo = e->srt;
    if ((r = dpo.ensure(sizeof(struct potential) * (n + 1)))) return r;
    HP_TRY(hipMemcpyAsync(dpo.p, po, sizeof(struct potential) * n, hipMemcpyHostToDevice, e->stream));
    // node ids a subtree reserves but does not use are never written: clear
    // the array so that the export shows them as empty nodes
    HP_TRY(hipMemsetAsync(e->nodes.p, 0, sizeof(Node) * e->cap_nodes, e->stream));
    e->st = hp_stats();
    e->st.halos = 1;
    int next = 0;
    if ((r = e->enqueue_roots(T, dpo.as<struct potential>())) || (r = e->finish_levels(1, &next)) ||
        (r = e->check_build()) || (r = e->enqueue_multipole((int)e->st.levels)) ||
        (r = e->read_counters()))
        return r;
    if (e->h_counters[kErrFlags])
        return set_error(HP_EINVAL, "tree build failed (flags %d)", e->h_counters[kErrFlags]);
    const int64_t nn = e->h_counters[kNodeCount];
    std::vector<Node> nodes(nn);
    std::vector<int32_t> par(nn);
    std::vector<float4> fin(n);
    HP_TRY(hipMemcpyAsync(nodes.data(), e->nodes.p, sizeof(Node) * nn, hipMemcpyDeviceToHost, e->stream));
    HP_TRY(hipMemcpyAsync(par.data(), e->parent.p, sizeof(int32_t) * nn, hipMemcpyDeviceToHost, e->stream));
    HP_TRY(hipMemcpyAsync(fin.data(), e->F.p, sizeof(float4) * n, hipMemcpyDeviceToHost, e->stream));
    HP_TRY(hipStreamSynchronize(e->stream));
    if (perm)
        for (int64_t i = 0; i < n; i++) {
            int32_t v;
            std::memcpy(&v, &fin[i].w, 4);
            perm[i] = v;
        }
    for (int64_t k = 0; k < nn && k < cap && out; k++) {
        hp_node_out &o = out[k];
        const Node &d = nodes[k];
        for (int j = 0; j < 3; j++) {
            o.lo[j] = d.lo[j];
            o.hi[j] = d.hi[j];
            o.com[j] = d.com[j];
        }
        o.mass = d.mass;
        o.dmin2 = d.dmin2;
        o.dmin = d.dmin;
        o.first = d.first;
        o.count = d.count;
        o.left = d.left;
        o.skip = d.skip;
        o.parent = par[k];
        o.pad = 0;
    }
    return nn;
}

}  // extern "C"
