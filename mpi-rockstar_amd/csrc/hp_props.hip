// hp_props.hip -- the halo property pass that follows the bound pass
// (src/properties.c:384-510, called for bound = 0 and bound = 1 by
// src/properties.c:552-553, with the defaults BOUND_PROPS = 1 and
// LIGHTCONE = 0), for every call of a batch at once.  calc_shape,
// calc_scale_radius and _calc_pseudo_evolution_masses are not part of it.
//
// The reference walks each call's records (sorted by r2) twice:
//   1. over every kept record (bound = 0) or every bound one, !(pe < ke)
//      (bound = 1), with the running count n: the density n / r^3 against
//      five thresholds and the virial one.  Each result is "n (or the index)
//      at the last record above a threshold", and vmax is the largest n / r
//      above the virial threshold from the first record above the main
//      threshold on (the first record reaching it wins ties).
//   2. over the bound records below dens_tot (the index of the last record
//      above the main threshold): angular momentum, running means of
//      position and velocity, the velocity dispersion, the half-mass index;
//      and estimate_total_energy (src/properties.c:328-349) over the records
//      with pe > ke below dens_tot.
// Here both walks are tile passes over the same scan tiles as the bound pass
// (2048 sorted records of one call, one wavefront each, rounds of 64 carried
// from lane 63), with per-call combines:
//   k_props_counts  bound and pe > ke counts per tile
//   k_props_prefix  per call: exclusive prefix of the tile counts
//   k_props_scan    walk 1 per tile: "last above" indices and counts as
//                   max-reductions (the running count is monotone), the
//                   vmax candidates as (value, first index) argmax
//   k_props_calls   per call: combine walk 1
//   k_props_sums    walk 2 per tile: f64 sums
//   k_props_final   per call: the halo fields
// Counts, indices, masses, radii and vmax are exact (the same IEEE
// operations on the same values).  The reference's running means, Welford
// dispersion and sequential f64 sums are restated as f64 sums of exact
// differences, which agree with them to f64 rounding.

#include "hp_internal.h"

#pragma clang fp contract(off)

namespace hp {

static_assert(sizeof(hp_halo_props) == 224 && offsetof(hp_halo_props, n_mdelta) == 192,
              "hp_halo_props layout");

constexpr double kVmaxConst = 6.55833746e-5;  // src/universal_constants.h:6
constexpr double kGc = 4.30117902e-9;         // src/universal_constants.h:4

// thresholds and constants of one property pass
struct PropArgs {
    double dens_thresh;   // particle_thresh_dens[0] * 4pi/3
    double d[4];          // particle_thresh_dens[1..4] * 4pi/3
    double rvir_thresh;   // particle_rvir_dens * 4pi/3
    double thresh0;       // particle_thresh_dens[0]
    double thresh3;       // particle_thresh_dens[3] (the second shape window)
    double rvir_dens;     // particle_rvir_dens
    double rvir_dens_z0;  // particle_rvir_dens_z0
    double force_res, particle_mass;
    float scale_now;
};

struct PropCount {
    int32_t nb;  // !(pe < ke)
    int32_t ns;  // pe > ke
};

// walk 1 partial of one tile
struct PropScan {
    int32_t last_all;      // last index with (j + 1) / r^3 > dens_thresh (all records)
    int32_t first0, last0; // first / last bound index with n / r^3 > dens_thresh
    int32_t n0;            // n at last0
    int32_t alt[4];        // n at the last bound index above d[k]
    int32_t vir;           // n at the last bound index above rvir_thresh
    int32_t j_all, j_after;
    double v_all;          // best n / r above rvir_thresh in the tile
    double v_after;        // the same from the tile's first index above dens_thresh on
    double pe_b;           // max of n^2 / r^1.5 (src/properties.c:369-373)
};

// per call, walk 1 combined
struct PropCall {
    int32_t kept;
    int32_t n_all;         // part_mdelta of the bound = 0 walk
    int32_t n0;            // part_mdelta of the bound = 1 walk
    int32_t dens_tot;
    int32_t alt[4];
    int32_t vir;
    int32_t j_vmax;        // -1: vmax stays 0
    double vmax;
    double pe_b;
};

// walk 2 partial of one tile
struct PropSums {
    double L[3];
    double sx[3];          // sum of x - centre (exact differences)
    double sv[3];          // sum of v - centre velocity
    double sv2[3];         // sum of (v - centre velocity)^2
    double ke;             // pe > ke records: sum of ke
    double phi_i;          // sum of particle_mass * i / r
    double phi_c;          // sum of (particle_mass / r) * (pe > ke records before i)
    int32_t n;             // bound records
    int32_t half;          // first bound index >= 1 with 2 n >= part_mdelta, else INT32_MAX
};

// a call's kept records, clamped to its record range (the plan's tables)
__device__ __forceinline__ int32_t call_kept(const hp_halo *halos, const int64_t *hcount,
                                             int32_t h) {
    const int64_t k = halos[h].kept;
    return (int32_t)(k < 0 ? 0 : (k > hcount[h] ? hcount[h] : k));
}

__device__ __forceinline__ double prop_radius(float r2, double force_res) {
    double r = sqrt((double)r2);
    return r < force_res ? force_res : r;
}

// (value, index): larger value wins, the smaller index on ties
__device__ __forceinline__ void arg_better(double &v, int32_t &j, double v2, int32_t j2) {
    if (v2 > v || (v2 == v && j2 < j)) {
        v = v2;
        j = j2;
    }
}

__device__ __forceinline__ void wave_arg_better(double &v, int32_t &j) {
    for (int o = 32; o > 0; o >>= 1) {
        const double v2 = __shfl_xor(v, o, 64);
        const int32_t j2 = __shfl_xor(j, o, 64);
        arg_better(v, j, v2, j2);
    }
}

__global__ __launch_bounds__(kWave) void k_props_counts(const struct potential *__restrict__ srt,
                                                        const ScanTile *__restrict__ tiles,
                                                        int32_t ntiles,
                                                        const hp_halo *__restrict__ halos,
                                                        const int64_t *__restrict__ hfirst,
                                                        const int64_t *__restrict__ hcount,
                                                        PropCount *__restrict__ cnt) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const int64_t first = hfirst[st.call];
    const int32_t end = min(st.off + kScanTile, call_kept(halos, hcount, st.call));
    int32_t nb = 0, ns = 0;
    for (int32_t j = st.off + lane_id(); j < end; j += kWave) {
        const double pe = srt[first + j].pe;
        const float ke = srt[first + j].ke;
        nb += !(pe < (double)ke) ? 1 : 0;
        ns += pe > (double)ke ? 1 : 0;
    }
    nb = wave_sum(nb);
    ns = wave_sum(ns);
    if (lane_id() == 0) cnt[t] = PropCount{nb, ns};
}

// per call: exclusive prefix of the tile counts, in tile order
__global__ void k_props_prefix(const int32_t *__restrict__ call_tile0,
                               const int32_t *__restrict__ call_ntiles, int32_t nh,
                               PropCount *__restrict__ cnt) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    int32_t nb = 0, ns = 0;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) {
        const PropCount c = cnt[t];
        cnt[t] = PropCount{nb, ns};
        nb += c.nb;
        ns += c.ns;
    }
}

// walk 1 (src/properties.c:402-431), one wavefront per tile
__global__ __launch_bounds__(kWave) void k_props_scan(const struct potential *__restrict__ srt,
                                                      const ScanTile *__restrict__ tiles,
                                                      int32_t ntiles,
                                                      const hp_halo *__restrict__ halos,
                                                      const int64_t *__restrict__ hfirst,
                                                      const int64_t *__restrict__ hcount,
                                                      const PropCount *__restrict__ pre,
                                                      PropArgs pa, PropScan *__restrict__ out) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const int64_t first = hfirst[st.call];
    const int32_t end = min(st.off + kScanTile, call_kept(halos, hcount, st.call));
    const int lane = lane_id();
    int32_t carry = pre[t].nb;
    PropScan s;
    s.last_all = -1;
    s.first0 = INT32_MAX;
    s.last0 = -1;
    s.n0 = 0;
    for (int k = 0; k < 4; k++) s.alt[k] = 0;
    s.vir = 0;
    s.v_all = s.v_after = -1.0;
    s.j_all = s.j_after = INT32_MAX;
    s.pe_b = 0.0;
    bool seen0 = false;  // an index above dens_thresh earlier in this tile
    for (int32_t base = st.off; base < end; base += kWave) {
        const int32_t j = base + lane;
        const bool v = j < end;
        float r2 = 0.0f, ke = 0.0f;
        double pe = 0.0;
        if (v) {
            r2 = srt[first + j].r2;
            pe = srt[first + j].pe;
            ke = srt[first + j].ke;
        }
        const bool b = v && !(pe < (double)ke);
        const int32_t n = wave_scan_add(b ? 1 : 0) + carry;
        carry = read_lane(n, kWave - 1);
        const double r = prop_radius(r2, pa.force_res);
        const double r3 = r * r * r;
        if (v && (double)(j + 1) / r3 > pa.dens_thresh) s.last_all = j;
        const double cur = (double)n / r3;
        const bool c0 = b && cur > pa.dens_thresh;
        const uint64_t m0 = ballot(c0);
        const bool after = seen0 || (m0 & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull))) != 0;
        seen0 = seen0 || m0 != 0;
        if (b) {
            // src/properties.c:365-373: unclamped radius
            double r32 = sqrt(sqrt((double)r2));
            r32 = r32 * r32 * r32;
            const double q = (double)((int64_t)n * (int64_t)n) / r32;
            s.pe_b = q > s.pe_b ? q : s.pe_b;
            if (c0) {
                s.first0 = min(s.first0, j);
                s.last0 = j;
                s.n0 = n;
            }
            for (int k = 0; k < 4; k++)
                if (cur > pa.d[k]) s.alt[k] = n;
            if (cur > pa.rvir_thresh) {
                s.vir = n;
                const double circ = (double)n / r;
                arg_better(s.v_all, s.j_all, circ, j);
                if (after) arg_better(s.v_after, s.j_after, circ, j);
            }
        }
    }
    // every field is monotone in j: max (or min) over the lanes
    s.last_all = wave_max(s.last_all);
    s.first0 = wave_min(s.first0);
    s.last0 = wave_max(s.last0);
    s.n0 = wave_max(s.n0);
    for (int k = 0; k < 4; k++) s.alt[k] = wave_max(s.alt[k]);
    s.vir = wave_max(s.vir);
    wave_arg_better(s.v_all, s.j_all);
    wave_arg_better(s.v_after, s.j_after);
    s.pe_b = wave_max(s.pe_b);
    if (lane == 0) out[t] = s;
}

__global__ void k_props_calls(const int32_t *__restrict__ call_tile0,
                              const int32_t *__restrict__ call_ntiles, int32_t nh,
                              const hp_halo *__restrict__ halos,
                              const int64_t *__restrict__ hcount,
                              const PropScan *__restrict__ ts, PropCall *__restrict__ pc) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const int32_t t0 = call_tile0[h], t1 = t0 + call_ntiles[h];
    PropCall c;
    c.kept = call_kept(halos, hcount, h);
    int32_t last_all = -1, last0 = -1, n0 = 0, tfirst = -1;
    double pe_b = 0.0;
    for (int k = 0; k < 4; k++) c.alt[k] = 0;
    c.vir = 0;
    for (int32_t t = t0; t < t1; t++) {
        const PropScan &s = ts[t];
        last_all = max(last_all, s.last_all);
        if (s.last0 >= 0) {
            if (tfirst < 0) tfirst = t;
            last0 = s.last0;
            n0 = s.n0;
        }
        for (int k = 0; k < 4; k++) c.alt[k] = max(c.alt[k], ts[t].alt[k]);
        c.vir = max(c.vir, s.vir);
        pe_b = s.pe_b > pe_b ? s.pe_b : pe_b;
    }
    c.pe_b = pe_b;
    // vmax counts from the first index above dens_thresh on: the part of
    // that tile after it, then whole tiles
    double vm = -1.0;
    int32_t jm = INT32_MAX;
    if (tfirst >= 0) {
        arg_better(vm, jm, ts[tfirst].v_after, ts[tfirst].j_after);
        for (int32_t t = tfirst + 1; t < t1; t++) arg_better(vm, jm, ts[t].v_all, ts[t].j_all);
    }
    c.n_all = last_all + 1;
    c.n0 = n0;
    c.dens_tot = last0 >= 0 ? last0 : 0;
    c.j_vmax = jm == INT32_MAX ? -1 : jm;
    c.vmax = jm == INT32_MAX ? 0.0 : vm;
    pc[h] = c;
}

// walk 2 (src/properties.c:433-450, 127-137) and estimate_total_energy
// (src/properties.c:328-349) over [0, dens_tot)
__global__ __launch_bounds__(kWave) void k_props_sums(const struct potential *__restrict__ srt,
                                                      const ScanTile *__restrict__ tiles,
                                                      int32_t ntiles,
                                                      const hp_halo *__restrict__ halos,
                                                      const int64_t *__restrict__ hfirst,
                                                      const PropCount *__restrict__ pre,
                                                      const PropCall *__restrict__ pc, PropArgs pa,
                                                      PropSums *__restrict__ out) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile st = tiles[t];
    const hp_halo &hh = halos[st.call];
    const PropCall c = pc[st.call];
    const int64_t first = hfirst[st.call];
    const int32_t end = min(st.off + kScanTile, c.dens_tot);
    const float cen[6] = {hh.pos[0], hh.pos[1], hh.pos[2], hh.pos[3], hh.pos[4], hh.pos[5]};
    const int lane = lane_id();
    int32_t carry_b = pre[t].nb, carry_s = pre[t].ns;
    double acc[15] = {};
    int32_t nb_tile = 0, half = INT32_MAX;
    for (int32_t base = st.off; base < end; base += kWave) {
        const int32_t j = base + lane;
        const bool v = j < end;
        struct potential rec = {};
        if (v) rec = srt[first + j];
        const bool b = v && !(rec.pe < (double)rec.ke);
        const bool sb = v && rec.pe > (double)rec.ke;
        const int32_t n = wave_scan_add(b ? 1 : 0) + carry_b;
        const int32_t ns = wave_scan_add(sb ? 1 : 0) + carry_s;  // inclusive
        carry_b = read_lane(n, kWave - 1);
        carry_s = read_lane(ns, kWave - 1);
        if (b) {
            const float *x = rec.pos;
            // src/properties.c:127-137: float differences and products
            acc[0] += (double)((x[1] - cen[1]) * (x[5] - cen[5]));
            acc[0] -= (double)((x[2] - cen[2]) * (x[4] - cen[4]));
            acc[1] += (double)((x[2] - cen[2]) * (x[3] - cen[3]));
            acc[1] -= (double)((x[0] - cen[0]) * (x[5] - cen[5]));
            acc[2] += (double)((x[0] - cen[0]) * (x[4] - cen[4]));
            acc[2] -= (double)((x[1] - cen[1]) * (x[3] - cen[3]));
            for (int k = 0; k < 3; k++) {
                const double dx = (double)x[k] - (double)cen[k];
                const double dv = (double)x[k + 3] - (double)cen[k + 3];
                acc[3 + k] += dx;
                acc[6 + k] += dv;
                acc[9 + k] += dv * dv;
            }
            nb_tile++;
            if (j >= 1 && 2 * (int64_t)n >= (int64_t)c.n0) half = min(half, j);
        }
        if (sb) {
            const double r = prop_radius(rec.r2, pa.force_res);
            acc[12] += (double)rec.ke;
            acc[13] += pa.particle_mass * (double)j / r;
            acc[14] += (pa.particle_mass / r) * (double)(ns - 1);  // pe > ke records before j
        }
    }
    PropSums s;
    for (int k = 0; k < 3; k++) {
        s.L[k] = wave_sum(acc[k]);
        s.sx[k] = wave_sum(acc[3 + k]);
        s.sv[k] = wave_sum(acc[6 + k]);
        s.sv2[k] = wave_sum(acc[9 + k]);
    }
    s.ke = wave_sum(acc[12]);
    s.phi_i = wave_sum(acc[13]);
    s.phi_c = wave_sum(acc[14]);
    s.n = wave_sum(nb_tile);
    s.half = wave_min(half);
    if (lane == 0) out[t] = s;
}

__global__ void k_props_final(const struct potential *__restrict__ srt,
                              const int32_t *__restrict__ call_tile0,
                              const int32_t *__restrict__ call_ntiles, int32_t nh,
                              const hp_halo *__restrict__ halos,
                              const int64_t *__restrict__ hfirst,
                              const PropCall *__restrict__ pc, const PropSums *__restrict__ ts,
                              PropArgs pa, hp_halo_props *__restrict__ props) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const hp_halo &hh = halos[h];
    const int64_t first = hfirst[h];
    const PropCall c = pc[h];
    const double PM = pa.particle_mass;
    const float scale_now = pa.scale_now;
    double L[3] = {0, 0, 0}, sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0}, sv2[3] = {0, 0, 0};
    double ke = 0, phi = 0;
    int64_t n = 0;
    int32_t half = INT32_MAX;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) {
        const PropSums &s = ts[t];
        for (int k = 0; k < 3; k++) {
            L[k] += s.L[k];
            sx[k] += s.sx[k];
            sv[k] += s.sv[k];
            sv2[k] += s.sv2[k];
        }
        ke += s.ke;
        phi += s.phi_i + s.phi_c;
        n += s.n;
        half = min(half, s.half);
    }
    hp_halo_props o;
    const double m = (double)c.n0 * PM;
    o.m = (float)((double)c.n_all * PM);
    o.mgrav = (float)m;
    // src/properties.c:460-475: float fields accumulated from double terms
    float xoff = 0.0f, voff = 0.0f;
    double vr = 0.0;
    for (int k = 0; k < 3; k++) {
        // the reference's running means start at 0 and stay there with no
        // record averaged
        const double dsx = n > 0 ? sx[k] / (double)n : -(double)hh.pos[k];
        const double dsv = n > 0 ? sv[k] / (double)n : -(double)hh.pos[k + 3];
        xoff = (float)((double)xoff + dsx * dsx);
        voff = (float)((double)voff + dsv * dsv);
        vr += n > 0 ? (sv2[k] - sv[k] * sv[k] / (double)n) / (double)n : 0.0;
    }
    for (int k = 0; k < 4; k++) o.alt_m[k] = (float)((double)c.alt[k] * PM);
    o.Xoff = (float)(sqrt((double)xoff) * 1e3);
    o.Voff = (float)sqrt((double)voff);
    o.vrms = (float)sqrt(vr);
    o.vmax = (float)(kVmaxConst * sqrt(c.vmax * (PM / (double)scale_now)));
    o.rvmax = c.j_vmax >= 0
                  ? (float)(prop_radius(srt[first + c.j_vmax].r2, pa.force_res) * 1e3)
                  : 0.0f;
    const int32_t ih = half == INT32_MAX ? 0 : half;
    o.halfmass_radius = c.kept > 0 ? (float)(sqrt((double)srt[first + ih].r2) * 1e3) : 0.0f;
    o.r = (float)(cbrt((3.0 / (4.0 * M_PI)) * (double)c.n0 / pa.thresh0) * 1e3);
    const double rvir = cbrt((3.0 / (4.0 * M_PI)) * (double)c.vir / pa.rvir_dens) * 1e3;
    const double mvir = (double)c.vir * PM;
    o.rvir = (float)rvir;
    for (int k = 0; k < 3; k++) o.J[k] = (float)(PM * (double)scale_now * L[k]);
    // src/properties.c:328-349
    const double total_phi = phi / 2.0;
    o.kin_to_pot = total_phi != 0.0 ? (float)(ke / total_phi) : 0.0f;
    o.energy = (float)((ke - total_phi) * PM * kGc / (double)scale_now);
    const double Jh = PM * (double)scale_now * sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    o.spin = (float)(m > 0 ? Jh * sqrt(fabs((double)o.energy)) / (kGc * pow(m, 2.5))
                           : 0.0);
    o.bullock_spin = (float)(m > 0 ? Jh / (mvir * sqrt(2.0 * kGc * mvir * rvir *
                                                        (double)scale_now / 1e3))
                                   : 0.0);
    // filled by the shape, scale-radius and pseudo-evolution kernels
    o.rs = o.klypin_rs = o.chi2 = 0.0f;
    o.b_to_a = o.c_to_a = o.b_to_a2 = o.c_to_a2 = 0.0f;
    for (int k = 0; k < 3; k++) o.A[k] = o.A2[k] = 0.0f;
    for (int k = 0; k < 6; k++) o.inertia_tensor[k] = o.inertia_tensor2[k] = 0.0f;
    o.m_pe_b = o.m_pe_d = 0.0f;
    o.n_mdelta = c.n0;
    o.dens_tot = c.dens_tot;
    o.n_vir = c.vir;
    o.n_half = ih;
    props[h] = o;
}

// ------------------------------------------------------------- shapes
//
// calc_shape (src/properties.c:216-326) twice per call: window 0 over the
// records [0, np_alt[2]) with h->r from MASS_DEFINITION4, window 1 over
// [0, dens_tot) with the final h->r.  Each iteration is a tile pass (the
// weighted second moments of the bound records inside the current
// ellipsoid) and a per-(call, window) step: normalise, Jacobi, axis ratios,
// convergence.  The host runs SHAPE_ITERATIONS iterations; a window that has
// stopped is skipped.  The records' offsets from the centre (float
// differences, as the reference takes them) and bound flags are compacted
// once into 16 bytes per record.

constexpr int kShapeIterations = 10;  // SHAPE_ITERATIONS, src/config.template.h:90

struct ShapeState {
    double orth[3][3];
    double eig[3];
    float r;           // h->r of the window, kpc/h
    float min_r;
    float b_to_a, c_to_a, A[3];
    float inertia[6];
    int32_t n;         // records [0, n)
    int32_t limit;     // iterations to run; -1 until the first pass counted
    int32_t done;
    int32_t assigned;  // inertia written by this window
    int32_t nbound;    // bound records in [0, n)
    int32_t pad;
};

struct ShapePart {
    double w;
    double t[6];       // xx, yy, zz, yx, zx, zy
    int32_t cnt;       // bound records of the window in the tile
    int32_t pad;
};

__global__ void k_shape_init(int32_t nh, const hp_halo *__restrict__ halos,
                             const int64_t *__restrict__ hcount, const PropCall *__restrict__ pc,
                             const hp_halo_props *__restrict__ props, PropArgs pa,
                             ShapeState *__restrict__ st) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const PropCall c = pc[h];
    for (int w = 0; w < 2; w++) {
        ShapeState s;
        s.r = w == 0 ? (float)(cbrt((3.0 / (4.0 * M_PI)) * (double)c.alt[2] / pa.thresh3) * 1e3)
                     : props[h].r;
        s.n = min(w == 0 ? c.alt[2] : c.dens_tot, c.kept);
        float min_r = (float)(pa.force_res * pa.force_res);
        min_r *= 1e6 / (s.r * s.r);
        s.min_r = min_r;
        s.b_to_a = s.c_to_a = 0.0f;
        for (int k = 0; k < 3; k++) s.A[k] = 0.0f;
        for (int k = 0; k < 6; k++) s.inertia[k] = 0.0f;
        for (int k = 0; k < 3; k++) {
            for (int l = 0; l < 3; l++) s.orth[k][l] = k == l ? 1.0 : 0.0;
            s.eig[k] = (double)(s.r * s.r) * 1e-6;
        }
        s.limit = -1;
        s.done = !(s.r > 0) || s.n <= 0;
        s.assigned = 0;
        s.nbound = 0;
        s.pad = 0;
        st[2 * h + w] = s;
    }
}

// x, y, z offsets from the centre and the bound flag of the records the
// windows cover; each window's bound records are counted
__global__ __launch_bounds__(kWave) void k_shape_prep(const struct potential *__restrict__ srt,
                                                      const ScanTile *__restrict__ tiles,
                                                      int32_t ntiles,
                                                      const hp_halo *__restrict__ halos,
                                                      const int64_t *__restrict__ hfirst,
                                                      ShapeState *__restrict__ st,
                                                      float4 *__restrict__ cd) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile tl = tiles[t];
    const int32_t n0 = st[2 * tl.call].n, n1 = st[2 * tl.call + 1].n;
    const int32_t n = max(n0, n1);
    const int64_t first = hfirst[tl.call];
    const float c0 = halos[tl.call].pos[0], c1 = halos[tl.call].pos[1], c2 = halos[tl.call].pos[2];
    int32_t b0 = 0, b1 = 0;
    for (int32_t j = tl.off + lane_id(); j < min(tl.off + kScanTile, n); j += kWave) {
        const struct potential &r = srt[first + j];
        const bool b = !(r.pe < (double)r.ke);
        cd[first + j] = make_float4(r.pos[0] - c0, r.pos[1] - c1, r.pos[2] - c2, b ? 1.0f : 0.0f);
        b0 += (b && j < n0) ? 1 : 0;
        b1 += (b && j < n1) ? 1 : 0;
    }
    b0 = wave_sum(b0);
    b1 = wave_sum(b1);
    if (lane_id() == 0) {
        if (b0) atomicAdd(&st[2 * tl.call].nbound, b0);
        if (b1) atomicAdd(&st[2 * tl.call + 1].nbound, b1);
    }
}

// Windows summed in the reference's order, record by record (the sums of a
// window of few records decide which records the next ellipsoid holds, and
// a change in their last bit can change the result): those of at most
// kShapeExactRecords records or fewer than kShapeExactBound bound records.
constexpr int32_t kShapeExactRecords = 256;
constexpr int32_t kShapeExactBound = 64;
__device__ __forceinline__ bool shape_exact(const ShapeState &s) {
    return s.n <= kShapeExactRecords || s.nbound < kShapeExactBound;
}

// src/properties.c:243-268 for one tile, both windows
__global__ __launch_bounds__(kWave) void k_shape_tiles(const ScanTile *__restrict__ tiles,
                                                       int32_t ntiles,
                                                       const int64_t *__restrict__ hfirst,
                                                       const ShapeState *__restrict__ st,
                                                       const float4 *__restrict__ cd,
                                                       ShapePart *__restrict__ part) {
    __shared__ double s_term[kWave][8];
    for (int32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const ScanTile tl = tiles[t];
    const int64_t first = hfirst[tl.call];
    for (int w = 0; w < 2; w++) {
        const ShapeState &s = st[2 * tl.call + w];
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        int32_t cnt = 0;
        if (!s.done && shape_exact(s)) {
            // the window's first tile sums the whole window in record order
            // (src/properties.c:243-268: r with the divisions by eig, then
            // the running sums); the other tiles add nothing
            if (tl.off == 0) {
                const int lane = lane_id();
                for (int32_t base = 0; base < s.n; base += kWave) {
                    const int32_t j = base + lane;
                    bool take = false;
                    if (j < s.n) {
                        const float4 q = cd[first + j];
                        if (q.w != 0.0f) {
                            cnt++;
                            const double d[3] = {(double)q.x, (double)q.y, (double)q.z};
                            double r = 0;
                            for (int k = 0; k < 3; k++) {
                                double dr = 0;
                                for (int l = 0; l < 3; l++) dr += s.orth[k][l] * d[l];
                                r += dr * dr / s.eig[k];
                            }
                            if (r < (double)s.min_r) r = s.min_r;
                            if (r > 0 && r <= 1) {
                                take = true;
                                const double tw = 1.0 / r;  // WEIGHTED_SHAPES = 1
                                s_term[lane][0] = tw;
                                s_term[lane][1] = d[0] * d[0] * tw;
                                s_term[lane][2] = d[1] * d[1] * tw;
                                s_term[lane][3] = d[2] * d[2] * tw;
                                s_term[lane][4] = d[0] * d[1] * tw;
                                s_term[lane][5] = d[0] * d[2] * tw;
                                s_term[lane][6] = d[1] * d[2] * tw;
                            }
                        }
                    }
                    uint64_t m = ballot(take);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    while (m) {  // every lane keeps the same running sums
                        const int i = __builtin_ctzll(m);
                        m &= m - 1;
                        for (int k = 0; k < 7; k++) acc[k] += s_term[i][k];
                    }
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            ShapePart pt;
            pt.w = acc[0];
            for (int k = 0; k < 6; k++) pt.t[k] = acc[k + 1];
            pt.cnt = wave_sum(cnt);
            pt.pad = 0;
            if (lane_id() == 0) part[2 * t + w] = pt;
            continue;
        }
        if (!s.done) {
            const int32_t end = min(tl.off + kScanTile, s.n);
            double o[3][3], ie[3];
            for (int k = 0; k < 3; k++) {
                ie[k] = 1.0 / s.eig[k];  // the reference divides per record; within 1 ulp
                for (int l = 0; l < 3; l++) o[k][l] = s.orth[k][l];
            }
            for (int32_t j = tl.off + lane_id(); j < end; j += kWave) {
                const float4 q = cd[first + j];
                if (q.w == 0.0f) continue;
                cnt++;
                const double d[3] = {(double)q.x, (double)q.y, (double)q.z};
                double r = 0;
                for (int k = 0; k < 3; k++) {
                    double dr = 0;
                    for (int l = 0; l < 3; l++) dr += o[k][l] * d[l];
                    r += dr * dr * ie[k];
                }
                if (r < (double)s.min_r) r = s.min_r;
                if (!(r > 0 && r <= 1)) continue;
                const double tw = 1.0 / r;  // WEIGHTED_SHAPES = 1
                acc[0] += tw;
                acc[1] += d[0] * d[0] * tw;
                acc[2] += d[1] * d[1] * tw;
                acc[3] += d[2] * d[2] * tw;
                acc[4] += d[0] * d[1] * tw;
                acc[5] += d[0] * d[2] * tw;
                acc[6] += d[1] * d[2] * tw;
            }
        }
        ShapePart pt;
        pt.w = wave_sum(acc[0]);
        for (int k = 0; k < 6; k++) pt.t[k] = wave_sum(acc[k + 1]);
        pt.cnt = wave_sum(cnt);
        pt.pad = 0;
        if (lane_id() == 0) part[2 * t + w] = pt;
    }
    }
}

// |a| + sqrt(x), the sum and the root taken in extended precision and
// rounded once (the reference evaluates both in x87 long double)
__device__ __forceinline__ double add_sqrt_ext(double a_abs, double x) {
    const double r = sqrt(x);
    if (!(r > 0.0)) return a_abs + r;
    const double e = fma(-r, r, x) / (2.0 * r);  // sqrt(x) = r + e to O(e^2)
    const double s = a_abs + r;
    const double bb = s - a_abs;
    const double err = (a_abs - (s - bb)) + (r - bb);
    return s + (err + e);
}

// The cyclic Jacobi eigenvalue iteration of src/jacobi.c:8-135 (largest
// off-diagonal element first, each row's largest column cached, stop when
// no eigenvalue changes by more than 1e-6 of itself).  orth's rows are the
// eigenvectors; m is left symmetric.  With three rows the pivot (k, l) is
// (0, 1), (0, 2) or (1, 2); each case is spelled out with constant indices
// so that the matrices stay in registers.
struct Jacobi3 {
    double m[3][3], eig[3], orth[3][3];
    int max_col[3], changed[3], state;
    double c, s;

    template <int W, int X, int Y, int Z, bool R>
    __device__ __forceinline__ void rot(double (&M)[3][3]) {
        const double tt = M[W][X], uu = M[Y][Z];
        M[W][X] = tt * c - s * uu;
        M[Y][Z] = s * tt + c * uu;
        if (R) {
            if (fabs(M[W][X]) > fabs(col(M, W))) max_col[W] = X;
            if (Y < 3 && fabs(M[Y][Z]) > fabs(col(M, Y))) max_col[Y] = Z;
        }
    }
    // M[w][max_col[w]] for the rows that hold a cached column (0 and 1)
    __device__ __forceinline__ double col(const double (&M)[3][3], int w) const {
        const int j = max_col[w];
        return w == 0 ? (j == 1 ? M[0][1] : M[0][2]) : M[1][2];
    }
    __device__ __forceinline__ void update(int x, double y) {
        double &e = x == 0 ? eig[0] : (x == 1 ? eig[1] : eig[2]);
        int &ch = x == 0 ? changed[0] : (x == 1 ? changed[1] : changed[2]);
        e += y;
        if (ch && fabs(y) < 1e-6 * fabs(e)) {
            ch = 0;
            state--;
        } else if (!ch && fabs(y) > 1e-6 * fabs(e)) {
            ch = 1;
            state++;
        }
    }
    template <int K, int L>
    __device__ __forceinline__ bool step() {
        const double v = m[K][L];
        if (v == 0.0) return false;
        const double a = (eig[L] - eig[K]) * 0.5;
        double t = add_sqrt_ext(fabs(a), v * v + a * a);
        s = add_sqrt_ext(0.0, v * v + t * t);
        c = t / s;
        s = v / s;
        t = v * v / t;
        if (a < 0) {
            s = -s;
            t = -t;
        }
        update(K, -t);
        update(L, t);
        m[K][L] = 0.0;
        // src/jacobi.c:113-124: rows before k, between k and l, after l
        if (K == 1) rot<0, 1, 0, L, true>(m);
        if (K == 0 && L == 2) rot<0, 1, 1, 2, true>(m);
        if (K == 0 && L == 1) rot<0, 2, 1, 2, true>(m);
        rot<K, 0, L, 0, false>(orth);
        rot<K, 1, L, 1, false>(orth);
        rot<K, 2, L, 2, false>(orth);
        return true;
    }
    __device__ void run() {
        max_col[0] = fabs(m[0][2]) > fabs(m[0][1]) ? 2 : 1;  // ties keep the first column
        max_col[1] = 2;
        max_col[2] = 3;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) orth[i][j] = i == j ? 1.0 : 0.0;
        state = 0;
        for (int i = 0; i < 3; i++) {
            eig[i] = m[i][i];
            changed[i] = 1;
        }
        if (m[0][1] == 0.0 && m[0][2] == 0.0) { state--; changed[0] = 0; }
        if (m[1][0] == 0.0 && m[1][2] == 0.0) { state--; changed[1] = 0; }
        if (m[2][0] == 0.0 && m[2][1] == 0.0) { state--; changed[2] = 0; }
        state += 3;
        for (int guard = 0; state && guard < 1000; guard++) {
            const bool row1 = fabs(m[1][2]) > fabs(col(m, 0));
            bool go;
            if (row1) go = step<1, 2>();
            else if (max_col[0] == 1) go = step<0, 1>();
            else go = step<0, 2>();
            if (!go) break;
        }
        m[0][1] = m[1][0];
        m[0][2] = m[2][0];
        m[1][2] = m[2][1];
    }
};

// src/properties.c:270-325 for each (call, window) still running
__global__ void k_shape_calls(int32_t nh, int32_t it, const int32_t *__restrict__ call_tile0,
                              const int32_t *__restrict__ call_ntiles,
                              const ShapePart *__restrict__ part, ShapeState *__restrict__ st) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * nh) return;
    ShapeState &s = st[i];
    if (s.done) return;
    const int32_t h = i >> 1, w = i & 1;
    double wt = 0.0, t6[6] = {0, 0, 0, 0, 0, 0};
    int64_t cnt = 0;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) {
        const ShapePart &pt = part[2 * t + w];
        wt += pt.w;
        for (int k = 0; k < 6; k++) t6[k] += pt.t[k];
        cnt += pt.cnt;
    }
    if (s.limit < 0) {  // src/properties.c:228-236
        if (cnt < 3) {
            s.done = 1;
            return;
        }
        s.limit = (int32_t)min((int64_t)kShapeIterations, cnt);
    }
    if (it >= s.limit || !(wt != 0.0)) {
        s.done = 1;
        return;
    }
    Jacobi3 jb;
    jb.m[0][0] = t6[0] / wt;
    jb.m[1][1] = t6[1] / wt;
    jb.m[2][2] = t6[2] / wt;
    jb.m[0][1] = jb.m[1][0] = t6[3] / wt;
    jb.m[0][2] = jb.m[2][0] = t6[4] / wt;
    jb.m[1][2] = jb.m[2][1] = t6[5] / wt;
    const float inert[6] = {(float)jb.m[0][0], (float)jb.m[1][1], (float)jb.m[2][2],
                            (float)jb.m[0][1], (float)jb.m[1][2], (float)jb.m[2][0]};
    jb.run();
    const double e0 = jb.eig[0], e1 = jb.eig[1], e2 = jb.eig[2];
    auto E = [&](int q) { return q == 0 ? e0 : (q == 1 ? e1 : e2); };
    int a = 0, b = 1, c = 2;
    if (e1 > e0) {
        b = 0;
        a = 1;
    }
    if (e2 > E(b)) {
        c = b;
        b = 2;
    }
    if (E(b) > E(a)) {
        const int q = a;
        a = b;
        b = q;
    }
    const double ea = E(a), eb = E(b), ec = E(c);
    for (int k = 0; k < 3; k++) {
        s.eig[k] = jb.eig[k];
        for (int l = 0; l < 3; l++) s.orth[k][l] = jb.orth[k][l];
    }
    if (!(ea != 0.0) || !(eb != 0.0) || !(ec != 0.0)) {
        s.done = 1;
        return;
    }
    const float b_to_a = (float)sqrt(eb / ea);
    const float c_to_a = (float)sqrt(ec / ea);
    if ((fabs((double)(b_to_a - s.b_to_a)) < 0.01 * (double)s.b_to_a) &&
        (fabs((double)(c_to_a - s.c_to_a)) < 0.01 * (double)s.c_to_a)) {
        s.done = 1;
        return;
    }
    s.b_to_a = b_to_a > 0 ? b_to_a : 0.0f;
    s.c_to_a = c_to_a > 0 ? c_to_a : 0.0f;
    const double r = sqrt(ea);
    const double oa[3] = {a == 0 ? jb.orth[0][0] : (a == 1 ? jb.orth[1][0] : jb.orth[2][0]),
                          a == 0 ? jb.orth[0][1] : (a == 1 ? jb.orth[1][1] : jb.orth[2][1]),
                          a == 0 ? jb.orth[0][2] : (a == 1 ? jb.orth[1][2] : jb.orth[2][2])};
    const double sc = ((double)(s.r * s.r) * 1e-6) / (r * r);
    for (int k = 0; k < 3; k++) {
        s.A[k] = (float)(1e3 * r * oa[k]);
        s.eig[k] = jb.eig[k] * sc;
    }
    for (int k = 0; k < 6; k++) s.inertia[k] = inert[k];
    s.assigned = 1;
}

// ------------------------------------------- scale radius, pseudo-evolution

constexpr int kScaleBins = 50;    // MAX_SCALE_BINS, src/nfw.c:9
constexpr int kMinPerBin = 15;    // MIN_PART_PER_BIN
constexpr int kMinScale = 100;    // MIN_SCALE_PART
constexpr double kRmaxToRs = 2.1626;       // src/universal_constants.h:7
constexpr double kRsConstant = 0.216216595;  // src/universal_constants.h:8

struct ScaleCall {
    int32_t ppbin;     // bound records per bin; 0: no fit (fewer than 100)
    int32_t nbins;
    int32_t end;       // bins come from records [0, end) = [0, dens_tot - 1)
    int32_t pad;
    float bin_r[kScaleBins + 1];
    float klypin, rs, chi2;  // results for the fitted calls
};

// src/nfw.c:134-152: the bound records of [0, dens_tot - 1) and the bin size
__global__ void k_scale_init(int32_t nh, const struct potential *__restrict__ srt,
                             const int64_t *__restrict__ hfirst,
                             const int32_t *__restrict__ call_tile0,
                             const int32_t *__restrict__ call_ntiles,
                             const PropCall *__restrict__ pc, const PropSums *__restrict__ ts,
                             ScaleCall *__restrict__ sc, int32_t *__restrict__ fit_list,
                             int32_t *__restrict__ fit_n) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    const PropCall c = pc[h];
    int64_t nb = 0;  // bound records in [0, dens_tot)
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) nb += ts[t].n;
    const int32_t end = c.dens_tot - 1;
    if (end >= 0) {
        const struct potential &r = srt[hfirst[h] + end];
        if (!(r.pe < (double)r.ke)) nb--;
    }
    ScaleCall s;
    s.end = end;
    s.pad = 0;
    s.nbins = 0;
    s.ppbin = 0;
    if (nb >= kMinScale) {
        int64_t pp = (int64_t)ceil((double)nb / (double)kScaleBins);
        if (pp < kMinPerBin) pp = nb / (nb / kMinPerBin);
        s.ppbin = (int32_t)pp;
        s.nbins = (int32_t)(nb / pp);
        fit_list[atomicAdd(fit_n, 1)] = h;  // the fits run one call per lane, together
    }
    for (int k = 0; k <= kScaleBins; k++) s.bin_r[k] = 0.0f;
    s.klypin = s.rs = s.chi2 = 0.0f;
    sc[h] = s;
}

// src/nfw.c:153-162: the outer edge of bin k is between the k * ppbin-th
// bound record and the record after it
__global__ __launch_bounds__(kWave) void k_scale_bins(const struct potential *__restrict__ srt,
                                                      const ScanTile *__restrict__ tiles,
                                                      int32_t ntiles,
                                                      const int64_t *__restrict__ hfirst,
                                                      const PropCount *__restrict__ pre,
                                                      ScaleCall *__restrict__ sc) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile tl = tiles[t];
    const int32_t pp = sc[tl.call].ppbin, end = min(tl.off + kScanTile, sc[tl.call].end);
    if (pp == 0 || tl.off >= end) return;
    const int64_t first = hfirst[tl.call];
    int32_t carry = pre[t].nb;
    for (int32_t base = tl.off; base < end; base += kWave) {
        const int32_t j = base + lane_id();
        const bool v = j < end;
        bool b = false;
        if (v) b = !(srt[first + j].pe < (double)srt[first + j].ke);
        const int32_t n = wave_scan_add(b ? 1 : 0) + carry;
        carry = read_lane(n, kWave - 1);
        if (b && n % pp == 0 && n / pp <= kScaleBins) {
            const float r = (float)(1e3 * 0.5 * (sqrt((double)srt[first + j].r2) +
                                                 sqrt((double)srt[first + j + 1].r2)));
            sc[tl.call].bin_r[n / pp] = r;
        }
    }
}

__device__ double c_to_f(double c) {
    const double cp1 = 1.0 + c;
    return c * cp1 / (log1p(c) * cp1 - c);
}

// concentration with c_to_f(c) = f, src/nfw.c:18-33
__device__ double f_to_c(double f) {
    double c = f, tc = c_to_f(c);
    for (int guard = 0; fabs((f - tc) / f) > 1e-7 && guard < 10000; guard++) {
        const double tc2 = c_to_f(c + 0.1);
        const double slope = (tc2 - tc) / 0.1;
        const double nc = c + (f - tc) / slope;
        if (nc < 0) c /= 2;
        else c = nc;
        tc = c_to_f(c);
    }
    return c;
}

// Klypin's rs from vmax, rvmax and the virial mass and radius, src/nfw.c:35-50
__device__ float klypin_rs(float mvir, float rvir, float vmax, float rvmax, float scale) {
    if (!(mvir != 0.0f) || !(rvir != 0.0f) || !(vmax != 0.0f))
        return (float)(rvmax / kRmaxToRs);
    float vm2 = (float)(vmax / kVmaxConst);
    vm2 = vm2 * vm2 * scale;
    const float f = (float)((rvir / 1.0e3) * vm2 / (mvir * kRsConstant));
    if (f < 4.625) return (float)(rvmax / kRmaxToRs);
    const float c = (float)f_to_c(f);
    if (c <= 0) return (float)(rvmax / kRmaxToRs);
    return rvir / c;
}

__device__ __forceinline__ float nfw_menc(float r, float rs) {
    return (float)(log((double)((rs + r) / rs)) - (double)(r / (rs + r)));
}

// bin weights, src/nfw.c:164-166: 0.1 for bins starting inside 3 FORCE_RES
__device__ __forceinline__ float bin_weight(const float *bin_r, int i, double force_res) {
    return bin_r[i] * 1e-3 < 3 * force_res ? 0.1f : 1.0f;
}

__device__ float chi2_rs(float rs, const float *bin_r, double force_res, int nb) {
    const float bin_mass = (float)(nfw_menc(bin_r[nb], rs) / (double)nb);
    float last = 0, chi2 = 0;
    for (int i = 0; i < nb; i++) {
        const float menc = nfw_menc(bin_r[i + 1], rs);
        const float dx = bin_weight(bin_r, i, force_res) * ((menc - last) - bin_mass) / bin_mass;
        chi2 += dx * dx;
        last = menc;
    }
    return chi2;
}

// parabolic steps on chi2(rs), src/nfw.c:75-128
__device__ float fit_rs(float rs, const float *bin_r, double force_res, int nb,
                       float *out_chi2) {
    if (!(rs > bin_r[0]) || !(rs < bin_r[nb])) rs = bin_r[(nb + 1) / 2];
    const float initial = rs;
    float chi2 = chi2_rs(rs, bin_r, force_res, nb);
    float drs = (float)(rs / 10.0);
    int iter = 0;
    float last = 1e30f;
    if (!(rs > 0)) return initial;
    for (int guard = 0; fabs((double)(chi2 - last)) > 0.005 * chi2 && iter < kMinPerBin &&
                        guard < 10000; guard++) {
        last = chi2;
        if (rs + drs > bin_r[nb]) drs = (float)(0.1 * (bin_r[nb] - rs));
        if (rs - drs < bin_r[0]) drs = (float)(0.1 * (rs - bin_r[0]));
        const float cr = chi2_rs(rs + drs, bin_r, force_res, nb);
        const float cl = chi2_rs(rs - drs, bin_r, force_res, nb);
        const float dx = (float)(0.5 * (cr - cl) / drs);
        const float dx2 = (float)((cr + cl - 2.0 * chi2) / (drs * drs));
        float move = 0;
        if (dx2 != 0) move = -dx / dx2;
        if (!(move != 0.0f)) return rs;
        if (rs + move > 4 * rs) move = (float)(3.0 * rs);
        if (rs + move < 0.25 * rs) move = (float)(-0.75 * rs);
        float nrs = rs + move;
        if (nrs > bin_r[nb]) nrs = (float)(rs + 0.8 * (bin_r[nb] - rs));
        else if (nrs < bin_r[0]) nrs = (float)(rs + 0.8 * (bin_r[0] - rs));
        float nchi2 = chi2_rs(nrs, bin_r, force_res, nb);
        if (cr < nchi2) {
            nchi2 = cr;
            nrs = rs + drs;
        }
        if (cl < nchi2) {
            nchi2 = cl;
            nrs = rs - drs;
        }
        if (last < nchi2) {
            drs *= 0.5f;
            continue;
        }
        drs = (float)(fabs((double)(rs - nrs)) / 10.0);
        rs = nrs;
        chi2 = nchi2;
        iter++;
    }
    *out_chi2 = chi2;
    return rs;
}

// the calls with a fit, one per lane: Klypin's estimate, then the fit
__global__ void k_scale_fit(const int32_t *__restrict__ fit_list,
                            const int32_t *__restrict__ fit_n, const PropCall *__restrict__ pc,
                            const hp_halo_props *__restrict__ props, PropArgs pa,
                            ScaleCall *__restrict__ sc) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *fit_n) return;
    const int32_t h = fit_list[i];
    const hp_halo_props &o = props[h];
    const double m = (double)pc[h].n0 * pa.particle_mass;
    ScaleCall &sk = sc[h];
    const float rs0 = klypin_rs((float)m, o.r, o.vmax, o.rvmax, pa.scale_now);
    float chi2 = 0.0f;
    const float rs = fit_rs(rs0, sk.bin_r, pa.force_res, sk.nbins, &chi2);
    sk.klypin = rs0;
    sk.rs = rs;
    sk.chi2 = chi2;
}

// per call: the shape fields, the scale radius (src/nfw.c:130-173), m_pe_b
// and the radius that bounds m_pe_d (src/properties.c:351-382)
__global__ void k_props_late(int32_t nh, const PropCall *__restrict__ pc,
                             const ShapeState *__restrict__ st, const ScaleCall *__restrict__ sc,
                             PropArgs pa, double *__restrict__ r_pe,
                             hp_halo_props *__restrict__ props) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    hp_halo_props o = props[h];
    const PropCall c = pc[h];
    const ShapeState &s0 = st[2 * h], &s1 = st[2 * h + 1];
    o.b_to_a2 = s0.b_to_a;
    o.c_to_a2 = s0.c_to_a;
    o.b_to_a = s1.b_to_a;
    o.c_to_a = s1.c_to_a;
    for (int k = 0; k < 3; k++) {
        o.A2[k] = s0.A[k];
        o.A[k] = s1.A[k];
    }
    // the second calc_shape keeps the first's tensor until it writes its own
    for (int k = 0; k < 6; k++) {
        o.inertia_tensor2[k] = s0.inertia[k];
        o.inertia_tensor[k] = s1.assigned ? s1.inertia[k] : s0.inertia[k];
    }
    const ScaleCall &sk = sc[h];
    float rs;
    if (sk.ppbin > 0) {  // fitted by k_scale_fit
        o.klypin_rs = sk.klypin;
        rs = sk.rs;
        o.chi2 = sk.chi2;
    } else {  // fewer than 100 bound records: Klypin's estimate
        const double m = (double)c.n0 * pa.particle_mass;
        rs = klypin_rs((float)m, o.r, o.vmax, o.rvmax, pa.scale_now);
        o.klypin_rs = rs;
    }
    o.rs = rs;
    // src/properties.c:351-382: m_pe_d counts the bound records with
    // r < r_pe_d; those come first (r2 is sorted): k_pe_tiles counts them
    double r_pe_d = rs * 4.0;
    if (r_pe_d < o.r / 5.0) r_pe_d = o.r / 5.0;
    r_pe_d *= 1e-3;
    r_pe[h] = r_pe_d;
    o.m_pe_b = (float)(pa.particle_mass * pow(c.pe_b, 2.0 / 3.0) /
                       cbrt(4.0 * M_PI * pa.rvir_dens_z0 / 3.0));
    props[h] = o;
}

// bound records with r < r_pe_d per tile
__global__ __launch_bounds__(kWave) void k_pe_tiles(const struct potential *__restrict__ srt,
                                                    const ScanTile *__restrict__ tiles,
                                                    int32_t ntiles,
                                                    const hp_halo *__restrict__ halos,
                                                    const int64_t *__restrict__ hfirst,
                                                    const int64_t *__restrict__ hcount,
                                                    const double *__restrict__ r_pe,
                                                    int32_t *__restrict__ cnt) {
    const int32_t t = blockIdx.x;
    if (t >= ntiles) return;
    const ScanTile tl = tiles[t];
    const int64_t first = hfirst[tl.call];
    const double lim = r_pe[tl.call];
    const int32_t end = min(tl.off + kScanTile, call_kept(halos, hcount, tl.call));
    int32_t n = 0;
    for (int32_t j = tl.off + lane_id(); j < end; j += kWave) {
        const struct potential &r = srt[first + j];
        n += !(r.pe < (double)r.ke) && sqrt((double)r.r2) < lim ? 1 : 0;
    }
    n = wave_sum(n);
    if (lane_id() == 0) cnt[t] = n;
}

__global__ void k_pe_calls(int32_t nh, const int32_t *__restrict__ call_tile0,
                           const int32_t *__restrict__ call_ntiles,
                           const int32_t *__restrict__ cnt, PropArgs pa,
                           hp_halo_props *__restrict__ props) {
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nh) return;
    int64_t n = 0;
    for (int32_t t = call_tile0[h]; t < call_tile0[h] + call_ntiles[h]; t++) n += cnt[t];
    props[h].m_pe_d = (float)((double)n * pa.particle_mass);
}

}  // namespace hp
