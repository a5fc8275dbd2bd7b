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

static_assert(sizeof(hp_halo_props) == 120 && offsetof(hp_halo_props, n_mdelta) == 88,
              "hp_halo_props layout");

constexpr double kVmaxConst = 6.55833746e-5;  // src/universal_constants.h:6
constexpr double kGc = 4.30117902e-9;         // src/universal_constants.h:4

// thresholds and constants of one property pass
struct PropArgs {
    double dens_thresh;   // particle_thresh_dens[0] * 4pi/3
    double d[4];          // particle_thresh_dens[1..4] * 4pi/3
    double rvir_thresh;   // particle_rvir_dens * 4pi/3
    double thresh0;       // particle_thresh_dens[0]
    double rvir_dens;     // particle_rvir_dens
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
    }
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
    o.pad = 0;
    o.n_mdelta = c.n0;
    o.dens_tot = c.dens_tot;
    o.n_vir = c.vir;
    o.n_half = ih;
    props[h] = o;
}

}  // namespace hp
