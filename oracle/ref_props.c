/*
 * oracle/ref_props.c -- a calling harness around the reference's own halo
 * property pass, for the checker library _ref/libref_props.so.
 *
 * TEST INFRASTRUCTURE ONLY (see restate.c).  Nothing here restates the
 * reference: the harness fills the reference's globals and `struct halo`,
 * calls its functions and copies their results out.
 *
 *   ref_halo_props      -> _calc_additional_halo_props(h, total_p, 0) and
 *                          (h, total_p, 1), as src/properties.c:552-553 calls
 *                          them (src/properties.c:384-510 with calc_shape,
 *                          calc_scale_radius, estimate_total_energy and
 *                          _calc_pseudo_evolution_masses inside)
 *   ref_mass_definition -> calc_mass_definition (src/groupies.c:124-142)
 *   ref_particle_radii  -> calc_particle_radii(h, h, h->pos, 0, 0, 0)
 *                          (src/groupies.c:235-268), the record set-up of
 *                          calc_additional_halo_props (src/properties.c:526)
 *   ref_bound_pass      -> the bound-pass steps of calc_additional_halo_props
 *                          (src/properties.c:537-550) on prepared records:
 *                          compute_potential, the potential-only strip,
 *                          qsort with dist_compare, calculate_corevel,
 *                          compute_kinetic_energy -- the reference's own
 *                          functions; the harness only sequences them
 *   ref_subfof          -> the phase-space sub-FOF step (_find_subfofs_better2,
 *                          src/groupies.c:287-314), with the intermediate
 *                          state its caller never sees (tree order, sample
 *                          distances, linking length); ref_next_closest and
 *                          ref_find_median_r expose its two search steps
 *   ref_halo_props_batch-> ref_halo_props for every call of a batch, OpenMP
 *                          dynamic over calls as the halo finder's loop over
 *                          halos (src/mpi_main.cpp:1377); the reference's
 *                          per-thread globals (po, particle_thresh_dens) are
 *                          threadprivate, so calls run concurrently
 *
 * The records must already be in the state src/properties.c:551 leaves them:
 * potential-only records dropped, sorted by r2, pe and ke set.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "config_vars.h"
#include "fof.h"
#include "groupies.h"
#include "halo.h"
#include "particle.h"
#include "potential.h"

extern struct potential *po; /* src/groupies.c:33 */
#pragma omp threadprivate(po)
extern double particle_thresh_dens[5]; /* src/groupies.c:56 */
extern double particle_rvir_dens, particle_rvir_dens_z0;
#pragma omp threadprivate(particle_thresh_dens, particle_rvir_dens, particle_rvir_dens_z0)

extern struct particle *copies; /* src/groupies.c:30 */
#pragma omp threadprivate(copies)

int64_t calc_particle_radii(struct halo *base_h, struct halo *h, float *cen, int64_t p_start,
                            int64_t level, int64_t potential_only,
                            const struct HaloInfo *haloinfo);
void _calc_additional_halo_props(struct halo *h, int64_t total_p, int64_t bound);
void calc_mass_definition(void);
void calculate_corevel(struct halo *h, struct potential *po, int64_t total_p);
int dist_compare(const void *a, const void *b);

typedef struct ref_props {
    float m, mgrav, r, vmax, rvmax, vrms, Xoff, Voff, alt_m[4], halfmass_radius;
    float J[3], energy, kin_to_pot, spin, bullock_spin;
    float rs, klypin_rs, b_to_a, c_to_a, A[3], b_to_a2, c_to_a2, A2[3];
    float m_pe_b, m_pe_d;
    float chi2, inertia_tensor[6], inertia_tensor2[6];
} ref_props;

void ref_mass_definition(double out[7]) {
    calc_mass_definition();
    for (int k = 0; k < 5; k++) out[k] = particle_thresh_dens[k];
    out[5] = particle_rvir_dens;
    out[6] = particle_rvir_dens_z0;
}

/* thresh: particle_thresh_dens[0..4] then particle_rvir_dens, as
   ref_mass_definition returns them (the halo finder computes them once per
   snapshot, src/groupies.c:124); particle_rvir_dens_z0 comes from this
   thread's own calc_mass_definition */
void ref_halo_props(struct potential *recs, int64_t total_p, const float pos[6],
                    const double thresh[7], ref_props *out) {
    struct halo h;
    memset(&h, 0, sizeof h);
    memcpy(h.pos, pos, sizeof h.pos);
    /* every thread's own copy of the mass-definition state, as the halo
       finder's threads hold it (particle_rvir_dens_z0 feeds the
       pseudo-evolution masses); then the thresholds given */
    calc_mass_definition();
    for (int k = 0; k < 5; k++) particle_thresh_dens[k] = thresh[k];
    particle_rvir_dens = thresh[5];
    po = recs;
    _calc_additional_halo_props(&h, total_p, 0);
    _calc_additional_halo_props(&h, total_p, 1);
    po = NULL;
    ref_props o;
    o.m = h.m;
    o.mgrav = h.mgrav;
    o.r = h.r;
    o.vmax = h.vmax;
    o.rvmax = h.rvmax;
    o.vrms = h.vrms;
    o.Xoff = h.Xoff;
    o.Voff = h.Voff;
    memcpy(o.alt_m, h.alt_m, sizeof o.alt_m);
    o.halfmass_radius = h.halfmass_radius;
    memcpy(o.J, h.J, sizeof o.J);
    o.energy = h.energy;
    o.kin_to_pot = h.kin_to_pot;
    o.spin = h.spin;
    o.bullock_spin = h.bullock_spin;
    o.rs = h.rs;
    o.klypin_rs = h.klypin_rs;
    o.b_to_a = h.b_to_a;
    o.c_to_a = h.c_to_a;
    memcpy(o.A, h.A, sizeof o.A);
    o.b_to_a2 = h.b_to_a2;
    o.c_to_a2 = h.c_to_a2;
    memcpy(o.A2, h.A2, sizeof o.A2);
    o.m_pe_b = h.m_pe_b;
    o.m_pe_d = h.m_pe_d;
    o.chi2 = h.chi2;
    memcpy(o.inertia_tensor, h.inertia_tensor, sizeof o.inertia_tensor);
    memcpy(o.inertia_tensor2, h.inertia_tensor2, sizeof o.inertia_tensor2);
    *out = o;
}

void ref_halo_props_batch(struct potential *recs, const int64_t *first, const int64_t *kept,
                          const float *pos6, int64_t nhalos, const double thresh[7],
                          ref_props *out, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t h = 0; h < nhalos; h++)
        ref_halo_props(recs + first[h], kept[h], pos6 + 6 * h, thresh, out + h);
}

/* halo fields the bound pass reads and writes (src/halo.h:8-16) */
typedef struct ref_bound_halo {
    float   pos[6], corevel[3], bulkvel[3];
    float   min_vel_err, min_bulkvel_err;
    int64_t n_core;
    int32_t is_sub, pad;
} ref_bound_halo;

/* src/properties.c:537-550 with the halo already resolved: is_sub selects
   the KE centre as extra_info[].sub_of > -1 does.  Returns total_p after
   the potential-only records are dropped. */
int64_t ref_bound_pass(struct potential *recs, int64_t total_p, ref_bound_halo *bh) {
    struct halo h;
    memset(&h, 0, sizeof h);
    memcpy(h.pos, bh->pos, sizeof h.pos);
    memcpy(h.corevel, bh->corevel, sizeof h.corevel);
    memcpy(h.bulkvel, bh->bulkvel, sizeof h.bulkvel);
    float scale_now = SCALE_NOW;
    calc_mass_definition(); /* particle_rvir_dens for calculate_corevel */
    if (total_p > 1) compute_potential(recs, total_p);
    for (int64_t j = 0; j < total_p; j++)
        if (recs[j].ke < 0) {
            total_p--;
            recs[j] = recs[total_p];
            j--;
        }
    qsort(recs, total_p, sizeof(struct potential), dist_compare);
    calculate_corevel(&h, recs, total_p);
    if (bh->is_sub)
        compute_kinetic_energy(recs, total_p, h.corevel, h.pos, scale_now);
    else
        compute_kinetic_energy(recs, total_p, h.bulkvel, h.pos, scale_now);
    memcpy(bh->pos, h.pos, sizeof h.pos);
    memcpy(bh->corevel, h.corevel, sizeof h.corevel);
    memcpy(bh->bulkvel, h.bulkvel, sizeof h.bulkvel);
    bh->min_vel_err = h.min_vel_err;
    bh->min_bulkvel_err = h.min_bulkvel_err;
    bh->n_core = h.n_core;
    return total_p;
}

/* one halo of the hierarchy (the product's hp_tree_halo layout) */
typedef struct ref_tree_halo {
    int64_t p_start, num_p, num_child_particles;
    int64_t child, next_cochild, sub_of;
    float   pos[6];
} ref_tree_halo;

/* po must hold every record the recursion writes (the caller sizes it) */
int64_t ref_particle_radii(struct particle *parts, const ref_tree_halo *tree, int64_t nh,
                           int64_t base, double ratio, struct potential *out) {
    struct halo *hs = calloc((size_t)nh, sizeof(struct halo));
    struct extra_halo_info *xs = calloc((size_t)nh, sizeof(struct extra_halo_info));
    for (int64_t i = 0; i < nh; i++) {
        hs[i].p_start = tree[i].p_start;
        hs[i].num_p = tree[i].num_p;
        hs[i].num_child_particles = tree[i].num_child_particles;
        memcpy(hs[i].pos, tree[i].pos, sizeof hs[i].pos);
        xs[i].child = tree[i].child;
        xs[i].next_cochild = tree[i].next_cochild;
        xs[i].sub_of = tree[i].sub_of;
    }
    struct HaloInfo hi = {hs, xs, nh};
    INCLUDE_HOST_POTENTIAL_RATIO = ratio;
    copies = parts;
    po = out;
    int64_t n = calc_particle_radii(hs + base, hs + base, hs[base].pos, 0, 0, 0, &hi);
    po = NULL;
    copies = NULL;
    free(hs);
    free(xs);
    return n;
}

/* ref_particle_radii for every listed halo, OpenMP dynamic over halos (the
   halo finder's per-thread copies / po / halo_ids are threadprivate); call
   b's records go to out + offset[b] */
void ref_particle_radii_batch(struct particle *parts, const ref_tree_halo *tree, int64_t nh,
                              const int64_t *base, const int64_t *offset, int64_t nbase,
                              double ratio, struct potential *out, int64_t *count, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t b = 0; b < nbase; b++)
        count[b] = ref_particle_radii(parts, tree, nh, base[b], ratio, out + offset[b]);
}

/* ------------------------------------------------------------ sub-FOF ---
 * The phase-space sub-FOF step of _find_subs (src/groupies.c:529):
 * _find_subfofs_better2 (src/groupies.c:287-314) with norm_sd
 * (src/groupies.c:788-830), the 6D GROUPIES fast3tree (leaf 40,
 * src/groupies.c:24-28), find_median_r (769-786) and _find_subfofs_at_r
 * (274-285) over src/fof.c. */
struct fast3tree;
struct fast3tree_results;
extern struct fast3tree *phasetree; /* src/groupies.c:44 */
extern struct fast3tree_results *res;
extern float *particle_r; /* src/groupies.c:32 */
extern int64_t num_copies;
#pragma omp threadprivate(phasetree, res, particle_r, num_copies)
struct fast3tree *GROUPIESfast3tree_init(int64_t n, struct particle *p);
void GROUPIESfast3tree_rebuild(struct fast3tree *t, int64_t n, struct particle *p);
struct fast3tree_results *GROUPIESfast3tree_results_init(void);
float GROUPIESfast3tree_find_next_closest_distance(struct fast3tree *t,
                                                   struct fast3tree_results *r, float c[6]);
void _find_subfofs_better2(struct fof *f, float thresh, struct FOFInfo *fofinfo);
void _find_subfofs_at_r(struct fof *f, float target_r, struct FOFInfo *fofinfo);

static void sub_setup(int64_t n, int top) {
    if (!phasetree) phasetree = GROUPIESfast3tree_init(0, NULL);
    if (!res) res = GROUPIESfast3tree_results_init();
    free(particle_r);
    particle_r = malloc(sizeof(float) * (size_t)(n > 10000 ? n : 10000));
    num_copies = top ? n : n + 1; /* norm_sd scales x when f->num_p == num_copies */
}

/*
 * One group, two ways:
 *  (1) the steps of _find_subfofs_better2 in its order, through the
 *      reference's own functions, keeping what it does not return:
 *      tree_order = the particles after norm_sd and fast3tree_rebuild (the
 *      order the linking loop walks), *target_r = find_median_r's result,
 *      sample_r = the next-closest distances of the sample, in sample order;
 *  (2) _find_subfofs_better2 itself on a fresh copy: final = its particle
 *      array (sorted by sub-FOF), fof_first / fof_num_p = its sub-FOFs.
 * Also (3) _find_subfofs_at_r at (1)'s target_r on (1)'s tree order, whose
 * sub-FOFs must equal (2)'s (returned in at_first / at_num_p). */
int64_t ref_subfof(const struct particle *parts, int64_t n, float thresh, int top,
                   float *target_r, struct particle *tree_order, float *sample_r,
                   int64_t *nsample, struct particle *final, int64_t *fof_first,
                   int64_t *fof_num_p, int64_t cap, int64_t *at_first, int64_t *at_num_p,
                   int64_t *nat) {
    sub_setup(n, top);
    /* (1) */
    struct particle *a = malloc(sizeof(struct particle) * (size_t)(n ? n : 1));
    memcpy(a, parts, sizeof(struct particle) * (size_t)n);
    struct fof fa = {n, a};
    int64_t num_test = 10000;
    if (EXACT_LL_CALC) num_test = n;
    norm_sd(&fa, thresh);
    GROUPIESfast3tree_rebuild(phasetree, n, a);
    memcpy(tree_order, a, sizeof(struct particle) * (size_t)n);
    if (num_test > n) num_test = n;
    unsigned int seed = (unsigned int)n;
    for (int64_t i = 0; i < num_test; i++) {
        int64_t j;
        if (num_test == n)
            j = i;
        else {
            j = rand_r(&seed);
            j <<= 31;
            j += rand_r(&seed);
            j %= n;
        }
        particle_r[i] = GROUPIESfast3tree_find_next_closest_distance(phasetree, res, a[j].pos);
    }
    memcpy(sample_r, particle_r, sizeof(float) * (size_t)num_test);
    *nsample = num_test;
    *target_r = find_median_r(particle_r, num_test, thresh, &seed);
    /* (3) */
    struct FOFInfo fi;
    init_fofinfo(&fi);
    _find_subfofs_at_r(&fa, *target_r, &fi);
    *nat = fi.num_fofs;
    for (int64_t i = 0; i < fi.num_fofs && i < cap; i++) {
        at_first[i] = fi.fofs[i].particles - a;
        at_num_p[i] = fi.fofs[i].num_p;
    }
    free_fofinfo(&fi);
    /* (2) */
    memcpy(final, parts, sizeof(struct particle) * (size_t)n);
    struct fof fb = {n, final};
    init_fofinfo(&fi);
    _find_subfofs_better2(&fb, thresh, &fi);
    int64_t nf = fi.num_fofs;
    for (int64_t i = 0; i < nf && i < cap; i++) {
        fof_first[i] = fi.fofs[i].particles - final;
        fof_num_p[i] = fi.fofs[i].num_p;
    }
    free_fofinfo(&fi);
    free(a);
    return nf;
}

/* fast3tree_find_next_closest_distance for each of the given points on a
   tree rebuilt over pts (in place, like src/groupies.c:294) */
void ref_next_closest(struct particle *pts, int64_t n, const float *q, int64_t nq, float *out) {
    sub_setup(n, 0);
    GROUPIESfast3tree_rebuild(phasetree, n, pts);
    for (int64_t i = 0; i < nq; i++) {
        float c[6];
        memcpy(c, q + 6 * i, sizeof c);
        out[i] = GROUPIESfast3tree_find_next_closest_distance(phasetree, res, c);
    }
}

/* find_median_r (src/groupies.c:769-786) on a copy of rad with the given seed */
float ref_find_median_r(const float *rad, int64_t n, float frac, unsigned int seed) {
    float *r = malloc(sizeof(float) * (size_t)(n ? n : 1));
    memcpy(r, rad, sizeof(float) * (size_t)n);
    float v = find_median_r(r, n, frac, &seed);
    free(r);
    return v;
}

/* _find_subfofs_better2 for every group of a batch (groups at parts +
   first[g], num_p[g]), OpenMP dynamic over groups as the halo finder's loop
   over FOF groups; the per-thread tree, results and sample buffer are the
   reference's threadprivate globals.  Returns the total sub-FOF count. */
int64_t ref_subfof_batch(struct particle *parts, const int64_t *first, const int64_t *num_p,
                         int64_t ng, float thresh, int top, int nthreads) {
    int64_t total = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
    for (int64_t g = 0; g < ng; g++) {
        sub_setup(num_p[g], top);
        struct fof f = {num_p[g], parts + first[g]};
        struct FOFInfo fi;
        init_fofinfo(&fi);
        _find_subfofs_better2(&f, thresh, &fi);
        total += fi.num_fofs;
        free_fofinfo(&fi);
    }
    return total;
}
