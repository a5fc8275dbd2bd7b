/*
 * halopot.h -- MI355X (gfx950) per-halo gravitational potential and bound
 * pass, behind MPI-Rockstar's own C signatures.
 *
 * libhalopot.so links in place of the reference's potential.o
 * (src/Makefile:34, MPI_ROCKSTAR_CORE).  The two reference entry points keep
 * their exact signatures, so src/properties.c:538 and :548,550 (compiled
 * inside src/groupies.c:272) call them unchanged.  The hp_* entry points are
 * additions: the batched path that the reference does not have, and
 * diagnostics.
 *
 * Every entry point runs on the GPU.  There is no CPU fallback: if no gfx950
 * device is usable the call fails through the error path below.
 */
#ifndef HALOPOT_H
#define HALOPOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ record
 * src/potential.h:7-21, built without POTENTIAL_COMPARISON / CALC_POTENTIALS
 * (the reference's default, src/Makefile:26).  48 bytes: pos[6] at 0, r2 at
 * 24, pe at 32, ke at 40, flags at 44. */
#ifndef _POTENTIAL_H_
struct potential {
    float   pos[6], r2;
    double  pe;
    float   ke;
    int32_t flags;
};
#endif

/* ------------------------------------------------- reference entry points */

/* Replaces compute_potential, src/potential.h:24 / src/potential.c:183-192.
 * Writes pe for every record from the Barnes-Hut potential of the reference's
 * k-d tree (10 points per leaf, dmin acceptance with ERR_TOL 1, monopoles).
 * Non-finite positions get pe = 0 and are left out of the tree
 * (src/fast3tree.c:875-887).  Like the reference, the call permutes po: the
 * records end in the tree's order, the order the reference's in-place
 * non-finite removal and partitions (src/fast3tree.c:762-773, 875-887)
 * leave, every field travelling with its record; so the caller's strip and
 * stable sort (src/properties.c:539-545) give the reference's order, ties
 * included.  Reads the globals FORCE_RES and PARTICLE_MASS at call time.  Safe to call from many host threads at once
 * (src/mpi_main.cpp:1369-1402); each thread gets its own stream and device
 * scratch. */
void compute_potential(struct potential *po, int64_t num_po);

/* Replaces compute_kinetic_energy, src/potential.h:22-23 /
 * src/potential.c:194-211.  Records with ke < 0 (potential-only) are left
 * alone.  ke = a/(2 Gc) |v - v_c + H(a) a (x - x_c)|^2, accumulated as the
 * reference does (float ke, double terms).  Reads Om, Ol, W0, WA
 * (src/hubble.c:8-19). */
void compute_kinetic_energy(struct potential *po, int64_t num_po,
                            float *vel_cen, float *pos_cen, float scale_now);

/* No counterpart in the reference: compute_potential for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same tree, decisions, values and
 * permutation of the records as compute_potential, on device-resident
 * records d_po[0..num_po), at most 32768 records (HP_EINVAL above).  It
 * calls no hipMalloc, hipFree, hipHostMalloc and creates no stream or event:
 * its scratch is one static device array of 20 MiB per GPU, held until the
 * process ends, whose address is looked up on the first call on each device.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (from several host threads or streams) race: serialise them.  Runs on
 * `stream` (a hipStream_t, or NULL for the null stream) and returns when the
 * potential is done (it reads the tree's level counters back to the host
 * between levels).  Returns 0 or an HP_E* code. */
int hp_compute_potential_noalloc(struct potential *d_po, int64_t num_po, void *stream);

/* Configuration globals read at call time (src/config.template.h:7,16,20,
 * 22-25 through src/config_vars.h).  libhalopot carries weak definitions with
 * the reference's defaults, so the program that defines the real ones (the
 * halo finder) overrides them and stand-alone callers can set them here.
 * SCALE_NOW is read only by the batched bound pass (the reference entry
 * point takes scale_now as an argument). */
extern double FORCE_RES, PARTICLE_MASS, Om, Ol, W0, WA, SCALE_NOW;

/* ----------------------------------------------------------- batched path */

/* One compute_potential call of a batch: a contiguous run of records in the
 * batch's record array, plus the halo fields the bound pass reads and
 * writes (src/halo.h:8-35: pos, corevel, bulkvel, min_vel_err,
 * min_bulkvel_err, n_core).  104 bytes. */
typedef struct hp_halo {
    int64_t first;          /* index of the call's first record */
    int64_t count;          /* records, potential-only ones (ke < 0) included */
    float   pos[6];         /* centre; pos[3..5] receives corevel
                               (src/properties.c:213) */
    float   corevel[3];     /* in: prior value; out: core velocity */
    float   bulkvel[3];     /* in: prior value; out: bulk velocity */
    float   min_vel_err, min_bulkvel_err;
    int64_t n_core;
    int32_t is_sub;         /* 1: KE about corevel (sub-halo), 0: about
                               bulkvel (src/properties.c:547-550) */
    int32_t status;         /* out: 0 */
    int64_t kept;           /* out: records left after dropping potential-
                               only ones (src/properties.c:539-544) */
    int64_t n_bound;        /* out: kept records with !(pe < ke)
                               (src/properties.c:403) */
} hp_halo;

enum {
    HP_ACC_COMPENSATED = 0, /* FP32 pair terms, two-sum compensated sums */
    HP_ACC_NAIVE = 1,       /* FP32 pair terms, plain FP32 sums */
    HP_ACC_F64 = 2          /* FP64 pair terms and sums (same tree) */
};

enum {
    HP_METHOD_TREE = 0,     /* the reference's Barnes-Hut walk; halos of at
                               most 10 finite points, where the tree is a
                               single leaf, go to the direct kernel */
    HP_METHOD_DIRECT = 1,   /* exact pairwise sum for every halo (not the
                               reference's answer above 10 points) */
    HP_METHOD_AUTO = 2      /* tree, or direct where the autotuner found it
                               faster AND the halo has at most 10 points */
};

enum {
    HP_WALK_PAIRS = 0,      /* both children of an opened node decided in one
                               step (a stack of pending right children) */
    HP_WALK_NODES = 1       /* one node per step on skip pointers (no stack) */
};

typedef struct hp_options {
    int32_t accumulate;     /* HP_ACC_* */
    int32_t method;         /* HP_METHOD_* */
    int32_t use_graph;      /* replay the launch sequence from a hipGraph */
    int32_t timing;         /* record HIP events around each phase */
    int32_t count_work;     /* count node visits / interactions (slower) */
    int32_t per_call;       /* 1: run the calls one after another, each with
                               its own launches (the plain per-halo path) */
    int32_t walk;           /* HP_WALK_PAIRS (default) or HP_WALK_NODES */
    int32_t reserved;
} hp_options;

typedef struct hp_plan hp_plan;

/* Timings of the last hp_plan_run / hp_plan_potential (milliseconds, HIP
 * events on the plan's stream) and work counters. */
typedef struct hp_stats {
    float   ms_total, ms_build, ms_multipole, ms_walk, ms_direct, ms_unbind;
    int64_t nodes, leaves, levels, max_depth;
    int64_t walk_launches, walk_groups;
    int64_t visits, p2p, p2m;  /* when count_work is set */
    int64_t records, halos;
} hp_stats;

/* Plans a batch from host copies of its descriptors: sizes the device
 * scratch (kept until hp_plan_destroy), builds the per-halo work tables.
 * halos[i].first / count must tile [0, total) without overlap. */
int hp_plan_create(hp_plan **plan, const hp_halo *halos, int64_t nhalos,
                   const hp_options *opt);

/* The bound pass of src/properties.c:537-550 for every call of the batch,
 * on device-resident data (d_po: the record array, d_halos: the
 * descriptors, d_bound: one byte per record or NULL).  The pass runs on the
 * plan's own stream, a blocking one (it follows the caller's null-stream
 * work).  stream: a hipStream_t of the caller, or NULL; the pass starts after
 * the work already queued on it, and work queued on it later starts after
 * the pass.  The call returns when the pass is done.  Afterwards, per call, records
 * [first, first + kept) are the non-potential-only records sorted by r2
 * in the reference's order (records of equal r2 in the order its
 * compute_potential and potential-only strip, src/properties.c:539-544,
 * leave them before its stable qsort) with pe and ke set, followed by the
 * potential-only records with pe set;
 * d_bound[first + i] = !(pe < ke) for i < kept.  Returns 0 or an HP_E*
 * code. */
int hp_plan_run(hp_plan *plan, struct potential *d_po, hp_halo *d_halos,
                uint8_t *d_bound, void *stream);

/* compute_potential over every call of the batch: pe, and each call's
 * records permuted as compute_potential permutes them (HP_METHOD_DIRECT
 * plans leave the records in place); stream as for hp_plan_run. */
int hp_plan_potential(hp_plan *plan, struct potential *d_po, void *stream);

int hp_plan_stats(const hp_plan *plan, hp_stats *out);
int hp_plan_destroy(hp_plan *plan);

/* Launch settings an autotuned plan uses.  None of them changes a tree, a
 * walk decision or a sum: they only move work between kernels. */
/* The launch setting is chosen per call-size bucket: calls of at most
 * 1,024, 8,192, 65,536 and more records. */
#define HP_TUNE_BUCKETS 4

typedef struct hp_tuning {
    int32_t big_threshold[HP_TUNE_BUCKETS];  /* per bucket: nodes above this many
                                                points are split by the level
                                                passes (128, 256, 512 or 1024) */
    int32_t wave_max[HP_TUNE_BUCKETS];       /* per bucket: subtrees of at most
                                                this many points are built by
                                                one wavefront (32 or 64) */
    int32_t direct_max;     /* HP_METHOD_AUTO: calls of at most this many
                               points use the direct kernel (0 or 10) */
    float   ms_potential;   /* the chosen setting's time for the potential */
} hp_tuning;

/* Autotuner: runs the plan's potential pass (compute_potential of every
 * call, pe overwritten) reps times per candidate setting on the caller's
 * device records and keeps the fastest for later runs of the plan: first one
 * setting for all buckets, then each bucket that holds calls in turn with
 * the others fixed.  No setting changes a tree, a decision or a sum order
 * of the reference's.  The chosen setting is returned in *out when out is
 * not NULL.  Per-call plans are not tuned (HP_EINVAL). */
int hp_plan_tune(hp_plan *plan, struct potential *d_po, int reps, hp_tuning *out);

/* Applies a setting hp_plan_tune returned (for this plan or another plan of
 * the same kind of batch) without timing anything: each bucket's
 * big_threshold one of 128, 256, 512, 1024 and wave_max 32 or 64;
 * direct_max as hp_plan_tune allows it.  HP_EINVAL for any other value and
 * for per-call plans. */
int hp_plan_set_tuning(hp_plan *plan, const hp_tuning *tuning);

/* ---------------------------------------------------- record assembly */

/* The reference's particle copies of a FOF group (src/particle.h:5-8),
 * 32 bytes: id at 0, pos[6] at 8. */
#ifndef PARTICLE_H
struct particle {
    int64_t id;
    float   pos[6];
};
#endif

/* One halo of the hierarchy: the fields of struct halo (src/halo.h:8-15)
 * and struct extra_halo_info (src/halo.h:37-41) that calc_particle_radii
 * reads.  Indices refer to the same array; -1: none. */
typedef struct hp_tree_halo {
    int64_t p_start;             /* first particle in the copies array */
    int64_t num_p;
    int64_t num_child_particles;
    int64_t child;               /* first child */
    int64_t next_cochild;        /* next sibling */
    int64_t sub_of;              /* parent, -1 for a host */
    float   pos[6];              /* centre (pos[0..2] is used) */
} hp_tree_halo;

/* calc_particle_radii (src/groupies.c:212-268; replaces the record set-up of
 * src/properties.c:526 and _reset_potentials, src/groupies.c:212-233) for
 * the halos base[0..nbase): each call's records are the halo's particles,
 * then its descendants' (depth first, children in sibling order), then --
 * when the halo is a sub-halo with num_child_particles above
 * include_host_ratio (INCLUDE_HOST_POTENTIAL_RATIO, src/config.template.h:
 * 116, default 0.3) times its parent's -- the parent's subtree as potential-
 * only records (ke = -1).  r2 is about the halo's pos[0..2]; flags = 1 for
 * the halo's own particles, 2 for those of descendants with fewer than 3% of
 * its particles.  A halo already on the current descent path is skipped, as
 * the reference does.
 *
 * d_particles: the group's copies (device); the calls' records are written
 * back to back into d_po (device, cap_records records).  calls[0..nbase) and
 * *nrecords are written on the host: first, count, pos, is_sub (sub_of >
 * -1); corevel and bulkvel are set to 0 and n_core, kept, n_bound to 0 -- the
 * caller sets the prior corevel / bulkvel (src/properties.c:71-126) before
 * hp_plan_run if it needs them.  Synchronous. */
int hp_gather_calls(const hp_tree_halo *tree, int64_t nhalos, const int64_t *base, int64_t nbase,
                    const struct particle *d_particles, int64_t nparticles,
                    double include_host_ratio, struct potential *d_po, int64_t cap_records,
                    hp_halo *calls, int64_t *nrecords);

/* ------------------------------------------------------ phase-space sub-FOF */

/* One FOF group of a batch: a struct fof (src/fof.h:7-10) given by index.
   first/num_p/top in; target_r, first_sub, num_sub out. */
typedef struct hp_fof_group {
    int64_t first;      /* first particle in d_particles */
    int64_t num_p;
    int32_t top;        /* the whole FOF (norm_sd's num_p == num_copies,
                           src/groupies.c:821): sigma_x *= INITIAL_METRIC_SCALING */
    float target_r;     /* the linking length found (src/groupies.c:312) */
    int64_t first_sub;  /* its sub-FOFs: subs[first_sub .. first_sub + num_sub) */
    int64_t num_sub;
} hp_fof_group;

/* A sub-FOF: num_p particles from d_particles[first] (struct fof). */
typedef struct hp_subfof {
    int64_t first;
    int64_t num_p;
} hp_subfof;

typedef struct hp_subfof_config {   /* src/config.template.h:96,112,114,118 */
    double fof_fraction;            /* FOF_FRACTION (0.7) */
    double initial_metric_scaling;  /* INITIAL_METRIC_SCALING (1) */
    int64_t min_halo_particles;     /* MIN_HALO_PARTICLES (10) */
    int32_t exact_ll_calc;          /* EXACT_LL_CALC (0) */
    int32_t ordered_search;         /* 1: every next-closest search walks the tree in the
                                       reference's order (a check of the default path,
                                       which gives the same distances) */
} hp_subfof_config;

/* Replaces _find_subfofs_better2(f, FOF_FRACTION, fofinfo)
   (src/groupies.c:287-314, called by _find_subs at src/groupies.c:529) for
   every group of a batch at once: norm_sd, the in-place 6D fast3tree
   rebuild, the sample's next-closest distances, find_median_r and the
   linking at that length.  On return each group's particles in d_particles
   are normalised (as norm_sd leaves them) and in the reference's order:
   sorted by its union-find roots with partition_sort_particles
   (src/fof.c:184-252), the particles with no neighbour first; the sub-FOFs
   of MIN_HALO_PARTICLES or more are listed in subs in the reference's root
   order (build_fullfofs).  Groups must not overlap.
   NULL config: the reference's defaults.  Returns HP_OK or an error code
   (non-finite coordinates, overlapping groups, too small a subs array). */
int hp_find_subfofs(struct particle *d_particles, int64_t nparticles, hp_fof_group *groups,
                    int64_t ngroups, const hp_subfof_config *config, hp_subfof *subs,
                    int64_t cap_subs, int64_t *nsubs);

/* ------------------------------------------------------- halo properties */

/* Density thresholds of the mass definitions, in particles per (Mpc/h)^3:
 * particle_thresh_dens[0..4], particle_rvir_dens and particle_rvir_dens_z0
 * as calc_mass_definition leaves them (src/groupies.c:102-142; float-rounded
 * values). */
typedef struct hp_mass_defs {
    double thresh_dens[5];  /* MASS_DEFINITION, MASS_DEFINITION2..5 */
    double rvir_dens;       /* "vir" */
    double rvir_dens_z0;    /* "vir" at a = 1 (pseudo-evolution masses) */
} hp_mass_defs;

/* calc_mass_definition (src/groupies.c:124-142) for the five definition
 * strings ("vir", "<x>b", "<x>c", optionally prefixed "m"; NULL entries take
 * the halo finder's defaults vir, 200b, 200c, 500c, 2500c,
 * src/config.template.h:9-13), with the current SCALE_NOW, PARTICLE_MASS,
 * Om, Ol, W0 and WA.  Host code only. */
int hp_mass_definitions(const char *const defs[5], hp_mass_defs *out);

/* One call's results of the halo property pass: src/properties.c:384-510
 * run for bound = 0 and then bound = 1 (src/properties.c:552-553), with the
 * defaults BOUND_PROPS = 1, LIGHTCONE = 0, SHAPE_ITERATIONS = 10 and
 * WEIGHTED_SHAPES = 1, calc_shape (src/properties.c:216-326),
 * calc_scale_radius (src/nfw.c:130-173) and _calc_pseudo_evolution_masses
 * (src/properties.c:351-382) included.  The fields are the values the
 * reference stores in struct halo (src/halo.h:8-35, built with
 * OUTPUT_NFW_CHI2 and OUTPUT_INERTIA_TENSOR).  Radii in kpc/h.  224 bytes. */
typedef struct hp_halo_props {
    float   m;               /* bound = 0 pass: part_mdelta * PARTICLE_MASS */
    float   mgrav;           /* bound = 1 pass */
    float   r;               /* final h->r, from part_mdelta */
    float   vmax, rvmax, vrms, Xoff, Voff;
    float   alt_m[4];
    float   halfmass_radius;
    float   J[3];
    float   energy, kin_to_pot, spin, bullock_spin;
    float   rvir;            /* the pass's virial radius (used by bullock_spin) */
    float   rs, klypin_rs, chi2;
    float   b_to_a, c_to_a, A[3];            /* within the main definition */
    float   b_to_a2, c_to_a2, A2[3];         /* within MASS_DEFINITION4 */
    float   inertia_tensor[6], inertia_tensor2[6];  /* xx, yy, zz, xy, yz, zx */
    float   m_pe_b, m_pe_d;
    int64_t n_mdelta;        /* part_mdelta: bound records within the main definition */
    int64_t dens_tot;        /* records [0, dens_tot) enter the averages */
    int64_t n_vir;           /* bound records within the virial definition */
    int64_t n_half;          /* index of the record behind halfmass_radius */
} hp_halo_props;

/* The property pass of every call of a plan, on the device results of the
 * last hp_plan_run of that plan: d_po and d_halos as hp_plan_run left them
 * (records sorted by r2 with pe and ke set; first, kept and the centre,
 * pos[3..5] = corevel).  d_props receives one hp_halo_props per call
 * (device memory).  Synchronous, with stream as for hp_plan_run. */
int hp_plan_props(hp_plan *plan, const struct potential *d_po, const hp_halo *d_halos,
                  const hp_mass_defs *defs, hp_halo_props *d_props, void *stream);

/* Host-buffer convenience: copies in, hp_plan_run, copies out. */
int hp_bound_pass_host(struct potential *po, int64_t nrecords, hp_halo *halos,
                       int64_t nhalos, uint8_t *bound, const hp_options *opt);

/* ------------------------------------------------------------- diagnostics */

/* Tree export for structural checks: for a single call of n records (host
 * buffers), fills perm[n] (record index of each position in tree order;
 * non-finite records last) and up to cap nodes in the layout of hp_node_out
 * below.  Returns the node count, or a negative HP_E* code. */
typedef struct hp_node_out {
    float   lo[3], hi[3];
    float   com[3], mass;
    float   dmin2, dmin;
    int32_t first, count;   /* range in tree order, relative to the call */
    int32_t left, skip;     /* -1: leaf / end of walk */
    int32_t parent, pad;
} hp_node_out;

int64_t hp_debug_tree(const struct potential *po, int64_t n, int32_t *perm,
                      hp_node_out *nodes, int64_t cap);

enum {
    HP_OK = 0,
    HP_EINVAL = -1,         /* bad argument */
    HP_ENODEV = -2,         /* no usable gfx950 device */
    HP_EHIP = -3,           /* a HIP call failed */
    HP_ENOMEM = -4          /* device allocation failed */
};

const char *hp_last_error(void);

/* Called on a fatal error inside the void reference entry points.  When the
 * program links src/error.cpp, rockstar_abort_impl (src/error.h:15) is
 * called instead, so the reference's exception path is kept.  With neither,
 * the library prints the error and aborts. */
typedef void (*hp_error_handler)(int code, const char *file, int line);
void hp_set_error_handler(hp_error_handler handler);

#ifdef __cplusplus
}
#endif

#endif /* HALOPOT_H */
