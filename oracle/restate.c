/*
 * oracle/restate.c -- CPU restatement of MPI-Rockstar's per-halo Barnes-Hut
 * potential and bound-particle pass, written from the reference's semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry and bench.py's cpu_baseline leg load it, and only to check or
 * to time the CPU path beside the GPU one.  The product (libhalopot) never
 * links, loads or calls anything here.
 *
 * What it restates (reference file:line):
 *   - struct potential record ............ src/potential.h:7-21
 *   - k-d tree build (POTENTIAL instance,
 *     10 points per leaf) ................. src/fast3tree.c:736-774 (split dim,
 *                                           midpoint, partition), 778-794 (tight
 *                                           bbox), 798-846 (recursion, failed
 *                                           splits), 865-908 (non-finite removal)
 *   - multipoles ......................... src/potential.c:71-82 (dmin),
 *                                           98-124 (mass centres)
 *   - group walk ......................... src/potential.c:126-168
 *   - pair terms ......................... src/potential.c:22-55
 *   - compute_potential .................. src/potential.c:183-192
 *   - compute_kinetic_energy ............. src/potential.c:194-211
 *   - hubble_scaling ..................... src/hubble.c:8-19
 *   - calculate_corevel .................. src/properties.c:165-214
 *   - bound-pass driver .................. src/properties.c:537-550 and the
 *                                           bound test at src/properties.c:403
 *   - vir_density / particle_rvir_dens ... src/groupies.c:97-142
 *   - record assembly .................... src/groupies.c:212-268
 *                                           (_reset_potentials,
 *                                           calc_particle_radii)
 *   - halo property pass ................. src/properties.c:127-137 (angular
 *                                           momentum), 216-326 (calc_shape),
 *                                           328-349 (energy), 351-382
 *                                           (pseudo-evolution masses),
 *                                           384-510; src/nfw.c:13-173
 *                                           (scale radius); src/jacobi.c:8-135
 *
 * The record order follows the reference's in-place loops: the non-finite
 * removal and the node partition swap with the current last slot
 * (src/fast3tree.c:762-773, 875-887), the potential-only strip overwrites
 * with the last record (src/properties.c:539-544), and the r2 sort is a
 * stable merge sort (glibc's qsort).  So compute_potential permutes po, and
 * the bound pass leaves it, exactly as the reference does.
 * Deliberate difference, which changes no node, decision, value or order:
 * nodes are numbered breadth-first, not depth-first.  The walk still visits
 * left before right, so every sink leaf sees the same sequence of
 * decisions.
 * Pinned by: the compiled reference (oracle/_ref) on the fixtures, the FP64
 * direct sum (N <= 10 is exact; the walk's mass bookkeeping is checked by an
 * interaction-count invariant) and closed forms for the kinetic energy.
 *
 * Built with -std=c99 -O3 -fno-math-errno (src/Makefile:8,23): no FP
 * contraction, so every float/double expression rounds as the reference's.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct potential {
    float   pos[6], r2;
    double  pe;
    float   ke;
    int32_t flags;
};

typedef struct orc_params {
    double force_res;          /* FORCE_RES, Mpc/h comoving */
    double particle_mass;      /* PARTICLE_MASS, Msun/h */
    double Om, Ol, W0, WA;     /* cosmology used by hubble_scaling */
    double scale_now;          /* SCALE_NOW */
    double particle_rvir_dens; /* groupies.c:135; particles per (Mpc/h)^3 */
} orc_params;

typedef struct orc_node {
    float   lo[3], hi[3];
    int64_t first, n;
    int64_t left; /* right child is left + 1; -1 for a leaf */
    int32_t dim;  /* split dimension; -1 for a leaf */
    int32_t depth;
    float   com[3], m, dmin;
} orc_node;

/* same layout as the product's per-call descriptor (104 bytes) */
typedef struct orc_halo {
    int64_t first, count;
    float   pos[6];
    float   corevel[3], bulkvel[3];
    float   min_vel_err, min_bulkvel_err;
    int64_t n_core;
    int32_t is_sub, status;
    int64_t kept, n_bound;
} orc_halo;

typedef struct orc_tree {
    orc_node *nodes;
    int64_t   nnodes, cap;
    int64_t   nfinite;
    int64_t   max_depth;
} orc_tree;

#define ORC_LEAF_POINTS 10 /* src/potential.c:57 */
#define ORC_DEGENERATE  20 /* src/potential.c:145: 2 * POINTS_PER_LEAF */
#define ORC_ERR_TOL     1.0 /* src/potential.c:10 */
#define ORC_GC          (4.30117902e-9) /* src/universal_constants.h:4 */

/* ---------------------------------------------------------------- helpers */

static void *orc_grow(void *p, size_t bytes) {
    void *q = realloc(p, bytes);
    if (!q && bytes) abort();
    return q;
}

/* squared distance with float differences and double squares
   (src/potential.c:22-29) */
static double orc_d2(const float *a, const float *b) {
    double s = 0.0;
    for (int k = 0; k < 3; k++) {
        double d = (float)(a[k] - b[k]);
        s += d * d;
    }
    return s;
}

static int orc_finite3(const float *p) {
    return isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]);
}

static void orc_bbox(const struct potential *po, int64_t first, int64_t n,
                     float lo[3], float hi[3]) {
    for (int k = 0; k < 3; k++) lo[k] = hi[k] = po[first].pos[k];
    for (int64_t i = first + 1; i < first + n; i++)
        for (int k = 0; k < 3; k++) {
            float x = po[i].pos[k];
            if (x < lo[k]) lo[k] = x;
            if (x > hi[k]) hi[k] = x;
        }
}

/* ------------------------------------------------------------------ build */

static int64_t orc_new_node(orc_tree *t) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 64;
        t->nodes = orc_grow(t->nodes, sizeof(orc_node) * (size_t)t->cap);
    }
    memset(t->nodes + t->nnodes, 0, sizeof(orc_node));
    t->nodes[t->nnodes].left = -1;
    t->nodes[t->nnodes].dim = -1;
    return t->nnodes++;
}

/* Split dimension: z first, x or y only if strictly wider
   (src/fast3tree.c:736-747). */
static int orc_split_dim(const float lo[3], const float hi[3]) {
    int dim = 2;
    float w = hi[2] - lo[2];
    for (int k = 0; k < 2; k++) {
        float e = hi[k] - lo[k];
        if (e > w) {
            w = e;
            dim = k;
        }
    }
    return dim;
}

static void orc_swap(struct potential *a, struct potential *b) {
    struct potential t = *a;
    *a = *b;
    *b = t;
}

/* Builds the tree over po[0..n); po is permuted so that every node owns a
   contiguous range, in the order the reference's in-place loops leave it.
   Returns the number of finite points. */
static int64_t orc_build(orc_tree *t, struct potential *po, int64_t n,
                         struct potential *scratch) {
    (void)scratch;
    /* non-finite points are swapped with the current last one and the slot
       is examined again (src/fast3tree.c:875-887 drops them from the tree) */
    int64_t nf = n;
    for (int64_t i = 0; i < nf; i++) {
        if (!orc_finite3(po[i].pos)) {
            nf--;
            orc_swap(po + i, po + nf);
            i--;
        }
    }

    t->nnodes = 0;
    t->nfinite = nf;
    t->max_depth = 0;
    int64_t r = orc_new_node(t);
    orc_node *root = t->nodes + r;
    root->first = 0;
    root->n = nf;
    if (nf) orc_bbox(po, 0, nf, root->lo, root->hi);

    /* breadth-first: nodes are appended in creation order, so a cursor over
       the node array visits every node exactly once */
    for (int64_t c = 0; c < t->nnodes; c++) {
        orc_node *nd = t->nodes + c;
        if (nd->depth > t->max_depth) t->max_depth = nd->depth;
        if (nd->n <= ORC_LEAF_POINTS) continue;
        int dim = orc_split_dim(nd->lo, nd->hi);
        if (nd->hi[dim] == nd->lo[dim]) continue; /* src/fast3tree.c:760-761 */
        /* float sum, halved exactly (src/fast3tree.c:757) */
        float lim = 0.5f * (nd->hi[dim] + nd->lo[dim]);
        /* in-place partition (src/fast3tree.c:762-773): a point above the
           plane is swapped with the current last slot, which is examined
           next; the loop ends where the two ends meet */
        struct potential *q = po + nd->first;
        int64_t i = 0, j = nd->n - 1;
        for (; i < j; i++) {
            if (q[i].pos[dim] > lim) {
                orc_swap(q + i, q + j);
                i--;
                j--;
            }
        }
        if (i == j && q[i].pos[dim] <= lim) i++;
        const int64_t nl = i;
        if (nl == 0 || nl == nd->n) continue; /* src/fast3tree.c:805-809 */
        int64_t a = orc_new_node(t);
        int64_t b = orc_new_node(t);
        nd = t->nodes + c; /* array may have moved */
        nd->dim = dim;
        nd->left = a;
        orc_node *L = t->nodes + a, *R = t->nodes + b;
        L->first = nd->first;
        L->n = nl;
        R->first = nd->first + nl;
        R->n = nd->n - nl;
        L->depth = R->depth = nd->depth + 1;
        orc_bbox(po, L->first, L->n, L->lo, L->hi);
        orc_bbox(po, R->first, R->n, R->lo, R->hi);
    }
    return nf;
}

/* ------------------------------------------------------------- multipoles */

/* dmin over every point below the node (src/potential.c:71-82) */
static void orc_dmin(orc_node *nd, const struct potential *po) {
    double sum = 0.0, big = 0.0;
    for (int64_t i = nd->first; i < nd->first + nd->n; i++) {
        double d2 = orc_d2(nd->com, po[i].pos);
        if (d2 > big) big = d2;
        sum += d2;
    }
    double b = sqrt(big) / 2.0;
    nd->dmin = (float)(b + sqrt(b * b + sum / ((double)nd->n * ORC_ERR_TOL)));
}

/* Leaves: f64 position sums; internal nodes: float combination of the two
   children, right term first (src/potential.c:98-124).  Children always have
   larger indices than their parent, so a reverse sweep is a post-order. */
static void orc_multipoles(orc_tree *t, const struct potential *po,
                           double particle_mass) {
    for (int64_t c = t->nnodes - 1; c >= 0; c--) {
        orc_node *nd = t->nodes + c;
        if (nd->n == 0) { /* empty root: nothing to compute */
            nd->m = 0;
            nd->com[0] = nd->com[1] = nd->com[2] = 0;
            nd->dmin = 0;
            continue;
        }
        if (nd->left < 0) {
            double s[3] = {0, 0, 0};
            nd->m = (float)(particle_mass * (double)nd->n);
            for (int64_t i = nd->first; i < nd->first + nd->n; i++)
                for (int k = 0; k < 3; k++) s[k] += po[i].pos[k];
            for (int k = 0; k < 3; k++)
                nd->com[k] = s[k] ? (float)(s[k] / (double)nd->n) : 0.0f;
        } else {
            const orc_node *L = t->nodes + nd->left, *R = L + 1;
            nd->m = L->m + R->m;
            for (int k = 0; k < 3; k++) {
                float rr = R->com[k] * R->m;
                float ll = L->com[k] * L->m;
                nd->com[k] = nd->m ? (rr + ll) / nd->m : 0.0f;
            }
        }
        orc_dmin(nd, po);
    }
}

/* ------------------------------------------------------------------- walk */

typedef struct orc_ctx {
    const orc_tree *t;
    struct potential *po;
    double force_res, particle_mass;
    int32_t *np2p, *np2m; /* optional per-particle interaction counters */
    int64_t p2p, p2m, visits;
    int64_t *nmass;       /* optional: points behind the accepted monopoles */
} orc_ctx;

/* Test hook: 1 makes every monopole unacceptable, so the walk reduces to the
   direct sum over leaf pairs. */
static int orc_never_accept = 0;
void orc_set_never_accept(int on) { orc_never_accept = on; }

static double orc_pair(const orc_ctx *x, const float *a, const float *b) {
    double r = sqrt(orc_d2(a, b));
    if (r < x->force_res) r = x->force_res; /* src/potential.c:31-36 */
    return x->particle_mass * (1.0 / r);
}

static void orc_self(orc_ctx *x, const orc_node *s) {
    struct potential *p = x->po + s->first;
    for (int64_t i = 0; i < s->n; i++)
        for (int64_t j = i + 1; j < s->n; j++) {
            double v = orc_pair(x, p[i].pos, p[j].pos);
            p[i].pe += v;
            p[j].pe += v;
        }
    x->p2p += s->n * (s->n - 1);
    if (x->np2p)
        for (int64_t i = 0; i < s->n; i++)
            x->np2p[s->first + i] += (int32_t)(s->n - 1);
}

static void orc_leaf_leaf(orc_ctx *x, const orc_node *s, const orc_node *b) {
    struct potential *p = x->po + s->first;
    const struct potential *q = x->po + b->first;
    for (int64_t i = 0; i < s->n; i++)
        for (int64_t j = 0; j < b->n; j++) p[i].pe += orc_pair(x, p[i].pos, q[j].pos);
    x->p2p += s->n * b->n;
    if (x->np2p)
        for (int64_t i = 0; i < s->n; i++) x->np2p[s->first + i] += (int32_t)b->n;
}

static void orc_monopole(orc_ctx *x, const orc_node *s, const orc_node *b) {
    struct potential *p = x->po + s->first;
    for (int64_t i = 0; i < s->n; i++) {
        float r = (float)sqrt(orc_d2(p[i].pos, b->com));
        p[i].pe += b->m / r; /* float quotient, unsoftened (potential.c:163-166) */
    }
    x->p2m += s->n;
    if (x->np2m)
        for (int64_t i = 0; i < s->n; i++) x->np2m[s->first + i] += 1;
    if (x->nmass)
        for (int64_t i = 0; i < s->n; i++) x->nmass[s->first + i] += b->n;
}

/* src/potential.c:126-131: float distance of the two centres against the
   float square of the source's dmin */
static int orc_accept(const orc_node *s, const orc_node *b) {
    if (orc_never_accept) return 0;
    float r2 = (float)orc_d2(s->com, b->com);
    return r2 > b->dmin * b->dmin;
}

static int orc_inside(const orc_node *s, const orc_node *b) {
    for (int k = 0; k < 3; k++)
        if (s->lo[k] < b->lo[k] || s->hi[k] > b->hi[k]) return 0;
    return 1;
}

static void orc_visit(orc_ctx *x, const orc_node *s, const orc_node *b) {
    x->visits++;
    if (s == b) {
        if (s->n <= ORC_DEGENERATE) orc_self(x, s);
        return;
    }
    if (orc_inside(s, b) || !orc_accept(s, b)) {
        if (b->left < 0) orc_leaf_leaf(x, s, b);
        else {
            orc_visit(x, s, x->t->nodes + b->left);
            orc_visit(x, s, x->t->nodes + b->left + 1);
        }
    } else
        orc_monopole(x, s, b);
}

/* --------------------------------------------------------- exported: tree */

/* optional per-record output of the next orc_compute_potential call: the
   number of points behind the monopoles each record received */
static int64_t *orc_nmass_out = NULL;
void orc_set_nmass_out(int64_t *nmass) { orc_nmass_out = nmass; }

typedef struct orc_walk_stats {
    int64_t p2p, p2m, visits, nodes, leaves, max_depth, nfinite;
} orc_walk_stats;

/* compute_potential (src/potential.c:183-192).  po is permuted in place into
   tree order.  np2p / np2m, when given, receive per-record interaction counts
   in the permuted order.  nodes_out, when given with room for cap_nodes,
   receives the tree.  Returns the node count. */
int64_t orc_compute_potential(struct potential *po, int64_t n,
                              const orc_params *prm, int32_t *np2p,
                              int32_t *np2m, orc_walk_stats *st,
                              orc_node *nodes_out, int64_t cap_nodes) {
    for (int64_t i = 0; i < n; i++) po[i].pe = 0;
    if (np2p) memset(np2p, 0, sizeof(int32_t) * (size_t)(n > 0 ? n : 0));
    if (np2m) memset(np2m, 0, sizeof(int32_t) * (size_t)(n > 0 ? n : 0));
    if (orc_nmass_out) memset(orc_nmass_out, 0, sizeof(int64_t) * (size_t)(n > 0 ? n : 0));
    if (n <= 0) return 0;
    orc_tree t = {0};
    struct potential *scratch = orc_grow(NULL, sizeof(struct potential) * (size_t)n);
    orc_build(&t, po, n, scratch);
    free(scratch);
    orc_multipoles(&t, po, prm->particle_mass);
    orc_ctx x = {&t, po, prm->force_res, prm->particle_mass, np2p, np2m, 0, 0, 0, orc_nmass_out};
    int64_t leaves = 0;
    if (t.nfinite > 0)
        for (int64_t c = 0; c < t.nnodes; c++)
            if (t.nodes[c].left < 0) {
                leaves++;
                orc_visit(&x, t.nodes + c, t.nodes);
            }
    if (st) {
        st->p2p = x.p2p;
        st->p2m = x.p2m;
        st->visits = x.visits;
        st->nodes = t.nnodes;
        st->leaves = leaves;
        st->max_depth = t.max_depth;
        st->nfinite = t.nfinite;
    }
    if (nodes_out)
        memcpy(nodes_out, t.nodes,
               sizeof(orc_node) * (size_t)(t.nnodes < cap_nodes ? t.nnodes : cap_nodes));
    int64_t nn = t.nnodes;
    free(t.nodes);
    return nn;
}

/* Tree only (build + multipoles), for structural comparison with the GPU
   build.  po is permuted into tree order. */
int64_t orc_build_tree(struct potential *po, int64_t n, const orc_params *prm,
                       orc_node *nodes_out, int64_t cap_nodes) {
    if (n <= 0) return 0;
    orc_tree t = {0};
    struct potential *scratch = orc_grow(NULL, sizeof(struct potential) * (size_t)n);
    orc_build(&t, po, n, scratch);
    free(scratch);
    orc_multipoles(&t, po, prm->particle_mass);
    if (nodes_out)
        memcpy(nodes_out, t.nodes,
               sizeof(orc_node) * (size_t)(t.nnodes < cap_nodes ? t.nnodes : cap_nodes));
    int64_t nn = t.nnodes;
    free(t.nodes);
    return nn;
}

/* The same tree and the same walk decisions, every term evaluated in f64:
   P2P as the reference, P2M as m / r in double.  This is the "FP64
   evaluation of the same Barnes-Hut tree" used to separate truncation error
   from arithmetic error.  po is permuted as by orc_compute_potential. */
static void orc_visit64(orc_ctx *x, const orc_node *s, const orc_node *b) {
    if (s == b) {
        if (s->n <= ORC_DEGENERATE) orc_self(x, s);
        return;
    }
    if (orc_inside(s, b) || !orc_accept(s, b)) {
        if (b->left < 0) orc_leaf_leaf(x, s, b);
        else {
            orc_visit64(x, s, x->t->nodes + b->left);
            orc_visit64(x, s, x->t->nodes + b->left + 1);
        }
    } else {
        struct potential *p = x->po + s->first;
        for (int64_t i = 0; i < s->n; i++) {
            double dd = 0.0;
            for (int k = 0; k < 3; k++) {
                double d = (double)p[i].pos[k] - (double)b->com[k];
                dd += d * d;
            }
            p[i].pe += (double)b->m / sqrt(dd);
        }
    }
}

void orc_compute_potential_bh64(struct potential *po, int64_t n,
                                const orc_params *prm) {
    for (int64_t i = 0; i < n; i++) po[i].pe = 0;
    if (n <= 0) return;
    orc_tree t = {0};
    struct potential *scratch = orc_grow(NULL, sizeof(struct potential) * (size_t)n);
    orc_build(&t, po, n, scratch);
    free(scratch);
    orc_multipoles(&t, po, prm->particle_mass);
    orc_ctx x = {&t, po, prm->force_res, prm->particle_mass, NULL, NULL, 0, 0, 0, NULL};
    if (t.nfinite > 0)
        for (int64_t c = 0; c < t.nnodes; c++)
            if (t.nodes[c].left < 0) orc_visit64(&x, t.nodes + c, t.nodes);
    free(t.nodes);
}

/* -------------------------------------------------- exported: FP64 direct */

/* pe_i = sum_{j != i, both finite} m / max(|x_i - x_j|, FORCE_RES), every
   operation in f64 (the differences of float inputs are exact in f64).
   Positions are read from stride-spaced float triples. */
void orc_direct64(const float *pos, int64_t stride, int64_t n,
                  double force_res, double particle_mass, double *pe) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        const float *a = pos + i * stride;
        double acc = 0.0;
        if (orc_finite3(a)) {
            double ax = a[0], ay = a[1], az = a[2];
            for (int64_t j = 0; j < n; j++) {
                if (j == i) continue;
                const float *b = pos + j * stride;
                if (!orc_finite3(b)) continue;
                double dx = ax - b[0], dy = ay - b[1], dz = az - b[2];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r < force_res) r = force_res;
                acc += 1.0 / r;
            }
        }
        pe[i] = acc * particle_mass;
    }
}

/* ------------------------------------------------------- kinetic energy */

/* src/hubble.c:8-19 */
static double orc_weff(const orc_params *p, double a) {
    if (a != 1.0) return p->W0 + p->WA - p->WA * (a - 1.0) / log(a);
    return p->W0;
}

double orc_hubble_scaling(const orc_params *p, double z) {
    double z1 = 1.0 + z;
    double a = 1.0 / z1;
    return sqrt(p->Om * (z1 * z1 * z1) + p->Ol * pow(a, -3.0 * (1.0 + orc_weff(p, a))));
}

/* src/potential.c:194-211: records with ke < 0 are potential-only and are
   left alone; the squared velocity is accumulated into the float ke. */
void orc_kinetic(struct potential *po, int64_t n, const float *vel_cen,
                 const float *pos_cen, float scale_now, const orc_params *p) {
    double conv = 0.5 * scale_now / ORC_GC;
    double z1 = 1.0 / scale_now;
    double hub = 100.0 * orc_hubble_scaling(p, z1 - 1.0);
    for (int64_t i = 0; i < n; i++) {
        if (po[i].ke < 0) continue;
        po[i].ke = 0;
        for (int k = 0; k < 3; k++) {
            float dv_pec = po[i].pos[k + 3] - vel_cen[k];
            float dx = po[i].pos[k] - pos_cen[k];
            double dv = (double)dv_pec + (hub * scale_now) * (double)dx;
            po[i].ke = (float)((double)po[i].ke + dv * dv);
        }
        po[i].ke = (float)((double)po[i].ke * conv);
    }
}

/* ---------------------------------------------------------- core velocity */

/* src/properties.c:165-214.  po is sorted by r2.  The running mean and
   variance are Welford's; the selection keeps the prefix whose variance per
   (i - 3) * i is lowest, the core restricted to the inner 10% in radius. */
void orc_corevel(orc_halo *h, const struct potential *po, int64_t n,
                 const orc_params *p) {
    const double four_thirds_pi = 4.0 * M_PI / 3.0;
    double thr = p->particle_rvir_dens * four_thirds_pi;
    int64_t j;
    for (j = n - 1; j >= 0; j--) {
        double r6 = po[j].r2 * po[j].r2 * po[j].r2; /* float product, promoted */
        if ((double)(j * j) > r6 * (thr * thr)) break;
    }
    int64_t rvir_max = j;
    if (rvir_max < 1) return;
    for (j = n - 1; j >= 0; j--)
        if (po[j].r2 * 100.0 < po[rvir_max].r2) break;
    int64_t core_max = j < 100 ? 100 : j;

    double mean[3] = {0, 0, 0}, m2[3] = {0, 0, 0}, best = 0;
    for (int64_t i = 0; i < rvir_max; i++) {
        for (int k = 0; k < 3; k++) {
            double v = po[i].pos[k + 3];
            double d = v - mean[k];
            mean[k] += d / ((double)(i + 1));
            m2[k] += d * (v - mean[k]);
        }
        double var = m2[0] + m2[1] + m2[2];
        if (i < 10 || var < best * (i - 3) * i) {
            best = (i > 3) ? var / (double)((i - 3) * i) : 0.0;
            if (i < core_max) {
                h->n_core = i;
                h->min_vel_err = (float)best;
                for (int k = 0; k < 3; k++) h->corevel[k] = (float)mean[k];
            }
            for (int k = 0; k < 3; k++) h->bulkvel[k] = (float)mean[k];
            h->min_bulkvel_err = (float)best;
        }
    }
    for (int k = 0; k < 3; k++) h->pos[k + 3] = h->corevel[k];
}

/* ------------------------------------------------------- bound-pass driver */

static void orc_merge_sort_r2(struct potential *a, struct potential *tmp,
                              int64_t n) {
    /* bottom-up stable merge sort by r2 (the reference's qsort is glibc's
       merge sort in practice, stable for equal keys) */
    for (int64_t w = 1; w < n; w *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * w) {
            int64_t mid = lo + w < n ? lo + w : n;
            int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
            int64_t i = lo, j = mid, k = lo;
            while (i < mid && j < hi) tmp[k++] = (a[j].r2 < a[i].r2) ? a[j++] : a[i++];
            while (i < mid) tmp[k++] = a[i++];
            while (j < hi) tmp[k++] = a[j++];
        }
        memcpy(a, tmp, sizeof(struct potential) * (size_t)n);
    }
}

/* Drops the potential-only records (ke < 0) the reference's way
   (src/properties.c:539-544): the slot is overwritten with the current last
   record and examined again.  Returns the records kept, in po[0..kept). */
static int64_t orc_strip(struct potential *po, int64_t n) {
    int64_t total = n;
    for (int64_t j = 0; j < total; j++)
        if (po[j].ke < 0) {
            total--;
            po[j] = po[total];
            j--;
        }
    return total;
}

/* src/properties.c:537-550: potential, drop potential-only records, sort by
   radius, core velocity, kinetic energy about the core (sub-halos) or bulk
   (hosts) velocity.  Returns the kept record count; bound[i] = !(pe < ke)
   (src/properties.c:403) for the kept, sorted records. */
int64_t orc_bound_pass(struct potential *po, int64_t n, orc_halo *h,
                       const orc_params *p, uint8_t *bound) {
    if (n > 1) orc_compute_potential(po, n, p, NULL, NULL, NULL, NULL, 0);
    int64_t kept = orc_strip(po, n);
    if (kept > 1) {
        struct potential *tmp = orc_grow(NULL, sizeof(struct potential) * (size_t)kept);
        orc_merge_sort_r2(po, tmp, kept);
        free(tmp);
    }
    orc_corevel(h, po, kept, p);
    orc_kinetic(po, kept, h->is_sub ? h->corevel : h->bulkvel, h->pos,
                (float)p->scale_now, p);
    int64_t nb = 0;
    for (int64_t i = 0; i < kept; i++) {
        int b = !(po[i].pe < po[i].ke);
        nb += b;
        if (bound) bound[i] = (uint8_t)b;
    }
    h->kept = kept;
    h->n_bound = nb;
    return kept;
}

/* The reference's concurrency context (src/mpi_main.cpp:1377): dynamic
   schedule over halos, one halo per thread.  first[h] / count[h] index the
   concatenated record array; kept[h] receives each halo's kept count. */
void orc_bound_pass_batch(struct potential *po, const int64_t *first,
                          const int64_t *count, orc_halo *halos,
                          int64_t nhalos, const orc_params *p, uint8_t *bound,
                          int64_t *kept, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t h = 0; h < nhalos; h++)
        kept[h] = orc_bound_pass(po + first[h], count[h], halos + h, p,
                                 bound ? bound + first[h] : NULL);
}

/* The same driver with caller-supplied potential and kinetic-energy
   functions: with the compiled reference's compute_potential and
   compute_kinetic_energy this is the reference's own CPU path for the bound
   pass (src/properties.c:537-550, src/mpi_main.cpp:1377), only the short
   strip / sort / calculate_corevel steps being restated. */
typedef void (*orc_cp_fn)(struct potential *, int64_t);
typedef void (*orc_ke_fn)(struct potential *, int64_t, float *, float *, float);

static int64_t orc_bound_pass_fn(struct potential *po, int64_t n, orc_halo *h,
                                 const orc_params *p, orc_cp_fn cp, orc_ke_fn ke) {
    if (n > 1) cp(po, n);
    int64_t kept = orc_strip(po, n);
    if (kept > 1) {
        struct potential *tmp = orc_grow(NULL, sizeof(struct potential) * (size_t)kept);
        orc_merge_sort_r2(po, tmp, kept);
        free(tmp);
    }
    orc_corevel(h, po, kept, p);
    ke(po, kept, h->is_sub ? h->corevel : h->bulkvel, h->pos, (float)p->scale_now);
    int64_t nb = 0;
    for (int64_t i = 0; i < kept; i++) nb += !(po[i].pe < po[i].ke);
    h->kept = kept;
    h->n_bound = nb;
    return kept;
}

void orc_bound_pass_batch_fn(struct potential *po, const int64_t *first,
                             const int64_t *count, orc_halo *halos, int64_t nhalos,
                             const orc_params *p, int64_t *kept, int nthreads,
                             orc_cp_fn cp, orc_ke_fn ke) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t h = 0; h < nhalos; h++)
        kept[h] = orc_bound_pass_fn(po + first[h], count[h], halos + h, p, cp, ke);
}

/* compute_potential over every call of a batch (src/potential.c:183-192
   per call, calls in parallel as src/mpi_main.cpp:1377 runs halos), with
   each call's P2P pair and P2M term counts of the reference's walk.
   Records are permuted within their calls as the restated build leaves
   them. */
void orc_potential_batch(struct potential *po, const int64_t *first,
                         const int64_t *count, int64_t nhalos, const orc_params *p,
                         int64_t *p2p, int64_t *p2m, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t h = 0; h < nhalos; h++) {
        orc_walk_stats st;
        memset(&st, 0, sizeof st);
        orc_compute_potential(po + first[h], count[h], p, NULL, NULL, &st, NULL, 0);
        p2p[h] = st.p2p;
        p2m[h] = st.p2m;
    }
}

/* ---------------------------------------------------- record assembly */

/* One halo of the hierarchy, in the layout of the product's hp_tree_halo. */
typedef struct orc_tree_halo {
    int64_t p_start, num_p, num_child_particles;
    int64_t child, next_cochild, sub_of;
    float   pos[6];
} orc_tree_halo;

typedef struct orc_particle { /* src/particle.h:5-8 */
    int64_t id;
    float   pos[6];
} orc_particle;

typedef struct orc_gather {
    const orc_particle  *parts;
    const orc_tree_halo *tree;
    int64_t              nh;
    double               ratio;
    struct potential    *po;
    int64_t              cap;
    int64_t             *path; /* halos on the current descent, by level */
    int64_t              npath;
    int                  err;
} orc_gather;

/* src/groupies.c:212-233 */
static void orc_reset_records(orc_gather *g, int64_t base, int64_t h, const float *cen,
                              int64_t p_start, int potential_only) {
    const orc_tree_halo *t = g->tree + h;
    if (p_start + t->num_p > g->cap) {
        g->err = 1;
        return;
    }
    memset(g->po + p_start, 0, sizeof(struct potential) * (size_t)t->num_p);
    for (int64_t j = 0; j < t->num_p; j++) {
        const orc_particle *q = g->parts + t->p_start + j;
        struct potential *r = g->po + p_start + j;
        float r2 = 0;
        for (int k = 0; k < 3; k++) {
            float dx = q->pos[k] - cen[k];
            r2 += dx * dx;
        }
        r->r2 = r2;
        memcpy(r->pos, q->pos, sizeof(float) * 6);
        if (potential_only) r->ke = -1;
        if (h == base) r->flags = 1;
        if (!potential_only && (t->num_p < g->tree[base].num_p * 0.03)) r->flags = 2;
    }
}

/* src/groupies.c:235-268 */
static int64_t orc_radii(orc_gather *g, int64_t base, int64_t h, const float *cen,
                         int64_t p_start, int64_t level, int potential_only) {
    if (g->err) return p_start;
    g->path[level] = h;
    for (int64_t j = 0; j < level; j++)
        if (g->path[j] == h) return p_start;
    orc_reset_records(g, base, h, cen, p_start, potential_only);
    int64_t total = p_start + g->tree[h].num_p;
    int64_t first_child = g->tree[h].child, child = first_child;
    while (child > -1) {
        total = orc_radii(g, base, child, cen, total, level + 1, potential_only);
        child = g->tree[child].next_cochild;
        if (child == first_child) { /* src/groupies.c:254 asserts against it */
            g->err = 2;
            return total;
        }
    }
    int64_t parent = g->tree[h].sub_of;
    if (h == base && parent > -1 &&
        g->tree[parent].num_child_particles * g->ratio < g->tree[h].num_child_particles)
        total = orc_radii(g, base, parent, cen, total, level + 1, 1);
    return total;
}

/* calc_particle_radii(h, h, h->pos, 0, 0, 0) for halo `base`: its records
   into po (at most cap).  Returns the record count, or -1 when the records
   do not fit or a sibling list returns to its first child. */
int64_t orc_particle_radii(const orc_particle *parts, const orc_tree_halo *tree, int64_t nh,
                           int64_t base, double ratio, struct potential *po, int64_t cap) {
    orc_gather g = {parts, tree, nh, ratio, po, cap, NULL, 0, 0};
    g.path = orc_grow(NULL, sizeof(int64_t) * (size_t)(nh + 2));
    int64_t n = orc_radii(&g, base, base, tree[base].pos, 0, 0, 0);
    free(g.path);
    return g.err ? -1 : n;
}

/* ------------------------------------------------------ halo properties */

/* Output of the property pass, in the layout of the product's
   hp_halo_props (include/halopot.h). */
typedef struct orc_props {
    float   m, mgrav, r, vmax, rvmax, vrms, Xoff, Voff;
    float   alt_m[4];
    float   halfmass_radius;
    float   J[3];
    float   energy, kin_to_pot, spin, bullock_spin;
    float   rvir;
    float   rs, klypin_rs, chi2;
    float   b_to_a, c_to_a, A[3], b_to_a2, c_to_a2, A2[3];
    float   inertia_tensor[6], inertia_tensor2[6];
    float   m_pe_b, m_pe_d;
    int64_t n_mdelta, dens_tot, n_vir, n_half;
} orc_props;

#define ORC_VMAX_CONST       6.55833746e-5 /* src/universal_constants.h:6 */
#define ORC_RMAX_TO_RS       2.1626        /* src/universal_constants.h:7 */
#define ORC_RS_CONSTANT      0.216216595   /* src/universal_constants.h:8 */
#define ORC_SHAPE_ITERATIONS 10            /* src/config.template.h:90 */
#define ORC_WEIGHTED_SHAPES  1             /* src/config.template.h:91 */

/* -------------------------------------------------- shape (calc_shape) */

/* The cyclic Jacobi eigenvalue iteration of src/jacobi.c:8-135: rotate away
   the largest off-diagonal element (rows cached by their largest column)
   until no eigenvalue changes by more than 1e-6 of itself; the rotation's
   t and s in long double as there.  m is left symmetric; orth's rows are
   the eigenvectors. */
static void orc_jacobi(double m[3][3], double eig[3], double orth[3][3]) {
    const int n = 3;
    int max_col[3], changed[3], state = 0;
    double c, s, t, u, a;
    for (int i = 0; i < n; i++) {
        double mx = 0;
        max_col[i] = i + 1;
        for (int j = i + 1; j < n; j++)
            if (fabs(m[i][j]) > mx) {
                max_col[i] = j;
                mx = fabs(m[i][j]);
            }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) orth[i][j] = (i == j) ? 1 : 0;
    for (int i = 0; i < n; i++) {
        eig[i] = m[i][i];
        changed[i] = 1;
        int j;
        for (j = 0; j < n; j++)
            if (j != i && m[i][j]) break;
        if (j == n) {
            state--;
            changed[i] = 0;
        }
    }
    state += n;
    while (state) {
        double mx = fabs(m[0][max_col[0]]);
        int k = 0;
        for (int i = 1; i < n - 1; i++)
            if (fabs(m[i][max_col[i]]) > mx) {
                mx = fabs(m[i][max_col[i]]);
                k = i;
            }
        int l = max_col[k];
        double v = m[k][l];
        if (v == 0) break;
        a = (eig[l] - eig[k]) * 0.5;
        t = fabsl(a) + sqrtl(v * v + a * a);
        s = sqrtl(v * v + t * t);
        c = t / s;
        s = v / s;
        t = v * v / t;
        if (a < 0) {
            s = -s;
            t = -t;
        }
        const int xs[2] = {k, l};
        const double ys[2] = {-t, t};
        for (int q = 0; q < 2; q++) {
            const int x = xs[q];
            const double y = ys[q];
            eig[x] += y;
            if (changed[x] && (fabs(y) < 1e-6 * fabs(eig[x]))) {
                changed[x] = 0;
                state--;
            } else if (!changed[x] && (fabs(y) > 1e-6 * fabs(eig[x]))) {
                changed[x] = 1;
                state++;
            }
        }
        m[k][l] = 0;
#define ORC_ROT(M, w, x, y, z, r)                                                   \
    do {                                                                            \
        t = M[w][x];                                                                \
        u = M[y][z];                                                                \
        M[w][x] = t * c - s * u;                                                    \
        M[y][z] = s * t + c * u;                                                    \
        if (r) {                                                                    \
            if (fabs(M[w][x]) > fabs(M[w][max_col[w]])) max_col[w] = x;             \
            if (y < 3 && fabs(M[y][z]) > fabs(M[y][max_col[y]])) max_col[y] = z;    \
        }                                                                           \
    } while (0)
        for (int i = 0; i < k; i++) ORC_ROT(m, i, k, i, l, 1);
        for (int i = k + 1; i < l; i++) ORC_ROT(m, k, i, i, l, 1);
        for (int i = l + 1; i < n; i++) ORC_ROT(m, k, i, l, i, 1);
        for (int i = 0; i < n; i++) ORC_ROT(orth, k, i, l, i, 0);
#undef ORC_ROT
    }
    for (int k = 0; k < n - 1; k++)
        for (int l = k + 1; l < n; l++) m[k][l] = m[l][k];
}

/* Test knob: walk the records of calc_shape's sums last to first.  Used to
   show how much of the shape of a window of few records is set by the
   summation order alone (the reference's own result moves with it). */
static int orc_shape_reverse = 0;
void orc_set_shape_reverse(int on) { orc_shape_reverse = on; }

typedef struct orc_shape { /* the struct halo fields calc_shape reads and writes */
    float r;               /* h->r, kpc/h */
    float b_to_a, c_to_a, A[3];
    float inertia[6];      /* kept from the previous call when this one stops early */
} orc_shape;

/* src/properties.c:216-326 over records [0, n) (bound: !(pe < ke)) */
static void orc_calc_shape(orc_shape *h, const struct potential *po, int64_t n,
                           const float cen[3], const orc_params *p) {
    int64_t iter = ORC_SHAPE_ITERATIONS, analyze_p = 0;
    float b_to_a, c_to_a, min_r = (float)(p->force_res * p->force_res);
    double mass_t[3][3], orth[3][3], eig[3] = {0, 0, 0}, r = 0, dr, dr2, weight = 0;
    h->b_to_a = h->c_to_a = 0;
    memset(h->A, 0, sizeof h->A);
    if (!(h->r > 0)) return;
    min_r *= 1e6 / (h->r * h->r);
    for (int64_t j = 0; j < n; j++)
        if (!(po[j].pe < po[j].ke)) analyze_p++;
    if (analyze_p < 3 || !(h->r > 0)) return;
    if (analyze_p < iter) iter = analyze_p;
    for (int i = 0; i < 3; i++) {
        memset(orth[i], 0, sizeof(double) * 3);
        orth[i][i] = 1;
        eig[i] = (h->r * h->r) * 1e-6;
    }
    for (int64_t it = 0; it < iter; it++) {
        for (int k = 0; k < 3; k++) memset(mass_t[k], 0, sizeof(double) * 3);
        weight = 0;
        for (int64_t jj = 0; jj < n; jj++) {
            const int64_t j = orc_shape_reverse ? n - 1 - jj : jj;
            if (po[j].pe < po[j].ke) continue;
            r = 0;
            for (int k = 0; k < 3; k++) {
                dr = 0;
                for (int l = 0; l < 3; l++) dr += orth[k][l] * (po[j].pos[l] - cen[l]);
                r += dr * dr / eig[k];
            }
            if (r < min_r) r = min_r;
            if (!(r > 0 && r <= 1)) continue;
            double tw = ORC_WEIGHTED_SHAPES ? 1.0 / r : 1.0;
            weight += tw;
            for (int k = 0; k < 3; k++) {
                dr = po[j].pos[k] - cen[k];
                mass_t[k][k] += dr * dr * tw;
                for (int l = 0; l < k; l++) {
                    dr2 = po[j].pos[l] - cen[l];
                    mass_t[k][l] += dr2 * dr * tw;
                    mass_t[l][k] = mass_t[k][l];
                }
            }
        }
        if (!weight) return;
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) mass_t[k][l] /= (double)weight;
        double mt2[3][3];
        memcpy(mt2, mass_t, sizeof mt2);
        orc_jacobi(mass_t, eig, orth);
        int a = 0, b = 1, c = 2;
        if (eig[1] > eig[0]) {
            b = 0;
            a = 1;
        }
        if (eig[2] > eig[b]) {
            c = b;
            b = 2;
        }
        if (eig[b] > eig[a]) {
            int q = a;
            a = b;
            b = q;
        }
        if (!eig[a] || !eig[b] || !eig[c]) return;
        b_to_a = sqrt(eig[b] / eig[a]);
        c_to_a = sqrt(eig[c] / eig[a]);
        if ((fabs(b_to_a - h->b_to_a) < 0.01 * h->b_to_a) &&
            (fabs(c_to_a - h->c_to_a) < 0.01 * h->c_to_a))
            return;
        h->b_to_a = (b_to_a > 0) ? b_to_a : 0;
        h->c_to_a = (c_to_a > 0) ? c_to_a : 0;
        r = sqrt(eig[a]);
        for (int k = 0; k < 3; k++) {
            h->A[k] = 1e3 * r * orth[a][k];
            eig[k] *= (h->r * h->r * 1e-6) / (r * r);
        }
        h->inertia[0] = mt2[0][0];
        h->inertia[1] = mt2[1][1];
        h->inertia[2] = mt2[2][2];
        h->inertia[3] = mt2[0][1];
        h->inertia[4] = mt2[1][2];
        h->inertia[5] = mt2[2][0];
    }
}

/* ------------------------------------- NFW scale radius (src/nfw.c:13-173) */

static double orc_c_to_f(double c) {
    double cp1 = 1.0 + c;
    return (c * cp1 / (log1p(c) * cp1 - c));
}

static double orc_f_to_c(double f) { /* Newton-like solve of c_to_f(c) = f */
    double c = f, tc = orc_c_to_f(c), nc;
    while (fabs((f - tc) / f) > 1e-7) {
        double tc2 = orc_c_to_f(c + 0.1);
        double slope = (tc2 - tc) / 0.1;
        nc = c + (f - tc) / slope;
        if (nc < 0) c /= 2;
        else c = nc;
        tc = orc_c_to_f(c);
    }
    return c;
}

/* Klypin's estimate from vmax, rvmax and the virial mass and radius */
static float orc_klypin_rs(float mvir, float rvir, float vmax, float rvmax, float scale) {
    float f, c, vm2;
    if (!mvir || !rvir || !vmax) return (rvmax / ORC_RMAX_TO_RS);
    vm2 = vmax / ORC_VMAX_CONST;
    vm2 = vm2 * vm2 * scale;
    f = (rvir / 1.0e3) * vm2 / (mvir * ORC_RS_CONSTANT);
    if (f < 4.625) return (rvmax / ORC_RMAX_TO_RS);
    c = orc_f_to_c(f);
    if (c <= 0) return (rvmax / ORC_RMAX_TO_RS);
    return (rvir / c);
}

static float orc_nfw_menc(float r, float rs) { return (log((rs + r) / rs) - r / (rs + r)); }

static float orc_chi2(float rs, const float *bin_r, const float *w, int64_t nb) {
    float bin_mass = orc_nfw_menc(bin_r[nb], rs) / ((double)nb);
    float last = 0, chi2 = 0, dx;
    for (int64_t i = 0; i < nb; i++) {
        float menc = orc_nfw_menc(bin_r[i + 1], rs);
        dx = w[i] * ((menc - last) - bin_mass) / bin_mass;
        chi2 += dx * dx;
        last = menc;
    }
    return chi2;
}

/* the minimisation of src/nfw.c:75-128: parabolic steps on chi2(rs) */
static float orc_fit_rs(float rs, const float *bin_r, const float *w, int64_t nb, float *out_chi2) {
    if (!(rs > bin_r[0]) || !(rs < bin_r[nb])) rs = bin_r[(nb + 1) / 2];
    float initial = rs;
    float chi2 = orc_chi2(rs, bin_r, w, nb);
    float drs = rs / 10.0;
    int64_t iter = 0;
    float last = 1e30;
    if (!(rs > 0)) return initial;
    while (fabs(chi2 - last) > 0.005 * chi2 && iter < 15) {
        last = chi2;
        if (rs + drs > bin_r[nb]) drs = 0.1 * (bin_r[nb] - rs);
        if (rs - drs < bin_r[0]) drs = 0.1 * (rs - bin_r[0]);
        float cr = orc_chi2(rs + drs, bin_r, w, nb);
        float cl = orc_chi2(rs - drs, bin_r, w, nb);
        float dx = 0.5 * (cr - cl) / drs;
        float dx2 = (cr + cl - 2.0 * chi2) / (drs * drs);
        float move = 0;
        if (dx2 != 0) move = -dx / dx2;
        if (!move) return rs;
        if (rs + move > 4 * rs) move = 3.0 * rs;
        if (rs + move < 0.25 * rs) move = -0.75 * rs;
        float nrs = rs + move;
        if (nrs > bin_r[nb]) nrs = rs + 0.8 * (bin_r[nb] - rs);
        else if (nrs < bin_r[0]) nrs = rs + 0.8 * (bin_r[0] - rs);
        float nchi2 = orc_chi2(nrs, bin_r, w, nb);
        if (cr < nchi2) {
            nchi2 = cr;
            nrs = rs + drs;
        }
        if (cl < nchi2) {
            nchi2 = cl;
            nrs = rs - drs;
        }
        if (last < nchi2) {
            drs *= 0.5;
            continue;
        }
        drs = fabs(rs - nrs) / 10.0;
        rs = nrs;
        chi2 = nchi2;
        iter++;
    }
    *out_chi2 = chi2;
    return rs;
}

/* src/nfw.c:130-173 over records [0, n): Klypin rs, then the fit to
   equal-count radial bins (50 at most, 15 records at least) */
static void orc_scale_radius(orc_props *out, float mvir, float rvir, float vmax, float rvmax,
                             float scale, const struct potential *po, int64_t n,
                             const orc_params *p) {
    float bin_r[51] = {0}, w[50] = {0};
    int64_t nb = 0, analyze_p = 0, ppbin;
    float rs = out->klypin_rs = orc_klypin_rs(mvir, rvir, vmax, rvmax, scale);
    for (int64_t i = 0; i < n - 1; i++)
        if (!(po[i].pe < po[i].ke)) analyze_p++;
    if (analyze_p < 100) {
        out->rs = rs;
        return;
    }
    ppbin = ceil(((double)analyze_p / (double)50));
    if (ppbin < 15) ppbin = (analyze_p) / ((int64_t)(analyze_p / 15));
    bin_r[0] = 0;
    for (int64_t i = 0, j = 0; i < n - 1; i++) {
        if (po[i].pe < po[i].ke) continue;
        if (++j == ppbin) {
            nb++;
            bin_r[nb] = 1e3 * 0.5 * (sqrt(po[i].r2) + sqrt(po[i + 1].r2));
            w[nb - 1] = 1;
            j = 0;
        }
    }
    for (int64_t i = 0; i < nb; i++)
        if (bin_r[i] * 1e-3 < 3 * p->force_res) w[i] = 0.1;
    float chi2 = 0.0;
    out->rs = orc_fit_rs(rs, bin_r, w, nb, &chi2);
    out->chi2 = chi2;
}

/* src/properties.c:351-382 over the bound records of [0, n) */
static void orc_pseudo_evolution(orc_props *out, const struct potential *po, int64_t n,
                                 float rs, float hr, double rvir_dens_z0, const orc_params *p) {
    int64_t num_part = 0, num_part_pe_d = 0;
    double r, r32, max_pe_b = 0;
    double r_pe_d = rs * 4.0;
    if (r_pe_d < hr / 5.0) r_pe_d = hr / 5.0;
    r_pe_d *= 1e-3;
    for (int64_t j = 0; j < n; j++) {
        if (po[j].pe < po[j].ke) continue;
        num_part++;
        r = sqrt(po[j].r2);
        r32 = sqrt(r);
        r32 = r32 * r32 * r32;
        if ((double)(num_part * num_part) / r32 > max_pe_b) max_pe_b = (double)(num_part * num_part) / r32;
        if (r < r_pe_d) num_part_pe_d = num_part;
    }
    out->m_pe_d = (float)(num_part_pe_d * p->particle_mass);
    out->m_pe_b = (float)(p->particle_mass * pow(max_pe_b, 2.0 / 3.0) /
                          cbrt(4.0 * M_PI * rvir_dens_z0 / 3.0));
}

/* src/properties.c:328-349 over records [0, n), highest index first */
static float orc_total_energy(const struct potential *po, int64_t n,
                              const orc_params *p, float *ratio) {
    double phi = 0, total_phi = 0, ke = 0;
    for (int64_t i = n - 1; i >= 0; i--) {
        if (!(po[i].pe > po[i].ke)) continue;
        ke += po[i].ke;
        double r = sqrt(po[i].r2);
        if (r < p->force_res) r = p->force_res;
        total_phi += p->particle_mass * i / r + phi;
        phi += p->particle_mass / r;
    }
    total_phi /= 2.0;
    *ratio = 0;
    if (total_phi) *ratio = (float)(ke / total_phi);
    float scale_now = (float)p->scale_now;
    return (float)((ke - total_phi) * p->particle_mass * ORC_GC / scale_now);
}

/* src/properties.c:384-510 with the defaults BOUND_PROPS = 1 and LIGHTCONE
   = 0, run for bound = 0 and then bound = 1 as src/properties.c:552-553
   does, calc_shape, calc_scale_radius and _calc_pseudo_evolution_masses
   included.  thresh[0..4] = particle_thresh_dens, thresh[5] =
   particle_rvir_dens, thresh[6] = particle_rvir_dens_z0.  po: the kept
   records sorted by r2 with pe and ke. */
void orc_halo_props(const struct potential *po, int64_t total_p, const float pos[6],
                    const orc_params *p, const double thresh[7], orc_props *out) {
    memset(out, 0, sizeof *out);
    const double four_thirds_pi = 4.0 * M_PI / 3.0;
    const float scale_now = (float)p->scale_now;
    for (int bound = 0; bound <= 1; bound++) {
        double dens_thresh = thresh[0] * four_thirds_pi;
        double d[4];
        for (int k = 0; k < 4; k++) d[k] = thresh[k + 1] * four_thirds_pi;
        double rvir_thresh = thresh[5] * four_thirds_pi;
        int64_t part_mdelta = 0, num_part = 0, np_alt[4] = {0, 0, 0, 0}, np_vir = 0;
        int64_t dens_tot = 0, parts_avgd = 0, num_part_half = 0;
        double vmax = 0, rvmax = 0, L[3] = {0, 0, 0};
        double xavg[3] = {0, 0, 0}, vavg[3] = {0, 0, 0}, vrms[3] = {0, 0, 0};
        for (int64_t j = 0; j < total_p; j++) {
            if (bound && po[j].pe < po[j].ke) continue;
            num_part++;
            double r = sqrt(po[j].r2);
            if (r < p->force_res) r = p->force_res;
            double cur = (double)num_part / (r * r * r);
            if (cur > dens_thresh) {
                part_mdelta = num_part;
                dens_tot = j;
            }
            for (int k = 0; k < 4; k++)
                if (cur > d[k]) np_alt[k] = num_part;
            if (cur > rvir_thresh) {
                double circ = (double)num_part / r;
                np_vir = num_part;
                if (part_mdelta && circ > vmax) {
                    vmax = circ;
                    rvmax = r;
                }
            }
        }
        for (int64_t j = 0; j < dens_tot; j++) {
            if (bound && po[j].pe < po[j].ke) continue;
            const float *x = po[j].pos;
            /* src/properties.c:127-137: float differences and products */
            L[0] += (float)((x[1] - pos[1]) * (x[5] - pos[5]));
            L[0] -= (float)((x[2] - pos[2]) * (x[4] - pos[4]));
            L[1] += (float)((x[2] - pos[2]) * (x[3] - pos[3]));
            L[1] -= (float)((x[0] - pos[0]) * (x[5] - pos[5]));
            L[2] += (float)((x[0] - pos[0]) * (x[4] - pos[4]));
            L[2] -= (float)((x[1] - pos[1]) * (x[3] - pos[3]));
            parts_avgd++;
            if (parts_avgd * 2 >= part_mdelta && !num_part_half) num_part_half = j;
            for (int k = 0; k < 3; k++) {
                xavg[k] += (x[k] - xavg[k]) / ((double)parts_avgd);
                double md = x[k + 3] - vavg[k];
                vavg[k] += md / (double)parts_avgd;
                vrms[k] += md * (x[k + 3] - vavg[k]);
            }
        }
        double m = part_mdelta * p->particle_mass;
        if (!bound) {
            out->m = (float)m;
            continue;
        }
        out->mgrav = (float)m;
        for (int k = 0; k < 3; k++) vrms[k] = parts_avgd > 0 ? vrms[k] / parts_avgd : 0;
        float xoff = 0, voff = 0; /* float fields, double terms */
        for (int k = 0; k < 3; k++) {
            double ds = xavg[k] - pos[k];
            xoff = (float)(xoff + ds * ds);
            ds = vavg[k] - pos[k + 3];
            voff = (float)(voff + ds * ds);
        }
        for (int k = 0; k < 4; k++) out->alt_m[k] = (float)(np_alt[k] * p->particle_mass);
        out->Xoff = (float)(sqrt(xoff) * 1e3);
        out->Voff = (float)sqrt(voff);
        out->vrms = (float)sqrt(vrms[0] + vrms[1] + vrms[2]);
        out->vmax = (float)(ORC_VMAX_CONST * sqrt(vmax * (p->particle_mass / scale_now)));
        out->rvmax = (float)(rvmax * 1e3);
        out->halfmass_radius = total_p > 0 ? (float)(sqrt(po[num_part_half].r2) * 1e3) : 0.0f;
        /* src/properties.c:476-492: shape within the 4th definition, kept
           as *2, then within the main one */
        orc_shape sh;
        memset(&sh, 0, sizeof sh);
        sh.r = (float)(cbrt((3.0 / (4.0 * M_PI)) * np_alt[2] / thresh[3]) * 1e3);
        orc_calc_shape(&sh, po, np_alt[2], pos, p);
        out->b_to_a2 = sh.b_to_a;
        out->c_to_a2 = sh.c_to_a;
        memcpy(out->A2, sh.A, sizeof out->A2);
        memcpy(out->inertia_tensor2, sh.inertia, sizeof out->inertia_tensor2);
        sh.r = out->r = (float)(cbrt((3.0 / (4.0 * M_PI)) * part_mdelta / thresh[0]) * 1e3);
        orc_calc_shape(&sh, po, dens_tot, pos, p);
        out->b_to_a = sh.b_to_a;
        out->c_to_a = sh.c_to_a;
        memcpy(out->A, sh.A, sizeof out->A);
        memcpy(out->inertia_tensor, sh.inertia, sizeof out->inertia_tensor);
        double rvir = cbrt((3.0 / (4.0 * M_PI)) * np_vir / thresh[5]) * 1e3;
        double mvir = np_vir * p->particle_mass;
        out->rvir = (float)rvir;
        orc_scale_radius(out, (float)m, out->r, out->vmax, out->rvmax, scale_now, po, dens_tot, p);
        for (int k = 0; k < 3; k++) out->J[k] = (float)(p->particle_mass * scale_now * L[k]);
        out->energy = orc_total_energy(po, dens_tot, p, &out->kin_to_pot);
        double Jh = p->particle_mass * scale_now * sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
        out->spin = (float)((m > 0) ? (Jh * sqrt(fabs(out->energy)) / (ORC_GC * pow(m, 2.5))) : 0);
        out->bullock_spin =
            (float)((m > 0) ? (Jh / (mvir * sqrt(2.0 * ORC_GC * mvir * rvir * scale_now / 1e3)))
                            : 0);
        out->n_mdelta = part_mdelta;
        out->dens_tot = dens_tot;
        out->n_vir = np_vir;
        out->n_half = num_part_half;
        orc_pseudo_evolution(out, po, total_p, out->rs, out->r, thresh[6], p);
    }
}

void orc_halo_props_batch(const struct potential *po, const orc_halo *halos, int64_t nhalos,
                          const orc_params *p, const double thresh[7], orc_props *out,
                          int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t h = 0; h < nhalos; h++)
        orc_halo_props(po + halos[h].first, halos[h].kept, halos[h].pos, p, thresh, out + h);
}

/* src/groupies.c:97-100 */
double orc_vir_density(const orc_params *p, double a) {
    double e = orc_hubble_scaling(p, 1.0 / a - 1.0);
    double x = (p->Om / pow(a, 3)) / pow(e, 2.0) - 1.0;
    return ((18 * M_PI * M_PI + 82.0 * x - 39 * x * x) / (1.0 + x));
}

int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int64_t orc_sizeof_node(void) { return (int64_t)sizeof(orc_node); }
int64_t orc_sizeof_halo(void) { return (int64_t)sizeof(orc_halo); }
int64_t orc_sizeof_potential(void) { return (int64_t)sizeof(struct potential); }
int64_t orc_sizeof_props(void) { return (int64_t)sizeof(orc_props); }
int64_t orc_sizeof_tree_halo(void) { return (int64_t)sizeof(orc_tree_halo); }

/* ================================================================ sub-FOF
 * The phase-space sub-FOF step of _find_subs: _find_subfofs_better2
 * (src/groupies.c:287-314) over one group -- norm_sd (788-830) with
 * calc_deviations (src/jacobi.c:137-175), the 6D GROUPIES fast3tree with 40
 * points per leaf (src/groupies.c:24-28; src/fast3tree.c:736-907) rebuilt in
 * place, next-closest distances of the sample (src/fast3tree.c:953-1030),
 * find_median_r (src/groupies.c:744-786) and the linking at that length
 * (_find_subfofs_at_r, src/groupies.c:274-285 with src/fof.c).
 *
 * Unlike the potential tree above, this one partitions in place exactly as
 * the reference does: the sample indices and the linking loop both walk the
 * array in the order the rebuild leaves it.  The linking result is given as
 * components (label = smallest tree position in the component, -1 for a
 * particle with no neighbour) and, by orc6_fofs, as the reference's own
 * output: the particles sorted by its union-find roots and the sub-FOFs in
 * root order. */
#define ORC_LEAF6 40

typedef struct orc_node6 {
    float min[6], max[6];
    int64_t start, n;
    int32_t div, left, right, parent;
} orc_node6;

typedef struct orc_tree6 {
    orc_node6 *nd;
    int64_t nn, cap;
    orc_particle *p;
} orc_tree6;

/* src/jacobi.c:137-151 */
static double orc_set_eig(double in[3][3]) {
    double out[3][3], e[3], m[3][3];
    memcpy(m, in, sizeof m);
    orc_jacobi(m, e, out);
    double r = e[0];
    if (e[1] < r) r = e[1];
    if (e[2] < r) r = e[2];
    if (r <= 0) return sqrt(in[0][0] + in[1][1] + in[2][2]);
    return sqrt(r);
}

/* src/groupies.c:788-830 */
static void orc_norm_sd(orc_particle *pp, int64_t n, int top, double scaling) {
    double mean[6] = {0}, corr[6][6] = {{0}};
    if (!n) return;
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 6; j++) mean[j] += pp[i].pos[j];
    for (int j = 0; j < 6; j++) mean[j] /= (double)n;
    for (int64_t i = 0; i < n; i++) {
        float *x = pp[i].pos;
        for (int j = 0; j < 6; j++) x[j] = (float)((double)x[j] - mean[j]);
        for (int j = 0; j < 6; j++)
            for (int k = j; k < 6; k++) {
                float prod = x[j] * x[k];
                corr[j][k] += prod;
            }
    }
    for (int j = 0; j < 6; j++)
        for (int k = j; k < 6; k++) corr[j][k] /= (double)n;
    for (int j = 0; j < 6; j++)
        for (int k = 0; k < j; k++) corr[j][k] = corr[k][j];
    double a[3][3], b[3][3];
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) {
            a[j][k] = corr[j][k];
            b[j][k] = corr[j + 3][k + 3];
        }
    double sx = orc_set_eig(a), sv = orc_set_eig(b);
    if (top) sx *= scaling;
    if (!sx || !sv) return;
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 6; j++)
            pp[i].pos[j] = (float)((double)pp[i].pos[j] / (j < 3 ? sx : sv));
}

/* src/fast3tree.c:778-794 */
static void orc6_bounds(orc_tree6 *t, orc_node6 *nd) {
    const orc_particle *p = t->p + nd->start;
    for (int j = 0; j < 6; j++) nd->min[j] = nd->max[j] = p[0].pos[j];
    for (int64_t i = 1; i < nd->n; i++)
        for (int j = 0; j < 6; j++) {
            float x = p[i].pos[j];
            if (x < nd->min[j]) nd->min[j] = x;
            else if (x > nd->max[j]) nd->max[j] = x;
        }
}

/* src/fast3tree.c:736-774: widest dimension (the last one first, a later one
   only when strictly wider), float midpoint, in-place swap partition */
static int64_t orc6_partition(orc_tree6 *t, orc_node6 *nd) {
    int dim = 5;
    float w = nd->max[5] - nd->min[5];
    for (int i = 0; i < 5; i++) {
        float w2 = nd->max[i] - nd->min[i];
        if (w2 > w) {
            w = w2;
            dim = i;
        }
    }
    nd->div = dim;
    if (nd->max[dim] == nd->min[dim]) return nd->n;
    const float lim = 0.5 * (nd->max[dim] + nd->min[dim]);
    orc_particle *p = t->p + nd->start;
    int64_t lo = 0, hi = nd->n - 1;
    while (lo < hi) {
        if (p[lo].pos[dim] > lim) {
            orc_particle tmp = p[hi];
            p[hi] = p[lo];
            p[lo] = tmp;
            hi--;
        } else {
            lo++;
        }
    }
    if (lo == hi && !(p[lo].pos[dim] > lim)) lo++;
    return lo;
}

/* src/fast3tree.c:798-846, left subtree before right */
static void orc6_split(orc_tree6 *t, int64_t ni) {
    int64_t nl = orc6_partition(t, t->nd + ni);
    if (nl == t->nd[ni].n || nl == 0) {
        t->nd[ni].div = -1;
        return;
    }
    if (t->nn + 2 > t->cap) {
        t->cap = t->cap * 2 + 16;
        t->nd = realloc(t->nd, sizeof(orc_node6) * (size_t)t->cap);
    }
    int64_t l = t->nn, r = t->nn + 1;
    t->nn += 2;
    orc_node6 *nd = t->nd + ni;
    nd->left = (int32_t)l;
    nd->right = (int32_t)r;
    orc_node6 *L = t->nd + l, *R = t->nd + r;
    memset(L, 0, 2 * sizeof(orc_node6));
    L->parent = R->parent = (int32_t)ni;
    L->div = R->div = -1;
    L->start = nd->start;
    L->n = nl;
    R->start = nd->start + nl;
    R->n = nd->n - nl;
    orc6_bounds(t, L);
    orc6_bounds(t, R);
    if (t->nd[l].n > ORC_LEAF6) orc6_split(t, l);
    if (t->nd[r].n > ORC_LEAF6) orc6_split(t, r);
}

static void orc6_build(orc_tree6 *t, orc_particle *p, int64_t n) {
    t->p = p;
    t->cap = 3 + n / (ORC_LEAF6 / 2);
    t->nd = calloc((size_t)t->cap, sizeof(orc_node6));
    t->nn = 1;
    t->nd[0].start = 0;
    t->nd[0].n = n;
    t->nd[0].div = -1;
    t->nd[0].parent = -1;
    if (n) orc6_bounds(t, t->nd);
    if (n > ORC_LEAF6) orc6_split(t, 0);
}

/* src/fast3tree.c:294-309 */
static int orc6_apart(const orc_node6 *nd, const float c[6], float r) {
    float d = 0, e;
    const float r2 = r * r;
    for (int i = 0; i < 6; i++) {
        if ((e = c[i] - nd->min[i]) < 0) {
            d += e * e;
            if (d >= r2) return 1;
        } else if ((e = c[i] - nd->max[i]) > 0) {
            d += e * e;
            if (d >= r2) return 1;
        }
    }
    return 0;
}

static float orc6_dist2(const float *a, const float *b) {
    float d = 0;
    for (int j = 0; j < 6; j++) {
        float dx = a[j] - b[j];
        d += dx * dx;
    }
    return d;
}

/* src/fast3tree.c:956-987 */
static float orc6_nc(const orc_tree6 *t, int64_t ni, const float c[6], float r, int64_t skip) {
    const orc_node6 *nd = t->nd + ni;
    if (orc6_apart(nd, c, r) || ni == skip) return r;
    if (nd->div < 0) {
        float r2 = r * r;
        for (int64_t i = 0; i < nd->n; i++) {
            float d = orc6_dist2(c, t->p[nd->start + i].pos);
            if (d < r2) r2 = d;
        }
        return (float)sqrt(r2);
    }
    int64_t near = nd->left, far = nd->right;
    if (c[nd->div] > 0.5 * (nd->min[nd->div] + nd->max[nd->div])) {
        near = nd->right;
        far = nd->left;
    }
    r = orc6_nc(t, near, c, r, skip);
    return orc6_nc(t, far, c, r, skip);
}

/* src/fast3tree.c:992-1030 */
static float orc6_next_closest(const orc_tree6 *t, const float c[6]) {
    int64_t ni = 0;
    while (t->nd[ni].div >= 0) {
        const orc_node6 *nd = t->nd + ni;
        ni = (c[nd->div] <= t->nd[nd->left].max[nd->div]) ? nd->left : nd->right;
    }
    while (ni != 0 && t->nd[ni].min[t->nd[t->nd[ni].parent].div] ==
                          t->nd[ni].max[t->nd[t->nd[ni].parent].div])
        ni = t->nd[ni].parent;
    const orc_node6 *nd = t->nd + ni;
    float md = 0;
    for (int j = 0; j < 6; j++) {
        float dx = nd->max[j] - nd->min[j];
        md += dx * dx;
    }
    for (int64_t i = 0; i < nd->n; i++) {
        float d = orc6_dist2(c, t->p[nd->start + i].pos);
        if (d && d < md) md = d;
    }
    return orc6_nc(t, 0, c, (float)sqrt(md), ni);
}

/* src/groupies.c:744-786: quickselect with rand_r pivots; it stops early when
   the ends of the current range are equal, so its answer is not always the
   k-th smallest value */
static float orc_random_unit(unsigned int *seed) {
    return ((float)(rand_r(seed) % (RAND_MAX)) / (float)(RAND_MAX));
}

static int64_t orc_rad_partition(float *v, int64_t left, int64_t right, int64_t pivot_at) {
    const float pv = v[pivot_at];
    float tmp;
    tmp = v[pivot_at];
    v[pivot_at] = v[right - 1];
    v[right - 1] = tmp;
    int64_t hi = right - 2, lo = left;
    while (lo < hi) {
        if (v[lo] > pv) {
            tmp = v[lo];
            v[lo] = v[hi];
            v[hi] = tmp;
            hi--;
        } else {
            lo++;
        }
    }
    /* the reference reads v[hi] here even when the range holds one value
       (hi = left - 1); the element before a range is never above its pivot */
    if (hi < left || v[hi] <= pv) hi++;
    tmp = v[right - 1];
    v[right - 1] = v[hi];
    v[hi] = tmp;
    return hi;
}

float orc_find_median_r(float *v, int64_t n, float frac, unsigned int *seed) {
    int64_t k = n * frac, left = 0, right = n;
    if (n < 2) return v[0];
    for (;;) {
        int64_t at = left + orc_random_unit(seed) * (right - left);
        int64_t p = orc_rad_partition(v, left, right, at);
        if (k == p || v[left] == v[right - 1]) return v[k];
        if (k < p)
            right = p;
        else
            left = p + 1;
    }
}

static int64_t orc_uf_find(int64_t *up, int64_t x) {
    while (up[x] != x) {
        up[x] = up[up[x]];
        x = up[x];
    }
    return x;
}

/* every pair closer than r (float distance squared < r * r) joins one
   component; label[i] = the smallest tree position of i's component, -1
   when i has no such neighbour */
static void orc6_link(const orc_tree6 *t, int64_t n, float r, int64_t *label) {
    int64_t *up = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    char *has = calloc((size_t)(n ? n : 1), 1);
    int64_t *stack = malloc(sizeof(int64_t) * (size_t)(2 * t->nn + 2));
    for (int64_t i = 0; i < n; i++) up[i] = i;
    const float r2 = r * r;
    for (int64_t i = 0; i < n; i++) {
        const float *c = t->p[i].pos;
        int64_t sp = 0;
        stack[sp++] = 0;
        while (sp) {
            const orc_node6 *nd = t->nd + stack[--sp];
            if (nd->start + nd->n <= i || orc6_apart(nd, c, r)) continue;
            if (nd->div >= 0) {
                stack[sp++] = nd->right;
                stack[sp++] = nd->left;
                continue;
            }
            for (int64_t j = nd->start > i + 1 ? nd->start : i + 1; j < nd->start + nd->n; j++) {
                if (!(orc6_dist2(c, t->p[j].pos) < r2)) continue;
                has[i] = has[j] = 1;
                int64_t a = orc_uf_find(up, i), b = orc_uf_find(up, j);
                if (a < b) up[b] = a;
                else if (b < a) up[a] = b;
            }
        }
    }
    for (int64_t i = 0; i < n; i++) label[i] = has[i] ? orc_uf_find(up, i) : -1;
    free(stack);
    free(has);
    free(up);
}

/* --- the reference's own output order: the linking loop of
 * _find_subfofs_at_r (src/groupies.c:274-285) and build_fullfofs
 * (src/fof.c:223-252), step by step.  A particle's sphere search
 * (fast3tree_find_sphere_skip, src/fast3tree.c:419-465) lists the particle
 * itself and then every later particle closer than r, in array order (the
 * depth-first walk meets the leaves in array order).  Small FOFs are
 * numbered as they are created; sf_up[] is the reference's root field. */
static void orc_sf_collapse(int64_t *sf_up, int64_t f) {  /* src/fof.c:58-70 */
    if (sf_up[sf_up[f]] == sf_up[f]) return;
    int64_t top = sf_up[f];
    while (sf_up[top] != top) top = sf_up[top];
    while (f != top) {
        const int64_t next = sf_up[f];
        sf_up[f] = top;
        f = next;
    }
}

/* merge_smallfofs(f1, f2), src/fof.c:72-89: f2's tree ends under f1's root */
static void orc_sf_merge(int64_t *sf_up, int64_t f1, int64_t f2) {
    if (sf_up[f1] == sf_up[f2]) return;
    orc_sf_collapse(sf_up, f1);
    const int64_t top = sf_up[f1];
    if (sf_up[f2] == f2) {
        sf_up[f2] = top;
        return;
    }
    for (;;) {  /* every link on f2's path to its root is pointed at top */
        const int64_t next = sf_up[f2];
        sf_up[f2] = top;
        if (next == top) break;
        f2 = next;
    }
}

/* partition_sort_particles, src/fof.c:184-221: recursive in-place
   partitions about the middle of each range's assignment values */
static void orc_sf_partition(orc_particle *p, int64_t *a, int64_t lo, int64_t hi) {
    while (hi - lo >= 2) {
        int64_t vmin = a[lo], vmax = a[lo];
        for (int64_t i = lo + 1; i < hi; i++) {
            if (a[i] > vmax) vmax = a[i];
            if (a[i] < vmin) vmin = a[i];
        }
        if (vmin == vmax) return;
        const int64_t pivot = vmin + (vmax - vmin) / 2;
        int64_t i = lo, last = hi - 1;
        while (i < last) {
            if (a[i] > pivot) {  /* swap with the last unexamined slot, look again */
                const orc_particle tp = p[i];
                p[i] = p[last];
                p[last] = tp;
                const int64_t ta = a[i];
                a[i] = a[last];
                a[last] = ta;
                last--;
            } else {
                i++;
            }
        }
        if (a[last] <= pivot) last++;
        orc_sf_partition(p, a, lo, last);
        lo = last;
    }
}

/* The linking at r on the tree order, then out = the tree order's particles
   sorted by root as the reference leaves them, and its sub-FOFs of at least
   min_p particles in root order; returns their count (at most cap are
   written). */
static int64_t orc6_fofs(const orc_tree6 *t, int64_t n, float r, int64_t min_p,
                         orc_particle *out, int64_t *fof_first, int64_t *fof_num, int64_t cap) {
    int64_t *sf = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));   /* particle_smallfofs */
    int64_t *sf_up = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    int64_t *nb = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    int64_t *stack = malloc(sizeof(int64_t) * (size_t)(2 * t->nn + 2));
    int64_t nsf = 0;
    const float r2 = r * r;
    for (int64_t i = 0; i < n; i++) sf[i] = -1;
    for (int64_t i = 0; i < n; i++) {
        /* the search result after the particle itself: later points, in order */
        int64_t k = 0, sp = 0;
        const float *c = t->p[i].pos;
        stack[sp++] = 0;
        while (sp) {
            const orc_node6 *nd = t->nd + stack[--sp];
            if (nd->start + nd->n <= i || orc6_apart(nd, c, r)) continue;
            if (nd->div >= 0) {
                stack[sp++] = nd->right;
                stack[sp++] = nd->left;
                continue;
            }
            for (int64_t j = nd->start > i + 1 ? nd->start : i + 1; j < nd->start + nd->n; j++)
                if (orc6_dist2(c, t->p[j].pos) < r2) nb[k++] = j;
        }
        if (!k) continue;  /* fewer than two points: src/fof.c:117 */
        /* link_particle_to_fof, src/fof.c:113-137 */
        int64_t f = sf[i];
        if (f < 0)
            for (int64_t q = 0; q < k && f < 0; q++) f = sf[nb[q]];
        if (f < 0) {
            sf_up[nsf] = nsf;
            f = nsf++;
        }
        for (int64_t q = -1; q < k; q++) {
            const int64_t x = q < 0 ? i : nb[q];
            if (sf[x] < 0) sf[x] = f;
            else orc_sf_merge(sf_up, sf[x], f);
        }
    }
    /* collapse_smallfofs, src/fof.c:159-170 */
    for (int64_t s = 0; s < nsf; s++) orc_sf_collapse(sf_up, s);
    for (int64_t i = 0; i < n; i++)
        if (sf[i] >= 0) sf[i] = sf_up[sf[i]];
    memcpy(out, t->p, sizeof(orc_particle) * (size_t)n);
    orc_sf_partition(out, sf, 0, n);
    /* build_fullfofs: one per run of equal roots, kept at min_p or more */
    int64_t nf = 0;
    for (int64_t i = 0; i < n;) {
        int64_t e = i + 1;
        while (e < n && sf[e] == sf[i]) e++;
        if (sf[i] >= 0 && e - i >= min_p) {
            if (nf < cap) {
                fof_first[nf] = i;
                fof_num[nf] = e - i;
            }
            nf++;
        }
        i = e;
    }
    free(stack);
    free(nb);
    free(sf_up);
    free(sf);
    return nf;
}

/* the whole step for one group, in place on pp (normalised, tree order);
   returns -1 (nothing done) when a coordinate is not finite.  With final
   given: final = the particles as the reference leaves them (sorted by
   union-find root) and its sub-FOFs (first, num_p) of at least min_p
   particles, *nfof of them (at most cap written). */
int orc_subfof(orc_particle *pp, int64_t n, float thresh, int top, double scaling,
               int64_t exact_ll, float *target_r, float *sample_r, int64_t *nsample,
               int64_t *label, orc_particle *final, int64_t min_p, int64_t *fof_first,
               int64_t *fof_num, int64_t cap, int64_t *nfof) {
    if (nfof) *nfof = 0;
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 6; j++)
            if (!isfinite(pp[i].pos[j])) return -1;
    *nsample = 0;
    *target_r = 0;
    if (!n) return 0;
    orc_norm_sd(pp, n, top, scaling);
    for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < 6; j++)
            if (!isfinite(pp[i].pos[j])) return -1;
    orc_tree6 t;
    orc6_build(&t, pp, n);
    int64_t ns = exact_ll ? n : 10000;
    if (ns > n) ns = n;
    unsigned int seed = (unsigned int)n;
    for (int64_t i = 0; i < ns; i++) {
        int64_t j = i;
        if (ns != n) {
            j = rand_r(&seed);
            j <<= 31;
            j += rand_r(&seed);
            j %= n;
        }
        sample_r[i] = orc6_next_closest(&t, pp[j].pos);
    }
    *nsample = ns;
    float *tmp = malloc(sizeof(float) * (size_t)ns);
    memcpy(tmp, sample_r, sizeof(float) * (size_t)ns);
    *target_r = orc_find_median_r(tmp, ns, thresh, &seed);
    free(tmp);
    orc6_link(&t, n, *target_r, label);
    if (final) {
        const int64_t nf = orc6_fofs(&t, n, *target_r, min_p, final, fof_first, fof_num, cap);
        if (nfof) *nfof = nf;
    }
    free(t.nd);
    return 0;
}

/* next-closest distances of the queries on the tree rebuilt (in place) over p */
void orc_next_closest(orc_particle *p, int64_t n, const float *q, int64_t nq, float *out) {
    orc_tree6 t;
    orc6_build(&t, p, n);
    for (int64_t i = 0; i < nq; i++) out[i] = orc6_next_closest(&t, q + 6 * i);
    free(t.nd);
}

/* orc_subfof over a batch of groups (first, num_p), OpenMP dynamic over
   groups; label[] is indexed like the particles (group-local positions) */
typedef struct orc_fof_group {
    int64_t first, num_p;
} orc_fof_group;

void orc_subfof_batch(orc_particle *pp, const orc_fof_group *g, int64_t ng, float thresh,
                      int top, double scaling, int64_t exact_ll, float *target_r,
                      int64_t *label, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < ng; i++) {
        int64_t ns, n = g[i].num_p;
        float *s = malloc(sizeof(float) * (size_t)(n > 10000 ? n : 10000));
        orc_subfof(pp + g[i].first, n, thresh, top, scaling, exact_ll, target_r + i, s, &ns,
                   label + g[i].first, NULL, 0, NULL, NULL, 0, NULL);
        free(s);
    }
}
