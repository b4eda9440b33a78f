/*
 * examd_oracle.c -- CPU restatement of exaMD's midpoint-method LJ force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product library never links it.
 *
 * Built with -ffp-contract=off so that every multiply and add rounds on its
 * own, as the reference's Fortran statements are written.  Citations are to
 * the exaMD v2.0 sources.
 */
#include "examd_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

/* SRC/exaMD_utils.f90:167-190: nblocks_xyz = Int(globalbox/block_length),
 * globalboxi = one/globalbox. */
int or_domain_create(const double box[3], double block_length, or_domain *d)
{
    if (!(block_length > 0.0)) return 1;
    d->bl = block_length;
    d->nblocks = 1;
    for (int a = 0; a < 3; ++a) {
        if (!(box[a] > 0.0)) return 1;
        d->box[a] = box[a];
        d->boxi[a] = 1.0 / box[a];
        d->nb[a] = (int32_t)(box[a] / block_length); /* Int() truncates */
        d->nblocks *= d->nb[a];
    }
    return 0;
}

/* SRC/exaMD_module.f90:101-125. */
int or_identify_block(const or_domain *d, const double pos[3],
                      int32_t blockxyz[3], int64_t *blocknum)
{
    for (int a = 0; a < 3; ++a) {
        double q = pos[a] / d->bl;          /* posxyz/block_length */
        q = q + 1.0;                        /* + one */
        int32_t b = (int32_t)q;             /* Int(): truncation toward 0 */
        if (b > d->nb[a]) b = d->nb[a];     /* Where (blockxyz > nblocks_xyz) */
        blockxyz[a] = b;
    }
    int64_t num = (int64_t)blockxyz[0]
                + (int64_t)(blockxyz[1] - 1) * d->nb[0]
                + (int64_t)(blockxyz[2] - 1) * ((int64_t)d->nb[0] * d->nb[1]);
    *blocknum = num;
    if (num < 1 || num > d->nblocks) return 109;
    return 0;
}

/* SRC/exaMD_module.f90:282-344: interactions(1:2, 1:63), stencil indices. */
static const int32_t OR_TABLE[63][2] = {
    {14, 14}, {14, 15}, {14, 16}, {14, 17}, {14, 18}, {14, 19}, {14, 20},
    {14, 21}, {14, 22}, {14, 23}, {14, 24}, {14, 25}, {14, 26}, {14, 27},
    {10, 18}, {10, 27}, {11, 17}, {11, 18}, {11, 25}, {11, 26}, {11, 27},
    {12, 16}, {12, 17}, {12, 25}, {13, 15}, {13, 18}, {13, 24}, {13, 27},
    {15, 16}, {15, 19}, {15, 22}, {15, 25}, {16, 21}, {16, 24}, {17, 19},
    {17, 20}, {17, 21}, {18, 19}, {19, 5},  {19, 9},  {20, 5},  {20, 7},
    {20, 8},  {20, 9},  {21, 5},  {21, 7},  {22, 5},  {22, 6},  {22, 9},
    {23, 5},  {24, 5},  {24, 4},  {24, 7},  {25, 5},  {25, 3},  {25, 6},
    {26, 5},  {26, 2},  {26, 3},  {27, 1},  {27, 2},  {27, 4},  {27, 5}};

void or_interactions(int32_t out[126])
{
    for (int r = 0; r < 63; ++r) {
        out[2 * r] = OR_TABLE[r][0];
        out[2 * r + 1] = OR_TABLE[r][1];
    }
}

/* SRC/exaMD_partitioning.f90:786-798: cnt runs over kk, jj, ii = -1..1 with
 * ii fastest. */
void or_stencil_offset(int32_t n, int32_t off[3])
{
    int32_t m = n - 1;
    off[0] = m % 3 - 1;
    off[1] = (m / 3) % 3 - 1;
    off[2] = m / 9 - 1;
}

/* Input-order binning: the reference appends each atom to its block's
 * linked list while reading (SRC/exaMD_input_formats.f90:169-188) and copies
 * the list in order into grid%posn (SRC/exaMD_partitioning.f90:957-962).  A
 * counting sort that keeps input order inside a block gives the same order. */
int or_bin(const or_domain *d, int64_t n, const double *pos, int64_t *start,
           int64_t *order, int64_t *blockid)
{
    int64_t nbk = d->nblocks;
    memset(start, 0, (size_t)(nbk + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        int32_t bxyz[3];
        int64_t num;
        int rc = or_identify_block(d, &pos[3 * i], bxyz, &num);
        if (rc) return rc;
        blockid[i] = num - 1;
        start[num] += 1;
    }
    for (int64_t b = 0; b < nbk; ++b) start[b + 1] += start[b];
    int64_t *fill = (int64_t *)malloc((size_t)nbk * sizeof(int64_t));
    if (!fill) return 3;
    memcpy(fill, start, (size_t)nbk * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) order[fill[blockid[i]]++] = i;
    free(fill);
    return 0;
}

/* user_func.f90:40-75, plus the pair energy 4 eps (d12 - d6). */
int or_lennardjones_12(const or_lj *p, double dx, double dy, double dz,
                       double d2, double f12[3], double *e12)
{
    double rc2 = p->rc * p->rc;             /* prmvdw(1,5)**2 */
    if (d2 > rc2) return 1;                 /* retval = 1: skip */
    double eps24 = 24.0 * p->eps;
    double sig2 = p->sig * p->sig;
    double d2i = 1.0 / d2;
    double s = sig2 * d2i;
    double d6 = s * s * s;                  /* (sig2*d2i)**3 */
    double d12 = d6 * d6;
    double dg = -(eps24 * (2.0 * d12 - d6) * d2i);
    f12[0] = dg * dx;
    f12[1] = dg * dy;
    f12[2] = dg * dz;
    *e12 = 4.0 * p->eps * (d12 - d6);
    return 0;
}

/* A tabulated pair function at d2 (within the cutoff): the interval's
 * cubic of or_table, the result in the form of user_func. */
static int or_table_pair(const or_table *tab, double rc, double dx, double dy,
                         double dz, double d2, double f12[3], double *e12,
                         int64_t *below)
{
    if (d2 > rc * rc) return 1;
    double u = (d2 - tab->r2_lo) * tab->inv_h;
    if (u < 0.0) {
        *below += 1;
        u = 0.0;
    }
    int32_t k = (int32_t)u;
    if (k > tab->n - 2) k = tab->n - 2;
    double tt = u - (double)k;
    const double *cg = tab->gc + 4 * (int64_t)k, *cv = tab->vc + 4 * (int64_t)k;
    double g = ((cg[3] * tt + cg[2]) * tt + cg[1]) * tt + cg[0];
    f12[0] = g * dx;
    f12[1] = g * dy;
    f12[2] = g * dz;
    *e12 = ((cv[3] * tt + cv[2]) * tt + cv[1]) * tt + cv[0];
    return 0;
}

/* SRC/exaMD_forces.f90:492-551.  With species, the pair's parameters are
 * the row of the species pair (type[i1], type[i2]); with a table, the pair
 * function is the tabulated one. */
static inline void or_particle_interaction(const or_domain *d, const or_lj *p,
                                           const or_species *sp,
                                           const or_table *tab,
                                           const double *pos, double *forces,
                                           int64_t i1, int64_t i2,
                                           int64_t home, int64_t *home_counts,
                                           or_tally *t)
{
    if (sp)
        p = &sp->pair[sp->type[i1] * sp->ntypes + sp->type[i2]];
    const double *x1 = &pos[3 * i1], *x2 = &pos[3 * i2];
    double dx = x2[0] - x1[0];
    double dy = x2[1] - x1[1];
    double dz = x2[2] - x1[2];
    double btx = round(dx * d->boxi[0]);    /* Nint: half away from zero */
    double bty = round(dy * d->boxi[1]);
    double btz = round(dz * d->boxi[2]);
    dx = dx - btx * d->box[0];
    dy = dy - bty * d->box[1];
    dz = dz - btz * d->box[2];
    double d2 = dx * dx + dy * dy + dz * dz;
    t->cand += 1;
    if (fabs(d2) < 0.001) t->zero += 1;     /* 'ZERO at' guard */
    double f12[3], e12;
    int rv = tab ? or_table_pair(tab, p->rc, dx, dy, dz, d2, f12, &e12, &t->below)
                 : or_lennardjones_12(p, dx, dy, dz, d2, f12, &e12);
    if (rv == 0) {
        double *f1 = &forces[3 * i1], *f2 = &forces[3 * i2];
        f1[0] = f1[0] + f12[0];
        f1[1] = f1[1] + f12[1];
        f1[2] = f1[2] + f12[2];
        f2[0] = f2[0] - f12[0];
        f2[1] = f2[1] - f12[1];
        f2[2] = f2[2] - f12[2];
        t->cut += 1;
        t->pe += e12;
        if (home_counts) home_counts[home] += 1;
    }
}

/* SRC/exaMD_forces.f90:308-390 on one rank: every block is interior
 * (SRC/exaMD_partitioning.f90:735-746) and interior_blks is filled with
 * k outermost, i fastest (:763-781). */
/* Checks of the arguments shared by the serial and threaded forms. */
static int forces_check(const or_domain *d, const or_lj *p,
                        const or_species *sp, int64_t n)
{
    if (sp) {
        if (sp->ntypes < 1) return 2;
        for (int32_t a = 0; a < sp->ntypes * sp->ntypes; ++a)
            if (sp->pair[a].rc > p->rc) return 2;
        for (int64_t i = 0; i < n; ++i)
            if (sp->type[i] < 0 || sp->type[i] >= sp->ntypes) return 2;
    }
    for (int a = 0; a < 3; ++a)
        if (d->nb[a] < 5 || !(p->rc < 0.5 * d->box[a])) return 4;
    return 0;
}

/* SRC/exaMD_forces.f90:308-390 for the home blocks of the z layers
 * [k0, k1) (1-based), in the reference's order: k outermost, i fastest
 * (the interior_blks order of SRC/exaMD_partitioning.f90:763-781), the
 * block's own triangle (:349-356), then table rows 2..63 (:359-378) over
 * the periodic neighbour references (neighrefs, :786-798). */
static void home_layers(const or_domain *d, const or_lj *p,
                        const or_species *sp, const or_table *tab,
                        const double *pos, double *forces,
                        int64_t *home_counts, or_tally *t,
                        const int64_t *start, const int64_t *order,
                        int32_t k0, int32_t k1)
{
    const int32_t nx = d->nb[0], ny = d->nb[1], nz = d->nb[2];
    int32_t offs[27][3];
    for (int s = 0; s < 27; ++s) or_stencil_offset(s + 1, offs[s]);

    for (int32_t k = k0; k < k1; ++k)
        for (int32_t j = 1; j <= ny; ++j)
            for (int32_t i = 1; i <= nx; ++i) {
                int64_t home = (int64_t)(i - 1) + (int64_t)(j - 1) * nx
                             + (int64_t)(k - 1) * nx * ny;
                /* neighrefs(:, cnt), periodic Mod (:788-790) */
                int64_t ref[27];
                for (int s = 0; s < 27; ++s) {
                    int32_t ii = (i + offs[s][0] + nx - 1) % nx;
                    int32_t jj = (j + offs[s][1] + ny - 1) % ny;
                    int32_t kk = (k + offs[s][2] + nz - 1) % nz;
                    ref[s] = (int64_t)ii + (int64_t)jj * nx
                           + (int64_t)kk * nx * ny;
                }
                /* self triangle (:349-356) */
                int64_t b0 = start[home], b1 = start[home + 1];
                for (int64_t a1 = b0; a1 < b1 - 1; ++a1)
                    for (int64_t a2 = a1 + 1; a2 < b1; ++a2)
                        or_particle_interaction(d, p, sp, tab, pos, forces, order[a1],
                                                order[a2], home, home_counts, t);
                /* pairs 2..63 (:359-378) */
                for (int r = 1; r < 63; ++r) {
                    int64_t g1 = ref[OR_TABLE[r][0] - 1];
                    int64_t g2 = ref[OR_TABLE[r][1] - 1];
                    int64_t s1 = start[g1], e1 = start[g1 + 1];
                    int64_t s2 = start[g2], e2 = start[g2 + 1];
                    for (int64_t a1 = s1; a1 < e1; ++a1)
                        for (int64_t a2 = s2; a2 < e2; ++a2)
                            or_particle_interaction(d, p, sp, tab, pos, forces,
                                                    order[a1], order[a2], home,
                                                    home_counts, t);
                }
            }
}

/* One z-slab of home layers per task of a phase of the threaded form. */
typedef struct slab_task {
    const or_domain *d;
    const or_lj *p;
    const or_species *sp;
    const or_table *tab;
    const double *pos;
    double *forces;
    int64_t *home_counts;
    const int64_t *start, *order;
    int32_t k0, k1;
    or_tally t;
} slab_task;

static void *slab_run(void *arg)
{
    slab_task *s = (slab_task *)arg;
    memset(&s->t, 0, sizeof(s->t));
    home_layers(s->d, s->p, s->sp, s->tab, s->pos, s->forces, s->home_counts,
                &s->t, s->start, s->order, s->k0, s->k1);
    return NULL;
}

/* The evaluation over nthreads threads.  The z layers are cut into an even
 * number of slabs at least two layers thick, and the slabs are run in two
 * phases (even, then odd slabs), every slab of a phase on its own thread.
 * A home block touches atoms of its own and the adjacent layers only, so
 * two slabs of a phase never touch the same atom and no force is written
 * by two threads at once.  The pairs and the arithmetic of each are the
 * serial form's; an atom near a slab face receives its sums in another
 * order (FP64 differences of a few ulps). */
static int forces_midpoint(const or_domain *d, const or_lj *p,
                           const or_species *sp, const or_table *tab, int64_t n,
                           const double *pos, double *forces,
                           int64_t *home_counts, or_tally *t, int nthreads)
{
    memset(t, 0, sizeof(*t));
    int rc = forces_check(d, p, sp, n);
    if (rc) return rc;
    int64_t nbk = d->nblocks;
    int64_t *start = (int64_t *)malloc((size_t)(nbk + 1) * sizeof(int64_t));
    int64_t *order = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t *bid = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!start || !order || !bid) {
        free(start); free(order); free(bid);
        return 3;
    }
    rc = or_bin(d, n, pos, start, order, bid);
    if (rc) {
        free(start); free(order); free(bid);
        return rc;
    }
    const int32_t nz = d->nb[2];
    int nslab = 2 * (nthreads < nz / 4 ? nthreads : nz / 4);
    if (nthreads <= 1 || nslab < 2) {
        home_layers(d, p, sp, tab, pos, forces, home_counts, t, start, order,
                    1, nz + 1);
    } else {
        slab_task *task = (slab_task *)calloc((size_t)nslab, sizeof(slab_task));
        pthread_t *th = (pthread_t *)calloc((size_t)nslab, sizeof(pthread_t));
        if (!task || !th) {
            free(task); free(th); free(start); free(order); free(bid);
            return 3;
        }
        for (int s = 0; s < nslab; ++s) {
            slab_task *k = &task[s];
            k->d = d; k->p = p; k->sp = sp; k->tab = tab; k->pos = pos;
            k->forces = forces; k->home_counts = home_counts;
            k->start = start; k->order = order;
            k->k0 = 1 + (int32_t)((int64_t)s * nz / nslab);
            k->k1 = 1 + (int32_t)((int64_t)(s + 1) * nz / nslab);
        }
        for (int phase = 0; phase < 2; ++phase) {
            for (int s = phase; s < nslab; s += 2)
                if (pthread_create(&th[s], NULL, slab_run, &task[s]) != 0)
                    slab_run(&task[s]), th[s] = 0;
            for (int s = phase; s < nslab; s += 2)
                if (th[s]) pthread_join(th[s], NULL);
        }
        for (int s = 0; s < nslab; ++s) {
            t->cand += task[s].t.cand;
            t->cut += task[s].t.cut;
            t->zero += task[s].t.zero;
            t->pe += task[s].t.pe;
            t->below += task[s].t.below;
        }
        free(task);
        free(th);
    }
    free(start);
    free(order);
    free(bid);
    return 0;
}

int or_forces_midpoint_all(const or_domain *d, const or_lj *p, int64_t n,
                           const double *pos, double *forces,
                           int64_t *home_counts, or_tally *t)
{
    return forces_midpoint(d, p, NULL, NULL, n, pos, forces, home_counts, t, 1);
}

int or_forces_midpoint_species(const or_domain *d, const or_lj *p,
                               const or_species *sp, int64_t n,
                               const double *pos, double *forces,
                               int64_t *home_counts, or_tally *t)
{
    return forces_midpoint(d, p, sp, NULL, n, pos, forces, home_counts, t, 1);
}

int or_forces_midpoint_table(const or_domain *d, const or_lj *p,
                             const or_table *tab, int64_t n,
                             const double *pos, double *forces,
                             int64_t *home_counts, or_tally *t)
{
    if (!tab || tab->n < 2) return 2;
    return forces_midpoint(d, p, NULL, tab, n, pos, forces, home_counts, t, 1);
}

double or_wrap(double x, double L)
{
    if (x < 0.0) x = x + L;
    if (x >= L) x = x - L;
    return x;
}

/* The velocity-Verlet sweeps over atoms [i0, i1): the opening half kick
 * and drift with the wrap (mode 0), or the closing half kick with the sum
 * of v^2 (mode 1). */
typedef struct vv_task {
    const or_domain *d;
    int64_t i0, i1;
    double h, dt;
    double *pos, *vel, *forces;
    int mode;
    double v2;
} vv_task;

static void *vv_run(void *arg)
{
    vv_task *k = (vv_task *)arg;
    k->v2 = 0.0;
    for (int64_t i = 3 * k->i0; i < 3 * k->i1; ++i) {
        int a = (int)(i % 3);
        if (k->mode == 0) {
            k->vel[i] = k->vel[i] + k->h * k->forces[i];
            k->pos[i] = k->pos[i] + k->dt * k->vel[i];
            k->pos[i] = or_wrap(k->pos[i], k->d->box[a]);
            k->forces[i] = 0.0;
        } else {
            k->vel[i] = k->vel[i] + k->h * k->forces[i];
            k->v2 = k->v2 + k->vel[i] * k->vel[i];
        }
    }
    return NULL;
}

/* One sweep on up to 64 threads; returns the sum of v^2 in thread order. */
static double vv_sweep(const or_domain *d, int64_t n, double h, double dt,
                       double *pos, double *vel, double *forces, int mode,
                       int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    vv_task k[64];
    pthread_t th[64];
    for (int t = 0; t < nthreads; ++t) {
        k[t] = (vv_task){d, n * t / nthreads, n * (t + 1) / nthreads, h, dt,
                         pos, vel, forces, mode, 0.0};
        th[t] = 0;
        if (t > 0 && pthread_create(&th[t], NULL, vv_run, &k[t]) != 0) th[t] = 0;
        if (t > 0 && !th[t]) vv_run(&k[t]);
    }
    vv_run(&k[0]);
    double v2 = 0.0;
    for (int t = 0; t < nthreads; ++t) {
        if (th[t]) pthread_join(th[t], NULL);
        v2 = v2 + k[t].v2;
    }
    return v2;
}

static int run_vv_impl(const or_domain *d, const or_lj *p, const or_species *sp,
                       int64_t n, double mass, double dt, int64_t nsteps,
                       double *pos, double *vel, double *forces, double *ke_out,
                       double *pe_out, int64_t *cut_out, int nthreads)
{
    double h = (0.5 * dt) / mass;
    for (int64_t s = 0; s < nsteps; ++s) {
        vv_sweep(d, n, h, dt, pos, vel, forces, 0, nthreads);
        or_tally t;
        int rc = forces_midpoint(d, p, sp, NULL, n, pos, forces, NULL, &t,
                                 nthreads);
        if (rc) return rc;
        double ke = 0.5 * mass * vv_sweep(d, n, h, dt, pos, vel, forces, 1,
                                          nthreads);
        if (ke_out) ke_out[s] = ke;
        if (pe_out) pe_out[s] = t.pe;
        if (cut_out) cut_out[s] = t.cut;
    }
    return 0;
}

int or_run_vv(const or_domain *d, const or_lj *p, int64_t n, double mass,
              double dt, int64_t nsteps, double *pos, double *vel,
              double *forces, double *ke_out, double *pe_out, int64_t *cut_out)
{
    return run_vv_impl(d, p, NULL, n, mass, dt, nsteps, pos, vel, forces,
                       ke_out, pe_out, cut_out, 1);
}

int or_run_vv_species(const or_domain *d, const or_lj *p, const or_species *sp,
                      int64_t n, double mass, double dt, int64_t nsteps,
                      double *pos, double *vel, double *forces, double *ke_out,
                      double *pe_out, int64_t *cut_out)
{
    return run_vv_impl(d, p, sp, n, mass, dt, nsteps, pos, vel, forces,
                       ke_out, pe_out, cut_out, 1);
}

int or_forces_midpoint_mt(const or_domain *d, const or_lj *p, int64_t n,
                          const double *pos, double *forces,
                          int64_t *home_counts, or_tally *t, int32_t nthreads)
{
    return forces_midpoint(d, p, NULL, NULL, n, pos, forces, home_counts, t,
                           nthreads);
}

int or_run_vv_mt(const or_domain *d, const or_lj *p, int64_t n, double mass,
                 double dt, int64_t nsteps, double *pos, double *vel,
                 double *forces, double *ke_out, double *pe_out,
                 int64_t *cut_out, int32_t nthreads)
{
    return run_vv_impl(d, p, NULL, n, mass, dt, nsteps, pos, vel, forces,
                       ke_out, pe_out, cut_out, nthreads);
}

/* SRC/exaMD_input_formats.f90:215-236: list-directed N, box, then x y z. */
int or_write_particles(const char *path, int64_t n, const double box[3],
                       const double *pos)
{
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fprintf(f, " %lld\n", (long long)n);
    fprintf(f, " %.17g %.17g %.17g\n", box[0], box[1], box[2]);
    for (int64_t i = 0; i < n; ++i)
        fprintf(f, " %.17g %.17g %.17g\n", pos[3 * i], pos[3 * i + 1],
                pos[3 * i + 2]);
    return fclose(f) ? 1 : 0;
}

/* SRC/exaMD_input_formats.f90:112-124. */
int or_read_params(const char *path, int64_t *n, double box[3])
{
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    long long nn;
    int ok = fscanf(f, "%lld", &nn) == 1 &&
             fscanf(f, "%lf %lf %lf", &box[0], &box[1], &box[2]) == 3;
    fclose(f);
    if (!ok) return 1;
    *n = nn;
    return 0;
}

/* SRC/exaMD_input_formats.f90:144-198 (positions only; binning is or_bin). */
int or_read_particles(const char *path, int64_t n, double *pos)
{
    FILE *f = fopen(path, "r");
    if (!f) return 1;
    long long nn;
    double box[3];
    int ok = fscanf(f, "%lld", &nn) == 1 &&
             fscanf(f, "%lf %lf %lf", &box[0], &box[1], &box[2]) == 3 &&
             nn >= n;
    for (int64_t i = 0; ok && i < n; ++i)
        ok = fscanf(f, "%lf %lf %lf", &pos[3 * i], &pos[3 * i + 1],
                    &pos[3 * i + 2]) == 3;
    fclose(f);
    return ok ? 0 : 1;
}
