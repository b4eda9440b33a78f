/*
 * examd_oracle.h -- CPU restatement of exaMD's midpoint-method Lennard-Jones
 * force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the parity tests, the
 * smoke check and the CPU-baseline leg of bench.py compare the HIP library
 * against.  Nothing in the product (examd_amd/) links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain C, FP64, with the
 * reference's own loop order and arithmetic order.  Citations are relative to
 * the exaMD v2.0 source tree.
 */
#ifndef EXAMD_ORACLE_H
#define EXAMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Domain of rc/2 midpoint blocks (SRC/exaMD_utils.f90:167-190). */
typedef struct or_domain {
    double box[3];      /* globalbox */
    double boxi[3];     /* globalboxi = 1/globalbox */
    double bl;          /* block_length = rc/2 (exaMD.f90:56) */
    int32_t nb[3];      /* nblocks_xyz = Int(globalbox/block_length) */
    int64_t nblocks;    /* Product(nblocks_xyz) */
} or_domain;

/* LJ parameters as in prmvdw(1,1:5) = (eps, sig, 0, 0, rc) (exaMD.f90:202). */
typedef struct or_lj {
    double eps, sig, rc;
} or_lj;

/* Several species: atom i has species type[i] in [0, ntypes), and the pair
 * (i, j) uses pair[type[i] * ntypes + type[j]] (eps, sig, rc; a symmetric
 * table, each rc at most the handle-wide cutoff that sizes the blocks).
 * The reference declares these tables for its user function
 * (user_func.f90:16-18: kktype, ltpvdw, prmvdw) but evaluates the first row
 * only (prmvdw(1,:), user_func.f90:55-61); this is lennardjones_12 with the
 * pair's own row. */
typedef struct or_species {
    int32_t ntypes;
    const int32_t *type;
    const or_lj *pair;
} or_species;

/* A tabulated pair function: n points uniform in r^2 from r2_lo to the
 * cutoff squared (inv_h = (n - 1) / (rc^2 - r2_lo)), g the factor of the
 * pair force f12 = g * (dx, dy, dz) and V the pair energy, as the
 * reference's user function returns them (user_func.f90:40-75).  Between
 * points: u = (d2 - r2_lo) * inv_h, k = min(floor(u), n - 2), t = u - k,
 * g = ((gc[4k+3] t + gc[4k+2]) t + gc[4k+1]) t + gc[4k] (V likewise from
 * vc): per interval the cubic Hermite interpolant of the two points' values
 * and slopes (the caller builds the coefficients; linear below five
 * points).  A pair closer than r2_lo is counted in `below` and uses the
 * first point. */
typedef struct or_table {
    int32_t n;
    double r2_lo, inv_h;
    const double *gc, *vc;   /* 4 (n - 1) coefficients each */
} or_table;

/* Per-evaluation tallies.  cand = candidate pairs visited by the midpoint
 * stencil, cut = pairs with d2 <= rc^2, zero = pairs flagged by the
 * "ZERO at" guard (|d2| < 0.001, SRC/exaMD_forces.f90:525-527), pe = sum
 * of 4 eps (s^12 - s^6) over in-cutoff pairs (unshifted; not in the
 * reference, whose e12 is always zero). */
typedef struct or_tally {
    int64_t cand, cut, zero;
    double pe;
    int64_t below;      /* tabulated function: pairs closer than its range */
} or_tally;

int  or_domain_create(const double box[3], double block_length, or_domain *d);

/* exaMD_identify_block (SRC/exaMD_module.f90:101-125).  blockxyz and
 * blocknum are 1-based as in the reference.  Returns 0, or 109 (the
 * reference's MPI_Abort code) when the block number is out of range. */
int  or_identify_block(const or_domain *d, const double pos[3],
                       int32_t blockxyz[3], int64_t *blocknum);

/* The 63-row midpoint interaction table (SRC/exaMD_module.f90:276-346):
 * out[2*r+0], out[2*r+1] = 1-based stencil indices (1..27). */
void or_interactions(int32_t out[126]);

/* Stencil index n (1..27) -> offset (ii fastest), as neighrefs is filled
 * in SRC/exaMD_partitioning.f90:786-798. */
void or_stencil_offset(int32_t n, int32_t off[3]);

/* Bin atoms into blocks in input order (linked list of
 * SRC/exaMD_input_formats.f90:169-188, scatter of
 * SRC/exaMD_partitioning.f90:932-968).  start has nblocks+1 entries
 * (0-based block ids), order receives atom indices grouped by block.
 * Returns 0 or 109. */
int  or_bin(const or_domain *d, int64_t n, const double *pos,
            int64_t *start, int64_t *order, int64_t *blockid);

/* exaMD_forces_midpoint_all at one rank (SRC/exaMD_forces.f90:308-390 with
 * SRC/exaMD_forces.f90:492-551 and user_func.f90:40-75).  pos and forces are
 * (3,n) column-major, indexed by input order; forces is ACCUMULATED (+=) as
 * in the reference.  home_counts (nblocks entries, may be NULL) receives the
 * number of in-cutoff pairs computed by each home block (0-based id).
 * Returns 0, 109 (atom out of box) or 4 (fewer than 5 blocks per side, or
 * rc >= L/2: the table would alias). */
int  or_forces_midpoint_all(const or_domain *d, const or_lj *p, int64_t n,
                            const double *pos, double *forces,
                            int64_t *home_counts, or_tally *t);

/* or_forces_midpoint_all with several species (p->rc sizes the blocks;
 * every pair uses its species row).  Returns 2 for a bad table or type. */
int  or_forces_midpoint_species(const or_domain *d, const or_lj *p,
                                const or_species *sp, int64_t n,
                                const double *pos, double *forces,
                                int64_t *home_counts, or_tally *t);

/* or_forces_midpoint_all with a tabulated pair function (p->rc sizes the
 * blocks and is the cutoff). */
int  or_forces_midpoint_table(const or_domain *d, const or_lj *p,
                              const or_table *tab, int64_t n,
                              const double *pos, double *forces,
                              int64_t *home_counts, or_tally *t);

/* LJ pair of user_func.f90:40-75 plus the pair energy.  Returns retval
 * (0 = apply, 1 = beyond cutoff). */
int  or_lennardjones_12(const or_lj *p, double dx, double dy, double dz,
                        double d2, double f12[3], double *e12);

/* Velocity Verlet with rebinning every step (not in the reference; the
 * build's MD loop).  State is (3,n) column-major.  On entry forces must
 * hold F(x); on exit x, v, forces are at the new time.  Energies of every
 * step (ke, pe, cut) are written to ke_out/pe_out/cut_out when non-NULL
 * (nsteps entries each).  h = dt/(2 m).  Returns 0 or an error code. */
int  or_run_vv(const or_domain *d, const or_lj *p, int64_t n, double mass,
               double dt, int64_t nsteps, double *pos, double *vel,
               double *forces, double *ke_out, double *pe_out,
               int64_t *cut_out);

/* The same evaluation and MD loop on nthreads threads (the CPU baseline
 * on all cores).  The z layers of home blocks are cut into an even number
 * of slabs, at least two layers each, run in two phases (even slabs, then
 * odd) so that no two threads write one atom's force; each slab keeps the
 * serial loop order, so the pairs and their arithmetic are the serial
 * form's and only the order of an atom's sums near slab faces differs. */
int  or_forces_midpoint_mt(const or_domain *d, const or_lj *p, int64_t n,
                           const double *pos, double *forces,
                           int64_t *home_counts, or_tally *t,
                           int32_t nthreads);
int  or_run_vv_mt(const or_domain *d, const or_lj *p, int64_t n, double mass,
                  double dt, int64_t nsteps, double *pos, double *vel,
                  double *forces, double *ke_out, double *pe_out,
                  int64_t *cut_out, int32_t nthreads);

int  or_run_vv_species(const or_domain *d, const or_lj *p,
                       const or_species *sp, int64_t n, double mass, double dt,
                       int64_t nsteps, double *pos, double *vel,
                       double *forces, double *ke_out, double *pe_out,
                       int64_t *cut_out);

/* Wrap one coordinate into [0, L): the build's own rule, shared with the
 * device integrator. */
double or_wrap(double x, double L);

/* Particle file exaMD_particles.dat (SRC/exaMD_input_formats.f90:112-124,
 * 144-198, 215-236): N, then box(3), then N lines of x y z, list-directed. */
int  or_write_particles(const char *path, int64_t n, const double box[3],
                        const double *pos);
int  or_read_params(const char *path, int64_t *n, double box[3]);
int  or_read_particles(const char *path, int64_t n, double *pos);

#ifdef __cplusplus
}
#endif
#endif
