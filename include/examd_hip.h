/*
 * examd_hip.h -- C ABI of the MI355X (gfx950) midpoint-method Lennard-Jones
 * force path.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns
 * EXAMD_OK (0) or a status code, and never aborts the process.  Host arrays
 * are borrowed for the duration of the call.  Position, velocity and force
 * arrays use the reference's Fortran layout posn(1:3, 1:n): x, y, z of atom
 * 1, then of atom 2, ... (column-major (3,n)), indexed in input order.
 *
 * Reference interfaces each entry point stands in for are cited as
 * file:line of exaMD v2.0.  A handle is not thread-safe; each handle owns
 * one HIP stream on one device.
 */
#ifndef EXAMD_HIP_H
#define EXAMD_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The numbers 101-109 are the reference's MPI_Abort codes
 * where the condition is the same one. */
enum {
    EXAMD_OK = 0,
    EXAMD_ERR_ARG = 1,          /* bad argument or size */
    EXAMD_ERR_HIP = 2,          /* HIP runtime error */
    EXAMD_ERR_NOMEM = 3,        /* allocation failed (ref. 103-108) */
    EXAMD_ERR_GRID = 4,         /* < 5 blocks per side or rc >= L/2: the
                                   63-row table would alias */
    EXAMD_ERR_TILE_OVERFLOW = 5,/* a cell block's region exceeded its LDS
                                   capacity */
    EXAMD_ERR_NO_DEVICE = 6,    /* no HIP device */
    EXAMD_ERR_NONFINITE = 7,    /* NaN/Inf force or position */
    EXAMD_ERR_NOT_READY = 8,    /* state not set */
    EXAMD_ERR_FIXED_RANGE = 9,  /* mixed path: a pair closer than r_fix,
                                   48 eps sig^12 / r_fix^13 = 2^14 (0.638
                                   sig at eps = 1), below which an atom's
                                   fixed-point force sum could leave its
                                   51-bit range */
    EXAMD_ERR_ZERO_DISTANCE = 10,/* a pair closer than sqrt(0.001): the
                                   reference prints 'ZERO at' */
    EXAMD_ERR_TABLE_RANGE = 11, /* a pair closer than the range of the
                                   tabulated pair function */
    EXAMD_ERR_STATIC_LIMIT = 12,/* examd_hip_forces_static: atoms, blocks
                                   or cell blocks beyond its static
                                   scratch limits */
    EXAMD_ERR_BIN_OVERFLOW = 13,/* a block held more atoms than its slots
                                   (the handle's entry points enlarge the
                                   slots and evaluate again; the
                                   allocation-free form returns this) */
    EXAMD_ERR_OUT_OF_BOX = 109  /* block number out of range (ref. 109) */
};

/* Arithmetic of one force evaluation. */
enum {
    EXAMD_PREC_FP64 = 0,  /* FP64 throughout, reference arithmetic order */
    EXAMD_PREC_MIXED = 1  /* FP32 pair arithmetic, int64 fixed-point (2^-32)
                             force and energy accumulation, FP64 integration */
};

typedef struct examd_handle examd_handle;

/* System parameters.  ε = prmvdw(1,1), σ = prmvdw(1,2), rc = prmvdw(1,5)
 * (exaMD.f90:202); block_length = rc/2 (exaMD.f90:56). */
typedef struct examd_params {
    int64_t natoms;
    double box[3];        /* globalbox (exaMD.f90:53) */
    double cutoff;        /* rc */
    double epsilon;
    double sigma;
    double mass;
    int32_t tiles[3];     /* cell blocks per dimension (workgroup tiles);
                             0 = choose */
    int32_t device;       /* HIP device ordinal */
    int32_t block_threads;/* force-kernel workgroup size (0 = choose;
                             256, 512 or 1024) */
    int32_t cap_slack_pct;/* LDS atom capacity of a cell block above the
                             largest region at set-up, in percent (0 = 15) */
    int32_t run_groups;   /* work items per region atom in the pair loop
                             (1-4; 0 = choose) */
    int32_t graph_chunk;  /* MD steps per captured hipGraph chunk (0 = launch
                             kernels one by one) */
    int32_t reserved[2];
} examd_params;

/* Tallies of one step or one force evaluation. */
typedef struct examd_stats {
    int64_t step;         /* step index (0 = the state given) */
    double ke;            /* 1/2 m sum v^2 (FP64) */
    double pe;            /* sum 4 eps (s^12 - s^6) over in-cutoff pairs */
    double pe_shifted;    /* pe - pairs_cut * V(rc); with several species,
                             pe minus the V(rc) of each in-cutoff pair's
                             own row (tallied in 2^-32 fixed point) */
    double etot;          /* ke + pe */
    int64_t pairs_cand;   /* candidate pairs visited by the midpoint stencil */
    int64_t pairs_cut;    /* pairs with r^2 <= rc^2 */
} examd_stats;

/* Block grid of exaMD_domain_create (SRC/exaMD_utils.f90:167-190):
 * nblocks_xyz = Int(box/block_length).  Pure host computation. */
int examd_hip_domain(const double box[3], double block_length,
                     int32_t nblocks_xyz[3], int64_t *nblocks);

/* Set up a handle (device buffers, stream, cell-block map).  Replaces
 * exaMD_Init + exaMD_domain_create + exaMD_partition_domain +
 * exaMD_setup_partitioned_grid + exaMD_setup_midpoint_interactions
 * (SRC/exaMD_utils.f90:38, :167; SRC/exaMD_partitioning.f90:65, :670;
 * SRC/exaMD_module.f90:276). */
int examd_hip_create(const examd_params *params, examd_handle **out);

/* One system of a batch. */
typedef struct examd_system {
    int64_t natoms;
    double box[3];
} examd_system;

/* Batching form (no reference counterpart: the reference runs one system
 * per MPI job): nsys independent periodic LJ systems of uneven sizes in one
 * handle.  Every kernel of a step covers all systems in one launch (the
 * force kernel walks one list of cell blocks across the systems), and
 * examd_hip_run captures multi-step chunks of the whole batch in one
 * hipGraph.  common supplies cutoff, epsilon, sigma, mass, device,
 * block_threads, cap_slack_pct, run_groups and graph_chunk for every system;
 * common->natoms, box and tiles must be 0.  The other entry points take the
 * batch's atoms concatenated in system order (n = total atoms), return one
 * examd_stats per system (forces: stats[nsys]; run: nsys entries per
 * reported step, system order), and number blocks per system
 * (block_ids: 1-based within the atom's system; home_pair_counts: the
 * systems' blocks concatenated). */
int examd_hip_batch_create(const examd_params *common,
                           const examd_system *systems, int32_t nsys,
                           examd_handle **out);

/* Systems, total atoms and total blocks of a handle (1 system for
 * examd_hip_create). */
int examd_hip_batch_info(examd_handle *h, int32_t *nsys, int64_t *natoms,
                         int64_t *nblocks);

/* Release everything (SRC/exaMD_utils.f90:90 exaMD_Finalize). */
int examd_hip_destroy(examd_handle *h);

/* Upload positions (and velocities when vel != NULL; zero otherwise).
 * Replaces the rank-0 scatter of exaMD_partition_work
 * (SRC/exaMD_partitioning.f90:889-1121).  Positions must lie in the box as
 * exaMD_identify_block accepts them. */
int examd_hip_set_state(examd_handle *h, const double *pos, const double *vel,
                        int64_t n);

/* Several species (no counterpart in the reference's evaluation, which reads
 * only the first row of its parameter table, prmvdw(1,:), user_func.f90:
 * 55-61; the table form follows the declarations of user_func.f90:16-18,
 * kktype / ltpvdw / prmvdw).  Atom i has species type[i] in [0, ntypes); the
 * pair (i, j) is the Lennard-Jones pair with eps, sigma and cutoff at index
 * type[i] * ntypes + type[j] of the ntypes x ntypes row-major tables, which
 * must be symmetric, with 0 < cutoff <= the handle's cutoff (which sizes the
 * blocks) and sigma > 0.  n is the handle's atom count (all systems of a
 * batch).  ntypes = 0 returns to one species (the handle's epsilon, sigma
 * and cutoff).  The species and the table hold for every later evaluation
 * and run of both precisions.  At most EXAMD_MAX_SPECIES species. */
#define EXAMD_MAX_SPECIES 8
#define EXAMD_MAX_TABLE_POINTS 65536
int examd_hip_set_species(examd_handle *h, int32_t ntypes,
                          const int32_t *type, int64_t n, const double *eps,
                          const double *sigma, const double *cutoff);

/* A tabulated pair function in place of the Lennard-Jones pair: the
 * general user function of exaMD_forces_midpoint_all
 * (SRC/exaMD_forces.f90:532, user_func.f90:40-75), which a GPU kernel cannot
 * call, sampled by the host.  g[k] and v[k] are the force factor (the pair
 * force is f12 = g * (dx, dy, dz), as user_func returns it) and the pair
 * energy at r^2 = r2_lo + k h, k = 0 .. npoints - 1, h = (cutoff^2 - r2_lo) /
 * (npoints - 1), with the handle's cutoff.  Between points both are cubic
 * in r^2: on each interval the Hermite cubic through the two points with
 * slopes from fourth-order differences of the samples, exact for cubics in
 * r^2 (linear below five points); a pair closer than sqrt(r2_lo) uses the
 * first point and the call returns EXAMD_ERR_TABLE_RANGE.  A user function's retval = 1 (skip) is
 * g = v = 0.  2 <= npoints <= EXAMD_MAX_TABLE_POINTS, 0 <= r2_lo <
 * cutoff^2; npoints = 0 returns to the Lennard-Jones pair.  Not combined
 * with several species.  The energy shift of pe_shifted is v[npoints - 1]. */
int examd_hip_set_pair_table(examd_handle *h, int32_t npoints, double r2_lo,
                             const double *g, const double *v);

/* Download the current state; any pointer may be NULL.  force is the last
 * evaluated force (FP64 or fixed point converted to FP64). */
int examd_hip_get_state(examd_handle *h, double *pos, double *vel,
                        double *force, int64_t n);

/* One force evaluation at the current positions:
 * exaMD_forces_midpoint_all(user_func = lennardjones_12, cutoff, noprocs)
 * (SRC/exaMD_forces.f90:308-390, :492-551; user_func.f90:40-75), with the
 * halo exchange of exaMD_start_comms_posn / exaMD_wait_comms_forces
 * (SRC/exaMD_comms.f90:44, :330) done as an on-device ghost gather and
 * ghost-force return.  forces (3,n) receives the force (overwritten, not
 * accumulated; may be NULL); stats may be NULL. */
/* Limits of examd_hip_forces_static and the size of its static device
 * array (one per GPU, held until the process ends). */
#define EXAMD_STATIC_MAX_ATOMS 65536
#define EXAMD_STATIC_MAX_BLOCKS 131072
#define EXAMD_STATIC_MAX_CELL_BLOCKS 4096
#define EXAMD_STATIC_SCRATCH_BYTES 12911104

/* No counterpart in the reference.  One force evaluation that allocates
 * nothing, for callers that cannot allocate device memory (code that must
 * not call hipMalloc: serving loops, harnesses that forbid it).  It is
 * examd_hip_forces without a handle: the same kernels, with every buffer
 * beyond the inputs and outputs in one static device array of
 * EXAMD_STATIC_SCRATCH_BYTES per GPU, whose address is looked up once per
 * device on the first call there.
 *
 * params: natoms, box, cutoff, epsilon, sigma, device, tiles (0 = the
 * default map, split evenly), block_threads, run_groups; the others are
 * ignored.  pos: DEVICE array (3, natoms) of positions, forces: DEVICE
 * array (3, natoms) that receives the forces (overwritten); stats (host,
 * may be NULL): pair counts and energy, ke = 0.  Runs on the null stream
 * and returns when the forces are written.  Returns EXAMD_ERR_STATIC_LIMIT
 * beyond EXAMD_STATIC_MAX_ATOMS atoms, EXAMD_STATIC_MAX_BLOCKS blocks or
 * EXAMD_STATIC_MAX_CELL_BLOCKS cell blocks, EXAMD_ERR_TILE_OVERFLOW when a
 * cell block's region does not fit the LDS (use more tiles),
 * EXAMD_ERR_BIN_OVERFLOW when a block holds more atoms than its list
 * slots (16 per block at EXAMD_STATIC_MAX_BLOCKS blocks, up to 256 with
 * fewer blocks), and the status flags of examd_hip_forces otherwise.  Every call on one GPU uses
 * that GPU's single array: calls that overlap on one GPU (several host
 * threads) race and must be serialised by the caller. */
int examd_hip_forces_static(const examd_params *params, int precision,
                            const double *pos, double *forces,
                            examd_stats *stats);

/* Chunks of examd_hip_run calls on this handle that ran again from their
 * start because a cell-block region outgrew the LDS capacity (no reference
 * counterpart: the reference's blocks live in host memory). */
int examd_hip_regrow_count(examd_handle *h, int64_t *regrows);

int examd_hip_forces(examd_handle *h, int precision, double *forces,
                     examd_stats *stats);

/* Velocity-Verlet MD for nsteps steps of dt with rebinning every step (no
 * reference counterpart: the reference evaluates forces once).  Energies
 * are tallied every energy_every steps (0 = only the last step) and written
 * to out[0..*n_out) when out != NULL (capacity out_cap). */
int examd_hip_run(examd_handle *h, int precision, int64_t nsteps, double dt,
                  int64_t energy_every, examd_stats *out, int64_t out_cap,
                  int64_t *n_out);

/* Steps per hipGraph chunk of examd_hip_run (0 = no graphs).  A chunk's
 * kernel sequence is captured once and replayed; the chunk length is
 * rounded to a multiple of energy_every. */
int examd_hip_set_graph_chunk(examd_handle *h, int64_t chunk);

/* Cell assignment of the current positions as exaMD_identify_block gives it
 * (SRC/exaMD_module.f90:101-125): 1-based block number per atom, and the
 * 1-based (i,j,k) when blockxyz (3,n) is not NULL. */
int examd_hip_block_ids(examd_handle *h, int64_t *blocknum, int32_t *blockxyz,
                        int64_t n);

/* Pair ownership: the number of in-cutoff pairs each home block computed in
 * a force evaluation at the current positions (0-based block id, nblocks
 * entries), as the 63-row table of SRC/exaMD_module.f90:276-346 assigns them
 * through SRC/exaMD_forces.f90:349-378. */
int examd_hip_home_pair_counts(examd_handle *h, int precision,
                               int64_t *counts, int64_t nblocks);

/* Launch shape of the force kernel for one precision: the cell-block map
 * (tiles per dimension: the METIS partition of
 * SRC/exaMD_partitioning.f90:65-455 becomes this map of cell blocks to
 * workgroups), the workgroup size and the run groups per atom. */
typedef struct examd_tune {
    int32_t tiles[3];     /* cell blocks per dimension (single system; of
                             system 0 in a batch) */
    int32_t block_threads;
    int32_t run_groups;
    int32_t lane_queue;   /* entries of the per-lane queues of deferred
                             in-cutoff pairs (an entry holds the hits among
                             one chunk of candidates, sixteen in the mixed
                             path, four in FP64): 0 (evaluate at once) to
                             16; the autotuner tries 0, 4 and 8 */
    int32_t tile_edge;    /* set_config: when tiles are 0, every system's map
                             is T = max(1, round(nblocks_xyz / tile_edge));
                             in a batch each map is then coarsened while its
                             largest region fits the LDS budget of one
                             workgroup (the whole LDS for 1024 threads, a
                             proportional share for fewer), in atoms and in
                             region blocks;
                             get_config: the edge the map came from (0 when
                             tiles were given) */
    int32_t workgroups;   /* persistent force grid: workgroups that loop over
                             the cell blocks (0 = one workgroup per block) */
    int32_t balance;      /* 1: cell-block boundaries along each axis at equal
                             shares of the atoms (a load-balanced map for
                             uneven density, set from the current positions);
                             0: even split of the blocks */
    int32_t rebalance_every; /* examd_hip_run re-partitions the map of this
                             precision from the current block counts once at
                             least this many steps have run since the last
                             time, at the next captured-chunk boundary (every
                             this many steps when no chunks are captured):
                             new balanced boundaries (balance = 1) and a new
                             LDS capacity, the old map kept when the new one
                             would not fit; 0 = never.  The autotuner keeps
                             the value it finds. */
    double force_ms;      /* mean force-kernel time measured by autotune */
} examd_tune;

/* Autotuner: time `trials` force evaluations of the current state for each
 * candidate launch shape, time the three fastest again in turn for six
 * rounds of `trials`, keep the fastest for `precision` and report it in best
 * (may be NULL).  The state is unchanged; forces must be recomputed. */
int examd_hip_autotune(examd_handle *h, int precision, int trials,
                       examd_tune *best);

/* Read or apply the launch shape of one precision (apply: e.g. a shape the
 * autotuner found earlier for the same system). */
int examd_hip_get_config(examd_handle *h, int precision, examd_tune *out);
int examd_hip_set_config(examd_handle *h, int precision,
                         const examd_tune *cfg);

/* Cell-block boundaries of one system's map in one precision, in blocks:
 * Tx + 1 values along x, then Ty + 1 along y, then Tz + 1 along z (first 0,
 * last nblocks_xyz of the axis), written to cuts[0 .. cap).  *n_out = the
 * number of values; EXAMD_ERR_ARG when cap is smaller.  The boundaries of
 * METIS's partition of the blocks over ranks in the reference
 * (SRC/exaMD_partitioning.f90:65-455) restated as a rectilinear map. */
int examd_hip_get_cuts(examd_handle *h, int precision, int32_t system,
                       int32_t *cuts, int32_t cap, int32_t *n_out);

/* Force-kernel timing over the calls since the last reset: enable, read
 * (total milliseconds of force-kernel launches measured with HIP events on
 * the handle's stream, and the number of launches), reset.  enable: 0 off,
 * 1 every force launch, k > 1 in examd_hip_run only the force launch of every
 * k-th step (steps k/2, k/2 + k, ... of each captured chunk or eager piece),
 * since each pair of event records costs the stream a few microseconds. */
int examd_hip_timing(examd_handle *h, int enable, double *force_ms,
                     int64_t *force_launches, int reset);

/* Diagnostic: one force evaluation with per-cell-block timestamps.
 * phase_cycles[0..3] receive the mean shader-clock cycles per cell block of
 * (0) region scan, (1) ghost gather into LDS, (2) pair evaluation, (3)
 * force return; phase_cycles[4] the mean total per cell block. */
int examd_hip_phase_profile(examd_handle *h, int precision,
                            double phase_cycles[5]);

/* Seeded initial states (not in the reference, whose generator uses an
 * unseeded RNG, SRC/exaMD_input_formats.f90:41-96).
 * FCC lattice of nc[0]*nc[1]*nc[2] cells of 4 atoms at number density rho,
 * atom order: cell x fastest, then y, then z, basis order (0,0,0),
 * (1/2,1/2,0), (1/2,0,1/2), (0,1/2,1/2); perturbed by uniform(-p, p) when
 * perturb > 0; box written to box[3].  Maxwell-Boltzmann velocities at
 * temperature T (k_B = 1), zero net momentum, rescaled so that
 * KE = (3n-3)/2 T exactly.  pos and vel are (3,n). */
int examd_hip_generate_fcc(const int32_t nc[3], double rho, double temperature,
                           double mass, double perturb, uint64_t seed,
                           double *pos, double *vel, double box[3]);

/* n atoms at number density rho in a cubic periodic box (box[0..2] = L =
 * (n/rho)^(1/3)): n sites drawn without replacement (seeded) from the
 * smallest FCC lattice of nc^3 cells with at least n sites, perturbed by
 * uniform(-p, p); velocities as examd_hip_generate_fcc.  The nearest sites
 * are L/(nc sqrt 2) apart, so no pair starts closer than that minus
 * 2 sqrt(3) p. */
int examd_hip_generate_vacancy_fcc(int64_t n, double rho, double temperature,
                                   double mass, double perturb, uint64_t seed,
                                   double *pos, double *vel, double box[3]);

/* Liquid-vapour slab in a periodic box: atoms uniform (seeded) in the slab
 * slab_lo <= x < slab_hi at density rho_slab and in the rest of the box at
 * rho_vapour, no pair closer than min_sep (random sequential addition with
 * up to 10^4 attempts per atom), slab atoms first; velocities as
 * examd_hip_generate_fcc.  The counts are round(rho * volume) per region;
 * *n receives the total.  With pos == NULL only *n is computed; otherwise
 * cap must be at least *n. */
int examd_hip_generate_slab(const double box[3], double slab_lo,
                            double slab_hi, double rho_slab,
                            double rho_vapour, double min_sep,
                            double temperature, double mass, uint64_t seed,
                            double *pos, double *vel, int64_t cap,
                            int64_t *n);

/* Human-readable text of a status code. */
const char *examd_hip_strerror(int status);

#ifdef __cplusplus
}
#endif
#endif
