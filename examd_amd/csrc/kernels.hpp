// kernels.hpp -- host-side launchers of the gfx950 kernels (kernels.hip).
#pragma once

#include "common.hpp"

namespace examd {

// Device state of one system.  Per-atom arrays are SoA in input order; the
// block-sorted order exists only as the index array `sorted`, which gives
// every block 2^bin_shift slots from b << bin_shift: block b's atoms are
// its valid slots (atom ids; -1 marks a free slot) and count[b] their
// number.  The lists persist from step to step and only the atoms that
// change block are moved: a leaving atom frees its slot, an arriving one
// takes the first free slot (FP64 lists are then closed up and put in
// input order).
struct DevState {
  int64_t n = 0;
  int64_t nblocks = 0;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *fx = nullptr, *fy = nullptr, *fz = nullptr;          // FP64 force
  unsigned long long *qx = nullptr, *qy = nullptr, *qz = nullptr;  // 2^-32
  int *cell = nullptr;     // 0-based block id per atom
  int *slot = nullptr;     // the atom's entry in `sorted` (-1: to insert)
  int *count = nullptr;    // atoms per block (nblocks)
  int *sorted = nullptr;   // atom ids by block (nblocks << bin_shift)
  unsigned char *bflag = nullptr;  // blocks whose list received atoms
  int bin_shift = 3;       // log2 of the slots per block
  StepLog *log = nullptr;  // per-step tallies
  unsigned long long *home = nullptr;  // home-block pair counts (diagnostic)
  Partial *part = nullptr; // per-workgroup partial tallies
  unsigned long long *prof = nullptr;  // phase timestamps (diagnostic)
  // several species (ntypes > 0): species of each atom and the species-pair
  // table, row t_i * ntypes + t_j, in both precisions
  unsigned char *type = nullptr;
  float4 *pair32 = nullptr;   // sig^2, 24 eps 2^32, 4 eps 2^32, rc^2
  Pair64 *pair64 = nullptr;
  long long *pair_shift = nullptr;  // V(rc) of each row, 2^-32 fixed point
  int ntypes = 0;
  double max_eps_sig12 = 0.0; // largest eps sig^12 of the table
  // tabulated pair function (tab_n > 0): tab_n - 1 intervals uniform in r^2
  // from tab_lo; per interval the cubics c0 + c1 t + c2 t^2 + c3 t^3 (t in
  // [0, 1] across the interval) of g, with f12 = g * d, and of V: FP64 as
  // (g c0..c3, V c0..c3), the mixed copy as two float4 (g, then V) scaled
  // by 2^32
  float4 *tab32 = nullptr;
  double *tab64 = nullptr;
  int tab_n = 0;
  double tab_lo = 0.0, tab_inv_h = 0.0, tab_v_last = 0.0;
  double tab_d2fix = 0.0;   // r^2 below which a pair force may reach 2^14
  float tab_lof = 0.f, tab_inv_hf = 0.f;
};

constexpr int kProfSlots = 5;   // timestamps per cell block
constexpr int kSysSlab = 256;   // atoms per entry of SysTable::slab_sys

// Partial slots needed by a launch sequence over n atoms and G tiles.
int64_t partial_slots(int64_t n, int64_t tiles);

// The systems of a launch (device tables).  A single system is the case
// nsys = 1, where tile_sys is not read.
struct SysTable {
  const SysDesc *sys = nullptr;        // nsys entries, tile map of one precision
  const int *tile_sys = nullptr;       // system of each tile (ntiles)
  const int *tile_seg = nullptr;       // first tile of each system (nsys + 1)
  // the force launch's tiles split into one contiguous range per XCD
  // (kXcds + 1 boundaries) and the workgroups' fetch counters (kXcds, then
  // a finish counter; all zero between launches); nullptr: the static
  // XCD-aware order
  const int *xcd_start = nullptr;
  int *xcd_next = nullptr;
  const long long *atom_seg = nullptr; // first atom of each system (nsys + 1)
  // system of the first atom of each slab of kSysSlab atoms (batches)
  const int *slab_sys = nullptr;
  const int *cuts = nullptr;           // cell-block boundaries of balanced maps
  int nsys = 1;
  int ntiles = 0;
};

struct LJParams {
  double eps, sig, rc, mass;
};

// Tunables of the force kernel.
struct ForceConfig {
  int threads = 256;  // workgroup size: 256, 512 or 1024
  int cap = 0;        // LDS atom capacity per cell block
  int maxrb = 0;      // largest region (home + shell) in blocks
  int groups = 1;     // run groups per atom (work items = groups x atoms)
  int lane_queue = 0; // entries per lane queue of deferred pairs (0..16);
                      // an entry holds the hits of one candidate chunk
  int workgroups = 0; // persistent grid: workgroups that loop over the tiles
                      // (0 = one workgroup per tile)
};

size_t force_lds_bytes(bool fp64, int cap, int maxrb, int threads,
                       int qdepth, bool typed);

// Binning / integration modes of launch_step_bin.
enum BinMode { kBinOnly = 0, kKickDriftBin = 1, kFinishKickDriftBin = 2 };

void launch_fill_i32(int *p, int v, int64_t n, hipStream_t s);
void launch_fill_log(StepLog *p, int n, hipStream_t s);
void launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s);

// ke_log (single system only): kinetic energy of the finished step.
// kBinOnly starts the block lists afresh (counts and flags cleared, every
// atom to be inserted by launch_bin_update); the other modes move the atoms
// that change block from their old block's list to their new one's.
void launch_step_bin(const DevState &d, const SysTable &t, bool mixed,
                     int mode, double h, double dt, StepLog *ke_log,
                     StepLog *flag_log, hipStream_t s);
// ke_log (single system only): kinetic energy after the kick.
void launch_kick(const DevState &d, bool mixed, double h, StepLog *ke_log,
                 hipStream_t s);
// Kinetic energy of every system into log[sys].ke (sum of v^2).
void launch_sys_ke(const DevState &d, const SysTable &t, StepLog *log,
                   hipStream_t s);
// Bytes of DevState::bflag for nblocks blocks (whole ints, so that it can
// be cleared by launch_fill_i32).
constexpr int64_t bin_flag_bytes(int64_t nblocks) { return (nblocks + 3) / 4 * 4; }
// The block lists after a binning pass: after a kBinOnly pass (fresh) every
// atom put into the first free slot of its block's list (kFlagBinOverflow
// into flag_log when a block has none), and with block_order each changed
// list closed up and put in input order.
void launch_bin_update(const DevState &d, bool fresh, bool block_order,
                       StepLog *flag_log, hipStream_t s);
// Tallies of system k go to log[k].
void launch_forces(const DevState &d, const SysTable &t, const LJParams &p,
                   const ForceConfig &c, bool mixed, bool energy,
                   bool homecount, StepLog *log, hipStream_t s);
// x y z per atom (the host interchange layout) to and from three arrays.
void launch_aos_to_soa(const double *aos, double *x, double *y, double *z,
                       int64_t n, hipStream_t s);
void launch_soa_to_aos(const double *x, const double *y, const double *z,
                       double *aos, int64_t n, hipStream_t s);
void launch_forces_to_fp64(const DevState &d, double *out_aos, bool mixed,
                           hipStream_t s);
// A system descriptor written by a kernel from its argument.
void launch_store_sys(const SysDesc &v, SysDesc *dst, hipStream_t s);

}  // namespace examd
