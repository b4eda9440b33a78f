// common.hpp -- shared definitions of the gfx950 midpoint LJ library.
//
// The midpoint pair table is the reference's interactions(1:2,1:63)
// (SRC/exaMD_module.f90:282-344) over the 27-entry stencil numbered as
// neighrefs is filled (SRC/exaMD_partitioning.f90:786-798, ii fastest).  The
// device code never walks it row by row: at compile time it is regrouped
// into 13 "runs" of block offsets d = o2 - o1 that share (dy, dz), so that the
// J blocks of one run are consecutive in x and their atoms form one
// contiguous range in a cell block's LDS image.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace examd {

// ---------------------------------------------------------------- table ----
struct Off3 { int x, y, z; };

// interactions(1:2, r), 1-based stencil indices (SRC/exaMD_module.f90:282-344)
constexpr int kTable[63][2] = {
    {14, 14}, {14, 15}, {14, 16}, {14, 17}, {14, 18}, {14, 19}, {14, 20},
    {14, 21}, {14, 22}, {14, 23}, {14, 24}, {14, 25}, {14, 26}, {14, 27},
    {10, 18}, {10, 27}, {11, 17}, {11, 18}, {11, 25}, {11, 26}, {11, 27},
    {12, 16}, {12, 17}, {12, 25}, {13, 15}, {13, 18}, {13, 24}, {13, 27},
    {15, 16}, {15, 19}, {15, 22}, {15, 25}, {16, 21}, {16, 24}, {17, 19},
    {17, 20}, {17, 21}, {18, 19}, {19, 5},  {19, 9},  {20, 5},  {20, 7},
    {20, 8},  {20, 9},  {21, 5},  {21, 7},  {22, 5},  {22, 6},  {22, 9},
    {23, 5},  {24, 5},  {24, 4},  {24, 7},  {25, 5},  {25, 3},  {25, 6},
    {26, 5},  {26, 2},  {26, 3},  {27, 1},  {27, 2},  {27, 4},  {27, 5}};

constexpr Off3 stencil_off(int n) {
  return Off3{(n - 1) % 3 - 1, ((n - 1) / 3) % 3 - 1, (n - 1) / 9 - 1};
}

// The 13 runs: (dy, dz) of J relative to I.  Run 0 is the self run
// (dy = dz = 0, dx = 0, 1, 2); dx = 0 is the block's own triangle.
constexpr int kNumRuns = 13;
constexpr int kRunDY[kNumRuns] = {0, 1, 2, -2, -1, 0, 1, 2, -2, -1, 0, 1, 2};
constexpr int kRunDZ[kNumRuns] = {0, 0, 0, 1, 1, 1, 1, 1, -2, -2, -2, -2, -2};

// Row of the table whose offset o2 - o1 equals d, or -1.
constexpr int row_of(int dx, int dy, int dz) {
  for (int r = 0; r < 63; ++r) {
    Off3 a = stencil_off(kTable[r][0]), b = stencil_off(kTable[r][1]);
    if (b.x - a.x == dx && b.y - a.y == dy && b.z - a.z == dz) return r;
  }
  return -1;
}

// 5-bit mask (bit k <-> dx = k - 2) of the dx in run `run` whose row has
// o1 component `axis` equal to o.  Ownership of a pair (a in I, b in J):
// the home block is I - o1, and it must lie in the cell block.
constexpr unsigned axis_mask(int run, int axis, int o) {
  unsigned m = 0;
  for (int k = 0; k < 5; ++k) {
    int r = row_of(k - 2, kRunDY[run], kRunDZ[run]);
    if (r < 0) continue;
    Off3 a = stencil_off(kTable[r][0]);
    int v = axis == 0 ? a.x : (axis == 1 ? a.y : a.z);
    if (v == o) m |= 1u << k;
  }
  return m;
}

// o1 of (run, dx) as a packed value, for the home-count diagnostic.
constexpr Off3 run_o1(int run, int k) {
  int r = row_of(k - 2, kRunDY[run], kRunDZ[run]);
  return r < 0 ? Off3{9, 9, 9} : stencil_off(kTable[r][0]);
}

constexpr bool table_runs_cover_all() {
  int n = 0;
  for (int run = 0; run < kNumRuns; ++run)
    for (int k = 0; k < 5; ++k)
      if (row_of(k - 2, kRunDY[run], kRunDZ[run]) >= 0) ++n;
  return n == 63;
}
static_assert(table_runs_cover_all(), "13 runs must cover the 63 rows");

// -------------------------------------------------------------- status ----
enum Flag : unsigned {
  kFlagOverflow = 1u,     // cell-block region larger than LDS capacity
  kFlagOutOfBox = 2u,     // block number out of range (reference 109)
  kFlagZero = 4u,         // pair with d2 < 0.001 ('ZERO at')
  kFlagFixedRange = 8u,   // pair force beyond the fixed-point range
  kFlagNonFinite = 16u,   // NaN/Inf position or force
  kFlagTableRange = 32u,  // pair closer than a tabulated function's range
  kFlagBinOverflow = 64u, // a block holds more atoms than its slots
};

// Per-step tallies written by device kernels.
struct StepLog {
  double ke;                 // sum v^2 (times m/2 on the host)
  double pe;                 // FP64 path pair energy
  unsigned long long pe_fix; // mixed path pair energy, 2^-32 fixed point
  unsigned long long cand;   // candidate pairs
  unsigned long long cut;    // in-cutoff pairs
  unsigned long long flags;  // Flag bits
  unsigned long long shift_fix;  // several species: sum of V(rc) of each
                                 // in-cutoff pair's species row, 2^-32
};

// Partial tallies of one workgroup, written with plain stores and summed in
// index order by a reduction kernel (no same-address atomics, and the sums
// are deterministic).
struct Partial {
  double ke;
  double pe;
  unsigned long long pe_fix;
  unsigned long long cand;
  unsigned long long cut;
  unsigned long long shift_fix;
};

constexpr double kFixScale = 4294967296.0;          // 2^32
constexpr double kFixInv = 2.3283064365386963e-10;   // 2^-32
constexpr double kFixMagic = 6755399441055744.0;     // 1.5 * 2^52

// Round f * 2^32 to the nearest integer (ties to even) as int64: one FMA
// into the 1.5*2^52 binade, then the integer is the low mantissa bits.
// Exact for |f| < 2^19; the sign is symmetric, so fix(-f) == -fix(f).
__device__ __forceinline__ long long fix32(float f) {
  double t = __builtin_fma((double)f, kFixScale, kFixMagic);
  return __double_as_longlong(t) - 0x4338000000000000LL;
}

// fix32 of a value already multiplied by 2^32 in FP32 (exact for |f| < 2^19:
// scaling by a power of two commutes with FP32 rounding), so the FP64 step
// is one add of the magic constant: fix32_scaled(f * 2^32) == fix32(f).
__device__ __forceinline__ long long fix32_scaled(float fs) {
  double t = (double)fs + kFixMagic;
  return __double_as_longlong(t) - 0x4338000000000000LL;
}

// Biased fixed point of the product f * x of an FP32 value x and a value f
// already multiplied by 2^32 (an FP32 value widened to FP64): the product
// of two FP32 values is exact in FP64, so one FMA into the 1.5*2^52 binade
// rounds it once, at 2^-32, and the magic constant's bits are left in (bits
// 51 and up of the constant; its bits below 51 are zero).  The low 51 bits
// of a sum of such values are the low 51 bits of the sum of the fixed-point
// values, whatever the count, so a sum of magnitude below 2^50 is recovered
// by fix_sext51.
__device__ __forceinline__ long long fix32_prod_raw(double f, float x) {
  return __double_as_longlong(__builtin_fma(f, (double)x, kFixMagic));
}
__device__ __forceinline__ long long fix_sext51(long long v) {
  return (long long)((unsigned long long)v << 13) >> 13;
}

// ------------------------------------------------------------- species ----
// Several species: at most kMaxSpecies, a pair (i, j) uses row
// type[i] * ntypes + type[j] of a symmetric table.
constexpr int kMaxSpecies = 8;
// Pair kinds of the force kernel: one Lennard-Jones row, a row per species
// pair, or a tabulated function of r^2.
enum : int { kPairLJ = 0, kPairSpecies = 1, kPairTable = 2 };
struct Pair64 {
  double sig2, eps24, eps4, rc2;
};

// ---------------------------------------------------------------- grid ----
struct Grid {
  int nb[3];        // nblocks_xyz
  int T[3];         // cell blocks (tiles) per dimension
  double L[3];      // box
  double bl;        // block length rc/2
};

// One independent periodic system of a launch.  Systems are concatenated:
// atoms [atom0, atom0 + n), global blocks [block0, block0 + nb^3), tiles
// [tile0, tile0 + T^3) of the tile map of one precision.
struct SysDesc {
  Grid g;
  long long atom0;
  int block0;
  int tile0;
  int cut[3];   // per axis: offset of the T + 1 cell-block boundaries in the
                // launch's boundary array, or -1 for the even split
};

// Home-block range [lo, hi) of tile t along one axis: nb blocks shared by T
// tiles as evenly as possible.
__host__ __device__ inline void tile_range(int t, int nb, int T, int &lo,
                                           int &hi) {
  lo = (int)(((long long)t * nb) / T);
  hi = (int)(((long long)(t + 1) * nb) / T);
}

// Spread consecutive tiles over the 8 XCDs in contiguous chunks: the
// hardware hands workgroup w to XCD w % 8, so workgroup w takes tile
// base(w % 8) + w / 8 and each XCD's L2 sees a compact slab of tiles.
constexpr int kXcds = 8;   // XCDs of an MI355X; workgroup w runs on XCD w % 8
__device__ inline int xcd_remap(int wg, int G) {
  const int nx = kXcds;
  int q = G / nx, r = G % nx;
  int x = wg % nx, k = wg / nx;
  int base = x * q + (x < r ? x : r);
  return base + k;
}

// exaMD_identify_block (SRC/exaMD_module.f90:109-113) on one axis: 1-based,
// IEEE division, add one, truncate, clamp to nb.
__device__ __forceinline__ int block_coord(double x, double bl, int nb) {
  double q = x / bl;
  q = q + 1.0;
  int b = (int)q;
  return b > nb ? nb : b;
}

__device__ __forceinline__ double wrap_coord(double x, double L) {
  if (x < 0.0) x = x + L;
  if (x >= L) x = x - L;
  return x;
}

}  // namespace examd
