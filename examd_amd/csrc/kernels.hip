// kernels.hip -- gfx950 kernels of the midpoint-method LJ path.
//
//   k_step_bin   velocity-Verlet kick/drift/wrap fused with cell assignment
//                (exaMD_identify_block, SRC/exaMD_module.f90:101-125) and
//                with the upkeep of the per-block atom lists kept from step
//                to step: an atom that changed block frees its slot and
//                takes the first free slot of its new block's list
//   k_bin_insert the lists built from scratch (after a kBinOnly pass);
//                (FP64 path) k_block_order closes up each changed list and
//                restores input order inside it
//                (SRC/exaMD_input_formats.f90:80-88)
//   k_forces     one workgroup per cell block: gathers the home blocks plus
//                the rc/2 ghost shell into LDS (the halo exchange of
//                SRC/exaMD_comms.f90:44-376 done on device), evaluates every
//                pair the 63-row table assigns to a home block of the cell
//                block (SRC/exaMD_forces.f90:308-390, :492-551;
//                user_func.f90:40-75) and returns ghost forces with atomics
//   k_kick       closing half kick + kinetic energy
#include <cmath>
#include <type_traits>

#include "kernels.hpp"

namespace examd {

namespace {

constexpr int kBinThreads = 256;
constexpr int kBinPer = 2;   // atoms per thread of the binning kernel
constexpr int kInsertPer = 4;  // atoms per thread of the list insertion

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ unsigned wave_or(unsigned v) {
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// Sum over the 64 lanes of a wave (every lane active), the same in every
// lane: an inclusive scan by shifts of 1, 2, 4 and 8 inside each row of 16
// lanes and row broadcasts of lanes 15 and 31, then lane 63's value.
__device__ __forceinline__ int wave_total(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ void add_partial(Partial &a, const Partial &b) {
  a.ke += b.ke;
  a.pe += b.pe;
  a.pe_fix += b.pe_fix;
  a.cand += b.cand;
  a.cut += b.cut;
  a.shift_fix += b.shift_fix;
}

// Workgroup sum of one Partial per thread, stored at out[slot] (all
// threads must call it; the order of the sums is fixed).
template <int NT>
__device__ void write_partial(Partial v, Partial *out, int slot) {
  __shared__ Partial s_p[NT / 64];
  v.ke = wave_sum(v.ke);
  v.pe = wave_sum(v.pe);
  v.pe_fix = wave_sum(v.pe_fix);
  v.cand = wave_sum(v.cand);
  v.cut = wave_sum(v.cut);
  v.shift_fix = wave_sum(v.shift_fix);
  if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    Partial t = s_p[0];
    for (int w = 1; w < NT / 64; ++w) add_partial(t, s_p[w]);
    out[slot] = t;
  }
}

enum : unsigned { kSumKE = 1u, kSumForce = 2u };

// Sum partials in index order into a step's log: one workgroup per
// segment k, partials [seg[k], seg[k+1]) into log[k] (seg == nullptr: one
// segment [0, n)).
constexpr int kReduceThreads = 256;
__global__ __launch_bounds__(kReduceThreads) void k_reduce_partials(
    const Partial *p, const int *seg, int n, StepLog *log, unsigned what) {
  const int k = blockIdx.x;
  const int i0 = seg ? seg[k] : 0, i1 = seg ? seg[k + 1] : n;
  Partial acc{0.0, 0.0, 0ull, 0ull, 0ull, 0ull};
  for (int i = i0 + (int)threadIdx.x; i < i1; i += kReduceThreads)
    add_partial(acc, p[i]);
  __shared__ Partial s_tot;
  write_partial<kReduceThreads>(acc, &s_tot, 0);
  __syncthreads();
  log += k;
  if (threadIdx.x == 0) {
    if (what & kSumKE) log->ke += s_tot.ke;
    if (what & kSumForce) {
      log->pe += s_tot.pe;
      log->pe_fix += s_tot.pe_fix;
      log->cand += s_tot.cand;
      log->cut += s_tot.cut;
      log->shift_fix += s_tot.shift_fix;
    }
  }
}

// ------------------------------------------------------------- fills ----
__global__ void k_fill_i32(int *p, int v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}
__global__ void k_fill_log(StepLog *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = StepLog{0.0, 0.0, 0ull, 0ull, 0ull, 0ull, 0ull};
}

// ---------------------------------------------------------- binning ----
// Atoms in input order, kBinPer per thread.  mode kFinishKickDriftBin closes
// the previous step (v += h F, kinetic energy) and opens the next one; all
// modes zero the force accumulators and assign the block.
// One atom of the binning pass: what it reads before any store (the
// force as loaded: FP64, or the fixed-point sums converted at the store,
// so that no load waits for another's conversion).
struct BinAtom {
  double x, y, z, vx, vy, vz;
  union { double f; long long q; } fx, fy, fz;
  int old, slot;
};

template <bool MIXED, int MODE>
__device__ __forceinline__ void bin_load(const DevState &d, int64_t i,
                                         BinAtom &a) {
  a.x = d.x[i];
  a.y = d.y[i];
  a.z = d.z[i];
  if constexpr (MODE != kBinOnly) {
    if constexpr (MIXED) {
      a.fx.q = (long long)d.qx[i];
      a.fy.q = (long long)d.qy[i];
      a.fz.q = (long long)d.qz[i];
    } else {
      a.fx.f = d.fx[i];
      a.fy.f = d.fy[i];
      a.fz.f = d.fz[i];
    }
    a.vx = d.vx[i];
    a.vy = d.vy[i];
    a.vz = d.vz[i];
    a.old = d.cell[i];
    a.slot = d.slot[i];
  }
}

// Integrate one loaded atom (mode permitting), store it, zero its force
// accumulators and return its global block id, assigned as
// exaMD_identify_block does (flags for a position the block map rejects);
// the caller stores the id where it changed.
template <bool MIXED, int MODE>
__device__ __forceinline__ int bin_store(const DevState &d, const Grid &g,
                                         int block0, double h, double dt,
                                         int64_t i, BinAtom &a, double &ke,
                                         unsigned &flags) {
  double X = a.x, Y = a.y, Z = a.z;
  if constexpr (MODE != kBinOnly) {
    const double Fx = MIXED ? (double)a.fx.q * kFixInv : a.fx.f;
    const double Fy = MIXED ? (double)a.fy.q * kFixInv : a.fy.f;
    const double Fz = MIXED ? (double)a.fz.q * kFixInv : a.fz.f;
    double Vx = a.vx, Vy = a.vy, Vz = a.vz;
    if constexpr (MODE == kFinishKickDriftBin) {
      Vx = Vx + h * Fx;
      Vy = Vy + h * Fy;
      Vz = Vz + h * Fz;
      ke += Vx * Vx + Vy * Vy + Vz * Vz;
    }
    Vx = Vx + h * Fx;
    Vy = Vy + h * Fy;
    Vz = Vz + h * Fz;
    X = X + dt * Vx;
    Y = Y + dt * Vy;
    Z = Z + dt * Vz;
    X = wrap_coord(X, g.L[0]);
    Y = wrap_coord(Y, g.L[1]);
    Z = wrap_coord(Z, g.L[2]);
    d.vx[i] = Vx;
    d.vy[i] = Vy;
    d.vz[i] = Vz;
    d.x[i] = X;
    d.y[i] = Y;
    d.z[i] = Z;
  }
  if constexpr (MIXED) {
    d.qx[i] = 0ull;
    d.qy[i] = 0ull;
    d.qz[i] = 0ull;
  } else {
    d.fx[i] = 0.0;
    d.fy[i] = 0.0;
    d.fz[i] = 0.0;
  }
  int cell = 0;
  if (!(__builtin_isfinite(X) && __builtin_isfinite(Y) &&
        __builtin_isfinite(Z)) ||
      !(X / g.bl < 2.0e9 && Y / g.bl < 2.0e9 && Z / g.bl < 2.0e9)) {
    flags |= kFlagNonFinite;
  } else {
    int bx = block_coord(X, g.bl, g.nb[0]);
    int by = block_coord(Y, g.bl, g.nb[1]);
    int bz = block_coord(Z, g.bl, g.nb[2]);
    long long num = (long long)bx + (long long)(by - 1) * g.nb[0] +
                    (long long)(bz - 1) * ((long long)g.nb[0] * g.nb[1]);
    long long nblk = (long long)g.nb[0] * g.nb[1] * g.nb[2];
    if (num < 1 || num > nblk) flags |= kFlagOutOfBox;
    else cell = block0 + (int)(num - 1);
  }
  return cell;
}

// Atom i into the first free slot of block c's list by an atomic
// compare-and-swap, so holes left by atoms that moved away are refilled and
// a list never needs closing up; the block's count rises by one and its
// list is flagged as changed.  The slots are read eight at a time as a
// hint: a slot read as free may have been taken since, and then the
// exchange fails and the search goes on.  Without a free slot the atom
// stays out of every list (slot -1) and false is returned.
__device__ __forceinline__ bool list_insert(const DevState &d, int64_t i, int c) {
  const int S = 1 << d.bin_shift;
  int *p = d.sorted + ((int64_t)c << d.bin_shift);
  int got = -1;
  for (int s0 = 0; s0 < S && got < 0; s0 += 8) {
    const int4 a = *reinterpret_cast<const int4 *>(p + s0);
    const int4 b = *reinterpret_cast<const int4 *>(p + s0 + 4);
    const int v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    for (int q = 0; q < 8 && got < 0; ++q)
      if (v[q] < 0 && atomicCAS(&p[s0 + q], -1, (int)i) == -1) got = s0 + q;
  }
  if (got < 0) {
    d.slot[i] = -1;
    return false;
  }
  d.slot[i] = (int)(((int64_t)c << d.bin_shift) + got);
  atomicAdd(&d.count[c], 1);
  d.bflag[c] = 1;
  return true;
}

// kBinPer atoms per thread (coalesced, kBinThreads apart).  Every atom's
// loads are issued before the first store, so they are in flight together
// (the stores could otherwise alias the next atom's loads as far as the
// compiler knows, and each atom would wait for the last).  One system
// (SINGLE) reads its grid once per thread.  kBinOnly starts the block
// lists afresh: every atom is inserted by k_bin_insert.  The other modes
// keep them: an atom that changes block frees its slot in its old block's
// list (the block's count drops by one) and takes a free slot in the new
// one's (list_insert); the others keep their slots.  Every change of a
// slot in this pass is an atomic operation (the free an exchange, the take
// a compare-and-swap), since other threads take slots of the same lists
// meanwhile.  A list whose last free slot is held by an atom that leaves
// in the same pass may look full to an arriving atom: that takes eight
// atoms in one block of rc/2, and it raises kFlagBinOverflow as a full
// list does.
template <bool MIXED, int MODE, bool SINGLE>
__global__ __launch_bounds__(kBinThreads) void k_step_bin(
    DevState d, SysTable t, double h, double dt, int want_ke,
    StepLog *flag_log) {
  const int64_t base =
      (int64_t)blockIdx.x * (kBinThreads * kBinPer) + threadIdx.x;
  // kinetic energy per slab of kBinThreads atoms, each slab's partial
  // written on its own: the partials and their sums are those of k_kick
  // (one atom per thread), so fused and separate closing kicks agree bit
  // for bit
  double ke[kBinPer];
  unsigned flags = 0;
  BinAtom at[kBinPer];
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    // past the last atom the thread reads the last one again (no branch
    // among the loads) and stores nothing
    const int64_t i = base + (int64_t)k * kBinThreads;
    bin_load<MIXED, MODE>(d, i < d.n ? i : d.n - 1, at[k]);
  }
  // nothing of the store phase is scheduled among the loads
  __builtin_amdgcn_sched_barrier(0);
  Grid g1;
  int b1 = 0;
  if constexpr (SINGLE) {
    g1 = t.sys[0].g;
    b1 = t.sys[0].block0;
  }
#pragma unroll
  for (int k = 0; k < kBinPer; ++k) {
    const int64_t i = base + (int64_t)k * kBinThreads;
    ke[k] = 0.0;
    if (i < d.n) {
      int c;
      if constexpr (SINGLE) {
        c = bin_store<MIXED, MODE>(d, g1, b1, h, dt, i, at[k], ke[k], flags);
      } else {
        // the system of the first atom of the wave's slab of kSysSlab atoms
        // from the slab table, then each lane steps forward (systems are
        // contiguous and hold hundreds of atoms or more: rarely more than
        // one step)
        const int64_t i0 = (int64_t)__builtin_amdgcn_readfirstlane((int)i);
        int sys = t.slab_sys[i0 / kSysSlab];
        while (sys + 1 < t.nsys && t.atom_seg[sys + 1] <= i) ++sys;
        // a wave inside one system reads its grid as one (scalar) copy
        const int s0 = __builtin_amdgcn_readfirstlane(sys);
        if (!__ballot(sys != s0)) {
          const SysDesc &sd = t.sys[s0];
          c = bin_store<MIXED, MODE>(d, sd.g, sd.block0, h, dt, i, at[k], ke[k], flags);
        } else {
          const SysDesc &sd = t.sys[sys];
          c = bin_store<MIXED, MODE>(d, sd.g, sd.block0, h, dt, i, at[k], ke[k], flags);
        }
      }
      if (MODE == kBinOnly) {
        d.cell[i] = c;
        d.slot[i] = -1;
      } else if (c != at[k].old) {
        d.cell[i] = c;
        // the atom leaves a hole in its old list (refilled by later
        // insertions) and the old block's count drops by one; the other
        // atoms keep their slots and their order
        if (at[k].slot >= 0) {
          atomicExch(&d.sorted[at[k].slot], -1);
          atomicSub(&d.count[at[k].old], 1);
        }
        if (!list_insert(d, i, c)) flags |= kFlagBinOverflow;
      }
    }
  }
  if constexpr (MODE == kFinishKickDriftBin) {
    if (want_ke) {
#pragma unroll
      for (int k = 0; k < kBinPer; ++k) {
        // slabs past the last atom have no partial slot
        const int64_t slab = (int64_t)blockIdx.x * kBinPer + k;
        if (slab * kBinThreads >= d.n) break;
        // the previous slab's sum has been read out of write_partial's LDS
        if (k > 0) __syncthreads();
        write_partial<kBinThreads>(Partial{ke[k], 0.0, 0ull, 0ull, 0ull, 0ull},
                                   d.part, (int)slab);
      }
    }
  }
  flags = wave_or(flags);
  if ((threadIdx.x & 63) == 0 && flags)
    atomicOr(&flag_log->flags, (unsigned long long)flags);
}

// Every atom into its block's list after a kBinOnly pass, which leaves the
// lists empty and the counts zero: the atom takes the next slot of the list
// from an atomic increment of the block's count.  kInsertPer atoms per
// thread.  A block with more atoms than slots keeps the first ones in its
// list and raises kFlagBinOverflow (whose recovery gives every list twice
// the slots and bins again).
__global__ __launch_bounds__(kBinThreads) void k_bin_insert(DevState d,
                                                            StepLog *flag_log) {
  const int64_t base =
      (int64_t)blockIdx.x * (kBinThreads * kInsertPer) + threadIdx.x;
  const int S = 1 << d.bin_shift;
  unsigned flags = 0;
#pragma unroll
  for (int k = 0; k < kInsertPer; ++k) {
    const int64_t i = base + (int64_t)k * kBinThreads;
    if (i >= d.n) continue;
    const int c = d.cell[i];
    const int s = atomicAdd(&d.count[c], 1);
    if (s < S) {
      const int64_t e = ((int64_t)c << d.bin_shift) + s;
      d.sorted[e] = (int)i;
      d.slot[i] = (int)e;
      d.bflag[c] = 1;
    } else {
      d.slot[i] = -1;
      flags |= kFlagBinOverflow;
    }
  }
  flags = wave_or(flags);
  if ((threadIdx.x & 63) == 0 && flags)
    atomicOr(&flag_log->flags, (unsigned long long)flags);
}

template <bool MIXED>
__global__ __launch_bounds__(kBinThreads) void k_kick(DevState d, double h,
                                                      int want_ke) {
  const int64_t i = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
  double ke = 0.0;
  if (i < d.n) {
    double Fx, Fy, Fz;
    if constexpr (MIXED) {
      Fx = (double)(long long)d.qx[i] * kFixInv;
      Fy = (double)(long long)d.qy[i] * kFixInv;
      Fz = (double)(long long)d.qz[i] * kFixInv;
    } else {
      Fx = d.fx[i];
      Fy = d.fy[i];
      Fz = d.fz[i];
    }
    double Vx = d.vx[i] + h * Fx;
    double Vy = d.vy[i] + h * Fy;
    double Vz = d.vz[i] + h * Fz;
    d.vx[i] = Vx;
    d.vy[i] = Vy;
    d.vz[i] = Vz;
    ke = Vx * Vx + Vy * Vy + Vz * Vz;
  }
  if (want_ke)
    write_partial<kBinThreads>(Partial{ke, 0.0, 0ull, 0ull, 0ull, 0ull}, d.part,
                               blockIdx.x);
}

// Sum of v^2 over the atoms of system blockIdx.x into log[blockIdx.x].ke,
// in a fixed order.
__global__ __launch_bounds__(kBinThreads) void k_sys_ke(DevState d,
                                                        SysTable t,
                                                        StepLog *log) {
  const int k = blockIdx.x;
  const long long a0 = t.atom_seg[k], a1 = t.atom_seg[k + 1];
  double ke = 0.0;
  for (long long i = a0 + threadIdx.x; i < a1; i += kBinThreads) {
    const double Vx = d.vx[i], Vy = d.vy[i], Vz = d.vz[i];
    ke += Vx * Vx + Vy * Vy + Vz * Vz;
  }
  __shared__ Partial s_tot;
  write_partial<kBinThreads>(Partial{ke, 0.0, 0ull, 0ull, 0ull, 0ull}, &s_tot, 0);
  __syncthreads();
  if (threadIdx.x == 0) log[k].ke += s_tot.ke;
}

// ------------------------------------------------------------- scan ----
// Inclusive scan of one int per thread across the workgroup; returns the
// inclusive prefix and writes the workgroup total to *total.
template <int NT>
__device__ __forceinline__ int block_inclusive_scan(int v, int *s_wave,
                                                    int *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) s_wave[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int w = 0; w < NT / 64; ++w) {
      int t = s_wave[w];
      s_wave[w] = acc;
      acc += t;
    }
    s_wave[NT / 64] = acc;
  }
  __syncthreads();
  int r = v + s_wave[wave];
  *total = s_wave[NT / 64];
  __syncthreads();
  return r;
}

// Sorting network for eight values (Batcher's odd-even merge sort, 19
// compare-exchanges), ascending.
__device__ __forceinline__ void sort8(unsigned (&v)[8]) {
  auto cx = [&](int a, int b) {
    const unsigned lo = min(v[a], v[b]), hi = max(v[a], v[b]);
    v[a] = lo;
    v[b] = hi;
  };
  cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
  cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
  cx(1, 2); cx(5, 6);
  cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
  cx(2, 4); cx(3, 5);
  cx(1, 2); cx(3, 4); cx(5, 6);
}

// Sixteen values, ascending: both halves sorted by sort8, the second
// reversed behind the first (a bitonic sequence), then merged by four
// stages of eight compare-exchanges.
__device__ __forceinline__ void sort16(unsigned (&v)[16]) {
  unsigned a[8], b[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = v[k];
    b[k] = v[8 + k];
  }
  sort8(a);
  sort8(b);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = a[k];
    v[15 - k] = b[k];
  }
#pragma unroll
  for (int j = 8; j >= 1; j >>= 1)
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (!(i & j)) {
        const unsigned lo = min(v[i], v[i + j]), hi = max(v[i], v[i + j]);
        v[i] = lo;
        v[i + j] = hi;
      }
}

// Input order inside each block (the order the reference's linked list
// gives, SRC/exaMD_input_formats.f90:80-88): each list that received atoms
// in this step (flagged where an atom took a slot) closed up (holes removed,
// the tail left free) and put in input order, the slots updated and the
// flag cleared, by an insertion sort in place.  A list that only lost
// atoms keeps the rest in order around its holes.  Only the FP64 path needs
// the order (its sums depend on the order of the pairs), and only for lists
// longer than sixteen slots (after repeated overflows): the force kernel's
// gather sorts a list of 8 or 16 slots itself.  The mixed path's
// fixed-point sums do not depend on it, and its lists keep their holes
// until insertions refill them.
__global__ void k_block_order(DevState d) {
  const int64_t w4 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w4 * 4 >= d.nblocks) return;
  unsigned *fw = reinterpret_cast<unsigned *>(d.bflag) + w4;
  const unsigned f = *fw;
  if (!f) return;
  *fw = 0u;
  const int S = 1 << d.bin_shift;
  for (int q = 0; q < 4; ++q) {
    if (!((f >> (8 * q)) & 0xffu)) continue;
    const int64_t b = w4 * 4 + q;
    const int base = (int)(b << d.bin_shift);
    int *p = d.sorted + base;
    int c = 0, last = -1;
    for (int r = 0; r < S; ++r) {
      const int v = p[r];
      if (v < 0) continue;
      p[c++] = v;
      last = r;
    }
    for (int r = c; r <= last; ++r) p[r] = -1;
    for (int k = 1; k < c; ++k) {
      int v = p[k];
      int m = k - 1;
      while (m >= 0 && p[m] > v) {
        p[m + 1] = p[m];
        --m;
      }
      p[m + 1] = v;
    }
    for (int k = 0; k < c; ++k) d.slot[p[k]] = base + k;
  }
}

// ----------------------------------------------------------- forces ----
struct ForceArgs {
  DevState d;
  SysTable t;
  StepLog *log;
  int cap, maxrb, groups, lane_queue;
  double rc2, eps24, eps4, sig2, d2zero, d2fix;
  float rc2f, eps24f, eps4f, sig2f, d2zerof, d2fixf;
  float eps24fs, eps48fs, eps4fs;   // eps24f, 2 eps24f, eps4f times 2^32
};

__host__ __device__ inline size_t align_up(size_t v, size_t a) {
  return (v + a - 1) / a * a;
}

// LDS image of one cell block: region-block offsets, then atoms.  The atom
// arrays hold cap + kChunkPad entries: the pair loop reads up to sixteen
// candidates at a time and may look up to fifteen entries past the last
// atom (masked off).
constexpr int kWorkBins = 256;  // work-order classes (own_bits keys by work)
constexpr int kMixedChunk = 16; // mixed candidates per step, at most
constexpr int kDrainMin = 2;    // partial drain once the wave's lanes hold
                                // this many pending pairs each on average
constexpr int kChunkPad = 16;   // atom entries past the last one a chunk may read
struct LdsLayout {
  size_t off_gst, off_wave, off_pos, off_f, off_id, off_bc, off_perm, off_hist,
      off_q, off_pair, off_ty, total;
};
__host__ __device__ inline LdsLayout lds_layout(bool fp64, int cap, int maxrb,
                                                int nwaves, int qdepth,
                                                bool typed) {
  cap += kChunkPad;
  LdsLayout L;
  size_t o = (size_t)(maxrb + 1) * 4;  // s_off
  // the global starts of the region blocks are read only by the gather, the
  // lane queues only by the pair phase: they share space when the queues
  // are large enough
  const size_t qbytes = (size_t)nwaves * 64 * qdepth * 4;
  const bool gst_in_q = qbytes >= (size_t)maxrb * 4;
  L.off_gst = o;
  if (!gst_in_q) o += (size_t)maxrb * 4;
  L.off_wave = o;
  o += 32 * 4;
  o = align_up(o, 16);
  L.off_pos = o;
  o += (size_t)cap * (fp64 ? 24 : 16);
  o = align_up(o, 8);
  L.off_f = o;
  o += (size_t)cap * 24;
  L.off_id = o;
  o += (size_t)cap * 4;
  L.off_bc = o;
  o += (size_t)cap * 4;
  o = align_up(o, 4);
  L.off_perm = o;
  o += (size_t)cap * 2;
  o = align_up(o, 16);   // the class counts are scanned as int4
  L.off_hist = o;
  o += (kWorkBins + 1) * 4;
  o = align_up(o, 16);
  L.off_q = o;
  if (gst_in_q) L.off_gst = o;
  o += qbytes;
  // several species: the species-pair table, then one species byte per atom
  o = align_up(o, 16);
  L.off_pair = o;
  if (typed) o += (size_t)kMaxSpecies * kMaxSpecies * (sizeof(Pair64) + 8);
  L.off_ty = o;
  if (typed) o += (size_t)cap;
  L.total = align_up(o, 16);
  return L;
}

// Bits of the o1 values (-1, 0, +1 -> bit 0, 1, 2) for which the home
// coordinate i - o1 of a region coordinate i lies in the tile's [1, t].
__device__ __forceinline__ unsigned own_bits(int i, int t) {
  return ((i + 1 >= 1 && i + 1 <= t) ? 1u : 0u) | ((i >= 1 && i <= t) ? 2u : 0u) |
         ((i - 1 >= 1 && i - 1 <= t) ? 4u : 0u);
}

// Per-run ownership masks for runtime run indices: kRunAxisMask[r][axis][o+1].
struct RunAxisMasks {
  unsigned char m[kNumRuns][3][3];
  constexpr RunAxisMasks() : m() {
    for (int r = 0; r < kNumRuns; ++r)
      for (int a = 0; a < 3; ++a)
        for (int o = -1; o <= 1; ++o)
          m[r][a][o + 1] = (unsigned char)axis_mask(r, a, o);
  }
};
// Ownership mask of run r (bit k <-> dx = k - 2) for an I block whose
// per-axis own_bits are the fields of key = ox + 8 oy + 64 oz: per axis the
// union over the allowed o1, then the intersection of the axes.
constexpr unsigned run_mask_of(const RunAxisMasks &M, int key, int r) {
  const unsigned ox = key & 7, oy = (key >> 3) & 7, oz = key >> 6;
  unsigned mx = 0, my = 0, mz = 0;
  for (int o = -1; o <= 1; ++o) {
    if (ox & (1u << (o + 1))) mx |= M.m[r][0][o + 1];
    if (oy & (1u << (o + 1))) my |= M.m[r][1][o + 1];
    if (oz & (1u << (o + 1))) mz |= M.m[r][2][o + 1];
  }
  return mx & my & mz & 31u;
}

// Owned J blocks of an atom as a function of its own_bits key: the sum over
// the 13 runs of the popcount of the ownership mask (the work order's
// primary key, below).
struct OwnedWork {
  unsigned char v[512];
  constexpr OwnedWork() : v() {
    const RunAxisMasks M;
    for (int i = 0; i < 512; ++i) {
      int n = 0;
      for (int r = 0; r < kNumRuns; ++r) {
        const unsigned m = run_mask_of(M, i, r);
        for (int k = 0; k < 5; ++k) n += (m >> k) & 1u;
      }
      v[i] = (unsigned char)n;
    }
  }
};

// The ownership masks of the 13 runs for each own_bits key: byte r of the
// four words holds run r's mask, so a work item loads its masks once and
// each run extracts one byte.
struct RunMaskTable {
  unsigned w[512][4];
  constexpr RunMaskTable() : w() {
    const RunAxisMasks M;
    for (int i = 0; i < 512; ++i)
      for (int r = 0; r < kNumRuns; ++r)
        w[i][r >> 2] |= run_mask_of(M, i, r) << ((r & 3) * 8);
  }
};
__constant__ __align__(16) RunMaskTable kRunMask;

// Work-order class of each own_bits key: the keys that own at least one J
// block ranked by owned blocks (most first), then by key, so that atoms
// with the same ownership masks are adjacent; 255 for keys that own
// nothing.  Per axis only the own_bits 1, 2, 3, 4, 6 and 7 occur, so at
// most 216 classes are used.
struct KeyRank {
  unsigned char r[512];
  constexpr KeyRank() : r() {
    const OwnedWork W;
    int next = 0;
    for (int w = 63; w >= 1; --w)
      for (int k = 0; k < 512; ++k) {
        const unsigned ox = k & 7, oy = (k >> 3) & 7, oz = k >> 6;
        const bool ok = ox != 0 && ox != 5 && oy != 0 && oy != 5 && oz != 0 && oz != 5;
        if (ok && W.v[k] == w) r[k] = (unsigned char)next++;
      }
    for (int k = 0; k < 512; ++k) {
      const unsigned ox = k & 7, oy = (k >> 3) & 7, oz = k >> 6;
      const bool ok = ox != 0 && ox != 5 && oy != 0 && oy != 5 && oz != 0 && oz != 5;
      if (!ok || W.v[k] == 0) r[k] = 255;
    }
  }
};
__constant__ KeyRank kKeyRank;

// Run groups: the 13 runs split into 1..4 groups of similar candidate
// counts (run 0 covers 2.5 blocks, the others 5).
__constant__ int kGroupStart[5][5] = {{0, 13, 13, 13, 13},
                                      {0, 13, 13, 13, 13},
                                      {0, 7, 13, 13, 13},
                                      {0, 5, 9, 13, 13},
                                      {0, 4, 7, 10, 13}};

// o1 of the table row with offset d = o2 - o1 (d in [-2, 2]^3, table
// orientation), packed (o1x+1) | (o1y+1) << 2 | (o1z+1) << 4; for the
// home-block diagnostic.
struct O1Table {
  unsigned char v[125];
  constexpr O1Table() : v() {
    for (int i = 0; i < 125; ++i) {
      int dx = i % 5 - 2, dy = (i / 5) % 5 - 2, dz = i / 25 - 2;
      int r = row_of(dx, dy, dz);
      Off3 o = r < 0 ? Off3{0, 0, 0} : stencil_off(kTable[r][0]);
      v[i] = (unsigned char)((o.x + 1) | ((o.y + 1) << 2) | ((o.z + 1) << 4));
    }
  }
};
__constant__ O1Table kO1;

__device__ __forceinline__ void lds_add(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(long long *p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_sub(long long *p, long long v) {
  __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Packed region coordinates of an LDS atom: 8 bits per axis, plus the
// periodic image code (s + 1, s in {-1, 0, 1}) per axis in bits 24..29.
__device__ __forceinline__ int bc_x(int bc) { return bc & 255; }
__device__ __forceinline__ int bc_y(int bc) { return (bc >> 8) & 255; }
__device__ __forceinline__ int bc_z(int bc) { return (bc >> 16) & 255; }
__device__ __forceinline__ int bc_s(int bc, int axis) {
  return ((bc >> (24 + 2 * axis)) & 3) - 1;
}

// One cell block (tile) of one system, evaluated by the whole workgroup.
template <bool FP64, bool ENERGY, bool HOMECNT, int PK, int NT>
__device__ __forceinline__ void force_tile(const ForceArgs &A, const int tile,
                                           unsigned char *smem) {
  // pair kind: one Lennard-Jones row, a row per species pair, or a table
  constexpr bool TYPED = PK == kPairSpecies;
  constexpr bool TABLE = PK == kPairTable;
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x;
  const int sys = A.t.nsys > 1 ? A.t.tile_sys[tile] : 0;
  const SysDesc &sd = A.t.sys[sys];
  const Grid g = sd.g;
  const int block0 = sd.block0;
  StepLog *const log = A.log + sys;
  const int lt = tile - sd.tile0;
  const int tI = lt % g.T[0];
  const int tJ = (lt / g.T[0]) % g.T[1];
  const int tK = lt / (g.T[0] * g.T[1]);
  int h0x, h1x, h0y, h1y, h0z, h1z;
  // home-block range of the tile per axis: the even split, or the
  // boundaries of a balanced map
  auto bounds = [&](int axis, int t, int &lo, int &hi) {
    const int c = sd.cut[axis];
    if (c < 0) {
      tile_range(t, g.nb[axis], g.T[axis], lo, hi);
    } else {
      lo = A.t.cuts[c + t];
      hi = A.t.cuts[c + t + 1];
    }
  };
  bounds(0, tI, h0x, h1x);
  bounds(1, tJ, h0y, h1y);
  bounds(2, tK, h0z, h1z);
  const int tx = h1x - h0x, ty = h1y - h0y, tz = h1z - h0z;
  const int RX = tx + 2, RY = ty + 2, RZ = tz + 2;
  const int RXY = RX * RY, NRB = RXY * RZ;
  const int nbx = g.nb[0], nby = g.nb[1], nbz = g.nb[2];
  // the LDS image holds maxrb region blocks: a map that does not match the
  // launch's tables is refused here, before the region scan writes s_off
  if (NRB > A.maxrb) {
    if (tid == 0) {
      atomicOr(&log->flags, (unsigned long long)kFlagOverflow);
      if (ENERGY) A.d.part[tile] = Partial{0.0, 0.0, 0ull, 0ull, 0ull, 0ull};
    }
    return;
  }

  const LdsLayout LL = lds_layout(FP64, A.cap, A.maxrb, NW, A.lane_queue, TYPED);
  // several species: species of each region atom, species-pair rows
  unsigned char *s_ty = smem + LL.off_ty;
  const float4 *s_pr32 = reinterpret_cast<const float4 *>(smem + LL.off_pair);
  const Pair64 *s_pr64 = reinterpret_cast<const Pair64 *>(smem + LL.off_pair);
  const long long *s_vrc = reinterpret_cast<const long long *>(
      smem + LL.off_pair + kMaxSpecies * kMaxSpecies * sizeof(Pair64));
  int *s_off = reinterpret_cast<int *>(smem);
  int *s_gst = reinterpret_cast<int *>(smem + LL.off_gst);
  int *s_wave = reinterpret_cast<int *>(smem + LL.off_wave);
  int *s_id = reinterpret_cast<int *>(smem + LL.off_id);
  int *s_bc = reinterpret_cast<int *>(smem + LL.off_bc);
  // FP64 image: SoA doubles + packed coords; mixed image: float4 with the
  // packed coords in w
  const int cap4 = A.cap + kChunkPad;
  double *s_x = reinterpret_cast<double *>(smem + LL.off_pos);
  double *s_y = s_x + cap4;
  double *s_z = s_y + cap4;
  float4 *s_p4 = reinterpret_cast<float4 *>(smem + LL.off_pos);
  // force sums interleaved per atom (x, y, z), so that a pair's three LDS
  // atomics share one address: FP64 (s_f) or fixed point (s_q)
  double *s_f = reinterpret_cast<double *>(smem + LL.off_f);
  long long *s_q = reinterpret_cast<long long *>(smem + LL.off_f);
  unsigned long long *prof =
      A.d.prof ? A.d.prof + (size_t)tile * kProfSlots : nullptr;
  if (prof && tid == 0) prof[0] = clock64();

  // Region coordinates of region block rb: x fastest.  The divisions by the
  // region's extents are multiplications by ceil(2^32 / d) taken high,
  // exact for n * d < 2^32: the maps keep region extents at most 255 per
  // axis (the 8-bit packed coordinates), and NRB <= maxrb, the region blocks
  // whose offsets fit the LDS, was checked above.
  // (the constants by an FP64 division, exact to far below one for these
  // divisors, 3 <= d <= 255, so that the ceiling is right and below 2^32)
  const unsigned mRX = (unsigned)__builtin_ceil(0x1p32 / (double)RX);
  const unsigned mRY = (unsigned)__builtin_ceil(0x1p32 / (double)RY);
  auto region_coords = [&](int rb, int &cx, int &cy, int &cz) {
    const int q = (int)__umulhi((unsigned)rb, mRX);
    cx = rb - q * RX;
    cz = (int)__umulhi((unsigned)q, mRY);
    cy = q - cz * RY;
  };

  // 1. region blocks: home blocks of the tile plus a one-block ghost shell,
  //    periodic (SRC/exaMD_partitioning.f90:788-790).
  for (int rb = tid; rb < NRB; rb += NT) {
    int cx, cy, cz;
    region_coords(rb, cx, cy, cz);
    int gx = h0x - 1 + cx, gy = h0y - 1 + cy, gz = h0z - 1 + cz;
    gx += gx < 0 ? nbx : 0;
    gx -= gx >= nbx ? nbx : 0;
    gy += gy < 0 ? nby : 0;
    gy -= gy >= nby ? nby : 0;
    gz += gz < 0 ? nbz : 0;
    gz -= gz >= nbz ? nbz : 0;
    const int gb = block0 + gx + nbx * (gy + nby * gz);
    s_off[rb] = min(A.d.count[gb], 1 << A.d.bin_shift);
    s_gst[rb] = gb << A.d.bin_shift;
  }
  __syncthreads();
  {
    // exclusive scan of the region counts, in place
    int per = (NRB + NT - 1) / NT;
    int b0 = tid * per, b1 = min(NRB, b0 + per);
    int sum = 0;
    for (int i = b0; i < b1; ++i) sum += s_off[i];
    int total;
    int inc = block_inclusive_scan<NT>(sum, s_wave, &total);
    int run = inc - sum;
    for (int i = b0; i < b1; ++i) {
      int c = s_off[i];
      s_off[i] = run;
      run += c;
    }
    if (tid == 0) s_off[NRB] = total;
    __syncthreads();
  }
  const int R = s_off[NRB];
  if (prof && tid == 0) prof[1] = clock64();
  if (R > A.cap) {
    if (tid == 0) {
      atomicOr(&log->flags, (unsigned long long)kFlagOverflow);
      if (ENERGY) A.d.part[tile] = Partial{0.0, 0.0, 0ull, 0ull, 0ull, 0ull};
      if (prof) prof[2] = prof[3] = prof[4] = prof[1];
    }
    return;
  }

  // 2. ghost gather: atoms of the region into LDS.  The mixed image is
  //    relative to the tile centre and pre-shifted by the periodic image;
  //    the FP64 image keeps raw coordinates (the minimum image is applied
  //    per pair, as the reference does).  First one thread per region
  //    block copies the block's atom ids from its list (consecutive slots)
  //    with the block's packed region coordinates and image codes, then
  //    one thread per slot gathers the positions.
  const bool wlx = h0x == 0, whx = h1x == nbx;
  const bool wly = h0y == 0, why = h1y == nby;
  const bool wlz = h0z == 0, whz = h1z == nbz;
  const bool wraps = wlx || whx || wly || why || wlz || whz;
  const double ox = 0.5 * (double)(h0x + h1x) * g.bl;
  const double oy = 0.5 * (double)(h0y + h1y) * g.bl;
  const double oz = 0.5 * (double)(h0z + h1z) * g.bl;
  // mixed candidate prefilter: cutoff widened by a margin far above its
  // FP32 rounding error (a few ulps of the largest |p|^2 in the region,
  // whose half-extent is t/2 + 1 blocks per axis)
  const double hxr = (0.5 * tx + 1.0) * g.bl, hyr = (0.5 * ty + 1.0) * g.bl,
               hzr = (0.5 * tz + 1.0) * g.bl;
  const float rc2pre =
      A.rc2f + (float)((hxr * hxr + hyr * hyr + hzr * hzr) * 0x1p-17);
  for (int rb = tid; rb < NRB; rb += NT) {
    int cx, cy, cz;
    region_coords(rb, cx, cy, cz);
    const int sx = (wlx && cx == 0) ? -1 : ((whx && cx == RX - 1) ? 1 : 0);
    const int sy = (wly && cy == 0) ? -1 : ((why && cy == RY - 1) ? 1 : 0);
    const int sz = (wlz && cz == 0) ? -1 : ((whz && cz == RZ - 1) ? 1 : 0);
    const int bc = cx | (cy << 8) | (cz << 16) | ((sx + 1) << 24) |
                   ((sy + 1) << 26) | ((sz + 1) << 28);
    const int o = s_off[rb], c = s_off[rb + 1] - o;
    // the block's c atoms are the valid entries of its slots (a list may
    // hold holes, -1, where atoms moved away); a list of S slots starts
    // 4*S-byte aligned (S at least 8) and comes in 16-byte loads issued
    // together, eight slots at a time
    const int *src = A.d.sorted + s_gst[rb];
    const int S = 1 << A.d.bin_shift;
    int w = 0;
    auto put = [&](int id) {
      if (id >= 0 && w < c) {
        s_id[o + w] = id;
        s_bc[o + w] = bc;
        ++w;
      }
    };
    if (FP64 && (S == 8 || S == 16)) {
      // the FP64 sums follow the order of the atoms in a block, which is
      // input order (SRC/exaMD_input_formats.f90:80-88): a list of 8 or 16
      // slots is sorted here, as unsigned values so that free slots (-1)
      // come last.  The second eight slots of a 16-slot list are read (and
      // all sixteen sorted) only when the first eight hold fewer than the
      // block's count.  A longer list (after repeated overflows) is kept in
      // input order by k_block_order.
      const int4 *q = reinterpret_cast<const int4 *>(src);
      const int4 q0 = q[0], q1 = q[1];
      unsigned u[8] = {(unsigned)q0.x, (unsigned)q0.y, (unsigned)q0.z, (unsigned)q0.w,
                       (unsigned)q1.x, (unsigned)q1.y, (unsigned)q1.z, (unsigned)q1.w};
      int v8 = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) v8 += (int)u[k] >= 0;
      if (S == 16 && v8 < c) {
        const int4 q2 = q[2], q3 = q[3];
        unsigned u16[16] = {u[0], u[1], u[2], u[3], u[4], u[5], u[6], u[7],
                            (unsigned)q2.x, (unsigned)q2.y, (unsigned)q2.z, (unsigned)q2.w,
                            (unsigned)q3.x, (unsigned)q3.y, (unsigned)q3.z, (unsigned)q3.w};
        sort16(u16);
#pragma unroll
        for (int k = 0; k < 16; ++k) put((int)u16[k]);
      } else {
        sort8(u);
#pragma unroll
        for (int k = 0; k < 8; ++k) put((int)u[k]);
      }
    } else {
      for (int s0 = 0; w < c && s0 < S; s0 += 8) {
        const int4 q0 = *reinterpret_cast<const int4 *>(src + s0);
        const int4 q1 = *reinterpret_cast<const int4 *>(src + s0 + 4);
        put(q0.x); put(q0.y); put(q0.z); put(q0.w);
        put(q1.x); put(q1.y); put(q1.z); put(q1.w);
      }
    }
    // a count above the list's valid entries would leave LDS entries
    // unset: they get the system's first atom (so that no read goes out of
    // bounds) and the list flag, whose recovery bins afresh
    if (w < c) {
      atomicOr(&log->flags, (unsigned long long)kFlagBinOverflow);
      for (; w < c; ++w) {
        s_id[o + w] = (int)sd.atom0;
        s_bc[o + w] = bc;
      }
    }
  }
  __syncthreads();
  for (int s = tid; s < R; s += NT) {
    const int id = s_id[s];
    const int bc = s_bc[s];
    const int sx = bc_s(bc, 0), sy = bc_s(bc, 1), sz = bc_s(bc, 2);
    const double X = A.d.x[id], Y = A.d.y[id], Z = A.d.z[id];
    if constexpr (TYPED) s_ty[s] = A.d.type[id];
    if constexpr (FP64) {
      s_x[s] = X;
      s_y[s] = Y;
      s_z[s] = Z;
      s_f[3 * s] = 0.0;
      s_f[3 * s + 1] = 0.0;
      s_f[3 * s + 2] = 0.0;
    } else {
      // w = |p|^2 for the candidate prefilter
      const float px = (float)((X + sx * g.L[0]) - ox);
      const float py = (float)((Y + sy * g.L[1]) - oy);
      const float pz = (float)((Z + sz * g.L[2]) - oz);
      s_p4[s] = make_float4(px, py, pz,
                            __builtin_fmaf(px, px, __builtin_fmaf(py, py, pz * pz)));
      s_q[3 * s] = 0;
      s_q[3 * s + 1] = 0;
      s_q[3 * s + 2] = 0;
    }
  }
  if constexpr (TYPED) {
    // the species-pair table (the mixed rows and the FP64 rows share the
    // slot; the FP64 rows are the larger)
    const int np = A.d.ntypes * A.d.ntypes;
    for (int k = tid; k < np; k += NT) {
      if constexpr (FP64) const_cast<Pair64 *>(s_pr64)[k] = A.d.pair64[k];
      else const_cast<float4 *>(s_pr32)[k] = A.d.pair32[k];
      if constexpr (ENERGY) const_cast<long long *>(s_vrc)[k] = A.d.pair_shift[k];
    }
  }
  __syncthreads();
  if (prof && tid == 0) prof[2] = clock64();

  // 3a. order the region atoms by the number of J blocks they own, most
  //     first, and atoms with the same ownership masks together, so that
  //     the lanes of a wave carry similar work and walk the same runs;
  //     atoms that own nothing (outer ghost layers) drop out.
  unsigned short *s_perm = reinterpret_cast<unsigned short *>(smem + LL.off_perm);
  int *s_hist = reinterpret_cast<int *>(smem + LL.off_hist);
  for (int b = tid; b <= kWorkBins; b += NT) s_hist[b] = 0;
  __syncthreads();
  auto class_of = [&](int a) {
    const int bc = s_bc[a];
    const unsigned key = own_bits(bc_x(bc), tx) | (own_bits(bc_y(bc), ty) << 3) |
                         (own_bits(bc_z(bc), tz) << 6);
    return (int)kKeyRank.r[key];
  };
  for (int a = tid; a < R; a += NT) {
    const int c = class_of(a);
    if (c != 255) atomicAdd(&s_hist[c], 1);
  }
  __syncthreads();
  // exclusive scan of the 256 class counts by the first wave, four per lane
  if (tid < 64) {
    const int4 c = reinterpret_cast<const int4 *>(s_hist)[tid];
    int v = c.x + c.y + c.z + c.w;
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(v, o, 64);
      if (tid >= o) v += t;
    }
    const int ex = v - (c.x + c.y + c.z + c.w);
    reinterpret_cast<int4 *>(s_hist)[tid] =
        make_int4(ex, ex + c.x, ex + c.x + c.y, ex + c.x + c.y + c.z);
    if (tid == 63) {
      s_hist[kWorkBins] = v;
      s_wave[31] = NW;   // next wave-item of the pair phase
    }
  }
  __syncthreads();
  // atoms that own at least one block (uniform: kept in a scalar register)
  const int RW = __builtin_amdgcn_readfirstlane(s_hist[kWorkBins]);
  for (int a = tid; a < R; a += NT) {
    const int c = class_of(a);
    if (c != 255) s_perm[atomicAdd(&s_hist[c], 1)] = (unsigned short)a;
  }
  __syncthreads();

  // 3. pairs.  Work item w = (run group g, atom perm[i]): the lane lists the
  //    J ranges atom a owns in the group's runs (home block I - o1 inside
  //    the tile) and tests them a chunk at a time.  Each run group
  //    takes whole waves, so the group and run indices are wave-uniform:
  //    the run tables are scalar loads and each run's row offset is scalar
  //    arithmetic; the lane loads its 13 ownership masks once per work
  //    item (kRunMask).  In-cutoff pairs are either
  //    evaluated at once or, with lane queues, deferred into a per-lane LDS
  //    queue (one entry per chunk with a hit: j << chunk | hit bits)
  //    that the wave empties together, one pair per lane per step.  The
  //    force on a accumulates in registers, the force on j with LDS
  //    atomics.
  unsigned cand = 0, cut = 0;   // this lane's counts in this cell block
  unsigned flags = 0;
  // mixed: smallest in-cutoff d^2 seen by the lane; the near-overlap flags
  // are set from it once, after the pair loop
  unsigned d2min32 = 0x7f7fffffu;   // bits of a non-negative float
  double pe64 = 0.0;
  long long pefix = 0;
  long long shiftfix = 0;   // several species, tally launches: sum of V(rc)
  // candidates tested per step (a chunk): up to sixteen (mixed: 4, 8, 12
  // or 16 by the wave's longest range) or four (FP64, whose differences
  // take twice the registers)
  constexpr int kChunk = FP64 ? 4 : kMixedChunk;
  constexpr unsigned kChunkMask = (1u << kChunk) - 1u;
  const int NG = A.groups;
  const int RWp = (RW + 63) & ~63;   // atoms per group, padded to whole waves
  const int W = NG * RWp;
  const int QD = A.lane_queue;   // queue entries per lane, 0 = evaluate at once
  const bool LQ = QD > 0;
  unsigned *my_q = reinterpret_cast<unsigned *>(smem + LL.off_q) +
                   (size_t)(tid >> 6) * 64 * QD + (tid & 63);
  // Wave-items are taken dynamically: wave w starts with item w, then
  // takes the next free one from a counter in LDS.  The work order puts
  // the heaviest items first, so the waves reach the barrier together.
  for (int wb = __builtin_amdgcn_readfirstlane(tid & ~63); wb < W;) {
    const int grp = __builtin_amdgcn_readfirstlane(wb / RWp);
    const int ia = wb - grp * RWp + (tid & 63);
    const bool active = ia < RW;
    const int a = active ? (int)s_perm[ia] : 0;
    int bc;
    double xa = 0.0, ya = 0.0, za = 0.0;
    float xf = 0.f, yf = 0.f, zf = 0.f;
    if constexpr (FP64) {
      xa = s_x[a];
      ya = s_y[a];
      za = s_z[a];
    } else {
      const float4 p = s_p4[a];
      xf = p.x;
      yf = p.y;
      zf = p.z;
    }
    bc = s_bc[a];
    // prefilter constants of atom a: 2a and rc'^2 - |a|^2
    const float txf = 2.0f * xf, tyf = 2.0f * yf, tzf = 2.0f * zf;
    const float kaf = rc2pre - __builtin_fmaf(xf, xf, __builtin_fmaf(yf, yf, zf * zf));
    const int ix = bc_x(bc), iy = bc_y(bc), iz = bc_z(bc);
    // an idle lane (group padding) owns nothing
    const unsigned okx = active ? own_bits(ix, tx) : 0u, oky = own_bits(iy, ty),
                   okz = own_bits(iz, tz);
    const uint4 mw = *reinterpret_cast<const uint4 *>(
        kRunMask.w[okx | (oky << 3) | (okz << 6)]);
    // region row of J block (ix - 2, iy, iz); a run adds its (dy, dz) rows
    const int RXs = __builtin_amdgcn_readfirstlane(RX);
    const int RYs = __builtin_amdgcn_readfirstlane(RY);
    const int rowbase = RX * (iy + RY * iz) + ix - 2;
    const int sax = bc_s(bc, 0), say = bc_s(bc, 1), saz = bc_s(bc, 2);
    double fax = 0.0, fay = 0.0, faz = 0.0;
    long long qax = 0, qay = 0, qaz = 0;
    // lane queue: a ring of QD entries from qh, the entry being consumed
    // in (cjb, cm), and the pairs pending in both
    int qn = 0;
    int qh = 0;
    int cjb = 0;
    unsigned cm = 0u;
    int pend = 0;
    // several species: first row of atom a's species in the pair table
    const int trow = TYPED ? (int)s_ty[a] * A.d.ntypes : 0;

    // in-cutoff pair (a, j) with the pair's parameters P, table
    // orientation: +f on a, -f on j (SRC/exaMD_forces.f90:538-542)
    auto pair64 = [&](int j, double dx, double dy, double dz, double d2,
                      const Pair64 &P) {
      // user_func.f90:58-69
      const double d2i = 1.0 / d2;
      const double sg = P.sig2 * d2i;
      const double d6 = sg * sg * sg;
      const double d12 = d6 * d6;
      const double dg = -(P.eps24 * (2.0 * d12 - d6) * d2i);
      const double f1 = dg * dx, f2 = dg * dy, f3 = dg * dz;
      fax += f1;
      fay += f2;
      faz += f3;
      double *fj = reinterpret_cast<double *>(
          reinterpret_cast<unsigned char *>(s_f) + __umul24((unsigned)j, 24u));
      lds_add(fj, -f1);
      lds_add(fj + 1, -f2);
      lds_add(fj + 2, -f3);
      if constexpr (ENERGY) pe64 += P.eps4 * (d12 - d6);
      if (d2 < A.d2zero) flags |= kFlagZero;
    };
    // P = (sig^2, 24 eps 2^32, 4 eps 2^32, rc^2) of the pair
    auto pair32 = [&](int j, float dx, float dy, float dz, float d2,
                      const float4 &P) {
      const float r2i = __builtin_amdgcn_rcpf(d2);
      const float s2 = P.x * r2i;
      const float s6 = s2 * s2 * s2;
      // force scalar 24 eps s6 (2 s6 - 1) / r^2, pre-multiplied by 2^32
      // (the fixed-point scale)
      const float fs = __builtin_fmaf(s6, TYPED ? 2.0f * P.y : A.eps48fs, -P.y) *
                       (s6 * r2i);
      // biased fixed point of the exact products -fs * d (the LDS sums are
      // sign-extended from bit 50 when the forces leave the tile)
      const double nfs = -(double)fs;
      const long long q1 = fix32_prod_raw(nfs, dx);
      const long long q2 = fix32_prod_raw(nfs, dy);
      const long long q3 = fix32_prod_raw(nfs, dz);
      qax += q1;
      qay += q2;
      qaz += q3;
      long long *qj = reinterpret_cast<long long *>(
          reinterpret_cast<unsigned char *>(s_q) + __umul24((unsigned)j, 24u));
      lds_sub(qj, q1);
      lds_sub(qj + 1, q2);
      lds_sub(qj + 2, q3);
      if constexpr (ENERGY) pefix += fix32_scaled(P.z * s6 * (s6 - 1.0f));
      d2min32 = min(d2min32, __float_as_uint(d2));
    };
    // a tabulated pair function, a cubic in r^2 on each interval between its
    // points (Hermite, from the points' values and slopes), evaluated by
    // Horner's rule in t across the interval (the pair force f12 = g d as
    // the reference's user function returns it)
    auto pair64tab = [&](int j, double dx, double dy, double dz, double d2) {
      double u = (d2 - A.d.tab_lo) * A.d.tab_inv_h;
      if (u < 0.0) {
        u = 0.0;
        flags |= kFlagTableRange;
      }
      int k = (int)u;
      if (k > A.d.tab_n - 2) k = A.d.tab_n - 2;
      const double t = u - (double)k;
      const double *e = A.d.tab64 + 8 * k;
      const double g = ((e[3] * t + e[2]) * t + e[1]) * t + e[0];
      const double f1 = g * dx, f2 = g * dy, f3 = g * dz;
      fax += f1;
      fay += f2;
      faz += f3;
      double *fj = reinterpret_cast<double *>(
          reinterpret_cast<unsigned char *>(s_f) + __umul24((unsigned)j, 24u));
      lds_add(fj, -f1);
      lds_add(fj + 1, -f2);
      lds_add(fj + 2, -f3);
      if constexpr (ENERGY) pe64 += ((e[7] * t + e[6]) * t + e[5]) * t + e[4];
      if (d2 < A.d2zero) flags |= kFlagZero;
    };
    auto pair32tab = [&](int j, float dx, float dy, float dz, float d2) {
      float u = (d2 - A.d.tab_lof) * A.d.tab_inv_hf;
      if (u < 0.0f) {
        u = 0.0f;
        flags |= kFlagTableRange;
      }
      const int k = min((int)u, A.d.tab_n - 2);
      const float t = u - (float)k;
      // g's cubic, times 2^32 (the V cubic follows it)
      const float4 e = A.d.tab32[2 * k];
      const double gs = (double)__builtin_fmaf(
          __builtin_fmaf(__builtin_fmaf(e.w, t, e.z), t, e.y), t, e.x);
      const long long q1 = fix32_prod_raw(gs, dx);
      const long long q2 = fix32_prod_raw(gs, dy);
      const long long q3 = fix32_prod_raw(gs, dz);
      qax += q1;
      qay += q2;
      qaz += q3;
      long long *qj = reinterpret_cast<long long *>(
          reinterpret_cast<unsigned char *>(s_q) + __umul24((unsigned)j, 24u));
      lds_sub(qj, q1);
      lds_sub(qj + 1, q2);
      lds_sub(qj + 2, q3);
      if constexpr (ENERGY) {
        const float4 ev = A.d.tab32[2 * k + 1];
        pefix += fix32_scaled(__builtin_fmaf(
            __builtin_fmaf(__builtin_fmaf(ev.w, t, ev.z), t, ev.y), t, ev.x));
      }
      d2min32 = min(d2min32, __float_as_uint(d2));
    };
    auto home_count = [&](int j) {
      if constexpr (HOMECNT) {
        const int bj = s_bc[j];
        const int di = (bc_x(bj) - ix + 2) + 5 * (bc_y(bj) - iy + 2) +
                       25 * (bc_z(bj) - iz + 2);
        const unsigned o = kO1.v[di];
        const int hx = h0x - 1 + ix - ((int)(o & 3) - 1);
        const int hy = h0y - 1 + iy - ((int)((o >> 2) & 3) - 1);
        const int hz = h0z - 1 + iz - ((int)((o >> 4) & 3) - 1);
        atomicAdd(&A.d.home[block0 + hx + nbx * (hy + nby * hz)], 1ull);
      }
    };
    // the pair's parameters: its species row, or the one species
    auto params64 = [&](int j) {
      if constexpr (TYPED) return s_pr64[trow + s_ty[j]];
      else return Pair64{A.sig2, A.eps24, A.eps4, A.rc2};
    };
    auto params32 = [&](int j) {
      if constexpr (TYPED) return s_pr32[trow + s_ty[j]];
      else return make_float4(A.sig2f, A.eps24fs, A.eps4fs, A.rc2f);
    };
    // an FP64 pair that passed the candidate test: with several species
    // it must also lie within its own species pair's cutoff
    auto hit64 = [&](int j, double dx, double dy, double dz, double d2) {
      const Pair64 P = params64(j);
      if (TYPED && d2 > P.rc2) return;
      if constexpr (TABLE) pair64tab(j, dx, dy, dz, d2);
      else pair64(j, dx, dy, dz, d2, P);
      if constexpr (TYPED && ENERGY) shiftfix += s_vrc[trow + s_ty[j]];
      home_count(j);
      ++cut;
    };
    // evaluate a queued pair (a, j) from positions
    auto eval_queued = [&](int j) {
      if constexpr (FP64) {
        double btx = 0.0, bty = 0.0, btz = 0.0;
        if (wraps) {
          const int bj = s_bc[j];
          btx = (double)(sax - bc_s(bj, 0)) * g.L[0];
          bty = (double)(say - bc_s(bj, 1)) * g.L[1];
          btz = (double)(saz - bc_s(bj, 2)) * g.L[2];
        }
        const double dx = (s_x[j] - xa) - btx;
        const double dy = (s_y[j] - ya) - bty;
        const double dz = (s_z[j] - za) - btz;
        hit64(j, dx, dy, dz, dx * dx + dy * dy + dz * dz);
      } else {
        const float4 p = s_p4[j];
        const float dx = p.x - xf, dy = p.y - yf, dz = p.z - zf;
        const float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
        // the exact cutoff test behind the candidate prefilter (the pair's
        // own species cutoff with several species)
        const float4 P = params32(j);
        if (d2 > P.w) return;
        if constexpr (TABLE) pair32tab(j, dx, dy, dz, d2);
        else pair32(j, dx, dy, dz, d2, P);
        if constexpr (TYPED && ENERGY) shiftfix += s_vrc[trow + s_ty[j]];
        home_count(j);
        ++cut;
      }
    };
    // empty the queues of the wave's lanes together, one pair per lane per
    // step: a lane takes its next entry when the hit bits of the current
    // one are used up
    auto drain_step = [&]() {
      if (cm == 0u && qn > 0) {
        const unsigned v = my_q[qh * 64];
        if (++qh == QD) qh = 0;
        --qn;
        cjb = (int)(v >> kChunk);
        cm = v & kChunkMask;
      }
      if (cm) {
        const int k = __builtin_ctz(cm);
        cm &= cm - 1u;
        --pend;
        eval_queued(cjb + k);
      }
    };
    // every pending pair of the wave's lanes
    auto drain = [&]() {
      while (__ballot(cm != 0u || qn > 0)) drain_step();
    };
    // K steps: about the mean pending count, so that nearly every lane
    // has a pair in each step; what remains stays queued
    auto drain_some = [&](int K) {
      for (int it = 0; it < K; ++it) drain_step();
    };
    // defer the hits among the chunk at j as one queue entry; a full ring
    // (rare) first has its oldest entry evaluated by its own lane
    auto push = [&](int j, unsigned hits) {
      if (qn == QD) {
        const unsigned v = my_q[qh * 64];
        if (++qh == QD) qh = 0;
        --qn;
        for (unsigned m = v & kChunkMask; m; m &= m - 1u) {
          --pend;
          eval_queued((int)(v >> kChunk) + __builtin_ctz(m));
        }
      }
      int tl = qh + qn;
      if (tl >= QD) tl -= QD;
      my_q[tl * 64] = ((unsigned)j << kChunk) | hits;
      ++qn;
      pend += __builtin_popcount(hits);
    };

    const int r0 = kGroupStart[NG][grp], r1 = kGroupStart[NG][grp + 1];
    for (int r = r0; r < r1; ++r) {
      const unsigned mwr = r < 4 ? mw.x : (r < 8 ? mw.y : (r < 12 ? mw.z : mw.w));
      unsigned mask = (mwr >> ((r & 3) * 8)) & 31u;
      // the run's row offset is uniform: computed on the scalar unit
      const int row0 = rowbase + RXs * (kRunDY[r] + RYs * kRunDZ[r]);
      double bty = 0.0, btz = 0.0;
      if constexpr (FP64) {
        const int jy = iy + kRunDY[r], jz = iz + kRunDZ[r];
        if (wraps) {
          const int sjy = (wly && jy == 0) ? -1 : ((why && jy == RY - 1) ? 1 : 0);
          const int sjz = (wlz && jz == 0) ? -1 : ((whz && jz == RZ - 1) ? 1 : 0);
          bty = (double)(say - sjy) * g.L[1];
          btz = (double)(saz - sjz) * g.L[2];
        }
      }
      while (mask) {
        const int lo = __builtin_ctz(mask);
        const int hi = lo + __builtin_ctz(~(mask >> lo)) - 1;
        mask &= ~((2u << hi) - 1u);
        // self run, dx = 0: the block's own triangle, j > a
        const int j0 = (r == 0 && lo == 2) ? a + 1 : s_off[row0 + lo];
        const int j1 = s_off[row0 + hi + 1];
        cand += (unsigned)(j1 > j0 ? j1 - j0 : 0);
        if constexpr (!FP64) {
          // mixed: up to sixteen candidates per step, in groups of four;
          // hits are evaluated (at once or from the queue) from the LDS
          // positions.  The step width (4, 8, 12 or 16) follows the longest
          // range among the wave's active lanes, so short segments do not
          // pay for sixteen slots.  The width is wave-uniform and the groups
          // beyond it are skipped by uniform branches inside one loop, so
          // the state carried through the loop (force sums, queue) keeps
          // one set of registers whatever the width.
          const int nj = j1 - j0;
          const int ng = __builtin_amdgcn_readfirstlane(
              !__ballot(nj > 4) ? 1 : (!__ballot(nj > 8) ? 2 : (!__ballot(nj > 12) ? 3 : 4)));
          const int CW = 4 * ng;
          for (int j = j0; j < j1; j += CW) {
            // prefilter: rc'^2 - d^2 = (rc'^2 - |a|^2) - |p|^2 + 2 a.p by
            // one subtraction and three FMAs; its sign bit (set beyond the
            // cutoff) is shifted into the miss mask, last candidate first,
            // so that bit k ends up holding candidate k
            unsigned miss = 0u;
            auto group = [&](int k0) {
#pragma unroll
              for (int k = k0 + 3; k >= k0; --k) {
                const float4 p = s_p4[j + k];
                const float t = __builtin_fmaf(
                    tzf, p.z, __builtin_fmaf(tyf, p.y, __builtin_fmaf(txf, p.x, kaf - p.w)));
                miss = __builtin_amdgcn_alignbit(miss, __float_as_uint(t), 31);
              }
            };
            if (ng > 3) group(12);
            if (ng > 2) group(8);
            if (ng > 1) group(4);
            group(0);
            // candidates past the range end are masked off
            const int lim = j1 - j;
            unsigned hits = ~miss & ((1u << min(lim, CW)) - 1u);
            if (LQ) {
              if (hits) push(j, hits);
            } else {
              while (hits) {
                const int k = __builtin_ctz(hits);
                hits &= hits - 1u;
                eval_queued(j + k);
              }
            }
          }
        } else {
          for (int j = j0; j < j1; j += kChunk) {
            unsigned hits;
            double ddx[4], ddy[4], ddz[4], dd2[4];
            double btx0 = 0.0, btx1 = 0.0, btx2 = 0.0, btx3 = 0.0;
            if (wraps) {
              auto btx_of = [&](int jj) {
                const int cj = bc_x(s_bc[jj]);
                const int sjx = (wlx && cj == 0) ? -1 : ((whx && cj == RX - 1) ? 1 : 0);
                return (double)(sax - sjx) * g.L[0];
              };
              btx0 = btx_of(j);
              btx1 = btx_of(j + 1);
              btx2 = btx_of(j + 2);
              btx3 = btx_of(j + 3);
            }
            // SRC/exaMD_forces.f90:509-523 with Nint(dx*globalboxi) = bt
            ddx[0] = (s_x[j] - xa) - btx0; ddy[0] = (s_y[j] - ya) - bty; ddz[0] = (s_z[j] - za) - btz;
            ddx[1] = (s_x[j + 1] - xa) - btx1; ddy[1] = (s_y[j + 1] - ya) - bty; ddz[1] = (s_z[j + 1] - za) - btz;
            ddx[2] = (s_x[j + 2] - xa) - btx2; ddy[2] = (s_y[j + 2] - ya) - bty; ddz[2] = (s_z[j + 2] - za) - btz;
            ddx[3] = (s_x[j + 3] - xa) - btx3; ddy[3] = (s_y[j + 3] - ya) - bty; ddz[3] = (s_z[j + 3] - za) - btz;
#pragma unroll
            for (int k = 0; k < 4; ++k)
              dd2[k] = ddx[k] * ddx[k] + ddy[k] * ddy[k] + ddz[k] * ddz[k];
            hits = ((dd2[0] <= A.rc2) ? 1u : 0u) | ((dd2[1] <= A.rc2) ? 2u : 0u) |
                   ((dd2[2] <= A.rc2) ? 4u : 0u) | ((dd2[3] <= A.rc2) ? 8u : 0u);
            // candidates past the range end are masked off
            const int lim = j1 - j;
            hits &= lim >= kChunk ? kChunkMask : ((1u << lim) - 1u);
            if (LQ) {
              if (hits) push(j, hits);
            } else {
              while (hits) {
                const int k = __builtin_ctz(hits);
                hits &= hits - 1;
                const double dx = k == 0 ? ddx[0] : k == 1 ? ddx[1] : k == 2 ? ddx[2] : ddx[3];
                const double dy = k == 0 ? ddy[0] : k == 1 ? ddy[1] : k == 2 ? ddy[2] : ddy[3];
                const double dz = k == 0 ? ddz[0] : k == 1 ? ddz[1] : k == 2 ? ddz[2] : ddz[3];
                const double d2 = k == 0 ? dd2[0] : k == 1 ? dd2[1] : k == 2 ? dd2[2] : dd2[3];
                hit64(j + k, dx, dy, dz, d2);
              }
            }
          }
        }
      }
      // after every second run (and the group's last): drain part of the
      // queues once the wave holds two pending pairs per lane on average.
      // The mixed path drains three quarters of the mean pending count
      // (lanes holding fewer pairs than that idle in the last steps of a
      // longer drain); the FP64 path, whose four-candidate entries fill
      // the rings sooner, the whole mean.
      if (LQ && (((r - r0) & 1) || r + 1 == r1)) {
        const int P = wave_total(pend);
        if (P >= 64 * kDrainMin) drain_some(FP64 ? (P >> 6) : ((3 * P) >> 8));
      }
    }
    if (LQ) drain();
    if (active) {
      if constexpr (FP64) {
        lds_add(&s_f[3 * a], fax);
        lds_add(&s_f[3 * a + 1], fay);
        lds_add(&s_f[3 * a + 2], faz);
      } else {
        lds_add(&s_q[3 * a], qax);
        lds_add(&s_q[3 * a + 1], qay);
        lds_add(&s_q[3 * a + 2], qaz);
      }
    }
    // the wave's next item: the first free one (heaviest first)
    int nx = 0;
    if ((tid & 63) == 0) nx = atomicAdd(&s_wave[31], 1);
    wb = __builtin_amdgcn_readlane(nx, 0) * 64;
  }
  __syncthreads();
  if (prof && tid == 0) prof[3] = clock64();

  // 4. return forces of home and ghost atoms (the force leg of the halo
  //    exchange, SRC/exaMD_comms.f90:263-376) with global atomics.
  for (int s = tid; s < R; s += NT) {
    int id = s_id[s];
    if constexpr (FP64) {
      double f1 = s_f[3 * s], f2 = s_f[3 * s + 1], f3 = s_f[3 * s + 2];
      if (f1 != 0.0) unsafeAtomicAdd(&A.d.fx[id], f1);
      if (f2 != 0.0) unsafeAtomicAdd(&A.d.fy[id], f2);
      if (f3 != 0.0) unsafeAtomicAdd(&A.d.fz[id], f3);
    } else {
      long long q1 = fix_sext51(s_q[3 * s]), q2 = fix_sext51(s_q[3 * s + 1]),
                q3 = fix_sext51(s_q[3 * s + 2]);
      if (q1) atomicAdd(&A.d.qx[id], (unsigned long long)q1);
      if (q2) atomicAdd(&A.d.qy[id], (unsigned long long)q2);
      if (q3) atomicAdd(&A.d.qz[id], (unsigned long long)q3);
    }
  }

  if constexpr (!FP64) {
    if (__uint_as_float(d2min32) < A.d2zerof) flags |= kFlagZero;
    if (__uint_as_float(d2min32) < A.d2fixf) flags |= kFlagFixedRange;
  }
  flags = wave_or(flags);
  if ((tid & 63) == 0 && flags)
    atomicOr(&log->flags, (unsigned long long)flags);
  // tallies (pair counts, energy) only on tally launches
  if constexpr (ENERGY) {
    Partial mine{0.0, 0.0, 0ull, (unsigned long long)cand,
                 (unsigned long long)cut, (unsigned long long)shiftfix};
    if constexpr (FP64) mine.pe = pe64;
    else mine.pe_fix = (unsigned long long)pefix;
    write_partial<NT>(mine, A.d.part, tile);
  }
  if (prof && tid == 0) prof[4] = clock64();
}

// Force kernel.  A persistent grid (fewer workgroups than tiles) with
// per-XCD tile ranges (xcd_start): each workgroup takes the next tile of its
// XCD's range from that XCD's counter (workgroup w runs on XCD w % 8), so
// the tiles of one XCD stay contiguous in launch order (shared ghost atoms
// in its L2) while the ranges, of about equal region atoms, and the dynamic
// take balance the load; a workgroup whose range is used up takes tiles of
// the next XCDs' ranges.  Every tile is taken exactly once whichever XCD a
// workgroup runs on, and the last workgroup to finish clears the counters
// for the next launch.  Otherwise (one workgroup per tile, whose launch the
// hardware already spreads as CUs free up, or the allocation-free form
// without ranges) workgroup b evaluates tiles b, b + gridDim, ... in the
// XCD-aware order.
template <bool FP64, bool ENERGY, bool HOMECNT, int PK, int NT>
__global__ __launch_bounds__(NT) void k_forces(ForceArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ntiles = A.t.ntiles;
  if (!A.t.xcd_start || (int)gridDim.x >= ntiles) {
    for (int wg = blockIdx.x; wg < ntiles; wg += gridDim.x) {
      if (wg != (int)blockIdx.x) __syncthreads();  // LDS of the last tile
      force_tile<FP64, ENERGY, HOMECNT, PK, NT>(A, xcd_remap(wg, ntiles), smem);
    }
    return;
  }
  __shared__ int s_tile;
  int x = (int)(blockIdx.x % kXcds), tried = 0;
  for (int first = 1;; first = 0) {
    if (!first) __syncthreads();  // LDS of the last tile, and s_tile read
    if (threadIdx.x == 0) {
      int tile = -1;
      while (tried < kXcds) {
        const int k = A.t.xcd_start[x] + atomicAdd(&A.t.xcd_next[x], 1);
        if (k < A.t.xcd_start[x + 1]) {
          tile = k;
          break;
        }
        x = (x + 1) % kXcds;
        ++tried;
      }
      s_tile = tile;
    }
    __syncthreads();
    const int tile = s_tile;
    if (tile < 0) break;
    force_tile<FP64, ENERGY, HOMECNT, PK, NT>(A, tile, smem);
  }
  if (threadIdx.x == 0) {
    // every fetch of this workgroup came before its finish
    __threadfence();
    if (atomicAdd(&A.t.xcd_next[kXcds], 1) == (int)gridDim.x - 1) {
      for (int i = 0; i < kXcds; ++i) A.t.xcd_next[i] = 0;
      A.t.xcd_next[kXcds] = 0;
    }
  }
}

// One system descriptor into device memory, from a kernel argument (no
// copy from host memory on the stream).
__global__ void k_store_sys(SysDesc v, SysDesc *dst) {
  if (threadIdx.x == 0) *dst = v;
}

__global__ void k_forces_to_fp64(DevState d, double *out, bool mixed) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.n) return;
  if (mixed) {
    out[3 * i + 0] = (double)(long long)d.qx[i] * kFixInv;
    out[3 * i + 1] = (double)(long long)d.qy[i] * kFixInv;
    out[3 * i + 2] = (double)(long long)d.qz[i] * kFixInv;
  } else {
    out[3 * i + 0] = d.fx[i];
    out[3 * i + 1] = d.fy[i];
    out[3 * i + 2] = d.fz[i];
  }
}

// Host interchange layout (x y z per atom) to and from the device SoA.
__global__ void k_aos_to_soa(const double *aos, double *x, double *y, double *z,
                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = aos[3 * i];
  y[i] = aos[3 * i + 1];
  z[i] = aos[3 * i + 2];
}
__global__ void k_soa_to_aos(const double *x, const double *y, const double *z,
                             double *aos, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  aos[3 * i] = x[i];
  aos[3 * i + 1] = y[i];
  aos[3 * i + 2] = z[i];
}

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <bool FP64, bool ENERGY, bool HOMECNT, int PK>
void launch_forces_t(const ForceArgs &A, int threads, size_t lds, int G,
                     hipStream_t s) {
  switch (threads) {
    case 512:
      k_forces<FP64, ENERGY, HOMECNT, PK, 512><<<G, 512, lds, s>>>(A);
      break;
    case 1024:
      k_forces<FP64, ENERGY, HOMECNT, PK, 1024><<<G, 1024, lds, s>>>(A);
      break;
    default:
      k_forces<FP64, ENERGY, HOMECNT, PK, 256><<<G, 256, lds, s>>>(A);
      break;
  }
}

}  // namespace

size_t force_lds_bytes(bool fp64, int cap, int maxrb, int threads,
                       int qdepth, bool typed) {
  return lds_layout(fp64, cap, maxrb, threads / 64, qdepth, typed).total;
}

void launch_fill_i32(int *p, int v, int64_t n, hipStream_t s) {
  if (n > 0) k_fill_i32<<<nblk(n, 256), 256, 0, s>>>(p, v, n);
}
void launch_fill_u64(unsigned long long *p, int64_t n, hipStream_t s) {
  if (n > 0) k_fill_u64<<<nblk(n, 256), 256, 0, s>>>(p, n);
}
void launch_fill_log(StepLog *p, int n, hipStream_t s) {
  if (n > 0) k_fill_log<<<nblk(n, 64), 64, 0, s>>>(p, n);
}

int64_t partial_slots(int64_t n, int64_t tiles) {
  int64_t b = (n + kBinThreads - 1) / kBinThreads;
  return b > tiles ? b : tiles;
}

void launch_step_bin(const DevState &d, const SysTable &t, bool mixed,
                     int mode, double h, double dt, StepLog *ke_log,
                     StepLog *flag_log, hipStream_t s) {
  unsigned nb = nblk(d.n, kBinThreads * kBinPer);
  if (nb == 0) return;
  if (mode == kBinOnly) {
    // empty lists: every slot free, every count zero
    launch_fill_i32(d.sorted, -1, d.nblocks << d.bin_shift, s);
    launch_fill_i32(d.count, 0, d.nblocks, s);
    launch_fill_i32(reinterpret_cast<int *>(d.bflag), 0,
                    bin_flag_bytes(d.nblocks) / 4, s);
  }
  const int want = (mode == kFinishKickDriftBin && ke_log) ? 1 : 0;
#define EXAMD_BIN(M, MD)                                                      \
  (t.nsys == 1                                                                \
       ? k_step_bin<M, MD, true><<<nb, kBinThreads, 0, s>>>(d, t, h, dt, want, \
                                                            flag_log)          \
       : k_step_bin<M, MD, false><<<nb, kBinThreads, 0, s>>>(d, t, h, dt, want, \
                                                             flag_log))
  if (mixed) {
    if (mode == kBinOnly) EXAMD_BIN(true, kBinOnly);
    else if (mode == kKickDriftBin) EXAMD_BIN(true, kKickDriftBin);
    else EXAMD_BIN(true, kFinishKickDriftBin);
  } else {
    if (mode == kBinOnly) EXAMD_BIN(false, kBinOnly);
    else if (mode == kKickDriftBin) EXAMD_BIN(false, kKickDriftBin);
    else EXAMD_BIN(false, kFinishKickDriftBin);
  }
#undef EXAMD_BIN
  if (want)
    k_reduce_partials<<<1, kReduceThreads, 0, s>>>(
        d.part, nullptr, (int)nblk(d.n, kBinThreads), ke_log, kSumKE);
}

void launch_kick(const DevState &d, bool mixed, double h, StepLog *ke_log,
                 hipStream_t s) {
  unsigned nb = nblk(d.n, kBinThreads);
  if (nb == 0) return;
  const int want = ke_log ? 1 : 0;
  if (mixed) k_kick<true><<<nb, kBinThreads, 0, s>>>(d, h, want);
  else k_kick<false><<<nb, kBinThreads, 0, s>>>(d, h, want);
  if (want)
    k_reduce_partials<<<1, kReduceThreads, 0, s>>>(d.part, nullptr, (int)nb,
                                                    ke_log, kSumKE);
}

void launch_sys_ke(const DevState &d, const SysTable &t, StepLog *log,
                   hipStream_t s) {
  if (t.nsys > 0) k_sys_ke<<<t.nsys, kBinThreads, 0, s>>>(d, t, log);
}

void launch_bin_update(const DevState &d, bool fresh, bool block_order,
                       StepLog *flag_log, hipStream_t s) {
  if (d.n <= 0) return;
  const unsigned gb = nblk(bin_flag_bytes(d.nblocks) / 4, 256);
  if (fresh)
    k_bin_insert<<<nblk(d.n, kBinThreads * kInsertPer), kBinThreads, 0, s>>>(d, flag_log);
  if (block_order) k_block_order<<<gb, 256, 0, s>>>(d);
}

void launch_forces(const DevState &d, const SysTable &t, const LJParams &p,
                   const ForceConfig &c, bool mixed, bool energy,
                   bool homecount, StepLog *log, hipStream_t s) {
  if (t.ntiles <= 0) return;
  ForceArgs A;
  A.d = d;
  A.t = t;
  A.log = log;
  A.cap = c.cap;
  A.maxrb = c.maxrb;
  A.groups = c.groups < 1 ? 1 : (c.groups > 4 ? 4 : c.groups);
  A.lane_queue = c.lane_queue;
  A.rc2 = p.rc * p.rc;
  A.eps24 = 24.0 * p.eps;
  A.eps4 = 4.0 * p.eps;
  A.sig2 = p.sig * p.sig;
  A.d2zero = 0.001;
  // The mixed path sums an atom's pair forces of one cell block in 51-bit
  // fixed point at 2^-32, exact while the sum stays below 2^18.  With no
  // pair closer than r_fix, |f| = 24 eps (2 s^12 - s^6)/r <= 48 eps
  // sig^12 / r^13 < 2^14 per pair, and an atom's pairs at r_fix and beyond
  // (first neighbour shell at most 12, further shells falling as r^-13)
  // stay below 16 such forces.
  // (several species: the largest 48 eps sig^12 of the pair table)
  double r13 = 48.0 * (d.ntypes > 0 ? d.max_eps_sig12 : p.eps * pow(p.sig, 12.0)) /
               16384.0;
  double rmin = pow(r13, 1.0 / 13.0);
  A.d2fix = rmin * rmin;
  // a table: the r^2 below which its pair force may reach 2^14
  if (d.tab_n > 0) A.d2fix = d.tab_d2fix;
  A.rc2f = (float)A.rc2;
  A.eps24f = (float)A.eps24;
  A.eps4f = (float)A.eps4;
  A.eps24fs = A.eps24f * 4294967296.0f;
  A.eps48fs = 2.0f * A.eps24fs;
  A.eps4fs = A.eps4f * 4294967296.0f;
  A.sig2f = (float)A.sig2;
  A.d2zerof = (float)A.d2zero;
  A.d2fixf = (float)A.d2fix;
  const bool fp64 = !mixed;
  const bool typed = d.ntypes > 0;
  const size_t lds =
      force_lds_bytes(fp64, c.cap, c.maxrb, c.threads, c.lane_queue, typed);
  // persistent grids are multiples of 8 (one tile order per XCD)
  int G = t.ntiles;
  if (c.workgroups > 0 && c.workgroups < G) G = (c.workgroups + 7) / 8 * 8;
  if (G > t.ntiles) G = t.ntiles;
#define EXAMD_F(F64, E, H)                                                  \
  (typed ? launch_forces_t<F64, E, H, kPairSpecies>(A, c.threads, lds, G, s) \
   : d.tab_n > 0 ? launch_forces_t<F64, E, H, kPairTable>(A, c.threads, lds, G, s) \
                 : launch_forces_t<F64, E, H, kPairLJ>(A, c.threads, lds, G, s))
  if (homecount) {
    if (fp64) EXAMD_F(true, false, true);
    else EXAMD_F(false, false, true);
  } else if (fp64) {
    if (energy) EXAMD_F(true, true, false);
    else EXAMD_F(true, false, false);
  } else {
    if (energy) EXAMD_F(false, true, false);
    else EXAMD_F(false, false, false);
  }
#undef EXAMD_F
  if (energy && !homecount)
    k_reduce_partials<<<t.nsys, kReduceThreads, 0, s>>>(
        d.part, t.nsys > 1 ? t.tile_seg : nullptr, t.ntiles, log, kSumForce);
}

void launch_store_sys(const SysDesc &v, SysDesc *dst, hipStream_t s) {
  k_store_sys<<<1, 64, 0, s>>>(v, dst);
}

void launch_forces_to_fp64(const DevState &d, double *out_aos, bool mixed,
                           hipStream_t s) {
  if (d.n > 0)
    k_forces_to_fp64<<<nblk(d.n, 256), 256, 0, s>>>(d, out_aos, mixed);
}

void launch_aos_to_soa(const double *aos, double *x, double *y, double *z,
                       int64_t n, hipStream_t s) {
  if (n > 0) k_aos_to_soa<<<nblk(n, 256), 256, 0, s>>>(aos, x, y, z, n);
}

void launch_soa_to_aos(const double *x, const double *y, const double *z,
                       double *aos, int64_t n, hipStream_t s) {
  if (n > 0) k_soa_to_aos<<<nblk(n, 256), 256, 0, s>>>(x, y, z, aos, n);
}

}  // namespace examd
