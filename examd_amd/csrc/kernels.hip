// kernels.hip -- gfx950 kernels of the midpoint-method LJ path.
//
//   k_step_bin   velocity-Verlet kick/drift/wrap fused with cell assignment
//                (exaMD_identify_block, SRC/exaMD_module.f90:101-125)
//   k_scan_*     exclusive scan of per-block counts
//   k_scatter    counting-sort scatter, then k_block_order restores input
//                order inside each block (SRC/exaMD_input_formats.f90:80-88)
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

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

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
  if (i < n) p[i] = StepLog{0.0, 0.0, 0ull, 0ull, 0ull, 0ull};
}

// ---------------------------------------------------------- binning ----
// One thread per atom (input order).  mode kFinishKickDriftBin closes the
// previous step (v += h F, kinetic energy) and opens the next one; all modes
// zero the force accumulators and assign the block.
template <bool MIXED, int MODE>
__global__ __launch_bounds__(kBinThreads) void k_step_bin(
    DevState d, Grid g, double h, double dt, StepLog *ke_log,
    StepLog *flag_log) {
  const int64_t i = (int64_t)blockIdx.x * kBinThreads + threadIdx.x;
  double ke = 0.0;
  unsigned flags = 0;
  if (i < d.n) {
    double X = d.x[i], Y = d.y[i], Z = d.z[i];
    if constexpr (MODE != kBinOnly) {
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
      double Vx = d.vx[i], Vy = d.vy[i], Vz = d.vz[i];
      if constexpr (MODE == kFinishKickDriftBin) {
        Vx = Vx + h * Fx;
        Vy = Vy + h * Fy;
        Vz = Vz + h * Fz;
        ke = Vx * Vx + Vy * Vy + Vz * Vz;
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
      else cell = (int)(num - 1);
    }
    d.cell[i] = cell;
    d.rank[i] = atomicAdd(&d.count[cell], 1);
  }
  if constexpr (MODE == kFinishKickDriftBin) {
    if (ke_log) {
      ke = wave_sum(ke);
      if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(&ke_log->ke, ke);
    }
  }
  flags = wave_or(flags);
  if ((threadIdx.x & 63) == 0 && flags)
    atomicOr(&flag_log->flags, (unsigned long long)flags);
}

template <bool MIXED>
__global__ __launch_bounds__(kBinThreads) void k_kick(DevState d, double h,
                                                      StepLog *ke_log) {
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
  if (ke_log) {
    ke = wave_sum(ke);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(&ke_log->ke, ke);
  }
}

// ------------------------------------------------------------- scan ----
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int kScanChunk = kScanThreads * kScanPer;  // 2048 counts

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

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(
    const int *count, int64_t n, int *chunk) {
  __shared__ int s_wave[kScanThreads / 64 + 1];
  int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k)
    if (base + k < n) s += count[base + k];
  int total;
  block_inclusive_scan<kScanThreads>(s, s_wave, &total);
  if (threadIdx.x == 0) chunk[blockIdx.x] = total;
}

// Exclusive scan of the chunk totals in one workgroup (any number of
// chunks, processed kScanThreads at a time).
__global__ __launch_bounds__(kScanThreads) void k_scan_chunks(int *chunk,
                                                              int nchunk) {
  __shared__ int s_wave[kScanThreads / 64 + 1];
  int carry = 0;
  for (int b = 0; b < nchunk; b += kScanThreads) {
    int i = b + threadIdx.x;
    int v = i < nchunk ? chunk[i] : 0;
    int total;
    int inc = block_inclusive_scan<kScanThreads>(v, s_wave, &total);
    if (i < nchunk) chunk[i] = carry + inc - v;
    carry += total;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(
    const int *count, int64_t n, const int *chunk, int *start) {
  __shared__ int s_wave[kScanThreads / 64 + 1];
  int64_t base = (int64_t)blockIdx.x * kScanChunk + threadIdx.x * kScanPer;
  int c[kScanPer];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    c[k] = base + k < n ? count[base + k] : 0;
    s += c[k];
  }
  int total;
  int inc = block_inclusive_scan<kScanThreads>(s, s_wave, &total);
  int run = chunk[blockIdx.x] + inc - s;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) start[base + k] = run;
    run += c[k];
  }
  if (base + kScanPer >= n && base < n) start[n] = run;
}

__global__ void k_scatter(DevState d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.n) d.sorted[d.start[d.cell[i]] + d.rank[i]] = (int)i;
}

// Input order inside each block (the order the reference's linked list
// gives, SRC/exaMD_input_formats.f90:80-88): insertion sort of the few
// atom ids of one block.  Also clears the block's count for the next
// binning pass (the force kernel reads counts from start[]).
__global__ void k_block_order(DevState d, int64_t nblocks) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  int c = d.count[b];
  d.count[b] = 0;
  if (c < 2) return;
  int *p = d.sorted + d.start[b];
  for (int k = 1; k < c; ++k) {
    int v = p[k];
    int m = k - 1;
    while (m >= 0 && p[m] > v) {
      p[m + 1] = p[m];
      --m;
    }
    p[m + 1] = v;
  }
}

// ----------------------------------------------------------- forces ----
struct ForceArgs {
  DevState d;
  Grid g;
  StepLog *log;
  int cap, maxrb;
  double rc2, eps24, eps4, sig2, d2zero, d2fix;
  float rc2f, eps24f, eps4f, sig2f, d2zerof, d2fixf;
};

__host__ __device__ inline size_t align_up(size_t v, size_t a) {
  return (v + a - 1) / a * a;
}

// LDS image of one cell block: region-block offsets, then atoms.
struct LdsLayout {
  size_t off_gst, off_wave, off_pos, off_f, off_id, off_bc, total;
};
__host__ __device__ inline LdsLayout lds_layout(bool fp64, int cap,
                                                int maxrb) {
  LdsLayout L;
  size_t o = (size_t)(maxrb + 1) * 4;  // s_off
  L.off_gst = o;
  o += (size_t)maxrb * 4;
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
  if (fp64) o += (size_t)cap * 4;
  L.total = align_up(o, 16);
  return L;
}

// Bits of the o1 values (-1, 0, +1 -> bit 0, 1, 2) for which the home
// coordinate i - o1 of a region coordinate i lies in the tile's [1, t].
__device__ __forceinline__ unsigned own_bits(int i, int t) {
  return ((i + 1 >= 1 && i + 1 <= t) ? 1u : 0u) | ((i >= 1 && i <= t) ? 2u : 0u) |
         ((i - 1 >= 1 && i - 1 <= t) ? 4u : 0u);
}

// Ownership mask of run R (bit k <-> dx = k - 2) for an I block with
// per-axis own_bits.  The per-axis masks are compile-time constants derived
// from the reference table.
template <int R>
__device__ __forceinline__ unsigned run_mask(unsigned ox, unsigned oy,
                                             unsigned oz) {
  constexpr unsigned x0 = axis_mask(R, 0, -1), x1 = axis_mask(R, 0, 0),
                     x2 = axis_mask(R, 0, 1);
  constexpr unsigned y0 = axis_mask(R, 1, -1), y1 = axis_mask(R, 1, 0),
                     y2 = axis_mask(R, 1, 1);
  constexpr unsigned z0 = axis_mask(R, 2, -1), z1 = axis_mask(R, 2, 0),
                     z2 = axis_mask(R, 2, 1);
  unsigned mx = ((ox & 1u) ? x0 : 0u) | ((ox & 2u) ? x1 : 0u) |
                ((ox & 4u) ? x2 : 0u);
  unsigned my = ((oy & 1u) ? y0 : 0u) | ((oy & 2u) ? y1 : 0u) |
                ((oy & 4u) ? y2 : 0u);
  unsigned mz = ((oz & 1u) ? z0 : 0u) | ((oz & 2u) ? z1 : 0u) |
                ((oz & 4u) ? z2 : 0u);
  return mx & my & mz;
}

template <int R>
__device__ __forceinline__ Off3 run_o1_rt(int k) {
  constexpr Off3 a = run_o1(R, 0), b = run_o1(R, 1), c = run_o1(R, 2),
                 e = run_o1(R, 3), f = run_o1(R, 4);
  return k == 0 ? a : k == 1 ? b : k == 2 ? c : k == 3 ? e : f;
}

__device__ __forceinline__ void lds_add(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(long long *p, long long v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_sub(long long *p, long long v) {
  __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <bool FP64, bool ENERGY, bool HOMECNT, int NT>
__global__ __launch_bounds__(NT) void k_forces(ForceArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int tid = threadIdx.x;
  const Grid &g = A.g;
  const int G = g.T[0] * g.T[1] * g.T[2];
  const int tile = xcd_remap(blockIdx.x, G);
  const int tI = tile % g.T[0];
  const int tJ = (tile / g.T[0]) % g.T[1];
  const int tK = tile / (g.T[0] * g.T[1]);
  int h0x, h1x, h0y, h1y, h0z, h1z;
  tile_range(tI, g.nb[0], g.T[0], h0x, h1x);
  tile_range(tJ, g.nb[1], g.T[1], h0y, h1y);
  tile_range(tK, g.nb[2], g.T[2], h0z, h1z);
  const int tx = h1x - h0x, ty = h1y - h0y, tz = h1z - h0z;
  const int RX = tx + 2, RY = ty + 2, RZ = tz + 2;
  const int RXY = RX * RY, NRB = RXY * RZ;
  const int nbx = g.nb[0], nby = g.nb[1], nbz = g.nb[2];

  const LdsLayout LL = lds_layout(FP64, A.cap, A.maxrb);
  int *s_off = reinterpret_cast<int *>(smem);
  int *s_gst = reinterpret_cast<int *>(smem + LL.off_gst);
  int *s_wave = reinterpret_cast<int *>(smem + LL.off_wave);
  int *s_id = reinterpret_cast<int *>(smem + LL.off_id);
  int *s_bc = reinterpret_cast<int *>(smem + LL.off_bc);
  // FP64 image: SoA doubles; mixed image: float4 (x, y, z, packed coords)
  double *s_x = reinterpret_cast<double *>(smem + LL.off_pos);
  double *s_y = s_x + A.cap;
  double *s_z = s_y + A.cap;
  float4 *s_p4 = reinterpret_cast<float4 *>(smem + LL.off_pos);
  double *s_fx = reinterpret_cast<double *>(smem + LL.off_f);
  double *s_fy = s_fx + A.cap;
  double *s_fz = s_fy + A.cap;
  long long *s_qx = reinterpret_cast<long long *>(smem + LL.off_f);
  long long *s_qy = s_qx + A.cap;
  long long *s_qz = s_qy + A.cap;

  // 1. region blocks: home blocks of the tile plus a one-block ghost shell,
  //    periodic (SRC/exaMD_partitioning.f90:788-790).
  for (int rb = tid; rb < NRB; rb += NT) {
    int cx = rb % RX, cy = (rb / RX) % RY, cz = rb / RXY;
    int gx = h0x - 1 + cx, gy = h0y - 1 + cy, gz = h0z - 1 + cz;
    gx += gx < 0 ? nbx : 0;
    gx -= gx >= nbx ? nbx : 0;
    gy += gy < 0 ? nby : 0;
    gy -= gy >= nby ? nby : 0;
    gz += gz < 0 ? nbz : 0;
    gz -= gz >= nbz ? nbz : 0;
    int gb = gx + nbx * (gy + nby * gz);
    int st = A.d.start[gb];
    s_off[rb] = A.d.start[gb + 1] - st;
    s_gst[rb] = st;
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
  if (R > A.cap) {
    if (tid == 0) atomicOr(&A.log->flags, (unsigned long long)kFlagOverflow);
    return;
  }

  // 2. ghost gather: atoms of the region into LDS.  The mixed image is
  //    relative to the tile centre and pre-shifted by the periodic image;
  //    the FP64 image keeps raw coordinates (minimum image is applied per
  //    pair, as the reference does).
  const bool wlx = h0x == 0, whx = h1x == nbx;
  const bool wly = h0y == 0, why = h1y == nby;
  const bool wlz = h0z == 0, whz = h1z == nbz;
  const double ox = 0.5 * (double)(h0x + h1x) * g.bl;
  const double oy = 0.5 * (double)(h0y + h1y) * g.bl;
  const double oz = 0.5 * (double)(h0z + h1z) * g.bl;
  for (int rb = tid; rb < NRB; rb += NT) {
    int b0 = s_off[rb], b1 = s_off[rb + 1];
    if (b0 == b1) continue;
    int cx = rb % RX, cy = (rb / RX) % RY, cz = rb / RXY;
    int gs = s_gst[rb];
    int bc = cx | (cy << 10) | (cz << 20);
    double shx = (wlx && cx == 0) ? -g.L[0] : ((whx && cx == RX - 1) ? g.L[0] : 0.0);
    double shy = (wly && cy == 0) ? -g.L[1] : ((why && cy == RY - 1) ? g.L[1] : 0.0);
    double shz = (wlz && cz == 0) ? -g.L[2] : ((whz && cz == RZ - 1) ? g.L[2] : 0.0);
    for (int k = 0; k < b1 - b0; ++k) {
      int id = A.d.sorted[gs + k];
      int s = b0 + k;
      double X = A.d.x[id], Y = A.d.y[id], Z = A.d.z[id];
      s_id[s] = id;
      if constexpr (FP64) {
        s_x[s] = X;
        s_y[s] = Y;
        s_z[s] = Z;
        s_bc[s] = bc;
        s_fx[s] = 0.0;
        s_fy[s] = 0.0;
        s_fz[s] = 0.0;
      } else {
        s_p4[s] = make_float4((float)((X + shx) - ox), (float)((Y + shy) - oy),
                              (float)((Z + shz) - oz), __int_as_float(bc));
        s_qx[s] = 0;
        s_qy[s] = 0;
        s_qz[s] = 0;
      }
    }
  }
  __syncthreads();

  // 3. pairs.  One lane per i-atom a in block I of the region; for each of
  //    the 13 runs the owned J blocks (home I - o1 inside the tile) form at
  //    most a few column segments, each one contiguous range of LDS atoms.
  unsigned long long cand = 0, cut = 0;
  unsigned flags = 0;
  double pe64 = 0.0;
  long long pefix = 0;

  for (int a = tid; a < R; a += NT) {
    int bc;
    double xa, ya, za;
    float xf, yf, zf;
    if constexpr (FP64) {
      xa = s_x[a];
      ya = s_y[a];
      za = s_z[a];
      bc = s_bc[a];
    } else {
      float4 p = s_p4[a];
      xf = p.x;
      yf = p.y;
      zf = p.z;
      bc = __float_as_int(p.w);
    }
    const int ix = bc & 1023, iy = (bc >> 10) & 1023, iz = bc >> 20;
    const unsigned okx = own_bits(ix, tx), oky = own_bits(iy, ty),
                   okz = own_bits(iz, tz);
    const int sIx = (wlx && ix == 0) ? -1 : ((whx && ix == RX - 1) ? 1 : 0);
    const int sIy = (wly && iy == 0) ? -1 : ((why && iy == RY - 1) ? 1 : 0);
    const int sIz = (wlz && iz == 0) ? -1 : ((whz && iz == RZ - 1) ? 1 : 0);
    double fax = 0.0, fay = 0.0, faz = 0.0;
    long long qax = 0, qay = 0, qaz = 0;

    // Pairs (a, j) for LDS atoms j in [j0, j1), all in the same periodic
    // image relative to a (btx/bty/btz = (sI - sJ) L; FP64 only).
    auto pair_range = [&](int j0, int j1, double btx, double bty, double btz,
                          int homeb) {
      cand += (unsigned long long)(j1 > j0 ? j1 - j0 : 0);
      for (int j = j0; j < j1; ++j) {
        if constexpr (FP64) {
          // SRC/exaMD_forces.f90:509-523 with Nint(dx*globalboxi) = bt
          double dx = (s_x[j] - xa) - btx;
          double dy = (s_y[j] - ya) - bty;
          double dz = (s_z[j] - za) - btz;
          double d2 = dx * dx + dy * dy + dz * dz;
          if (d2 <= A.rc2) {  // user_func.f90:52
            double d2i = 1.0 / d2;
            double s = A.sig2 * d2i;
            double d6 = s * s * s;
            double d12 = d6 * d6;
            double dg = -(A.eps24 * (2.0 * d12 - d6) * d2i);
            double f1 = dg * dx, f2 = dg * dy, f3 = dg * dz;
            fax += f1;
            fay += f2;
            faz += f3;
            lds_add(&s_fx[j], -f1);
            lds_add(&s_fy[j], -f2);
            lds_add(&s_fz[j], -f3);
            ++cut;
            if constexpr (ENERGY) pe64 += A.eps4 * (d12 - d6);
            if (d2 < A.d2zero) flags |= kFlagZero;
            if constexpr (HOMECNT) atomicAdd(&A.d.home[homeb], 1ull);
          }
        } else {
          float4 p = s_p4[j];
          float dx = p.x - xf, dy = p.y - yf, dz = p.z - zf;
          float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
          if (d2 <= A.rc2f) {
            float r2i = __builtin_amdgcn_rcpf(d2);
            float s2 = A.sig2f * r2i;
            float s6 = s2 * s2 * s2;
            float fs = A.eps24f * s6 * __builtin_fmaf(2.0f, s6, -1.0f) * r2i;
            long long q1 = fix32(-fs * dx);
            long long q2 = fix32(-fs * dy);
            long long q3 = fix32(-fs * dz);
            qax += q1;
            qay += q2;
            qaz += q3;
            lds_sub(&s_qx[j], q1);
            lds_sub(&s_qy[j], q2);
            lds_sub(&s_qz[j], q3);
            ++cut;
            if constexpr (ENERGY)
              pefix += fix32(A.eps4f * s6 * (s6 - 1.0f));
            if (d2 < A.d2zerof) flags |= kFlagZero;
            if (d2 < A.d2fixf) flags |= kFlagFixedRange;
            if constexpr (HOMECNT) atomicAdd(&A.d.home[homeb], 1ull);
          }
        }
      }
    };

    static_for<0, kNumRuns>([&](auto RUN) {
      constexpr int r = decltype(RUN)::value;
      unsigned mask = run_mask<r>(okx, oky, okz);
      if (!mask) return;
      const int jy = iy + kRunDY[r], jz = iz + kRunDZ[r];
      const int row0 = RX * (jy + RY * jz);
      double bty = 0.0, btz = 0.0;
      if constexpr (FP64) {
        int sJy = (wly && jy == 0) ? -1 : ((why && jy == RY - 1) ? 1 : 0);
        int sJz = (wlz && jz == 0) ? -1 : ((whz && jz == RZ - 1) ? 1 : 0);
        bty = (double)(sIy - sJy) * g.L[1];
        btz = (double)(sIz - sJz) * g.L[2];
      }
      // columns [c0, c1] of J's row, one periodic image
      auto cols = [&](int c0, int c1, int sJx) {
        double btx = 0.0;
        if constexpr (FP64) btx = (double)(sIx - sJx) * g.L[0];
        if constexpr (HOMECNT) {
          for (int c = c0; c <= c1; ++c) {
            Off3 o = run_o1_rt<r>(c - ix + 2);
            int hx = h0x - 1 + ix - o.x, hy = h0y - 1 + iy - o.y,
                hz = h0z - 1 + iz - o.z;
            int homeb = hx + nbx * (hy + nby * hz);
            int j0 = (r == 0 && c == ix) ? a + 1 : s_off[row0 + c];
            pair_range(j0, s_off[row0 + c + 1], btx, bty, btz, homeb);
          }
        } else {
          int j0 = (r == 0 && c0 == ix) ? a + 1 : s_off[row0 + c0];
          pair_range(j0, s_off[row0 + c1 + 1], btx, bty, btz, 0);
        }
      };
      while (mask) {
        int lo = __builtin_ctz(mask);
        int hi = lo + __builtin_ctz(~(mask >> lo)) - 1;
        mask &= ~((2u << hi) - 1u);
        int c0 = ix + lo - 2, c1 = ix + hi - 2;
        if constexpr (FP64) {
          // split at wrapped shell columns so each piece has one image
          if (wlx && c0 == 0) {
            cols(0, 0, -1);
            c0 = 1;
          }
          bool tail = whx && c1 == RX - 1;
          if (tail) c1 = RX - 2;
          if (c0 <= c1) cols(c0, c1, 0);
          if (tail) cols(RX - 1, RX - 1, 1);
        } else {
          cols(c0, c1, 0);
        }
      }
    });

    if constexpr (FP64) {
      lds_add(&s_fx[a], fax);
      lds_add(&s_fy[a], fay);
      lds_add(&s_fz[a], faz);
    } else {
      lds_add(&s_qx[a], qax);
      lds_add(&s_qy[a], qay);
      lds_add(&s_qz[a], qaz);
    }
  }
  __syncthreads();

  // 4. return forces of home and ghost atoms (the force leg of the halo
  //    exchange, SRC/exaMD_comms.f90:263-376) with global atomics.
  for (int s = tid; s < R; s += NT) {
    int id = s_id[s];
    if constexpr (FP64) {
      double f1 = s_fx[s], f2 = s_fy[s], f3 = s_fz[s];
      if (f1 != 0.0) unsafeAtomicAdd(&A.d.fx[id], f1);
      if (f2 != 0.0) unsafeAtomicAdd(&A.d.fy[id], f2);
      if (f3 != 0.0) unsafeAtomicAdd(&A.d.fz[id], f3);
    } else {
      long long q1 = s_qx[s], q2 = s_qy[s], q3 = s_qz[s];
      if (q1) atomicAdd(&A.d.qx[id], (unsigned long long)q1);
      if (q2) atomicAdd(&A.d.qy[id], (unsigned long long)q2);
      if (q3) atomicAdd(&A.d.qz[id], (unsigned long long)q3);
    }
  }

  cand = wave_sum(cand);
  cut = wave_sum(cut);
  flags = wave_or(flags);
  if constexpr (ENERGY) {
    if constexpr (FP64) pe64 = wave_sum(pe64);
    else pefix = (long long)wave_sum((unsigned long long)pefix);
  }
  if ((tid & 63) == 0) {
    atomicAdd(&A.log->cand, cand);
    atomicAdd(&A.log->cut, cut);
    if (flags) atomicOr(&A.log->flags, (unsigned long long)flags);
    if constexpr (ENERGY) {
      if constexpr (FP64) unsafeAtomicAdd(&A.log->pe, pe64);
      else atomicAdd(&A.log->pe_fix, (unsigned long long)pefix);
    }
  }
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

inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <bool FP64, bool ENERGY, bool HOMECNT>
void launch_forces_t(const ForceArgs &A, int threads, size_t lds, int G,
                     hipStream_t s) {
  switch (threads) {
    case 512:
      k_forces<FP64, ENERGY, HOMECNT, 512><<<G, 512, lds, s>>>(A);
      break;
    case 1024:
      k_forces<FP64, ENERGY, HOMECNT, 1024><<<G, 1024, lds, s>>>(A);
      break;
    default:
      k_forces<FP64, ENERGY, HOMECNT, 256><<<G, 256, lds, s>>>(A);
      break;
  }
}

}  // namespace

size_t force_lds_bytes(bool fp64, int cap, int maxrb) {
  return lds_layout(fp64, cap, maxrb).total;
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

void launch_step_bin(const DevState &d, const Grid &g, bool mixed, int mode,
                     double h, double dt, StepLog *ke_log, StepLog *flag_log,
                     hipStream_t s) {
  unsigned nb = nblk(d.n, kBinThreads);
  if (nb == 0) return;
#define EXAMD_BIN(M, MD)                                               \
  k_step_bin<M, MD><<<nb, kBinThreads, 0, s>>>(d, g, h, dt, ke_log, flag_log)
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
}

void launch_kick(const DevState &d, bool mixed, double h, StepLog *ke_log,
                 hipStream_t s) {
  unsigned nb = nblk(d.n, kBinThreads);
  if (nb == 0) return;
  if (mixed) k_kick<true><<<nb, kBinThreads, 0, s>>>(d, h, ke_log);
  else k_kick<false><<<nb, kBinThreads, 0, s>>>(d, h, ke_log);
}

void launch_sort(const DevState &d, int64_t nblocks, hipStream_t s) {
  int nchunk = (int)((nblocks + kScanChunk - 1) / kScanChunk);
  k_scan_reduce<<<nchunk, kScanThreads, 0, s>>>(d.count, nblocks, d.chunk);
  k_scan_chunks<<<1, kScanThreads, 0, s>>>(d.chunk, nchunk);
  k_scan_down<<<nchunk, kScanThreads, 0, s>>>(d.count, nblocks, d.chunk,
                                              d.start);
  if (d.n > 0) k_scatter<<<nblk(d.n, 256), 256, 0, s>>>(d);
  k_block_order<<<nblk(nblocks, 256), 256, 0, s>>>(d, nblocks);
}

void launch_forces(const DevState &d, const Grid &g, const LJParams &p,
                   const ForceConfig &c, bool mixed, bool energy,
                   bool homecount, StepLog *log, hipStream_t s) {
  ForceArgs A;
  A.d = d;
  A.g = g;
  A.log = log;
  A.cap = c.cap;
  A.maxrb = c.maxrb;
  A.rc2 = p.rc * p.rc;
  A.eps24 = 24.0 * p.eps;
  A.eps4 = 4.0 * p.eps;
  A.sig2 = p.sig * p.sig;
  A.d2zero = 0.001;
  // |f| = 24 eps (2 s^12 - s^6)/r < 2^19 for r^2 above d2fix: with
  // s = sig/r, |f| <= 48 eps sig^12 / r^13, so r^13 >= 48 eps sig^12 / 2^18
  // keeps a factor-two margin.
  double r13 = 48.0 * p.eps * pow(p.sig, 12.0) / 262144.0;
  double rmin = pow(r13, 1.0 / 13.0);
  A.d2fix = rmin * rmin;
  A.rc2f = (float)A.rc2;
  A.eps24f = (float)A.eps24;
  A.eps4f = (float)A.eps4;
  A.sig2f = (float)A.sig2;
  A.d2zerof = (float)A.d2zero;
  A.d2fixf = (float)A.d2fix;
  const bool fp64 = !mixed;
  const size_t lds = force_lds_bytes(fp64, c.cap, c.maxrb);
  const int G = g.T[0] * g.T[1] * g.T[2];
#define EXAMD_F(F64, E, H) launch_forces_t<F64, E, H>(A, c.threads, lds, G, s)
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
}

void launch_forces_to_fp64(const DevState &d, double *out_aos, bool mixed,
                           hipStream_t s) {
  if (d.n > 0)
    k_forces_to_fp64<<<nblk(d.n, 256), 256, 0, s>>>(d, out_aos, mixed);
}

}  // namespace examd
