// api.hip -- C ABI of the gfx950 midpoint LJ library (include/examd_hip.h).
//
// Host orchestration only: device buffers, the cell-block (tile) map that
// replaces the METIS partition (SRC/exaMD_partitioning.f90:65-455), LDS
// sizing, the per-step kernel sequence, captured multi-step chunks and the
// seeded generators.  A handle holds one system (examd_hip_create) or a
// batch of independent systems (examd_hip_batch_create); the kernels see
// both through one table of systems, the single system being a batch of
// one.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/examd_hip.h"
#include "host_maps.hpp"
#include "kernels.hpp"

using namespace examd;

using Tri = std::array<int, 3>;

// One system of a handle, host side.
struct SysHost {
  int64_t atom0 = 0, n = 0;       // atom range in the concatenated order
  int64_t block0 = 0, nblocks = 0;
  Grid g{};                       // nb, L, bl (T unused here)
};

// Cell-block maps and force-kernel launch shape of one precision.
struct TileConfig {
  std::vector<Tri> T;     // cell blocks per dimension, per system
  int edge = 0;           // tile edge the maps were derived from (0 = given)
  int threads = 256;      // workgroup size
  int groups = 1;         // run groups per region atom
  int lane_queue = 0;     // per-lane queue entries of deferred pairs (0 = off)
  int workgroups = 0;     // persistent force grid (0 = one per tile)
  bool balance = false;   // boundaries at equal atom counts along each axis
  std::vector<Cuts> cuts; // balanced boundaries per system (when balance)
  int cap = 0;            // LDS atom capacity
  int maxrb = 0;          // largest region in blocks
  bool fixed_tiles = false;
  int fixed_groups = 0;
  int rebalance = 0;      // steps between re-partitions in a run (0 = never)
  // tiles of each XCD's share of the force launch (9 boundaries: contiguous
  // ranges in launch order of about equal region atoms)
  std::vector<int> xcd;
};

// What a captured chunk depends on.
struct GraphKey {
  int prec = -1;
  int64_t chunk = 0, every = 0;
  double dt = 0.0;
  bool timing = false;
  int timing_every = 1;
  TileConfig cfg;
};

// Device copy of the system table of one precision.
struct DevTables {
  SysDesc *sys = nullptr;
  int *tile_sys = nullptr;
  int *tile_seg = nullptr;
  int *xcd_start = nullptr;   // kXcds + 1 tile boundaries
  int *xcd_next = nullptr;    // kXcds fetch counters and a finish counter
  int *cuts = nullptr;
  int tile_cap = 0;
  int cut_cap = 0;
  int ntiles = 0;
};

struct examd_handle {
  examd_params p{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int nsys = 1;
  std::vector<SysHost> sys;
  long long *atom_seg = nullptr;  // device, nsys + 1
  int *slab_sys = nullptr;        // device, system of each kSysSlab-atom slab (batches)
  DevTables tab[2];               // per precision
  int64_t nblocks = 0;            // all systems
  int64_t n = 0;                  // all systems
  DevState d{};
  StepLog *log = nullptr;
  int64_t log_cap = 0;            // StepLog entries
  TileConfig tc[2];               // per precision
  size_t lds_limit = 0;
  int64_t part_cap = 0;
  // captured multi-step chunk
  int64_t graph_chunk = 0;        // steps per captured chunk (0 = eager)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  GraphKey graph_key{};
  size_t graph_events = 0;        // timing events recorded inside the graph
  bool state_set = false;
  int forces_valid = -1;          // precision whose forces match the positions
  // timing
  bool timing = false;
  int timing_every = 1;           // a run times every this many steps' force
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double force_ms = 0.0;
  int64_t force_launches = 0;
  int64_t since_repartition[2] = {0, 0};  // steps run since, per precision
  int *pinned_counts = nullptr;   // pinned copy of the block counts (re-partitions)
  double *dev_stage = nullptr;    // 3 n doubles: x y z per atom, host layout
  double *chunk_start = nullptr;  // 6 n doubles: x y z vx vy vz at the start
                                  // of the chunk a run is replaying
  int64_t regrows = 0;            // chunks run again with a larger capacity
  double *host_stage = nullptr;   // pinned 3 n doubles, the same layout
  hipEvent_t piece_ev[4] = {};    // the pieces of a download through the stage
};

namespace {

// Slots per block of the atom lists, at least (more when set_state finds a
// fuller block, or after a block ran out of slots).
constexpr int kMinBinShift = 3;
constexpr int64_t kMinBinSlots = (int64_t)1 << kMinBinShift;

#define HIPCHK(expr)                          \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) return EXAMD_ERR_HIP; \
  } while (0)

int flags_to_status(unsigned long long f) {
  if (f & kFlagOutOfBox) return EXAMD_ERR_OUT_OF_BOX;
  if (f & kFlagNonFinite) return EXAMD_ERR_NONFINITE;
  if (f & kFlagOverflow) return EXAMD_ERR_TILE_OVERFLOW;
  if (f & kFlagBinOverflow) return EXAMD_ERR_BIN_OVERFLOW;
  // a pair beyond the fixed-point range makes the mixed forces invalid; a
  // 'ZERO at' pair alone is reported and the forces kept, as the reference
  // prints and goes on
  if (f & kFlagFixedRange) return EXAMD_ERR_FIXED_RANGE;
  if (f & kFlagZero) return EXAMD_ERR_ZERO_DISTANCE;
  if (f & kFlagTableRange) return EXAMD_ERR_TABLE_RANGE;
  return EXAMD_OK;
}

template <class T>
int dalloc(T **p, int64_t n) {
  size_t bytes = (size_t)std::max<int64_t>(n, 1) * sizeof(T);
  if (hipMalloc((void **)p, bytes) != hipSuccess) {
    *p = nullptr;
    return EXAMD_ERR_NOMEM;
  }
  return EXAMD_OK;
}

void drop_graph(examd_handle *h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_events = 0;
}

void free_tables(DevTables &t) {
  if (t.sys) (void)hipFree(t.sys);
  if (t.tile_sys) (void)hipFree(t.tile_sys);
  if (t.tile_seg) (void)hipFree(t.tile_seg);
  if (t.xcd_start) (void)hipFree(t.xcd_start);
  if (t.xcd_next) (void)hipFree(t.xcd_next);
  if (t.cuts) (void)hipFree(t.cuts);
  t = DevTables{};
}

void free_state(examd_handle *h) {
  DevState &d = h->d;
  void *ptrs[] = {d.x,  d.y,  d.z,    d.vx,    d.vy,     d.vz,    d.fx,
                  d.fy, d.fz, d.qx,   d.qy,    d.qz,     d.cell,  d.slot,
                  d.count, d.sorted, d.bflag, d.home, d.part, h->log,
                  h->atom_seg, h->slab_sys, d.type, d.pair32, d.pair64, d.pair_shift,
                  d.tab32, d.tab64};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  d = DevState{};
  h->log = nullptr;
  h->atom_seg = nullptr;
  h->slab_sys = nullptr;
  h->log_cap = 0;
  h->part_cap = 0;
  free_tables(h->tab[0]);
  free_tables(h->tab[1]);
  if (h->pinned_counts) (void)hipHostFree(h->pinned_counts);
  h->pinned_counts = nullptr;
  if (h->dev_stage) (void)hipFree(h->dev_stage);
  if (h->chunk_start) (void)hipFree(h->chunk_start);
  h->chunk_start = nullptr;
  h->dev_stage = nullptr;
  if (h->host_stage) (void)hipHostFree(h->host_stage);
  h->host_stage = nullptr;
}

// Device buffer of x y z per atom for transfers in the host layout (one
// contiguous copy, transposed on the device).
int ensure_stage(examd_handle *h) {
  const int64_t m = 3 * std::max<int64_t>(h->n, 1);
  if (!h->dev_stage && dalloc(&h->dev_stage, m)) return EXAMD_ERR_NOMEM;
  if (!h->host_stage &&
      hipHostMalloc(reinterpret_cast<void **>(&h->host_stage),
                    m * sizeof(double)) != hipSuccess) {
    h->host_stage = nullptr;
    return EXAMD_ERR_NOMEM;
  }
  return EXAMD_OK;
}

int64_t tiles_of(const TileConfig &c) {
  int64_t t = 0;
  for (const Tri &T : c.T) t += (int64_t)T[0] * T[1] * T[2];
  return t;
}

Grid grid_of(const examd_handle *h, int prec, int k) {
  Grid g = h->sys[k].g;
  for (int a = 0; a < 3; ++a) g.T[a] = h->tc[prec].T[k][a];
  return g;
}

// The kernels' view of the systems with the tile maps of one precision.
SysTable table(const examd_handle *h, int prec) {
  SysTable t;
  t.sys = h->tab[prec].sys;
  t.tile_sys = h->tab[prec].tile_sys;
  t.tile_seg = h->tab[prec].tile_seg;
  t.xcd_start = h->tab[prec].xcd_start;
  t.xcd_next = h->tab[prec].xcd_next;
  t.atom_seg = h->atom_seg;
  t.slab_sys = h->slab_sys;
  t.cuts = h->tab[prec].cuts;
  t.nsys = h->nsys;
  t.ntiles = h->tab[prec].ntiles;
  return t;
}

int ensure_partials(examd_handle *h);

// Upload the system table of one precision (after its tile maps changed).
int upload_tables(examd_handle *h, int prec) {
  drop_graph(h);   // a captured chunk holds the old tables
  HIPCHK(hipStreamSynchronize(h->stream));
  DevTables &t = h->tab[prec];
  const int nsys = h->nsys;
  std::vector<SysDesc> sd((size_t)nsys);
  std::vector<int> seg((size_t)nsys + 1), tsys, cuts;
  int tiles = 0;
  const TileConfig &c = h->tc[prec];
  for (int k = 0; k < nsys; ++k) {
    sd[k].g = grid_of(h, prec, k);
    sd[k].atom0 = (long long)h->sys[k].atom0;
    sd[k].block0 = (int)h->sys[k].block0;
    sd[k].tile0 = tiles;
    for (int a = 0; a < 3; ++a) {
      sd[k].cut[a] = -1;
      if (c.balance && (size_t)k < c.cuts.size() && !c.cuts[k][a].empty()) {
        sd[k].cut[a] = (int)cuts.size();
        cuts.insert(cuts.end(), c.cuts[k][a].begin(), c.cuts[k][a].end());
      }
    }
    seg[k] = tiles;
    const Tri &T = h->tc[prec].T[k];
    const int nt = T[0] * T[1] * T[2];
    for (int i = 0; i < nt; ++i) tsys.push_back(k);
    tiles += nt;
  }
  seg[nsys] = tiles;
  if (!t.sys && dalloc(&t.sys, nsys)) return EXAMD_ERR_NOMEM;
  if (!t.tile_seg && dalloc(&t.tile_seg, nsys + 1)) return EXAMD_ERR_NOMEM;
  if (!t.xcd_start && dalloc(&t.xcd_start, kXcds + 1)) return EXAMD_ERR_NOMEM;
  if (!t.xcd_next) {
    // the counters start at zero; each force launch leaves them at zero
    if (dalloc(&t.xcd_next, kXcds + 1)) return EXAMD_ERR_NOMEM;
    HIPCHK(hipMemset(t.xcd_next, 0, (kXcds + 1) * sizeof(int)));
  }
  // each XCD's share of the tiles: the planned equal-cost ranges, or equal
  // counts for a map planned without block counts
  std::vector<int> xr = c.xcd;
  if ((int)xr.size() != kXcds + 1 || xr[kXcds] != tiles) {
    xr.assign(kXcds + 1, 0);
    for (int x = 0; x <= kXcds; ++x) xr[x] = (int)((int64_t)tiles * x / kXcds);
  }
  HIPCHK(hipMemcpy(t.xcd_start, xr.data(), (kXcds + 1) * sizeof(int),
                   hipMemcpyHostToDevice));
  if (t.tile_cap < tiles) {
    if (t.tile_sys) (void)hipFree(t.tile_sys);
    t.tile_sys = nullptr;
    t.tile_cap = 0;
    if (dalloc(&t.tile_sys, tiles)) return EXAMD_ERR_NOMEM;
    t.tile_cap = tiles;
  }
  HIPCHK(hipMemcpy(t.sys, sd.data(), nsys * sizeof(SysDesc),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t.tile_seg, seg.data(), (nsys + 1) * sizeof(int),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(t.tile_sys, tsys.data(), tiles * sizeof(int),
                   hipMemcpyHostToDevice));
  if (!cuts.empty()) {
    if (t.cut_cap < (int)cuts.size()) {
      if (t.cuts) (void)hipFree(t.cuts);
      t.cuts = nullptr;
      t.cut_cap = 0;
      if (dalloc(&t.cuts, (int64_t)cuts.size())) return EXAMD_ERR_NOMEM;
      t.cut_cap = (int)cuts.size();
    }
    HIPCHK(hipMemcpy(t.cuts, cuts.data(), cuts.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  }
  t.ntiles = tiles;
  // the force kernel writes one partial tally per tile: every path that
  // changes the maps comes through here, so the slots are sized here
  return ensure_partials(h);
}

// Partial-tally slots for the current tile maps.
int ensure_partials(examd_handle *h) {
  int64_t need = partial_slots(h->n, std::max(tiles_of(h->tc[0]), tiles_of(h->tc[1])));
  if (h->part_cap >= need) return EXAMD_OK;
  drop_graph(h);   // a captured chunk holds the old address
  if (h->d.part) (void)hipFree(h->d.part);
  h->d.part = nullptr;
  h->part_cap = 0;
  if (dalloc(&h->d.part, need)) return EXAMD_ERR_NOMEM;
  h->part_cap = need;
  return EXAMD_OK;
}

// StepLog entries for `steps` + 1 steps of every system.
int ensure_log(examd_handle *h, int64_t steps) {
  const int64_t n = (steps + 1) * h->nsys;
  if (h->log_cap >= n) return EXAMD_OK;
  drop_graph(h);   // a captured chunk holds the old address
  if (h->log) (void)hipFree(h->log);
  h->log = nullptr;
  h->log_cap = 0;
  if (dalloc(&h->log, n)) return EXAMD_ERR_NOMEM;
  h->log_cap = n;
  return EXAMD_OK;
}

// Boundaries of the even split of nb blocks into T tiles.
std::vector<int> even_cuts(int nb, int T) {
  std::vector<int> b((size_t)T + 1);
  for (int t = 0; t < T; ++t) {
    int lo, hi;
    tile_range(t, nb, T, lo, hi);
    b[t] = lo;
    b[t + 1] = hi;
  }
  return b;
}

// Cell-block boundaries of system k's map in one precision.
Cuts cuts_of(const examd_handle *h, int prec, int k, const Grid &g) {
  const TileConfig &c = h->tc[prec];
  Cuts b;
  for (int a = 0; a < 3; ++a) {
    if (c.balance && (size_t)k < c.cuts.size() && !c.cuts[k][a].empty())
      b[a] = c.cuts[k][a];
    else
      b[a] = even_cuts(g.nb[a], g.T[a]);
  }
  return b;
}

// f(i0, i1) over [0, n) in contiguous pieces on up to 16 host threads, one
// thread when n is below min_n.
template <class F>
void host_parallel_for(int64_t n, F f, int64_t min_n = (int64_t)1 << 16) {
  unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < min_n) t = 1;
  if ((int64_t)t > n) t = (unsigned)std::max<int64_t>(n, 1);
  std::vector<std::thread> th;
  for (unsigned k = 1; k < t; ++k)
    th.emplace_back(f, n * k / t, n * (k + 1) / t);
  f((int64_t)0, n / t);
  for (std::thread &x : th) x.join();
}

// x y z per atom between a caller's array and dev_stage, through the pinned
// staging buffer; the caller's array is copied on several threads.  The
// stream is synchronised before the staging buffer is refilled.
int upload_aos(examd_handle *h, const double *src) {
  const int64_t m = 3 * h->n;
  HIPCHK(hipStreamSynchronize(h->stream));
  host_parallel_for(m, [&](int64_t i0, int64_t i1) {
    std::memcpy(h->host_stage + i0, src + i0, (size_t)(i1 - i0) * sizeof(double));
  });
  HIPCHK(hipMemcpyAsync(h->dev_stage, h->host_stage, m * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
  return EXAMD_OK;
}
// The download goes in pieces, each followed by an event: the host copies
// piece k out of the stage while the device still sends the later ones.
int download_aos(examd_handle *h, double *dst) {
  const int64_t m = 3 * h->n;
  constexpr int kPieces = 4;
  int64_t off[kPieces + 1];
  for (int k = 0; k <= kPieces; ++k) off[k] = m * k / kPieces;
  for (int k = 0; k < kPieces; ++k) {
    if (!h->piece_ev[k])
      HIPCHK(hipEventCreateWithFlags(&h->piece_ev[k], hipEventDisableTiming));
    if (off[k + 1] > off[k])
      HIPCHK(hipMemcpyAsync(h->host_stage + off[k], h->dev_stage + off[k],
                            (off[k + 1] - off[k]) * sizeof(double),
                            hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipEventRecord(h->piece_ev[k], h->stream));
  }
  for (int k = 0; k < kPieces; ++k) {
    HIPCHK(hipEventSynchronize(h->piece_ev[k]));
    host_parallel_for(off[k + 1] - off[k], [&](int64_t i0, int64_t i1) {
      std::memcpy(dst + off[k] + i0, h->host_stage + off[k] + i0,
                  (size_t)(i1 - i0) * sizeof(double));
    });
  }
  return EXAMD_OK;
}

// Atoms per block of every system for host positions, in the block order
// of the binning kernel (x fastest, systems one after another), in one pass
// on several threads: each atom's block found and counted (relaxed atomic
// increments: atoms of one block are rare in one thread's range).  With a
// staging buffer the positions are copied into it in the same pass, and a
// non-finite coordinate makes the call return false (its atom is not
// counted).
bool scan_positions(const examd_handle *h, const double *pos, double *stage,
                    std::vector<int> &cnt) {
  cnt.assign((size_t)std::max<int64_t>(h->nblocks, 1), 0);
  std::atomic<bool> finite(true);
  host_parallel_for(h->n, [&](int64_t i0, int64_t i1) {
    if (stage)
      std::memcpy(stage + 3 * i0, pos + 3 * i0, (size_t)(i1 - i0) * 3 * sizeof(double));
    bool ok = true;
    for (const SysHost &S : h->sys) {
      const int64_t a0 = std::max(i0, S.atom0), a1 = std::min(i1, S.atom0 + S.n);
      const Grid &g = S.g;
      for (int64_t i = a0; i < a1; ++i) {
        const double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        if (!(std::isfinite(x) && std::isfinite(y) && std::isfinite(z))) {
          ok = false;
          continue;
        }
        const int64_t b = S.block0 + host_block(x, g.bl, g.nb[0]) +
                          (int64_t)g.nb[0] * (host_block(y, g.bl, g.nb[1]) +
                                              (int64_t)g.nb[1] * host_block(z, g.bl, g.nb[2]));
        __atomic_fetch_add(&cnt[(size_t)b], 1, __ATOMIC_RELAXED);
      }
    }
    if (!ok) finite.store(false, std::memory_order_relaxed);
  });
  return finite.load();
}

std::vector<int> host_counts(const examd_handle *h, const double *pos) {
  std::vector<int> cnt;
  scan_positions(h, pos, nullptr, cnt);
  return cnt;
}

// Cell-block map of a system from a tile edge in blocks.
Tri tiles_from_edge(const Grid &g, double edge) {
  Tri T;
  for (int a = 0; a < 3; ++a)
    T[a] = std::max(1, std::min(g.nb[a], (int)std::lround(g.nb[a] / edge)));
  return T;
}

// Capacity a region of mra atoms needs at the least: 5% headroom.
int64_t needed_cap(int64_t mra) {
  return (int64_t)std::ceil((double)mra * 1.05) + 16;
}

// Pick the LDS capacity of a cell block for one precision from the largest
// region of the given block counts (over all systems), with cap_slack_pct
// headroom for density fluctuations during a run (at least 5%).  When the
// maps were chosen automatically and a region does not fit the device's
// LDS, refine the map of the system with the fullest region along its
// coarsest axis.  On success uploads the system table (when `upload`) and
// reports the largest region in *mra_out.
int choose_caps_counts(examd_handle *h, int prec, const int *cnt,
                       int64_t *mra_out, bool upload) {
  TileConfig &c = h->tc[prec];
  const bool fp64 = prec == EXAMD_PREC_FP64;
  const bool typed = h->d.ntypes > 0;
  const int nsys = h->nsys;
  // maps from a tile edge are planned from the edge every time (a batch's
  // coarsened maps would otherwise only ever get coarser), so the result
  // depends on the counts alone, not on the calls before
  if (c.edge > 0)
    for (int k = 0; k < nsys; ++k)
      c.T[k] = tiles_from_edge(h->sys[k].g, (double)c.edge);
  // the region atoms of every tile of the current maps in launch order
  // (systems in order, x fastest), kept for the split of the launch over the
  // XCDs; returns system k's largest
  std::vector<int64_t> cost;
  auto add_costs = [&](int k, const Grid &g, const Cuts &cb) {
    const size_t t0 = cost.size();
    cost.resize(t0 + (size_t)g.T[0] * g.T[1] * g.T[2]);
    int64_t m = 0;
    for_region_atoms(g, cb, cnt + h->sys[k].block0,
                     [&](int tx, int ty, int tz, int64_t v) {
                       cost[t0 + tx + (size_t)g.T[0] * (ty + (size_t)g.T[1] * tz)] = v;
                       m = std::max(m, v);
                     });
    return m;
  };
  for (int iter = 0; iter < 4096; ++iter) {
    bool ok = true;
    int64_t mra = 0;
    int worst = 0;
    c.maxrb = 0;
    cost.clear();
    if (c.balance) c.cuts.resize((size_t)nsys);
    for (int k = 0; k < nsys; ++k) {
      const Grid g = grid_of(h, prec, k);
      const SysHost &S = h->sys[k];
      if (c.balance) c.cuts[k] = balanced_cuts(g, cnt + S.block0, S.n);
      const Cuts cb = cuts_of(h, prec, k, g);
      c.maxrb = std::max(c.maxrb, max_region_blocks(cb));
      // region coordinates are packed 8 bits per axis on the device
      for (int a = 0; a < 3; ++a)
        for (int t = 0; t < g.T[a]; ++t)
          if (cb[a][t + 1] - cb[a][t] + 2 > 255) ok = false;
      if (!ok) break;
      const int64_t m = add_costs(k, g, cb);
      if (m > mra) {
        mra = m;
        worst = k;
      }
    }
    const double slack =
        (h->p.cap_slack_pct > 0 ? h->p.cap_slack_pct : 15) / 100.0;
    int64_t want = ((int64_t)std::ceil((double)mra * (1.0 + slack)) + 32 + 15) / 16 * 16;
    int64_t need = needed_cap(mra);
    int64_t cap = want;
    if (ok) {
      while (cap > need &&
             force_lds_bytes(fp64, (int)cap, c.maxrb, c.threads, c.lane_queue,
                             typed) > h->lds_limit)
        cap -= 16;
      if (force_lds_bytes(fp64, (int)cap, c.maxrb, c.threads, c.lane_queue,
                          typed) > h->lds_limit ||
          cap > 65535)
        ok = false;
    }
    if (ok) {
      c.cap = (int)cap;
      // a batch's maps from a tile edge: every system's map coarsened while
      // its largest region fits the LDS budget of one workgroup (the whole
      // LDS for 1024 threads, one workgroup per CU; a proportional share
      // for smaller workgroups, several per CU; never less than the largest
      // region now needs).  The limits in atoms and in region blocks come
      // from that budget, not from the largest region itself: a small
      // system whose one cell block spans it, periodic images included,
      // has a region of twice its atoms and far more blocks than its
      // neighbours' tiles, and would otherwise hold every map back.
      if (c.edge > 0 && nsys > 1) {
        const size_t now = force_lds_bytes(fp64, (int)cap, c.maxrb, c.threads,
                                           c.lane_queue, typed);
        const size_t budget =
            std::max(now, h->lds_limit * (size_t)c.threads / 1024);
        auto cap_of = [&](int64_t a) {
          return ((int64_t)std::ceil((double)a * (1.0 + slack)) + 32 + 15) / 16 * 16;
        };
        auto fits = [&](int64_t a, int64_t b) {
          const int64_t ca = cap_of(a);
          return ca <= 65535 &&
                 force_lds_bytes(fp64, (int)ca, (int)b, c.threads, c.lane_queue,
                                 typed) <= budget;
        };
        int64_t atom_lim = mra, blk_lim = c.maxrb;
        if (fits(atom_lim, blk_lim)) {
          // at least one atom per region block for the denser systems, then
          // whatever the LDS leaves for the blocks of the dilute ones
          while (fits(atom_lim + 16, std::max(blk_lim, atom_lim + 16))) {
            atom_lim += 16;
            blk_lim = std::max(blk_lim, atom_lim);
          }
          while (fits(atom_lim, blk_lim + 16)) blk_lim += 16;
        }
        host_parallel_for(nsys, [&](int64_t k0, int64_t k1) {
          for (int64_t k = k0; k < k1; ++k) {
            const SysHost &S = h->sys[k];
            const int *sc = cnt + S.block0;
            Grid g = grid_of(h, prec, (int)k);
            coarsen_map(
                g,
                [&](const Grid &t) {
                  if (c.balance) return balanced_cuts(t, sc, S.n);
                  Cuts b;
                  for (int a = 0; a < 3; ++a) b[a] = even_cuts(t.nb[a], t.T[a]);
                  return b;
                },
                sc, atom_lim, (int)blk_lim);
            for (int a = 0; a < 3; ++a) c.T[k][a] = g.T[a];
            if (c.balance) c.cuts[k] = balanced_cuts(g, sc, S.n);
          }
        }, 1);
        // the launch's largest region after coarsening sets the capacity
        c.maxrb = 0;
        mra = 0;
        cost.clear();
        for (int k = 0; k < nsys; ++k) {
          const Grid g = grid_of(h, prec, k);
          const Cuts cb = cuts_of(h, prec, k, g);
          c.maxrb = std::max(c.maxrb, max_region_blocks(cb));
          mra = std::max(mra, add_costs(k, g, cb));
        }
        int64_t cap2 = std::max<int64_t>(cap, cap_of(mra));
        while (cap2 > std::max<int64_t>(cap, needed_cap(mra)) &&
               force_lds_bytes(fp64, (int)cap2, c.maxrb, c.threads, c.lane_queue,
                               typed) > h->lds_limit)
          cap2 -= 16;
        c.cap = (int)cap2;
      }
      // the tiles in launch order split into eight contiguous ranges of
      // about equal region atoms, one per XCD
      c.xcd = equal_cost_ranges(cost, kXcds);
      // about three work items per lane: run groups from the mean region
      double reg_atoms = 0.0, reg_tiles = 0.0;
      for (int k = 0; k < nsys; ++k) {
        const Grid g = grid_of(h, prec, k);
        double reg = (double)h->sys[k].n / (g.L[0] * g.L[1] * g.L[2]);
        for (int a = 0; a < 3; ++a) reg *= ((double)g.nb[a] / g.T[a] + 2.0) * g.bl;
        const double nt = (double)g.T[0] * g.T[1] * g.T[2];
        reg_atoms += reg * nt;
        reg_tiles += nt;
      }
      const double reg = reg_tiles > 0.0 ? reg_atoms / reg_tiles : 1.0;
      int ng = c.fixed_groups > 0
                   ? c.fixed_groups
                   : (int)std::lround(3.0 * c.threads / std::max(reg, 1.0));
      c.groups = std::max(1, std::min(4, ng));
      if (mra_out) *mra_out = mra;
      return upload ? upload_tables(h, prec) : EXAMD_OK;
    }
    if (c.fixed_tiles) return EXAMD_ERR_TILE_OVERFLOW;
    Tri &T = c.T[worst];
    const Grid &g = h->sys[worst].g;
    int ax = -1;
    double best = 0.0;
    for (int a = 0; a < 3; ++a) {
      double size = (double)g.nb[a] / T[a];
      if (T[a] < g.nb[a] && size > best) {
        best = size;
        ax = a;
      }
    }
    if (ax < 0) return EXAMD_ERR_TILE_OVERFLOW;
    T[ax] += 1;
  }
  return EXAMD_ERR_TILE_OVERFLOW;
}

int choose_caps(examd_handle *h, int prec, const double *pos) {
  const std::vector<int> cnt = host_counts(h, pos);
  return choose_caps_counts(h, prec, cnt.data(), nullptr, true);
}

// The block counts of the last binning pass on the host (4 B per block,
// through a pinned buffer).
int device_block_counts(examd_handle *h, std::vector<int> &cnt) {
  if (!h->pinned_counts &&
      hipHostMalloc(reinterpret_cast<void **>(&h->pinned_counts),
                    std::max<int64_t>(h->nblocks, 1) * sizeof(int)) != hipSuccess) {
    h->pinned_counts = nullptr;
    return EXAMD_ERR_NOMEM;
  }
  HIPCHK(hipMemcpyAsync(h->pinned_counts, h->d.count, h->nblocks * sizeof(int),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  cnt.assign(h->pinned_counts, h->pinned_counts + h->nblocks);
  if (cnt.empty()) cnt.push_back(0);
  return EXAMD_OK;
}

// Slots per block for blocks of up to `need` atoms: a power of two of at
// least 8 and need + 2 (a block that fills up later is given twice the
// slots during the run).  The list array only grows; a new one starts
// empty, so the next evaluation bins afresh.
int ensure_bin_slots(examd_handle *h, int64_t need) {
  int shift = kMinBinShift;
  while (((int64_t)1 << shift) < need + 2) ++shift;
  if (shift <= h->d.bin_shift) return EXAMD_OK;
  if ((h->nblocks << shift) >= ((int64_t)1 << 31)) return EXAMD_ERR_BIN_OVERFLOW;
  HIPCHK(hipStreamSynchronize(h->stream));
  int *p = nullptr;
  if (dalloc(&p, h->nblocks << shift)) return EXAMD_ERR_NOMEM;
  (void)hipFree(h->d.sorted);
  h->d.sorted = p;
  h->d.bin_shift = shift;
  drop_graph(h);
  h->forces_valid = -1;
  return EXAMD_OK;
}

// Re-partition the map of one precision from the block counts of the last
// binning pass: new balanced
// boundaries (balanced maps) and a new LDS capacity.  The old map stays when
// the new one would not fit, and when the new one has the same boundaries
// and the old capacity still holds the fullest region with 5% to spare, so
// that captured chunks are not captured again for nothing.  *changed: the
// map was replaced.
int repartition(examd_handle *h, int prec, bool *changed) {
  *changed = false;
  std::vector<int> cnt;
  const int stc = device_block_counts(h, cnt);
  if (stc != EXAMD_OK) return stc;
  const TileConfig saved = h->tc[prec];
  TileConfig &c = h->tc[prec];
  int64_t mra = 0;
  if (choose_caps_counts(h, prec, cnt.data(), &mra, false) != EXAMD_OK) {
    c = saved;
    return EXAMD_OK;
  }
  if (c.T == saved.T && c.cuts == saved.cuts && c.maxrb == saved.maxrb &&
      c.groups == saved.groups && saved.cap >= needed_cap(mra)) {
    c = saved;
    return EXAMD_OK;
  }
  int st = upload_tables(h, prec);
  if (st == EXAMD_OK) st = ensure_partials(h);
  *changed = st == EXAMD_OK;
  return st;
}

// After a chunk whose largest cell-block region outgrew the LDS capacity:
// a new capacity planned from the block counts of the chunk's last binning
// pass (with the configured slack) and at least 5/4 of the old one per
// attempt, or a finer map when the map was chosen automatically.  Leaves the
// map unchanged and returns EXAMD_ERR_TILE_OVERFLOW when no capacity fits.
int regrow(examd_handle *h, int prec, int attempt) {
  std::vector<int> cnt;
  const int stc = device_block_counts(h, cnt);
  if (stc != EXAMD_OK) return stc;
  const TileConfig saved = h->tc[prec];
  TileConfig &c = h->tc[prec];
  if (choose_caps_counts(h, prec, cnt.data(), nullptr, false) != EXAMD_OK) {
    c = saved;
    return EXAMD_ERR_TILE_OVERFLOW;
  }
  // a grown capacity when the map is the same one
  if (c.T == saved.T && c.cuts == saved.cuts) {
    int64_t want = (int64_t)saved.cap;
    for (int a = 0; a < attempt; ++a) want = want * 5 / 4;
    want = (want + 15) / 16 * 16;
    int64_t cap = std::max<int64_t>(c.cap, want);
    const bool fp64 = prec == EXAMD_PREC_FP64;
    while (cap > c.cap &&
           (cap > 65535 ||
            force_lds_bytes(fp64, (int)cap, c.maxrb, c.threads, c.lane_queue,
                            h->d.ntypes > 0) > h->lds_limit))
      cap -= 16;
    if (cap <= saved.cap) {
      c = saved;
      return EXAMD_ERR_TILE_OVERFLOW;
    }
    c.cap = (int)cap;
  }
  h->regrows += 1;
  return upload_tables(h, prec);
}

ForceConfig force_config(const examd_handle *h, int prec) {
  ForceConfig c;
  c.threads = h->tc[prec].threads;
  c.cap = h->tc[prec].cap;
  c.maxrb = h->tc[prec].maxrb;
  c.groups = h->tc[prec].groups;
  c.lane_queue = h->tc[prec].lane_queue;
  c.workgroups = h->tc[prec].workgroups;
  return c;
}

LJParams lj(const examd_handle *h) {
  return LJParams{h->p.epsilon, h->p.sigma, h->p.cutoff, h->p.mass};
}

hipEvent_t next_event(examd_handle *h) {
  if (h->ev_used == h->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->ev_pool.push_back(e);
  }
  return h->ev_pool[h->ev_used++];
}

// Record an event on the handle's stream.  Under stream capture HIP does
// not turn hipEventRecord into a graph node, so the node is added to the
// capture graph explicitly and made the stream's new dependency.
void record_event(examd_handle *h, hipEvent_t e) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess ||
      cs != hipStreamCaptureStatusActive) {
    (void)hipEventRecord(e, h->stream);
    return;
  }
  hipGraph_t g = nullptr;
  const hipGraphNode_t *deps = nullptr;
  size_t nd = 0;
  if (hipStreamGetCaptureInfo_v2(h->stream, &cs, nullptr, &g, &deps, &nd) !=
      hipSuccess || !g)
    return;
  hipGraphNode_t node = nullptr;
  if (hipGraphAddEventRecordNode(&node, g, deps, nd, e) != hipSuccess) return;
  (void)hipStreamUpdateCaptureDependencies(h->stream, &node, 1,
                                           hipStreamSetCaptureDependencies);
}

// Force launch (tallies of system k into log[k]), bracketed by events when
// timing is on.
void forces_timed(examd_handle *h, int prec, bool energy, bool homecount,
                  StepLog *log, bool sample = true) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (h->timing && sample) {
    e0 = next_event(h);
    e1 = next_event(h);
    if (e0) record_event(h, e0);
  }
  launch_forces(h->d, table(h, prec), lj(h), force_config(h, prec),
                prec == EXAMD_PREC_MIXED, energy, homecount, log, h->stream);
  if (e1) record_event(h, e1);
}

// Accumulate the recorded event pairs (stream already synchronised).
void collect_timing(examd_handle *h) {
  for (size_t k = 0; k + 1 < h->ev_used; k += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev_pool[k], h->ev_pool[k + 1]) ==
        hipSuccess) {
      h->force_ms += ms;
      h->force_launches += 1;
    }
  }
  h->ev_used = 0;
}

// Bin + sort + force evaluation at the current positions into log[0..nsys).
void evaluate(examd_handle *h, int prec, bool energy, bool homecount,
              StepLog *log) {
  launch_step_bin(h->d, table(h, prec), prec == EXAMD_PREC_MIXED, kBinOnly,
                  0.0, 0.0, nullptr, log, h->stream);
  launch_bin_update(h->d, true, prec == EXAMD_PREC_FP64 && h->d.bin_shift > 4, log, h->stream);
  forces_timed(h, prec, energy, homecount, log);
}

// Kinetic energy of the current velocities into log[0..nsys) (h = 0 leaves
// them unchanged).
void kinetic(examd_handle *h, int prec, StepLog *log) {
  if (h->nsys == 1)
    launch_kick(h->d, prec == EXAMD_PREC_MIXED, 0.0, log, h->stream);
  else
    launch_sys_ke(h->d, table(h, prec), log, h->stream);
}

// Enqueue `count` velocity-Verlet steps on the handle's stream, tallies of
// chunk step c in log[c * nsys + k]: kick-drift-bin, sort, force; then
// (finish-kick-drift-bin, sort, force) x (count - 1); then the closing kick.
// A batch takes its per-system kinetic energies between a separate closing
// kick and the next kick-drift-bin (the same arithmetic as the fused form).
// Chunks of any length compose into the same arithmetic as one long run.
template <class IsEnergy>
void enqueue_steps(examd_handle *h, int prec, int64_t count, double dt,
                   IsEnergy is_energy) {
  const bool mixed = prec == EXAMD_PREC_MIXED;
  const double hk = (0.5 * dt) / h->p.mass;
  const int ns = h->nsys;
  const SysTable t = table(h, prec);
  launch_fill_log(h->log + ns, (int)(count * ns), h->stream);
  for (int64_t c = 1; c <= count; ++c) {
    StepLog *Lc = h->log + c * ns;
    StepLog *Lp = h->log + (c - 1) * ns;
    if (c == 1) {
      launch_step_bin(h->d, t, mixed, kKickDriftBin, hk, dt, nullptr, Lc,
                      h->stream);
    } else if (ns == 1 || !is_energy(c - 1)) {
      launch_step_bin(h->d, t, mixed, kFinishKickDriftBin, hk, dt,
                      is_energy(c - 1) ? Lp : nullptr, Lc, h->stream);
    } else {
      launch_kick(h->d, mixed, hk, nullptr, h->stream);
      launch_sys_ke(h->d, t, Lp, h->stream);
      launch_step_bin(h->d, t, mixed, kKickDriftBin, hk, dt, nullptr, Lc,
                      h->stream);
    }
    if (!mixed && h->d.bin_shift > 4) launch_bin_update(h->d, false, true, Lc, h->stream);
    // with timing on, every timing_every-th step's force launch is
    // bracketed by events (mid-stride, away from the tally steps)
    const int k = h->timing_every;
    forces_timed(h, prec, is_energy(c), false, Lc, c % k == (k / 2) % k);
  }
  if (count > 0) {
    StepLog *Ln = h->log + count * ns;
    if (ns == 1) {
      launch_kick(h->d, mixed, hk, is_energy(count) ? Ln : nullptr, h->stream);
    } else {
      launch_kick(h->d, mixed, hk, nullptr, h->stream);
      if (is_energy(count)) launch_sys_ke(h->d, t, Ln, h->stream);
    }
  }
}

bool same_key(const GraphKey &a, const GraphKey &b) {
  if (a.prec != b.prec || a.chunk != b.chunk || a.every != b.every ||
      a.dt != b.dt || a.timing != b.timing ||
      a.timing_every != b.timing_every)
    return false;
  const TileConfig &x = a.cfg, &y = b.cfg;
  return x.T == y.T && x.threads == y.threads && x.groups == y.groups &&
         x.lane_queue == y.lane_queue && x.workgroups == y.workgroups &&
         x.balance == y.balance && x.cuts == y.cuts && x.cap == y.cap &&
         x.maxrb == y.maxrb;
}

// Capture one chunk of `key.chunk` steps into an executable graph (reused
// while the key matches).  Event records of the force timing are captured
// too, so replays report each force launch.
int ensure_graph(examd_handle *h, const GraphKey &key) {
  if (h->graph_exec && same_key(h->graph_key, key)) return EXAMD_OK;
  drop_graph(h);
  h->ev_used = 0;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  const int64_t every = key.every;
  const int64_t chunk = key.chunk;
  enqueue_steps(h, key.prec, chunk, key.dt, [&](int64_t c) {
    return c == chunk || (every > 0 && c % every == 0);
  });
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    h->ev_used = 0;
    return EXAMD_ERR_HIP;
  }
  hipGraphExec_t ge = nullptr;
  if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    h->ev_used = 0;
    return EXAMD_ERR_HIP;
  }
  h->graph = g;
  h->graph_exec = ge;
  h->graph_key = key;
  h->graph_events = h->ev_used;
  return EXAMD_OK;
}

void fill_stats(const examd_handle *h, const StepLog &L, int prec,
                int64_t step, examd_stats *s) {
  double pe = prec == EXAMD_PREC_MIXED
                  ? (double)(long long)L.pe_fix * kFixInv
                  : L.pe;
  double rc = h->p.cutoff, sr2 = h->p.sigma * h->p.sigma / (rc * rc);
  double sr6 = sr2 * sr2 * sr2;
  double vrc = 4.0 * h->p.epsilon * (sr6 * sr6 - sr6);
  s->step = step;
  s->ke = 0.5 * h->p.mass * L.ke;
  s->pe = pe;
  // several species: the V(rc) of each pair's own row, tallied on device
  if (h->d.tab_n > 0) vrc = h->d.tab_v_last;   // the table's V(rc)
  s->pe_shifted = h->d.ntypes > 0 ? pe - (double)(long long)L.shift_fix * kFixInv
                                  : pe - (double)L.cut * vrc;
  s->etot = s->ke + pe;
  s->pairs_cand = (int64_t)L.cand;
  s->pairs_cut = (int64_t)L.cut;
}

int download_positions(examd_handle *h, std::vector<double> &aos) {
  const int64_t n = h->n;
  aos.assign((size_t)std::max<int64_t>(3 * n, 1), 0.0);
  if (ensure_stage(h)) return EXAMD_ERR_NOMEM;
  launch_soa_to_aos(h->d.x, h->d.y, h->d.z, h->dev_stage, n, h->stream);
  return download_aos(h, aos.data());
}

// Apply a launch shape to one precision: the cell-block map (explicit T of
// a single system, or every system's map from a tile edge), threads,
// groups, queue depth and persistent grid.
int apply_config(examd_handle *h, int prec, const int *T, double edge,
                 int threads, int groups, int lane_queue, int workgroups,
                 bool balance, const double *pos) {
  TileConfig saved = h->tc[prec];
  TileConfig &c = h->tc[prec];
  c.lane_queue = std::max(0, std::min(16, lane_queue));
  if (T) {
    for (int a = 0; a < 3; ++a)
      c.T[0][a] = std::max(1, std::min(T[a], h->sys[0].g.nb[a]));
    c.edge = 0;
  } else {
    for (int k = 0; k < h->nsys; ++k) c.T[k] = tiles_from_edge(h->sys[k].g, edge);
    c.edge = (int)std::lround(edge);
  }
  c.threads = threads;
  c.workgroups = std::max(0, workgroups);
  c.balance = balance;
  c.cuts.clear();
  c.fixed_tiles = true;
  c.fixed_groups = groups;
  int st = choose_caps(h, prec, pos);
  if (st == EXAMD_OK) st = ensure_partials(h);
  if (st != EXAMD_OK) {
    h->tc[prec] = saved;
    (void)upload_tables(h, prec);
  }
  if (h->forces_valid == prec) h->forces_valid = -1;
  return st;
}

uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t &s) {  // (0, 1)
  return ((double)(splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Maxwell-Boltzmann velocities at temperature T, zero net momentum,
// rescaled so that sum v^2 = (3n - 3) T / m exactly.
void mb_velocities(int64_t n, double temperature, double mass, uint64_t &s,
                   double *vel) {
  const double sd = std::sqrt(temperature / mass);
  for (int64_t t = 0; t < 3 * n; t += 2) {
    // Box-Muller pairs
    double u1 = uniform01(s), u2 = uniform01(s);
    double r = std::sqrt(-2.0 * std::log(u1));
    vel[t] = sd * r * std::cos(2.0 * M_PI * u2);
    if (t + 1 < 3 * n) vel[t + 1] = sd * r * std::sin(2.0 * M_PI * u2);
  }
  double mom[3] = {0, 0, 0};
  for (int64_t t = 0; t < n; ++t)
    for (int a = 0; a < 3; ++a) mom[a] += vel[3 * t + a];
  for (int a = 0; a < 3; ++a) mom[a] /= (double)n;
  double v2 = 0.0;
  for (int64_t t = 0; t < n; ++t)
    for (int a = 0; a < 3; ++a) {
      vel[3 * t + a] -= mom[a];
      v2 += vel[3 * t + a] * vel[3 * t + a];
    }
  double target = (3.0 * n - 3.0) * temperature / mass;  // sum v^2
  double scale = (v2 > 0.0 && n > 1) ? std::sqrt(target / v2) : 0.0;
  for (int64_t t = 0; t < 3 * n; ++t) vel[t] *= scale;
}

// Shared set-up of examd_hip_create and examd_hip_batch_create.
int create_handle(const examd_params *prm, const examd_system *systems,
                  int nsys, bool allow_tiles, examd_handle **out) {
  if (!prm || !out || !systems || nsys <= 0) return EXAMD_ERR_ARG;
  *out = nullptr;
  if (!(prm->cutoff > 0.0) || !(prm->mass > 0.0) || !(prm->sigma > 0.0) ||
      !(prm->epsilon >= 0.0))
    return EXAMD_ERR_ARG;
  std::vector<SysHost> sys((size_t)nsys);
  int64_t atoms = 0, blocks = 0;
  for (int k = 0; k < nsys; ++k) {
    if (systems[k].natoms < 0) return EXAMD_ERR_ARG;
    int32_t nbxyz[3];
    int64_t nb;
    int st = examd_hip_domain(systems[k].box, 0.5 * prm->cutoff, nbxyz, &nb);
    if (st) return st;
    for (int a = 0; a < 3; ++a)
      if (nbxyz[a] < 5 || !(prm->cutoff < 0.5 * systems[k].box[a]))
        return EXAMD_ERR_GRID;
    SysHost &S = sys[k];
    S.atom0 = atoms;
    S.n = systems[k].natoms;
    S.block0 = blocks;
    S.nblocks = nb;
    for (int a = 0; a < 3; ++a) {
      S.g.nb[a] = nbxyz[a];
      S.g.T[a] = 1;
      S.g.L[a] = systems[k].box[a];
    }
    S.g.bl = 0.5 * prm->cutoff;
    atoms += S.n;
    blocks += nb;
    // every block's slots are addressed with 32-bit ints
    if (atoms > (int64_t)2000000000 || blocks > ((int64_t)1 << 31) / kMinBinSlots - 1)
      return EXAMD_ERR_ARG;
  }
  const bool fixed = prm->tiles[0] > 0 || prm->tiles[1] > 0 || prm->tiles[2] > 0;
  if (fixed && !allow_tiles) return EXAMD_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return EXAMD_ERR_NO_DEVICE;
  if (prm->device < 0 || prm->device >= ndev) return EXAMD_ERR_ARG;

  examd_handle *h = new (std::nothrow) examd_handle();
  if (!h) return EXAMD_ERR_NOMEM;
  h->p = *prm;
  h->dev = prm->device;
  h->nsys = nsys;
  h->sys = sys;
  h->n = atoms;
  h->nblocks = blocks;
  const int thr = prm->block_threads;
  for (int prec = 0; prec < 2; ++prec) {
    TileConfig &c = h->tc[prec];
    c.fixed_tiles = fixed;
    c.fixed_groups = prm->run_groups;
    c.threads = (thr == 256 || thr == 512 || thr == 1024) ? thr : 512;
    c.T.resize((size_t)nsys);
    for (int k = 0; k < nsys; ++k)
      for (int a = 0; a < 3; ++a) {
        const int nb = sys[k].g.nb[a];
        int T = prm->tiles[a] > 0 ? prm->tiles[a] : (nb + 6) / 7;
        c.T[k][a] = std::max(1, std::min(T, nb));
      }
  }
  h->graph_chunk = prm->graph_chunk > 0 ? prm->graph_chunk : 0;

  if (hipSetDevice(h->dev) != hipSuccess) {
    delete h;
    return EXAMD_ERR_HIP;
  }
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            h->dev) != hipSuccess || lds <= 0)
    lds = 65536;
  // keep room for the kernels' static LDS (reduction scratch)
  h->lds_limit = (size_t)lds > 4096 ? (size_t)lds - 4096 : (size_t)lds;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) !=
      hipSuccess) {
    delete h;
    return EXAMD_ERR_HIP;
  }
  DevState &d = h->d;
  d.n = h->n;
  const int64_t n = h->n;
  int rc = EXAMD_OK;
  rc |= dalloc(&d.x, n) | dalloc(&d.y, n) | dalloc(&d.z, n);
  rc |= dalloc(&d.vx, n) | dalloc(&d.vy, n) | dalloc(&d.vz, n);
  rc |= dalloc(&d.fx, n) | dalloc(&d.fy, n) | dalloc(&d.fz, n);
  rc |= dalloc(&d.qx, n) | dalloc(&d.qy, n) | dalloc(&d.qz, n);
  d.nblocks = blocks;
  d.bin_shift = kMinBinShift;
  rc |= dalloc(&d.cell, n) | dalloc(&d.slot, n);
  rc |= dalloc(&d.sorted, blocks << kMinBinShift) | dalloc(&d.count, blocks);
  rc |= dalloc(&d.bflag, bin_flag_bytes(blocks));
  rc |= dalloc(&h->atom_seg, nsys + 1);
  const int64_t nslab = (std::max<int64_t>(n, 1) + kSysSlab - 1) / kSysSlab;
  if (nsys > 1) rc |= dalloc(&h->slab_sys, nslab);
  if (rc == EXAMD_OK) rc = ensure_log(h, 1);
  if (rc != EXAMD_OK) {
    free_state(h);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return EXAMD_ERR_NOMEM;
  }
  std::vector<long long> seg((size_t)nsys + 1);
  for (int k = 0; k < nsys; ++k) seg[k] = (long long)sys[k].atom0;
  seg[nsys] = (long long)atoms;
  hipError_t e = hipMemcpy(h->atom_seg, seg.data(), seg.size() * sizeof(long long),
                           hipMemcpyHostToDevice);
  if (nsys > 1 && e == hipSuccess) {
    std::vector<int> slab((size_t)nslab);
    for (int64_t k = 0, q = 0; k < nslab; ++k) {
      while (q + 1 < nsys && seg[q + 1] <= k * kSysSlab) ++q;
      slab[k] = (int)q;
    }
    e = hipMemcpy(h->slab_sys, slab.data(), slab.size() * sizeof(int),
                  hipMemcpyHostToDevice);
  }
  launch_fill_i32(d.count, 0, blocks, h->stream);
  launch_fill_i32(reinterpret_cast<int *>(d.bflag), 0, bin_flag_bytes(blocks) / 4,
                  h->stream);
  if (e != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) {
    free_state(h);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return EXAMD_ERR_HIP;
  }
  *out = h;
  return EXAMD_OK;
}

// ------------------------------------------------ allocation-free form ----
// examd_hip_forces_static keeps everything beyond its inputs and outputs in
// one static device array per GPU: the atoms' SoA positions, force sums,
// block ids and list slots, the block lists (16 slots per block at the
// block limit, more for fewer blocks), counts and flags, the per-cell-block
// partial tallies, the step log and the system descriptor.
constexpr int64_t kStaticAtoms = EXAMD_STATIC_MAX_ATOMS;
constexpr int64_t kStaticBlocks = EXAMD_STATIC_MAX_BLOCKS;
constexpr int64_t kStaticTiles = EXAMD_STATIC_MAX_CELL_BLOCKS;
// the block lists hold 16 slots per block at kStaticBlocks blocks; a call
// with fewer blocks gives each block the largest power of two that fits
// (at most 256)
constexpr int kStaticListShift = 4;
constexpr int64_t kStaticListInts = kStaticBlocks << kStaticListShift;
constexpr size_t scratch_align(size_t v) { return (v + 255) / 256 * 256; }
struct StaticLayout {
  size_t xyz, f, ints, sorted, count, bflag, part, log, sys, total;
};
constexpr StaticLayout static_layout() {
  StaticLayout L{};
  size_t o = 0;
  L.xyz = o;
  o = scratch_align(o + 3 * (size_t)kStaticAtoms * sizeof(double));
  L.f = o;   // FP64 sums or fixed-point sums, one precision per call
  o = scratch_align(o + 3 * (size_t)kStaticAtoms * sizeof(double));
  L.ints = o;   // cell, slot
  o = scratch_align(o + 2 * (size_t)kStaticAtoms * sizeof(int));
  L.sorted = o;
  o = scratch_align(o + (size_t)kStaticListInts * sizeof(int));
  L.count = o;
  o = scratch_align(o + (size_t)kStaticBlocks * sizeof(int));
  L.bflag = o;
  o = scratch_align(o + (size_t)bin_flag_bytes(kStaticBlocks));
  L.part = o;
  o = scratch_align(o + (size_t)std::max<int64_t>(kStaticTiles, kStaticAtoms / 256 + 1) *
                            sizeof(Partial));
  L.log = o;
  o = scratch_align(o + sizeof(StepLog));
  L.sys = o;
  o = scratch_align(o + sizeof(SysDesc));
  L.total = o;
  return L;
}
constexpr StaticLayout kStatic = static_layout();
static_assert(kStatic.total == EXAMD_STATIC_SCRATCH_BYTES,
              "EXAMD_STATIC_SCRATCH_BYTES must match the scratch layout");

// The array itself.  In host code its name stands for a host-side symbol,
// not a device address: host code reaches it only through the per-device
// addresses hipGetSymbolAddress returns (scratch_of).
static __device__ __align__(256) unsigned char scratch[kStatic.total];

constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_dev[kScratchDevices] = {};
size_t g_scratch_lds[kScratchDevices] = {};

// The device address of `scratch` on device dev and that device's LDS
// limit for the force kernel, looked up once per device with the device
// current (each device holds its own copy of the array).
int scratch_of(int dev, unsigned char **ptr, size_t *lds_limit) {
  if (dev < 0 || dev >= kScratchDevices) return EXAMD_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    HIPCHK(hipSetDevice(dev));
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return EXAMD_ERR_HIP;
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                              dev) != hipSuccess || lds <= 0)
      lds = 65536;
    g_scratch_lds[dev] = (size_t)lds > 4096 ? (size_t)lds - 4096 : (size_t)lds;
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  *ptr = g_scratch_dev[dev];
  *lds_limit = g_scratch_lds[dev];
  return EXAMD_OK;
}

}  // namespace

extern "C" {

const char *examd_hip_strerror(int s) {
  switch (s) {
    case EXAMD_OK: return "ok";
    case EXAMD_ERR_ARG: return "invalid argument";
    case EXAMD_ERR_HIP: return "HIP runtime error";
    case EXAMD_ERR_NOMEM: return "device allocation failed";
    case EXAMD_ERR_GRID:
      return "fewer than 5 blocks per dimension or cutoff >= box/2";
    case EXAMD_ERR_TILE_OVERFLOW:
      return "cell-block region exceeds its LDS capacity";
    case EXAMD_ERR_NO_DEVICE: return "no HIP device";
    case EXAMD_ERR_NONFINITE: return "non-finite position";
    case EXAMD_ERR_NOT_READY: return "state not set";
    case EXAMD_ERR_FIXED_RANGE:
      return "pair force beyond the fixed-point range";
    case EXAMD_ERR_ZERO_DISTANCE: return "pair distance^2 below 0.001";
    case EXAMD_ERR_TABLE_RANGE:
      return "pair closer than the tabulated pair function's range";
    case EXAMD_ERR_OUT_OF_BOX: return "block number out of range";
    case EXAMD_ERR_BIN_OVERFLOW:
      return "a block holds more atoms than its slots";
    case EXAMD_ERR_STATIC_LIMIT:
      return "system beyond the static scratch limits of "
             "examd_hip_forces_static";
    default: return "unknown status";
  }
}

int examd_hip_domain(const double box[3], double block_length,
                     int32_t nblocks_xyz[3], int64_t *nblocks) {
  if (!box || !nblocks_xyz || !nblocks || !(block_length > 0.0))
    return EXAMD_ERR_ARG;
  int64_t nb = 1;
  for (int a = 0; a < 3; ++a) {
    if (!(box[a] > 0.0)) return EXAMD_ERR_ARG;
    double q = box[a] / block_length;
    if (!(q < 2.0e9)) return EXAMD_ERR_ARG;
    nblocks_xyz[a] = (int32_t)q;  // Int(globalbox/block_length)
    nb *= nblocks_xyz[a];
  }
  *nblocks = nb;
  return EXAMD_OK;
}

int examd_hip_create(const examd_params *prm, examd_handle **out) {
  if (!prm || !out) return EXAMD_ERR_ARG;
  if (prm->natoms < 0 || prm->natoms > (int64_t)2000000000) return EXAMD_ERR_ARG;
  examd_system one;
  one.natoms = prm->natoms;
  for (int a = 0; a < 3; ++a) one.box[a] = prm->box[a];
  return create_handle(prm, &one, 1, true, out);
}

int examd_hip_batch_create(const examd_params *common,
                           const examd_system *systems, int32_t nsys,
                           examd_handle **out) {
  if (!common || !out) return EXAMD_ERR_ARG;
  if (common->natoms != 0 || common->box[0] != 0.0 || common->box[1] != 0.0 ||
      common->box[2] != 0.0)
    return EXAMD_ERR_ARG;
  return create_handle(common, systems, nsys, false, out);
}

int examd_hip_batch_info(examd_handle *h, int32_t *nsys, int64_t *natoms,
                         int64_t *nblocks) {
  if (!h) return EXAMD_ERR_ARG;
  if (nsys) *nsys = h->nsys;
  if (natoms) *natoms = h->n;
  if (nblocks) *nblocks = h->nblocks;
  return EXAMD_OK;
}

int examd_hip_destroy(examd_handle *h) {
  if (!h) return EXAMD_ERR_ARG;
  (void)hipSetDevice(h->dev);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  free_state(h);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  for (hipEvent_t e : h->piece_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return EXAMD_OK;
}

int examd_hip_set_state(examd_handle *h, const double *pos, const double *vel,
                        int64_t n) {
  if (!h || !pos || n != h->n) return EXAMD_ERR_ARG;
  HIPCHK(hipSetDevice(h->dev));
  if (ensure_stage(h)) return EXAMD_ERR_NOMEM;
  // one pass over the caller's positions: finiteness, the block counts and
  // the copy into the pinned staging buffer (which an earlier copy may
  // still be reading)
  HIPCHK(hipStreamSynchronize(h->stream));
  std::vector<int> cnt;
  if (!scan_positions(h, pos, h->host_stage, cnt)) return EXAMD_ERR_NONFINITE;
  // the positions travel to the device while the host plans the maps; the
  // device state changes only once the plans have succeeded
  HIPCHK(hipMemcpyAsync(h->dev_stage, h->host_stage, 3 * n * sizeof(double),
                        hipMemcpyHostToDevice, h->stream));
  {
    // both precisions' maps planned on two threads (each writes only its
    // own TileConfig), then their tables uploaded here.  A plan that fails
    // leaves both maps as they were (the device tables and the state on the
    // device still agree with them).
    const int maxocc = *std::max_element(cnt.begin(), cnt.end());
    const int stb = ensure_bin_slots(h, maxocc);
    if (stb != EXAMD_OK) return stb;
    const TileConfig saved[2] = {h->tc[0], h->tc[1]};
    int st1 = EXAMD_OK;
    std::thread other([&] {
      st1 = choose_caps_counts(h, 1, cnt.data(), nullptr, false);
    });
    const int st0 = choose_caps_counts(h, 0, cnt.data(), nullptr, false);
    other.join();
    if (st0 != EXAMD_OK || st1 != EXAMD_OK) {
      h->tc[0] = saved[0];
      h->tc[1] = saved[1];
      return st0 != EXAMD_OK ? st0 : st1;
    }
    for (int prec = 0; prec < 2; ++prec) {
      const int st = upload_tables(h, prec);
      if (st) {
        h->state_set = false;
        return st;
      }
    }
  }
  if (ensure_partials(h)) return EXAMD_ERR_NOMEM;
  drop_graph(h);
  // one copy per array in the host layout, transposed on the device
  launch_aos_to_soa(h->dev_stage, h->d.x, h->d.y, h->d.z, n, h->stream);
  int st = EXAMD_OK;
  if (vel) {
    if ((st = upload_aos(h, vel)) != EXAMD_OK) return st;
    launch_aos_to_soa(h->dev_stage, h->d.vx, h->d.vy, h->d.vz, n, h->stream);
  } else {
    HIPCHK(hipMemsetAsync(h->d.vx, 0, n * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(h->d.vy, 0, n * sizeof(double), h->stream));
    HIPCHK(hipMemsetAsync(h->d.vz, 0, n * sizeof(double), h->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->state_set = true;
  h->forces_valid = -1;
  h->since_repartition[0] = h->since_repartition[1] = 0;
  return EXAMD_OK;
}

int examd_hip_set_species(examd_handle *h, int32_t ntypes,
                          const int32_t *type, int64_t n, const double *eps,
                          const double *sigma, const double *cutoff) {
  if (!h || ntypes < 0 || ntypes > EXAMD_MAX_SPECIES) return EXAMD_ERR_ARG;
  DevState &d = h->d;
  if (ntypes > 0 && d.tab_n > 0) return EXAMD_ERR_ARG;   // not with a table
  std::vector<float4> p32;
  std::vector<Pair64> p64;
  std::vector<long long> vrc;
  std::vector<unsigned char> ty;
  double max_es12 = 0.0;
  if (ntypes > 0) {
    if (!type || n != h->n || !eps || !sigma || !cutoff) return EXAMD_ERR_ARG;
    const int nt = ntypes;
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        const int k = a * nt + b, kt = b * nt + a;
        if (!(std::isfinite(eps[k]) && eps[k] >= 0.0) ||
            !(std::isfinite(sigma[k]) && sigma[k] > 0.0) ||
            !(cutoff[k] > 0.0 && cutoff[k] <= h->p.cutoff) ||
            eps[k] != eps[kt] || sigma[k] != sigma[kt] ||
            cutoff[k] != cutoff[kt])
          return EXAMD_ERR_ARG;
      }
    ty.resize((size_t)std::max<int64_t>(n, 1));
    for (int64_t i = 0; i < n; ++i) {
      if (type[i] < 0 || type[i] >= nt) return EXAMD_ERR_ARG;
      ty[i] = (unsigned char)type[i];
    }
    p32.resize(kMaxSpecies * kMaxSpecies);
    p64.resize(kMaxSpecies * kMaxSpecies);
    vrc.resize(kMaxSpecies * kMaxSpecies);
    for (int k = 0; k < nt * nt; ++k) {
      // the same roundings as the one-species constants of launch_forces
      const double sig2 = sigma[k] * sigma[k], rc2 = cutoff[k] * cutoff[k];
      const double e24 = 24.0 * eps[k], e4 = 4.0 * eps[k];
      p64[k] = Pair64{sig2, e24, e4, rc2};
      p32[k] = make_float4((float)sig2, (float)e24 * 4294967296.0f,
                           (float)e4 * 4294967296.0f, (float)rc2);
      max_es12 = std::max(max_es12, eps[k] * std::pow(sigma[k], 12.0));
      const double sr2 = sig2 / rc2, sr6 = sr2 * sr2 * sr2;
      vrc[k] = std::llround(4.0 * eps[k] * (sr6 * sr6 - sr6) * 4294967296.0);
    }
  }
  HIPCHK(hipSetDevice(h->dev));
  HIPCHK(hipStreamSynchronize(h->stream));
  drop_graph(h);   // a captured chunk holds the old species state
  // the species bytes and the pair table take LDS: with positions set,
  // re-check the maps first, so that a failure changes nothing
  const int old_ntypes = d.ntypes;
  if (h->state_set && (ntypes > 0) != (old_ntypes > 0)) {
    std::vector<double> pos;
    const TileConfig saved[2] = {h->tc[0], h->tc[1]};
    int st = download_positions(h, pos);
    d.ntypes = ntypes;
    for (int prec = 0; st == EXAMD_OK && prec < 2; ++prec)
      st = choose_caps(h, prec, pos.data());
    if (st != EXAMD_OK) {
      // the maps and device tables as they were before this call
      d.ntypes = old_ntypes;
      for (int prec = 0; prec < 2; ++prec) {
        h->tc[prec] = saved[prec];
        (void)upload_tables(h, prec);
      }
      return st;
    }
  }
  if (ntypes > 0) {
    int st = EXAMD_OK;
    if (!d.type) st |= dalloc(&d.type, h->n);
    if (!d.pair32) st |= dalloc(&d.pair32, kMaxSpecies * kMaxSpecies);
    if (!d.pair64) st |= dalloc(&d.pair64, kMaxSpecies * kMaxSpecies);
    if (!d.pair_shift) st |= dalloc(&d.pair_shift, kMaxSpecies * kMaxSpecies);
    hipError_t e = hipSuccess;
    if (st == EXAMD_OK) {
      e = hipMemcpy(d.type, ty.data(), (size_t)n, hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(d.pair32, p32.data(), p32.size() * sizeof(float4),
                      hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(d.pair64, p64.data(), p64.size() * sizeof(Pair64),
                      hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = hipMemcpy(d.pair_shift, vrc.data(), vrc.size() * sizeof(long long),
                      hipMemcpyHostToDevice);
    }
    if (st != EXAMD_OK || e != hipSuccess) {
      // the tables may be half written: fall back to one species
      d.ntypes = 0;
      d.max_eps_sig12 = 0.0;
      h->forces_valid = -1;
      return st != EXAMD_OK ? EXAMD_ERR_NOMEM : EXAMD_ERR_HIP;
    }
  }
  d.ntypes = ntypes;
  d.max_eps_sig12 = max_es12;
  h->forces_valid = -1;
  return EXAMD_OK;
}

int examd_hip_set_pair_table(examd_handle *h, int32_t npoints, double r2_lo,
                             const double *g, const double *v) {
  if (!h || npoints < 0 || npoints == 1 || npoints > EXAMD_MAX_TABLE_POINTS)
    return EXAMD_ERR_ARG;
  DevState &d = h->d;
  const double rc2 = h->p.cutoff * h->p.cutoff;
  std::vector<float4> t32;
  std::vector<double> t64;
  double d2fix = 0.0;   // (mixed path) see launch_forces: 2^14 per pair
  if (npoints > 0) {
    if (d.ntypes > 0 || !g || !v || !(r2_lo >= 0.0 && r2_lo < rc2))
      return EXAMD_ERR_ARG;
    for (int k = 0; k < npoints; ++k) {
      if (!std::isfinite(g[k]) || !std::isfinite(v[k])) return EXAMD_ERR_ARG;
      const double r2 = r2_lo + (rc2 - r2_lo) * k / (npoints - 1);
      // a point with |f| >= 2^14: pairs below the next point are flagged
      if (std::fabs(g[k]) * std::sqrt(r2) >= 16384.0)
        d2fix = std::max(d2fix, r2_lo + (rc2 - r2_lo) * (k + 1) / (npoints - 1));
    }
    // per interval the cubic (in t, the position between the points) of
    // g and of v through the two points with the points' slopes: Hermite
    // interpolation, exact for cubics in r^2; slopes per point spacing by
    // the fourth-order differences (centred inside, one-sided at the two
    // ends), exact for quartics.  Fewer than five points: linear.
    const int ni = npoints - 1;
    auto slopes = [&](const double *y) {
      std::vector<double> m((size_t)npoints, 0.0);
      const int n = npoints;
      if (n < 5) return m;   // unused: the intervals are linear
      for (int k = 2; k < n - 2; ++k)
        m[k] = (y[k - 2] - 8.0 * y[k - 1] + 8.0 * y[k + 1] - y[k + 2]) / 12.0;
      m[0] = (-25.0 * y[0] + 48.0 * y[1] - 36.0 * y[2] + 16.0 * y[3] - 3.0 * y[4]) / 12.0;
      m[1] = (-3.0 * y[0] - 10.0 * y[1] + 18.0 * y[2] - 6.0 * y[3] + y[4]) / 12.0;
      m[n - 2] = (-y[n - 5] + 6.0 * y[n - 4] - 18.0 * y[n - 3] + 10.0 * y[n - 2] +
                  3.0 * y[n - 1]) / 12.0;
      m[n - 1] = (3.0 * y[n - 5] - 16.0 * y[n - 4] + 36.0 * y[n - 3] - 48.0 * y[n - 2] +
                  25.0 * y[n - 1]) / 12.0;
      return m;
    };
    auto cubic = [&](const double *y, const std::vector<double> &m, int k, double c[4]) {
      if (npoints < 5) {
        c[0] = y[k];
        c[1] = y[k + 1] - y[k];
        c[2] = c[3] = 0.0;
        return;
      }
      c[0] = y[k];
      c[1] = m[k];
      c[2] = 3.0 * (y[k + 1] - y[k]) - 2.0 * m[k] - m[k + 1];
      c[3] = 2.0 * (y[k] - y[k + 1]) + m[k] + m[k + 1];
    };
    const std::vector<double> mg = slopes(g), mv = slopes(v);
    t32.resize(2 * (size_t)ni);
    t64.resize(8 * (size_t)ni);
    constexpr double kS = 4294967296.0;   // the mixed path's 2^32 scale
    for (int k = 0; k < ni; ++k) {
      double *cg = &t64[8 * (size_t)k], *cv = cg + 4;
      cubic(g, mg, k, cg);
      cubic(v, mv, k, cv);
      t32[2 * (size_t)k] = make_float4((float)(cg[0] * kS), (float)(cg[1] * kS),
                                       (float)(cg[2] * kS), (float)(cg[3] * kS));
      t32[2 * (size_t)k + 1] = make_float4((float)(cv[0] * kS), (float)(cv[1] * kS),
                                           (float)(cv[2] * kS), (float)(cv[3] * kS));
    }
  }
  HIPCHK(hipSetDevice(h->dev));
  HIPCHK(hipStreamSynchronize(h->stream));
  drop_graph(h);   // a captured chunk holds the old pair function
  if (d.tab32) (void)hipFree(d.tab32);
  if (d.tab64) (void)hipFree(d.tab64);
  d.tab32 = nullptr;
  d.tab64 = nullptr;
  d.tab_n = 0;
  h->forces_valid = -1;
  if (npoints == 0) return EXAMD_OK;
  if (dalloc(&d.tab32, 2 * (npoints - 1)) || dalloc(&d.tab64, 8 * (npoints - 1))) {
    if (d.tab32) (void)hipFree(d.tab32);
    d.tab32 = nullptr;
    return EXAMD_ERR_NOMEM;
  }
  if (hipMemcpy(d.tab32, t32.data(), t32.size() * sizeof(float4),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d.tab64, t64.data(), t64.size() * sizeof(double),
                hipMemcpyHostToDevice) != hipSuccess)
    return EXAMD_ERR_HIP;
  d.tab_n = npoints;
  d.tab_lo = r2_lo;
  d.tab_inv_h = (double)(npoints - 1) / (rc2 - r2_lo);
  d.tab_lof = (float)r2_lo;
  d.tab_inv_hf = (float)d.tab_inv_h;
  d.tab_v_last = v[npoints - 1];
  d.tab_d2fix = d2fix;
  return EXAMD_OK;
}

int examd_hip_get_state(examd_handle *h, double *pos, double *vel,
                        double *force, int64_t n) {
  if (!h || n != h->n) return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  if (force && h->forces_valid < 0) return EXAMD_ERR_NOT_READY;
  if (ensure_stage(h)) return EXAMD_ERR_NOMEM;
  // each array: transposed into the host layout on the device, one copy
  int st = EXAMD_OK;
  if (pos) {
    launch_soa_to_aos(h->d.x, h->d.y, h->d.z, h->dev_stage, n, h->stream);
    if ((st = download_aos(h, pos)) != EXAMD_OK) return st;
  }
  if (vel) {
    launch_soa_to_aos(h->d.vx, h->d.vy, h->d.vz, h->dev_stage, n, h->stream);
    if ((st = download_aos(h, vel)) != EXAMD_OK) return st;
  }
  if (force) {
    launch_forces_to_fp64(h->d, h->dev_stage, h->forces_valid == EXAMD_PREC_MIXED,
                          h->stream);
    if ((st = download_aos(h, force)) != EXAMD_OK) return st;
  }
  HIPCHK(hipGetLastError());
  return EXAMD_OK;
}

int examd_hip_forces(examd_handle *h, int precision, double *forces,
                     examd_stats *stats) {
  if (!h || (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  const int ns = h->nsys;
  std::vector<StepLog> L((size_t)ns);
  unsigned long long fl = 0;
  // a block that runs out of slots is given twice the slots and the
  // evaluation repeated
  for (int attempt = 0;; ++attempt) {
    launch_fill_log(h->log, ns, h->stream);
    evaluate(h, precision, true, false, h->log);
    kinetic(h, precision, h->log);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(L.data(), h->log, ns * sizeof(StepLog),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->timing) collect_timing(h);
    fl = 0;
    for (int k = 0; k < ns; ++k) fl |= L[k].flags;
    if (!(fl & kFlagBinOverflow) || attempt >= 4) break;
    const int sb = ensure_bin_slots(h, ((int64_t)1 << h->d.bin_shift) + 1);
    if (sb != EXAMD_OK) return sb;
  }
  h->forces_valid = precision;
  if (stats)
    for (int k = 0; k < ns; ++k) fill_stats(h, L[k], precision, 0, &stats[k]);
  int st = flags_to_status(fl);
  if (st != EXAMD_OK && st != EXAMD_ERR_ZERO_DISTANCE) {
    h->forces_valid = -1;
    return st;
  }
  if (forces) {
    if (ensure_stage(h)) return EXAMD_ERR_NOMEM;
    launch_forces_to_fp64(h->d, h->dev_stage, precision == EXAMD_PREC_MIXED,
                          h->stream);
    const int s2 = download_aos(h, forces);
    if (s2 != EXAMD_OK) return s2;
  }
  return st;
}

int examd_hip_forces_static(const examd_params *prm, int precision,
                            const double *pos, double *forces,
                            examd_stats *stats) {
  if (!prm || !pos || !forces ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  const int64_t n = prm->natoms;
  if (n < 1 || !(prm->cutoff > 0.0) || !(prm->sigma > 0.0) ||
      !(prm->epsilon >= 0.0))
    return EXAMD_ERR_ARG;
  if (n > kStaticAtoms) return EXAMD_ERR_STATIC_LIMIT;
  int32_t nb[3];
  int64_t nblocks = 0;
  int st = examd_hip_domain(prm->box, 0.5 * prm->cutoff, nb, &nblocks);
  if (st) return st;
  for (int a = 0; a < 3; ++a)
    if (nb[a] < 5 || !(prm->cutoff < 0.5 * prm->box[a])) return EXAMD_ERR_GRID;
  if (nblocks > kStaticBlocks) return EXAMD_ERR_STATIC_LIMIT;
  // the cell-block map: the given tiles or the default, split evenly
  Grid g{};
  int64_t ntiles = 1;
  int wmax[3];
  for (int a = 0; a < 3; ++a) {
    g.nb[a] = nb[a];
    g.L[a] = prm->box[a];
    const int T = prm->tiles[a] > 0 ? prm->tiles[a] : (nb[a] + 6) / 7;
    g.T[a] = std::max(1, std::min(T, nb[a]));
    ntiles *= g.T[a];
    const std::vector<int> c = even_cuts(nb[a], g.T[a]);
    wmax[a] = 0;
    for (int t = 0; t < g.T[a]; ++t) wmax[a] = std::max(wmax[a], c[t + 1] - c[t]);
    if (wmax[a] + 2 > 255) return EXAMD_ERR_TILE_OVERFLOW;
  }
  g.bl = 0.5 * prm->cutoff;
  if (ntiles > kStaticTiles) return EXAMD_ERR_STATIC_LIMIT;
  unsigned char *base = nullptr;
  size_t lds_limit = 0;
  if ((st = scratch_of(prm->device, &base, &lds_limit)) != EXAMD_OK) return st;
  HIPCHK(hipSetDevice(prm->device));
  const bool mixed = precision == EXAMD_PREC_MIXED;
  // launch shape: the LDS holds the largest region the limit allows (a
  // region beyond it is reported as TILE_OVERFLOW by the kernel); run
  // groups as a handle chooses them
  ForceConfig cfg;
  const int thr = prm->block_threads;
  cfg.threads = (thr == 256 || thr == 512 || thr == 1024) ? thr : 512;
  cfg.maxrb = (wmax[0] + 2) * (wmax[1] + 2) * (wmax[2] + 2);
  cfg.lane_queue = 0;
  cfg.workgroups = 0;
  int64_t cap = 65520;
  while (cap >= 32 && force_lds_bytes(!mixed, (int)cap, cfg.maxrb, cfg.threads,
                                      0, false) > lds_limit)
    cap -= 16;
  if (cap < 32) return EXAMD_ERR_TILE_OVERFLOW;
  cfg.cap = (int)cap;
  double reg = (double)n / (g.L[0] * g.L[1] * g.L[2]);
  for (int a = 0; a < 3; ++a) reg *= ((double)g.nb[a] / g.T[a] + 2.0) * g.bl;
  const int ng = prm->run_groups > 0
                     ? prm->run_groups
                     : (int)std::lround(3.0 * cfg.threads / std::max(reg, 1.0));
  cfg.groups = std::max(1, std::min(4, ng));
  // the device state, every array in the scratch
  DevState d{};
  d.n = n;
  double *xyz = reinterpret_cast<double *>(base + kStatic.xyz);
  d.x = xyz;
  d.y = xyz + n;
  d.z = xyz + 2 * n;
  double *f = reinterpret_cast<double *>(base + kStatic.f);
  d.fx = f;
  d.fy = f + n;
  d.fz = f + 2 * n;
  d.qx = reinterpret_cast<unsigned long long *>(f);
  d.qy = d.qx + n;
  d.qz = d.qx + 2 * n;
  d.nblocks = nblocks;
  d.bin_shift = kStaticListShift;
  while (d.bin_shift < 8 && (nblocks << (d.bin_shift + 1)) <= kStaticListInts) ++d.bin_shift;
  int *ints = reinterpret_cast<int *>(base + kStatic.ints);
  d.cell = ints;
  d.slot = ints + n;
  d.sorted = reinterpret_cast<int *>(base + kStatic.sorted);
  d.count = reinterpret_cast<int *>(base + kStatic.count);
  d.bflag = base + kStatic.bflag;
  d.part = reinterpret_cast<Partial *>(base + kStatic.part);
  StepLog *log = reinterpret_cast<StepLog *>(base + kStatic.log);
  SysDesc *sysd = reinterpret_cast<SysDesc *>(base + kStatic.sys);
  SysDesc sd{};
  sd.g = g;
  sd.atom0 = 0;
  sd.block0 = 0;
  sd.tile0 = 0;
  sd.cut[0] = sd.cut[1] = sd.cut[2] = -1;
  SysTable t;
  t.sys = sysd;
  t.nsys = 1;
  t.ntiles = (int)ntiles;
  const LJParams p{prm->epsilon, prm->sigma, prm->cutoff, prm->mass};
  hipStream_t s = nullptr;
  // every call starts from cleared counts and tallies (fill kernels: the
  // array keeps the last call's contents)
  launch_store_sys(sd, sysd, s);
  launch_aos_to_soa(pos, d.x, d.y, d.z, n, s);
  launch_fill_log(log, 1, s);
  // a kBinOnly pass clears the counts and flags itself
  launch_step_bin(d, t, mixed, kBinOnly, 0.0, 0.0, nullptr, log, s);
  launch_bin_update(d, true, !mixed && d.bin_shift > 4, log, s);
  launch_forces(d, t, p, cfg, mixed, true, false, log, s);
  launch_forces_to_fp64(d, forces, mixed, s);
  HIPCHK(hipGetLastError());
  StepLog L{};
  HIPCHK(hipMemcpy(&L, log, sizeof(StepLog), hipMemcpyDeviceToHost));
  if (stats) {
    const double pe = mixed ? (double)(long long)L.pe_fix * kFixInv : L.pe;
    const double sr2 = prm->sigma * prm->sigma / (prm->cutoff * prm->cutoff);
    const double sr6 = sr2 * sr2 * sr2;
    const double vrc = 4.0 * prm->epsilon * (sr6 * sr6 - sr6);
    stats->step = 0;
    stats->ke = 0.0;
    stats->pe = pe;
    stats->pe_shifted = pe - (double)L.cut * vrc;
    stats->etot = pe;
    stats->pairs_cand = (int64_t)L.cand;
    stats->pairs_cut = (int64_t)L.cut;
  }
  return flags_to_status(L.flags);
}

int examd_hip_run(examd_handle *h, int precision, int64_t nsteps, double dt,
                  int64_t energy_every, examd_stats *out, int64_t out_cap,
                  int64_t *n_out) {
  if (!h || nsteps < 0 || !std::isfinite(dt) ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  if (n_out) *n_out = 0;
  HIPCHK(hipSetDevice(h->dev));
  const int ns = h->nsys;
  auto is_energy = [&](int64_t s) {
    return s == nsteps || (energy_every > 0 && s % energy_every == 0);
  };
  // chunks of C steps replayed from a captured graph; C a multiple of the
  // tally stride so every chunk has the same tally pattern
  int64_t C = 0;
  if (h->graph_chunk > 0 && nsteps > 0) {
    C = h->graph_chunk;
    if (energy_every > 0) C = std::max<int64_t>(energy_every, C / energy_every * energy_every);
    if (C > nsteps) C = 0;
  }
  const int64_t nchunks = C > 0 ? nsteps / C : 0;
  const int64_t done_by_graph = nchunks * C;
  const int64_t rem = nsteps - done_by_graph;
  if (ensure_log(h, std::max<int64_t>(C, rem))) return EXAMD_ERR_NOMEM;
  // log[0..ns): flags of the evaluation at the starting positions, if one
  // runs
  launch_fill_log(h->log, ns, h->stream);
  // forces at the current positions under the current map; only the step
  // launches are timed
  auto initial_forces = [&]() {
    const bool t = h->timing;
    h->timing = false;
    evaluate(h, precision, true, false, h->log);
    h->timing = t;
  };
  if (h->forces_valid != precision) initial_forces();
  unsigned long long fl = 0;
  int64_t k = 0;
  std::vector<StepLog> L;
  // read back the tallies of steps base+1..base+count and check flags
  auto harvest = [&](int64_t base, int64_t count) -> int {
    L.resize((size_t)((count + 1) * ns));
    HIPCHK(hipMemcpyAsync(L.data(), h->log, (count + 1) * ns * sizeof(StepLog),
                          hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->timing) collect_timing(h);
    for (int q = 0; q < ns; ++q) fl |= L[q].flags;
    for (int64_t c = 1; c <= count; ++c) {
      const int64_t s = base + c;
      for (int q = 0; q < ns; ++q) {
        const StepLog &Lq = L[(size_t)(c * ns + q)];
        fl |= Lq.flags;
        if (out && is_energy(s) && k < out_cap)
          fill_stats(h, Lq, precision, s, &out[k++]);
      }
    }
    return EXAMD_OK;
  };
  const unsigned long long fatal =
      kFlagOverflow | kFlagBinOverflow | kFlagOutOfBox | kFlagNonFinite;
  // periodic re-partition of the map, between chunks; a new map
  // re-evaluates the forces, so the run continues exactly as one that
  // stopped here, applied the map and started again
  const int64_t R = h->tc[precision].rebalance;
  int64_t &since = h->since_repartition[precision];
  auto after_steps = [&](int64_t count) -> int {
    since += count;
    if (R <= 0 || since < R || (fl & fatal)) return EXAMD_OK;
    since = 0;
    bool changed = false;
    const int st = repartition(h, precision, &changed);
    if (st == EXAMD_OK && changed) {
      launch_fill_log(h->log, ns, h->stream);
      initial_forces();
      std::vector<StepLog> L0((size_t)ns);
      HIPCHK(hipMemcpyAsync(L0.data(), h->log, ns * sizeof(StepLog),
                            hipMemcpyDeviceToHost, h->stream));
      HIPCHK(hipStreamSynchronize(h->stream));
      for (const StepLog &q : L0) fl |= q.flags;
    }
    return st;
  };
  // A chunk whose largest cell-block region outgrew the LDS capacity is
  // run again from its start with a larger capacity (or a finer automatic
  // map): positions and velocities are copied aside before each chunk, and
  // after an overflow they are put back, the forces are evaluated again
  // under the new map and the chunk's tallies are dropped.  The result is
  // that of a run started with the larger capacity.
  const int64_t n = h->n;
  if (!h->chunk_start && dalloc(&h->chunk_start, 6 * std::max<int64_t>(n, 1)))
    return EXAMD_ERR_NOMEM;
  double *const saved_arrays[6] = {h->d.x, h->d.y, h->d.z, h->d.vx, h->d.vy, h->d.vz};
  auto save_chunk_start = [&]() -> int {
    for (int a = 0; a < 6; ++a)
      HIPCHK(hipMemcpyAsync(h->chunk_start + a * n, saved_arrays[a], n * sizeof(double),
                            hipMemcpyDeviceToDevice, h->stream));
    return EXAMD_OK;
  };
  constexpr int kMaxRegrow = 4;
  // after a harvest: 1 = the chunk is to be run again, 0 = go on
  // after a harvest: 1 = the chunk is to be run again, 0 = go on, -1 = a
  // HIP error
  auto regrown = [&](int &attempt, int64_t k0, unsigned long long fl0) -> int {
    constexpr unsigned long long kGrow = kFlagOverflow | kFlagBinOverflow;
    if (!(fl & kGrow) || (fl0 & kGrow) || attempt >= kMaxRegrow) return 0;
    const bool lds = (fl & kFlagOverflow) != 0, bins = (fl & kFlagBinOverflow) != 0;
    for (int a = 0; a < 6; ++a)
      if (hipMemcpyAsync(saved_arrays[a], h->chunk_start + a * n, n * sizeof(double),
                         hipMemcpyDeviceToDevice, h->stream) != hipSuccess)
        return -1;
    k = k0;
    fl = fl0;
    ++attempt;
    // a block that ran out of slots: twice the slots (the state is the
    // chunk's start either way)
    if (bins && ensure_bin_slots(h, ((int64_t)1 << h->d.bin_shift) + 1) != EXAMD_OK) {
      fl |= kFlagBinOverflow;
      return 0;
    }
    if (lds && regrow(h, precision, attempt) != EXAMD_OK) {
      fl |= kFlagOverflow;
      return 0;
    }
    if (!lds) h->regrows++;
    launch_fill_log(h->log, ns, h->stream);
    initial_forces();
    std::vector<StepLog> L0((size_t)ns);
    if (hipMemcpyAsync(L0.data(), h->log, ns * sizeof(StepLog),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
      return -1;
    for (const StepLog &q : L0) fl |= q.flags;
    return 1;
  };
  for (int64_t c = 0; c < nchunks; ++c) {
    int attempt = 0;
    for (;;) {
      // re-captured only when a re-partition or a regrow changed the map
      GraphKey key{precision, C, energy_every > 0 ? energy_every : 0, dt,
                   h->timing, h->timing_every, h->tc[precision]};
      int st = ensure_graph(h, key);
      if (st) return st;
      if ((st = save_chunk_start()) != EXAMD_OK) return st;
      const int64_t k0 = k;
      const unsigned long long fl0 = fl;
      HIPCHK(hipGraphLaunch(h->graph_exec, h->stream));
      h->ev_used = h->graph_events;   // the events the replay recorded
      st = harvest(c * C, C);
      if (st) return st;
      const int again = regrown(attempt, k0, fl0);
      if (again < 0) return EXAMD_ERR_HIP;
      if (!again) break;
    }
    if (fl & fatal) break;
    const int st = after_steps(C);
    if (st != EXAMD_OK) return st;
  }
  // the rest eagerly, in pieces that end where a re-partition is due
  for (int64_t base = done_by_graph; base < nsteps && !(fl & fatal);) {
    int64_t count = nsteps - base;
    if (R > 0) count = std::min(count, std::max<int64_t>(1, R - since));
    int attempt = 0;
    for (;;) {
      int st = save_chunk_start();
      if (st) return st;
      const int64_t k0 = k;
      const unsigned long long fl0 = fl;
      enqueue_steps(h, precision, count, dt,
                    [&](int64_t c) { return is_energy(base + c); });
      HIPCHK(hipGetLastError());
      st = harvest(base, count);
      if (st) return st;
      const int again = regrown(attempt, k0, fl0);
      if (again < 0) return EXAMD_ERR_HIP;
      if (!again) break;
    }
    base += count;
    const int st = after_steps(count);
    if (st != EXAMD_OK) return st;
  }
  h->forces_valid = precision;
  if (n_out) *n_out = k;
  int st = flags_to_status(fl);
  if (st != EXAMD_OK) h->forces_valid = -1;
  return st;
}

int examd_hip_regrow_count(examd_handle *h, int64_t *regrows) {
  if (!h || !regrows) return EXAMD_ERR_ARG;
  *regrows = h->regrows;
  return EXAMD_OK;
}

int examd_hip_set_graph_chunk(examd_handle *h, int64_t chunk) {
  if (!h || chunk < 0) return EXAMD_ERR_ARG;
  h->graph_chunk = chunk;
  return EXAMD_OK;
}

int examd_hip_block_ids(examd_handle *h, int64_t *blocknum, int32_t *blockxyz,
                        int64_t n) {
  if (!h || !blocknum || n != h->n) return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  launch_fill_log(h->log, 1, h->stream);
  // a pass that starts the block lists afresh and leaves them empty: the
  // next evaluation bins again (forces_valid below)
  launch_step_bin(h->d, table(h, 0), false, kBinOnly, 0.0, 0.0, nullptr,
                  h->log, h->stream);
  HIPCHK(hipGetLastError());
  std::vector<int> cell((size_t)std::max<int64_t>(n, 1));
  StepLog L;
  HIPCHK(hipMemcpyAsync(cell.data(), h->d.cell, n * sizeof(int),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(&L, h->log, sizeof(L), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->forces_valid = -1;  // force accumulators were cleared
  for (int k = 0; k < h->nsys; ++k) {
    const SysHost &S = h->sys[k];
    const int64_t nx = S.g.nb[0], ny = S.g.nb[1];
    for (int64_t i = S.atom0; i < S.atom0 + S.n; ++i) {
      const int64_t b = cell[i] - S.block0;
      blocknum[i] = b + 1;
      if (blockxyz) {
        blockxyz[3 * i + 0] = (int32_t)(b % nx + 1);
        blockxyz[3 * i + 1] = (int32_t)((b / nx) % ny + 1);
        blockxyz[3 * i + 2] = (int32_t)(b / (nx * ny) + 1);
      }
    }
  }
  return flags_to_status(L.flags & (kFlagOutOfBox | kFlagNonFinite));
}

int examd_hip_home_pair_counts(examd_handle *h, int precision, int64_t *counts,
                               int64_t nblocks) {
  if (!h || !counts || nblocks != h->nblocks ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  if (!h->d.home && dalloc(&h->d.home, h->nblocks)) return EXAMD_ERR_NOMEM;
  const int ns = h->nsys;
  launch_fill_u64(h->d.home, h->nblocks, h->stream);
  launch_fill_log(h->log, ns, h->stream);
  evaluate(h, precision, false, true, h->log);
  HIPCHK(hipGetLastError());
  std::vector<StepLog> L((size_t)ns);
  HIPCHK(hipMemcpyAsync(counts, h->d.home, nblocks * sizeof(int64_t),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipMemcpyAsync(L.data(), h->log, ns * sizeof(StepLog),
                        hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->timing) collect_timing(h);
  h->forces_valid = precision;
  unsigned long long fl = 0;
  for (int k = 0; k < ns; ++k) fl |= L[k].flags;
  return flags_to_status(fl & ~(unsigned long long)kFlagZero);
}

int examd_hip_get_config(examd_handle *h, int precision, examd_tune *out) {
  if (!h || !out ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  const TileConfig &c = h->tc[precision];
  for (int a = 0; a < 3; ++a) out->tiles[a] = c.T[0][a];
  out->block_threads = c.threads;
  out->run_groups = c.groups;
  out->lane_queue = c.lane_queue;
  out->tile_edge = c.edge;
  out->workgroups = c.workgroups;
  out->balance = c.balance ? 1 : 0;
  out->rebalance_every = c.rebalance;
  out->force_ms = 0.0;
  return EXAMD_OK;
}

int examd_hip_set_config(examd_handle *h, int precision,
                         const examd_tune *cfg) {
  if (!h || !cfg ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  const int thr = cfg->block_threads;
  if (thr != 256 && thr != 512 && thr != 1024) return EXAMD_ERR_ARG;
  if (cfg->run_groups < 0 || cfg->run_groups > 4) return EXAMD_ERR_ARG;
  if (cfg->lane_queue < 0 || cfg->lane_queue > 16) return EXAMD_ERR_ARG;
  if (cfg->workgroups < 0 || cfg->balance < 0 || cfg->balance > 1 ||
      cfg->rebalance_every < 0)
    return EXAMD_ERR_ARG;
  const bool explicit_tiles =
      cfg->tiles[0] > 0 && cfg->tiles[1] > 0 && cfg->tiles[2] > 0;
  if (explicit_tiles) {
    if (h->nsys != 1) return EXAMD_ERR_ARG;
  } else if (cfg->tile_edge <= 0 || cfg->tiles[0] != 0 || cfg->tiles[1] != 0 ||
             cfg->tiles[2] != 0) {
    return EXAMD_ERR_ARG;
  }
  HIPCHK(hipSetDevice(h->dev));
  std::vector<double> pos;
  int st = download_positions(h, pos);
  if (st) return st;
  st = apply_config(h, precision, explicit_tiles ? cfg->tiles : nullptr,
                    (double)cfg->tile_edge, thr, cfg->run_groups,
                    cfg->lane_queue, cfg->workgroups, cfg->balance != 0,
                    pos.data());
  if (st == EXAMD_OK) {
    h->tc[precision].rebalance = cfg->rebalance_every;
    h->since_repartition[precision] = 0;
  }
  return st;
}

int examd_hip_get_cuts(examd_handle *h, int precision, int32_t system,
                       int32_t *cuts, int32_t cap, int32_t *n_out) {
  if (!h || !cuts || !n_out || system < 0 || system >= h->nsys ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  const Cuts cb = cuts_of(h, precision, system, grid_of(h, precision, system));
  const int32_t n = (int32_t)(cb[0].size() + cb[1].size() + cb[2].size());
  *n_out = n;
  if (cap < n) return EXAMD_ERR_ARG;
  int32_t k = 0;
  for (int a = 0; a < 3; ++a)
    for (int v : cb[a]) cuts[k++] = v;
  return EXAMD_OK;
}

int examd_hip_autotune(examd_handle *h, int precision, int trials,
                       examd_tune *best) {
  if (!h || trials <= 0 ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  std::vector<double> pos;
  int st = download_positions(h, pos);
  if (st) return st;
  // candidate cell-block sizes (home blocks per side) and launch shapes; a
  // single system takes fractional edges as explicit maps, a batch integer
  // edges for every system
  static const double sizes1[] = {3.5, 4.0, 4.5, 5.0, 5.5, 6.0, 6.7, 7.5,
                                  8.5, 9.0, 10.0};
  static const double sizesN[] = {3.0, 4.0, 5.0, 6.0, 7.0, 8.0};
  const bool single = h->nsys == 1;
  const double *sizes = single ? sizes1 : sizesN;
  const int nsizes = single ? 11 : 6;
  static const int threads[] = {256, 512, 1024};
  const bool save_timing = h->timing;
  const double save_ms = h->force_ms;
  const int64_t save_n = h->force_launches;
  const TileConfig start = h->tc[precision];
  const int ns = h->nsys;
  TileConfig bestc = start;
  double best_ms = 1e300;
  // the fastest few shapes of the first pass, refined in the second
  constexpr int kRefine = 3;
  std::vector<std::pair<double, TileConfig>> top;
  // time `reps` evaluations of the current shape (1e300 if it overflowed)
  auto time_shape = [&](double *ms, int reps) -> int {
    launch_fill_log(h->log, ns, h->stream);
    evaluate(h, precision, false, false, h->log);  // warm-up
    h->timing = true;
    h->ev_used = 0;
    h->force_ms = 0.0;
    h->force_launches = 0;
    for (int t = 0; t < reps; ++t) {
      launch_fill_log(h->log, ns, h->stream);
      evaluate(h, precision, false, false, h->log);
    }
    hipError_t e = hipStreamSynchronize(h->stream);
    collect_timing(h);
    std::vector<StepLog> L((size_t)ns);
    if (e == hipSuccess)
      e = hipMemcpy(L.data(), h->log, ns * sizeof(StepLog), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return EXAMD_ERR_HIP;
    unsigned long long fl = 0;
    for (int k = 0; k < ns; ++k) fl |= L[k].flags;
    *ms = (fl & (kFlagOverflow | kFlagBinOverflow | kFlagOutOfBox | kFlagNonFinite))
              ? 1e300
              : h->force_ms / std::max<int64_t>(h->force_launches, 1);
    return EXAMD_OK;
  };
  auto consider = [&](const int *T, double edge, int thr, int grp, int lq,
                      int wgs, bool bal) -> int {
    if (apply_config(h, precision, T, edge, thr, grp, lq, wgs, bal,
                     pos.data()) != EXAMD_OK)
      return EXAMD_OK;
    double ms;
    int e = time_shape(&ms, trials);
    if (e) return e;
    if (ms < best_ms) {
      best_ms = ms;
      bestc = h->tc[precision];
    }
    if (ms < 1e300) {
      top.emplace_back(ms, h->tc[precision]);
      std::sort(top.begin(), top.end(),
                [](const auto &x, const auto &y) { return x.first < y.first; });
      if ((int)top.size() > kRefine) top.resize(kRefine);
    }
    return EXAMD_OK;
  };
  for (int si = 0; si < nsizes; ++si)
    for (int thr : threads)
      for (int lq = 0; lq <= 8; lq += 4)
        for (int grp = 0; grp <= (thr == 1024 ? 3 : 1); ++grp) {
          if (grp == 1) continue;  // 0 = automatic choice covers one group
          int T[3];
          if (single) {
            for (int a = 0; a < 3; ++a)
              T[a] = std::max(1, (int)std::lround(h->sys[0].g.nb[a] / sizes[si]));
          }
          st = consider(single ? T : nullptr, sizes[si], thr, grp, lq, 0, false);
          if (st) {
            h->timing = save_timing;
            return st;
          }
        }
  // for the fastest few shapes: balanced boundaries, then persistent
  // grids of a few workgroups per CU (a shape can win only in one of them)
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            h->dev) != hipSuccess || cus <= 0)
    cus = 256;
  const std::vector<std::pair<double, TileConfig>> firsts = top;
  for (const auto &cand : firsts) {
    const TileConfig &b = cand.second;
    for (int bal = 0; bal <= 1; ++bal)
      for (int per = bal ? 0 : 1; per <= 4; per = per ? per * 2 : 1) {
        const int wgs = cus * per;
        if (per > 0 && wgs >= (int)tiles_of(b)) break;
        st = consider(b.edge == 0 ? b.T[0].data() : nullptr, (double)b.edge,
                      b.threads, b.fixed_groups, b.lane_queue, wgs, bal != 0);
        if (st) {
          h->timing = save_timing;
          return st;
        }
      }
  }
  // the fastest shapes of both passes, timed again in turn (A, B, C, A, B,
  // C, ...) for six rounds of `trials` launches each: the first timings are
  // a few launches each, the leaders are often within their noise of one
  // another, and interleaving keeps slow drifts of the clock from favouring
  // whichever shape is timed first
  const std::vector<std::pair<double, TileConfig>> finals = top;
  if (finals.size() > 1) {
    std::vector<double> total(finals.size(), 0.0);
    for (int round = 0; round < 6; ++round)
      for (size_t k = 0; k < finals.size(); ++k) {
        if (total[k] >= 1e300) continue;
        h->tc[precision] = finals[k].second;
        if ((st = upload_tables(h, precision)) != EXAMD_OK ||
            (st = ensure_partials(h)) != EXAMD_OK) {
          h->timing = save_timing;
          return st;
        }
        double ms;
        if ((st = time_shape(&ms, trials)) != EXAMD_OK) {
          h->timing = save_timing;
          return st;
        }
        total[k] = ms >= 1e300 ? 1e300 : total[k] + ms;
      }
    size_t kbest = 0;
    for (size_t k = 1; k < finals.size(); ++k)
      if (total[k] < total[kbest]) kbest = k;
    if (total[kbest] < 1e300) {
      best_ms = total[kbest] / 6.0;
      bestc = finals[kbest].second;
    }
  }
  h->timing = save_timing;
  h->force_ms = save_ms;
  h->force_launches = save_n;
  h->ev_used = 0;
  if (best_ms >= 1e300) {
    h->tc[precision] = start;
    (void)upload_tables(h, precision);
    return EXAMD_ERR_TILE_OVERFLOW;
  }
  st = apply_config(h, precision, bestc.edge == 0 ? bestc.T[0].data() : nullptr,
                    (double)bestc.edge, bestc.threads, bestc.fixed_groups,
                    bestc.lane_queue, bestc.workgroups, bestc.balance,
                    pos.data());
  h->forces_valid = -1;
  if (st) return st;
  if (best) {
    examd_hip_get_config(h, precision, best);
    best->force_ms = best_ms;
  }
  return EXAMD_OK;
}

int examd_hip_phase_profile(examd_handle *h, int precision,
                            double phase_cycles[5]) {
  if (!h || !phase_cycles ||
      (precision != EXAMD_PREC_FP64 && precision != EXAMD_PREC_MIXED))
    return EXAMD_ERR_ARG;
  if (!h->state_set) return EXAMD_ERR_NOT_READY;
  HIPCHK(hipSetDevice(h->dev));
  const int64_t G = tiles_of(h->tc[precision]);
  unsigned long long *buf = nullptr;
  if (dalloc(&buf, G * kProfSlots)) return EXAMD_ERR_NOMEM;
  h->d.prof = buf;
  launch_fill_log(h->log, h->nsys, h->stream);
  evaluate(h, precision, false, false, h->log);
  h->d.prof = nullptr;
  std::vector<unsigned long long> t((size_t)(G * kProfSlots));
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(t.data(), buf, t.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, h->stream);
  hipError_t e2 = hipStreamSynchronize(h->stream);
  (void)hipFree(buf);
  if (e != hipSuccess || e2 != hipSuccess) return EXAMD_ERR_HIP;
  if (h->timing) collect_timing(h);
  h->forces_valid = precision;
  for (int k = 0; k < 5; ++k) phase_cycles[k] = 0.0;
  for (int64_t b = 0; b < G; ++b) {
    const unsigned long long *p = &t[(size_t)(b * kProfSlots)];
    for (int k = 0; k < 4; ++k) phase_cycles[k] += (double)(p[k + 1] - p[k]);
    phase_cycles[4] += (double)(p[4] - p[0]);
  }
  for (int k = 0; k < 5; ++k) phase_cycles[k] /= (double)G;
  return EXAMD_OK;
}

int examd_hip_timing(examd_handle *h, int enable, double *force_ms,
                     int64_t *force_launches, int reset) {
  if (!h) return EXAMD_ERR_ARG;
  if (force_ms) *force_ms = h->force_ms;
  if (force_launches) *force_launches = h->force_launches;
  if (reset) {
    h->force_ms = 0.0;
    h->force_launches = 0;
  }
  if (enable < 0) return EXAMD_ERR_ARG;
  h->timing = enable != 0;
  h->timing_every = std::max(1, enable);
  return EXAMD_OK;
}

int examd_hip_generate_fcc(const int32_t nc[3], double rho, double temperature,
                           double mass, double perturb, uint64_t seed,
                           double *pos, double *vel, double box[3]) {
  if (!nc || !pos || !box || !(rho > 0.0) || !(mass > 0.0) ||
      !(temperature >= 0.0) || !(perturb >= 0.0))
    return EXAMD_ERR_ARG;
  for (int a = 0; a < 3; ++a)
    if (nc[a] <= 0) return EXAMD_ERR_ARG;
  const double cell = std::cbrt(4.0 / rho);
  for (int a = 0; a < 3; ++a) box[a] = nc[a] * cell;
  static const double basis[4][3] = {
      {0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
  uint64_t s = seed;
  int64_t i = 0;
  for (int k = 0; k < nc[2]; ++k)
    for (int j = 0; j < nc[1]; ++j)
      for (int m = 0; m < nc[0]; ++m)
        for (int b = 0; b < 4; ++b, ++i) {
          const int c[3] = {m, j, k};
          for (int a = 0; a < 3; ++a) {
            double x = (c[a] + basis[b][a]) * cell;
            if (perturb > 0.0) x += perturb * (2.0 * uniform01(s) - 1.0);
            double L = box[a];
            x = x - std::floor(x / L) * L;
            if (x >= L) x -= L;
            if (x < 0.0) x = 0.0;
            pos[3 * i + a] = x;
          }
        }
  if (vel) mb_velocities(i, temperature, mass, s, vel);
  return EXAMD_OK;
}

int examd_hip_generate_vacancy_fcc(int64_t n, double rho, double temperature,
                                   double mass, double perturb, uint64_t seed,
                                   double *pos, double *vel, double box[3]) {
  if (n <= 0 || n > (int64_t)2000000000 || !pos || !box || !(rho > 0.0) ||
      !(mass > 0.0) || !(temperature >= 0.0) || !(perturb >= 0.0))
    return EXAMD_ERR_ARG;
  const double L = std::cbrt((double)n / rho);
  int64_t nc = 1;
  while (4 * nc * nc * nc < n) ++nc;
  const int64_t m = 4 * nc * nc * nc;
  const double a = L / (double)nc;
  for (int d = 0; d < 3; ++d) box[d] = L;
  uint64_t s = seed;
  // partial Fisher-Yates: the first n entries of a seeded shuffle of sites
  std::vector<int64_t> site((size_t)m);
  for (int64_t i = 0; i < m; ++i) site[i] = i;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t j = i + (int64_t)(uniform01(s) * (double)(m - i));
    std::swap(site[i], site[std::min(j, m - 1)]);
  }
  std::sort(site.begin(), site.begin() + n);   // lattice order
  static const double basis[4][3] = {
      {0.0, 0.0, 0.0}, {0.5, 0.5, 0.0}, {0.5, 0.0, 0.5}, {0.0, 0.5, 0.5}};
  for (int64_t i = 0; i < n; ++i) {
    const int64_t q = site[i];
    const int b = (int)(q % 4);
    const int64_t c = q / 4;
    const int64_t cc[3] = {c % nc, (c / nc) % nc, c / (nc * nc)};
    for (int d = 0; d < 3; ++d) {
      double x = ((double)cc[d] + basis[b][d]) * a;
      if (perturb > 0.0) x += perturb * (2.0 * uniform01(s) - 1.0);
      x = x - std::floor(x / L) * L;
      if (x >= L) x -= L;
      if (x < 0.0) x = 0.0;
      pos[3 * i + d] = x;
    }
  }
  if (vel) mb_velocities(n, temperature, mass, s, vel);
  return EXAMD_OK;
}

int examd_hip_generate_slab(const double box[3], double slab_lo,
                            double slab_hi, double rho_slab,
                            double rho_vapour, double min_sep,
                            double temperature, double mass, uint64_t seed,
                            double *pos, double *vel, int64_t cap,
                            int64_t *n) {
  if (!box || !n || !(rho_slab >= 0.0) || !(rho_vapour >= 0.0) ||
      !(min_sep >= 0.0) || !(mass > 0.0) || !(temperature >= 0.0))
    return EXAMD_ERR_ARG;
  for (int d = 0; d < 3; ++d)
    if (!(box[d] > 0.0)) return EXAMD_ERR_ARG;
  if (!(slab_lo >= 0.0 && slab_lo < slab_hi && slab_hi <= box[0]))
    return EXAMD_ERR_ARG;
  const double area = box[1] * box[2];
  const double vs = (slab_hi - slab_lo) * area;
  const double vv = box[0] * area - vs;
  const int64_t ns = (int64_t)std::llround(rho_slab * vs);
  const int64_t nv = (int64_t)std::llround(rho_vapour * vv);
  *n = ns + nv;
  if (*n > (int64_t)2000000000) return EXAMD_ERR_ARG;
  if (!pos) return EXAMD_OK;
  if (cap < *n) return EXAMD_ERR_ARG;
  // cell list of accepted atoms, cells at least min_sep wide
  const double cw = std::max(min_sep, 1e-3);
  int nc[3];
  for (int d = 0; d < 3; ++d)
    nc[d] = std::max(1, std::min(1 << 20, (int)(box[d] / cw)));
  const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
  if (ncell > (int64_t)1 << 31) return EXAMD_ERR_ARG;
  std::vector<int64_t> head((size_t)ncell, -1), next((size_t)*n, -1);
  auto cell_of = [&](const double *x, int *c) {
    for (int d = 0; d < 3; ++d) {
      int v = (int)(x[d] / box[d] * nc[d]);
      c[d] = v < 0 ? 0 : (v >= nc[d] ? nc[d] - 1 : v);
    }
  };
  const double ms2 = min_sep * min_sep;
  auto fits = [&](const double *x) {
    int c[3];
    cell_of(x, c);
    const int r = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3 ? 1 : 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (r == 0 && (dx || dy || dz)) continue;
          const int cx = (c[0] + dx + nc[0]) % nc[0];
          const int cy = (c[1] + dy + nc[1]) % nc[1];
          const int cz = (c[2] + dz + nc[2]) % nc[2];
          for (int64_t j = head[(size_t)cx + (size_t)nc[0] * (cy + (size_t)nc[1] * cz)];
               j >= 0; j = next[(size_t)j]) {
            double d2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              double u = pos[3 * j + d] - x[d];
              u -= box[d] * std::nearbyint(u / box[d]);
              d2 += u * u;
            }
            if (d2 < ms2) return false;
          }
        }
    return true;
  };
  if (nc[0] < 3 || nc[1] < 3 || nc[2] < 3) {
    // tiny boxes: check every accepted atom
    for (int d = 0; d < 3; ++d) nc[d] = 1;
  }
  uint64_t s = seed;
  const double wv = box[0] - (slab_hi - slab_lo);
  for (int64_t i = 0; i < *n; ++i) {
    const bool slab = i < ns;
    bool placed = false;
    for (int attempt = 0; attempt < 10000 && !placed; ++attempt) {
      double x[3];
      if (slab) {
        x[0] = slab_lo + uniform01(s) * (slab_hi - slab_lo);
      } else {
        // vapour: the complement of the slab along x, periodic
        double u = uniform01(s) * wv;
        x[0] = slab_hi + u;
        if (x[0] >= box[0]) x[0] -= box[0];
      }
      x[1] = uniform01(s) * box[1];
      x[2] = uniform01(s) * box[2];
      for (int d = 0; d < 3; ++d) {
        if (x[d] >= box[d]) x[d] = std::nextafter(box[d], 0.0);
        if (x[d] < 0.0) x[d] = 0.0;
      }
      if (min_sep > 0.0 && !fits(x)) continue;
      for (int d = 0; d < 3; ++d) pos[3 * i + d] = x[d];
      int c[3];
      cell_of(x, c);
      const size_t ci = (size_t)c[0] + (size_t)nc[0] * (c[1] + (size_t)nc[1] * c[2]);
      next[(size_t)i] = head[ci];
      head[ci] = i;
      placed = true;
    }
    if (!placed) return EXAMD_ERR_ARG;
  }
  if (vel) mb_velocities(*n, temperature, mass, s, vel);
  return EXAMD_OK;
}

}  // extern "C"
