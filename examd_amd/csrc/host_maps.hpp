// host_maps.hpp -- host-side arithmetic of the cell-block maps: the block
// of a position, balanced boundaries from block counts, and the largest
// region (a cell block plus one block on each side, periodic) of a map, and
// the coarsening of a map under limits on that region.
// Plain C++ (no HIP), so that it compiles with a host compiler alone; the
// grid type G needs int nb[3] (blocks per axis) and int T[3] (cell blocks
// per axis).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <vector>

namespace examd {

using Cuts = std::array<std::vector<int>, 3>;   // T + 1 boundaries per axis

// Block coordinate of a position along one axis (exaMD_identify_block).
// The clamp is applied before the conversion to int, so that a finite
// position far outside the box converts without overflow.
inline int host_block(double x, double bl, int nb) {
  double q = x / bl + 1.0;
  if (q > (double)nb) q = (double)nb;
  if (q < 1.0) q = 1.0;
  return (int)q - 1;
}

// Balanced map: along each axis, boundaries at equal shares of the atoms
// counted per block slab (the weights METIS gets from atom counts,
// SRC/exaMD_partitioning.f90:199-203, restated for a rectilinear map), each
// tile at least one block wide.  cnt: the system's atoms per block.
template <class G>
Cuts balanced_cuts(const G &g, const int *cnt, int64_t n) {
  // slab sums in integers (each row of x summed first, so that the y and z
  // slabs take one addition per row), then as doubles
  std::vector<int64_t> isl[3];
  for (int a = 0; a < 3; ++a) isl[a].assign((size_t)g.nb[a], 0);
  const int nx = g.nb[0];
  for (int z = 0; z < g.nb[2]; ++z)
    for (int y = 0; y < g.nb[1]; ++y) {
      const int *row = cnt + (size_t)nx * (y + (size_t)g.nb[1] * z);
      int64_t r = 0;
      for (int x = 0; x < nx; ++x) {
        isl[0][(size_t)x] += row[x];
        r += row[x];
      }
      isl[1][(size_t)y] += r;
      isl[2][(size_t)z] += r;
    }
  std::vector<double> slab[3];
  for (int a = 0; a < 3; ++a) slab[a].assign(isl[a].begin(), isl[a].end());
  Cuts b;
  for (int a = 0; a < 3; ++a) {
    const int nb = g.nb[a], T = g.T[a];
    std::vector<double> &w = slab[a];
    // a small floor so empty slabs still spread over tiles
    const double floor_w = std::max(1e-3, 0.02 * (double)n / nb);
    double total = 0.0;
    for (double &x : w) total += (x += floor_w);
    std::vector<int> &c = b[a];
    c.assign((size_t)T + 1, 0);
    c[T] = nb;
    double acc = 0.0;
    int blk = 0;
    for (int t = 1; t < T; ++t) {
      const double target = total * t / T;
      while (blk < nb && acc + w[blk] * 0.5 < target) acc += w[blk++];
      int v = std::max(blk, c[t - 1] + 1);
      v = std::min(v, nb - (T - t));
      while (blk < v) acc += w[blk++];
      c[t] = v;
    }
  }
  return b;
}

// Largest region (tile plus one ghost block per side) in blocks.
inline int max_region_blocks(const Cuts &b) {
  int m[3];
  for (int a = 0; a < 3; ++a) {
    int mx = 0;
    for (size_t t = 0; t + 1 < b[a].size(); ++t)
      mx = std::max(mx, b[a][t + 1] - b[a][t] + 2);
    m[a] = mx;
  }
  return m[0] * m[1] * m[2];
}

// Sums of n values v(0..n) over the periodic ranges [c[t] - 1, c[t+1] + 1)
// of the T tiles of one axis (the tile and one block on each side), from a
// prefix sum over the line with one periodic block added at each end.
template <class V>
void region_line(int n, const std::vector<int> &c, int T, V v,
                 std::vector<int64_t> &pre, int64_t *out, size_t ostride) {
  pre.resize((size_t)n + 3);
  int64_t *q = pre.data();
  q[0] = 0;
  q[1] = v(n - 1);
  for (int i = 0; i < n; ++i) q[i + 2] = q[i + 1] + v(i);
  q[n + 2] = q[n + 1] + v(0);
  for (int t = 0; t < T; ++t) out[(size_t)t * ostride] = q[c[t + 1] + 2] - q[c[t]];
}

// The atoms of every tile's region (the tile plus one block on each side,
// periodic), passed to f(tx, ty, tz, atoms).  cnt: the system's atoms per
// block, x fastest.  On a rectilinear map the region sums separate: along x
// to one sum per (tile column, y, z), then along y, then along z.
template <class G, class F>
void for_region_atoms(const G &g, const Cuts &cb, const int *cnt, F f) {
  const int nx = g.nb[0], ny = g.nb[1], nz = g.nb[2];
  const int Tx = g.T[0], Ty = g.T[1], Tz = g.T[2];
  std::vector<int64_t> pre;
  // A[(z * ny + y) * Tx + tx]: x-region sums of row (y, z)
  std::vector<int64_t> A((size_t)Tx * ny * nz);
  pre.resize((size_t)nx + 3);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      // the one pass over every block: the row's prefix sum in place
      const int *row = cnt + (size_t)nx * (y + (size_t)ny * z);
      int64_t *q = pre.data();
      q[0] = 0;
      q[1] = row[nx - 1];
      int64_t acc = q[1];
      for (int i = 0; i < nx; ++i) q[i + 2] = (acc += row[i]);
      q[nx + 2] = acc + row[0];
      int64_t *out = &A[((size_t)z * ny + y) * Tx];
      for (int t = 0; t < Tx; ++t) out[t] = q[cb[0][t + 1] + 2] - q[cb[0][t]];
    }
  // B[(z * Ty + ty) * Tx + tx]: then over y
  std::vector<int64_t> B((size_t)Tx * Ty * nz);
  for (int z = 0; z < nz; ++z)
    for (int tx = 0; tx < Tx; ++tx) {
      const int64_t *col = &A[(size_t)z * ny * Tx + tx];
      region_line(ny, cb[1], Ty, [&](int y) { return col[(size_t)y * Tx]; }, pre,
                  &B[(size_t)z * Ty * Tx + tx], (size_t)Tx);
    }
  // then over z
  std::vector<int64_t> C((size_t)Tz);
  for (int ty = 0; ty < Ty; ++ty)
    for (int tx = 0; tx < Tx; ++tx) {
      const int64_t *col = &B[(size_t)ty * Tx + tx];
      region_line(nz, cb[2], Tz, [&](int z) { return col[(size_t)z * Ty * Tx]; },
                  pre, C.data(), 1);
      for (int tz = 0; tz < Tz; ++tz) f(tx, ty, tz, C[(size_t)tz]);
    }
}

// Largest number of atoms any tile's region holds.
template <class G>
int64_t max_region_atoms(const G &g, const Cuts &cb, const int *cnt) {
  int64_t best = 0;
  for_region_atoms(g, cb, cnt,
                   [&](int, int, int, int64_t v) { best = std::max(best, v); });
  return best;
}

// Boundaries of nparts contiguous ranges of the tiles (in launch order) of
// about equal total cost: part x is [b[x], b[x + 1]).
inline std::vector<int> equal_cost_ranges(const std::vector<int64_t> &cost,
                                          int nparts) {
  const int n = (int)cost.size();
  std::vector<int> b((size_t)nparts + 1, n);
  b[0] = 0;
  int64_t total = 0;
  for (int64_t v : cost) total += v;
  int64_t acc = 0;
  int x = 1;
  for (int i = 0; i < n && x < nparts; ++i) {
    acc += cost[(size_t)i];
    // part x starts after the tile that brings the sum to x / nparts
    while (x < nparts && acc * nparts >= total * x) b[(size_t)x++] = i + 1;
  }
  return b;
}

// Coarsen a map while its largest region stays within a launch's limits:
// take one cell block off the axis whose cell blocks are the thinnest (ties
// to the lower axis), else off the next thinnest, until no single axis can
// give one up.  A larger cell block owns a larger share of its region.
// cuts_for(g): the boundaries of the map g.T (even or balanced); the limits
// are the atoms and the blocks of the largest region (a launch sizes its LDS
// for both) and 255 blocks per axis of a region.
template <class G, class F>
void coarsen_map(G &g, F cuts_for, const int *cnt, int64_t max_atoms,
                 int max_blocks) {
  for (;;) {
    int order[3] = {0, 1, 2};
    std::stable_sort(order, order + 3, [&](int p, int q) {
      return (int64_t)g.nb[p] * g.T[q] < (int64_t)g.nb[q] * g.T[p];
    });
    bool merged = false;
    for (int a : order) {
      if (g.T[a] <= 1) continue;
      G t = g;
      t.T[a] -= 1;
      const Cuts cb = cuts_for(t);
      bool ok = max_region_blocks(cb) <= max_blocks;
      for (int b = 0; b < 3 && ok; ++b)
        for (size_t i = 0; i + 1 < cb[b].size(); ++i)
          if (cb[b][i + 1] - cb[b][i] + 2 > 255) ok = false;
      if (ok && max_region_atoms(t, cb, cnt) <= max_atoms) {
        g = t;
        merged = true;
        break;
      }
    }
    if (!merged) return;
  }
}

}  // namespace examd
