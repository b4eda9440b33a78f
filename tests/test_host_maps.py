"""Host-side arithmetic of the cell-block maps (examd_amd/csrc/host_maps.hpp),
compiled with the host compiler alone:

- the largest region of a map (a cell block plus one block on each side,
  periodic), which sizes the force kernel's LDS, against a brute-force
  numpy sum over every cell block of random maps and block counts;
- balanced boundaries: ordered, every cell block at least one block wide,
  the even split for uniform counts, and every cell block within one block
  slab of an equal share of the atoms;
- the block of a position (exaMD_identify_block: x/bl + 1 truncated, clamped
  to the grid) on exact faces, one ulp below them, and past the last face;
- the coarsening of a batch system's map: within the limits on the largest
  region (atoms, blocks), and no axis can give up one more cell block;
- every cell block's region atoms (the per-XCD cost split of a force launch)
  against the same brute force, cell block by cell block;
- the split of a launch's tiles into contiguous ranges of about equal cost:
  ordered, covering, each range's cost within one tile of an equal share.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include <vector>
#include "examd_amd/csrc/host_maps.hpp"
struct G { int nb[3]; int T[3]; };
int main() {
  int cases;
  if (std::scanf("%d", &cases) != 1) return 1;
  for (int c = 0; c < cases; ++c) {
    int mode;
    G g;
    std::scanf("%d %d %d %d %d %d %d", &mode, &g.nb[0], &g.nb[1], &g.nb[2],
               &g.T[0], &g.T[1], &g.T[2]);
    const int nblk = g.nb[0] * g.nb[1] * g.nb[2];
    if (mode == 0) {            // largest region of a given map
      examd::Cuts cb;
      for (int a = 0; a < 3; ++a) {
        cb[a].resize(g.T[a] + 1);
        for (int &v : cb[a]) std::scanf("%d", &v);
      }
      std::vector<int> cnt(nblk);
      for (int &v : cnt) std::scanf("%d", &v);
      std::printf("%lld\n", (long long)examd::max_region_atoms(g, cb, cnt.data()));
    } else if (mode == 1) {     // balanced boundaries
      std::vector<int> cnt(nblk);
      long long n = 0;
      for (int &v : cnt) { std::scanf("%d", &v); n += v; }
      const examd::Cuts cb = examd::balanced_cuts(g, cnt.data(), n);
      for (int a = 0; a < 3; ++a) {
        for (int v : cb[a]) std::printf("%d ", v);
        std::printf("\n");
      }
    } else if (mode == 3) {     // coarsened balanced map: limits, counts
      long long lim_atoms;
      int lim_blocks;
      std::scanf("%lld %d", &lim_atoms, &lim_blocks);
      std::vector<int> cnt(nblk);
      long long n = 0;
      for (int &v : cnt) { std::scanf("%d", &v); n += v; }
      examd::coarsen_map(
          g, [&](const G &t) { return examd::balanced_cuts(t, cnt.data(), n); },
          cnt.data(), lim_atoms, lim_blocks);
      std::printf("%d %d %d\n", g.T[0], g.T[1], g.T[2]);
    } else if (mode == 4) {     // every tile's region atoms, x fastest
      examd::Cuts cb;
      for (int a = 0; a < 3; ++a) {
        cb[a].resize(g.T[a] + 1);
        for (int &v : cb[a]) std::scanf("%d", &v);
      }
      std::vector<int> cnt(nblk);
      for (int &v : cnt) std::scanf("%d", &v);
      std::vector<long long> out((size_t)g.T[0] * g.T[1] * g.T[2]);
      examd::for_region_atoms(g, cb, cnt.data(), [&](int x, int y, int z, int64_t v) {
        out[x + (size_t)g.T[0] * (y + (size_t)g.T[1] * z)] = v;
      });
      for (long long v : out) std::printf("%lld ", v);
      std::printf("\n");
    } else if (mode == 5) {     // equal-cost ranges: nparts, n, costs
      int parts, n;
      std::scanf("%d %d", &parts, &n);
      std::vector<int64_t> cost(n);
      for (auto &v : cost) { long long t; std::scanf("%lld", &t); v = t; }
      for (int b : examd::equal_cost_ranges(cost, parts)) std::printf("%d ", b);
      std::printf("\n");
    } else {                    // block of positions: bl, count, values
      double bl;
      int m;
      std::scanf("%lf %d", &bl, &m);
      for (int i = 0; i < m; ++i) {
        double x;
        std::scanf("%lf", &x);
        std::printf("%d ", examd::host_block(x, bl, g.nb[0]));
      }
      std::printf("\n");
    }
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ absent")
    d = tmp_path_factory.mktemp("host_maps")
    src = d / "driver.cpp"
    src.write_text(DRIVER)
    exe = d / "driver"
    subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-I", ROOT, str(src),
                    "-o", str(exe)], check=True, capture_output=True,
                   timeout=120)

    def run(text):
        r = subprocess.run([str(exe)], input=text, capture_output=True,
                           text=True, timeout=120, check=True)
        return r.stdout.split("\n")
    return run


def _random_cuts(rng, nb, T):
    inner = np.sort(rng.choice(np.arange(1, nb), T - 1, replace=False))
    return [0, *inner.tolist(), nb]


def _region_max(cnt, cuts):
    """Brute force: every cell block's region, periodic, summed by numpy."""
    nz, ny, nx = cnt.shape
    best = 0
    for tz in range(len(cuts[2]) - 1):
        iz = np.arange(cuts[2][tz] - 1, cuts[2][tz + 1] + 1) % nz
        for ty in range(len(cuts[1]) - 1):
            iy = np.arange(cuts[1][ty] - 1, cuts[1][ty + 1] + 1) % ny
            for tx in range(len(cuts[0]) - 1):
                ix = np.arange(cuts[0][tx] - 1, cuts[0][tx + 1] + 1) % nx
                best = max(best, int(cnt[np.ix_(iz, iy, ix)].sum()))
    return best


def test_largest_region_matches_brute_force(driver):
    rng = np.random.default_rng(7)
    cases, want = [], []
    for _ in range(60):
        nb = rng.integers(5, 19, 3)
        T = [int(rng.integers(1, min(n, 6) + 1)) for n in nb]
        cuts = [_random_cuts(rng, int(n), t) for n, t in zip(nb, T)]
        cnt = rng.integers(0, 6, (nb[2], nb[1], nb[0]))
        want.append(_region_max(cnt, cuts))
        cases.append(" ".join(map(str, [0, *nb, *T])) + "\n" +
                     "\n".join(" ".join(map(str, c)) for c in cuts) + "\n" +
                     " ".join(map(str, cnt.ravel())))
    out = driver(f"{len(cases)}\n" + "\n".join(cases) + "\n")
    got = [int(v) for v in out[:len(cases)]]
    assert got == want


def test_balanced_cuts_properties(driver):
    rng = np.random.default_rng(8)
    specs = []
    for k in range(40):
        nb = rng.integers(5, 25, 3)
        T = [int(rng.integers(1, min(n, 8) + 1)) for n in nb]
        if k < 10:
            cnt = np.full((nb[2], nb[1], nb[0]), 3)        # uniform
        else:
            cnt = rng.integers(0, 8, (nb[2], nb[1], nb[0]))
            cnt[:, :, : nb[0] // 3] *= 5                   # a dense slab
        specs.append((nb, T, cnt))
    text = f"{len(specs)}\n" + "\n".join(
        " ".join(map(str, [1, *nb, *T])) + "\n" + " ".join(map(str, cnt.ravel()))
        for nb, T, cnt in specs) + "\n"
    out = driver(text)
    for k, (nb, T, cnt) in enumerate(specs):
        for a in range(3):
            c = [int(v) for v in out[3 * k + a].split()]
            assert c[0] == 0 and c[-1] == nb[a] and len(c) == T[a] + 1
            assert all(c[t + 1] > c[t] for t in range(T[a]))
            slab = cnt.sum(axis=tuple(i for i in range(3) if i != 2 - a))
            if k < 10:
                even = [t * nb[a] / T[a] for t in range(T[a] + 1)]
                assert max(abs(x - y) for x, y in zip(c, even)) <= 1
            else:
                # each cell block within one slab of an equal share (the
                # floor weight the balancer adds to empty slabs included)
                w = slab + max(1e-3, 0.02 * cnt.sum() / nb[a])
                share = w.sum() / T[a]
                for t in range(T[a]):
                    tile = w[c[t]:c[t + 1]].sum()
                    if c[t + 1] - c[t] > 1:
                        assert tile <= share + w.max() + 1e-9


def test_block_of_position_faces(driver):
    nb, bl = 13, 1.25
    L = 16.9                                  # the last block is enlarged
    xs = [0.0, bl, np.nextafter(bl, 0.0), 5 * bl, np.nextafter(5 * bl, 0.0),
          12 * bl, np.nextafter(L, 0.0), 14 * bl + 0.3, -1e-12]
    want = [0, 1, 0, 5, 4, 12, 12, 12, 0]
    out = driver(f"1\n2 {nb} {nb} {nb} 1 1 1\n{bl!r} {len(xs)} " +
                 " ".join(repr(float(x)) for x in xs) + "\n")
    assert [int(v) for v in out[0].split()] == want


def _balanced(driver, nb, T, cnt):
    out = driver(f"1\n{' '.join(map(str, [1, *nb, *T]))}\n" +
                 " ".join(map(str, cnt.ravel())) + "\n")
    return [[int(v) for v in out[a].split()] for a in range(3)]


def _region_blocks(cuts):
    return int(np.prod([max(c[t + 1] - c[t] + 2 for t in range(len(c) - 1))
                        for c in cuts]))


def test_coarsened_map_within_limits_and_maximal(driver):
    rng = np.random.default_rng(9)
    coarsened = 0
    for k in range(12):
        nb = [int(v) for v in rng.integers(6, 16, 3)]
        T0 = [int(rng.integers(2, n // 2 + 1)) for n in nb]
        cnt = rng.integers(0, 4, (nb[2], nb[1], nb[0]))
        if k % 2:
            cnt[:, :, : nb[0] // 2] *= 3                      # a dense half
        c0 = _balanced(driver, nb, T0, cnt)
        # limits a little above the starting map's own largest region
        lim_a = int(_region_max(cnt, c0) * rng.uniform(1.0, 1.6))
        lim_b = int(_region_blocks(c0) * rng.uniform(1.0, 1.6))
        out = driver(f"1\n{' '.join(map(str, [3, *nb, *T0]))}\n{lim_a} {lim_b}\n" +
                     " ".join(map(str, cnt.ravel())) + "\n")
        T = [int(v) for v in out[0].split()]
        assert all(1 <= t <= t0 for t, t0 in zip(T, T0))
        coarsened += T != T0
        c = _balanced(driver, nb, T, cnt)
        assert _region_max(cnt, c) <= lim_a and _region_blocks(c) <= lim_b
        for a in range(3):
            if T[a] > 1:
                Tm = list(T)
                Tm[a] -= 1
                cm = _balanced(driver, nb, Tm, cnt)
                assert (_region_max(cnt, cm) > lim_a or
                        _region_blocks(cm) > lim_b), (k, a, T)
    assert coarsened >= 6


def _region_all(cnt, cuts):
    """Brute force: every cell block's region atoms, x fastest."""
    nz, ny, nx = cnt.shape
    out = []
    for tz in range(len(cuts[2]) - 1):
        iz = np.arange(cuts[2][tz] - 1, cuts[2][tz + 1] + 1) % nz
        for ty in range(len(cuts[1]) - 1):
            iy = np.arange(cuts[1][ty] - 1, cuts[1][ty + 1] + 1) % ny
            for tx in range(len(cuts[0]) - 1):
                ix = np.arange(cuts[0][tx] - 1, cuts[0][tx + 1] + 1) % nx
                out.append(int(cnt[np.ix_(iz, iy, ix)].sum()))
    return out


def test_region_atoms_of_every_cell_block(driver):
    """for_region_atoms (the costs of the per-XCD split) equals a numpy sum
    over each cell block's periodic region, on random maps and counts,
    including one cell block per axis (the region wraps onto itself)."""
    rng = np.random.default_rng(12)
    lines, want = [], []
    for case in range(30):
        nb = rng.integers(5, 14, 3)
        T = [int(rng.integers(1, max(2, v // 2) + 1)) for v in nb]
        if case % 7 == 0:
            T = [1, 1, 1]
        cuts = [_random_cuts(rng, int(nb[a]), T[a]) for a in range(3)]
        cnt = rng.integers(0, 6, (nb[2], nb[1], nb[0]))
        lines.append(" ".join(str(v) for v in
                              [4, *nb.tolist(), *T, *sum(cuts, []), *cnt.ravel().tolist()]))
        want.append(_region_all(cnt, cuts))
    out = driver(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    for k, w in enumerate(want):
        assert [int(v) for v in out[k].split()] == w, k


def test_equal_cost_ranges(driver):
    """Ranges of tiles for the XCDs: boundaries ordered from 0 to n, the
    smallest ones at which the prefix cost reaches each x / parts of the
    total (so every range's cost is within the largest tile of an equal
    share), and equal counts for equal costs."""
    rng = np.random.default_rng(13)
    lines, cases = [], []
    for case in range(40):
        n = int(rng.integers(1, 300))
        parts = 8
        cost = (rng.integers(1, 3000, n) if case % 5 else np.full(n, 7)).tolist()
        lines.append(" ".join(str(v) for v in [5, 1, 1, 1, 1, 1, 1, parts, n, *cost]))
        cases.append((parts, cost))
    out = driver(f"{len(lines)}\n" + "\n".join(lines) + "\n")
    for k, (parts, cost) in enumerate(cases):
        b = [int(v) for v in out[k].split()]
        n = len(cost)
        assert len(b) == parts + 1 and b[0] == 0 and b[-1] == n, k
        assert all(b[i] <= b[i + 1] for i in range(parts)), k
        pre = np.concatenate([[0], np.cumsum(cost)])
        total = pre[-1]
        for x in range(1, parts):
            # the first boundary whose prefix reaches x / parts of the total
            i = int(np.argmax(pre * parts >= total * x))
            assert b[x] == min(i, n), (k, x)
        share = total / parts
        for x in range(parts):
            assert pre[b[x + 1]] - pre[b[x]] <= share + max(cost) + 1e-9, (k, x)
        if len(set(cost)) == 1:
            sizes = np.diff(b)
            assert sizes.max() - sizes.min() <= 1, k
