"""The non-uniform-density workload (liquid-vapour slab, rc 3.0) against the
oracle on the GPU, through the C ABI.

Workload: box 3:1:1, the central third of x at density 0.80 and the rest at
0.05, uniform with a 0.8 sigma minimum separation, T* 0.85, cutoff 3.0
(block length rc/2 = 1.5, exaMD.f90:56).  The map is the one the reference
would take for it: cell-block boundaries at equal atom shares along each
axis (its METIS partition weighted by atom counts,
SRC/exaMD_partitioning.f90:199-203).

Bars: FP64 per-atom forces within 1e-10 relative (normwise) and energy within
1e-10 of the oracle, candidate and in-cutoff pair counts and per-home-block
ownership equal; the mixed path within its FP32 bars, the same candidate
count and the sum of the fixed-point forces exactly zero.  At the full
4,000,000-atom size: the same FP64 bars against the threaded oracle, the
candidate count equal to the block-occupancy formula, and the fixed-point
sum exactly zero.
"""
import os

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

RC = 3.0
FP64_TOL = 1e-10
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


@pytest.fixture(scope="module")
def slab(E):
    return H.slab_system(E, 120000, seed=41)


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def _fixed_sum_zero(f):
    q = np.round(f * 2.0 ** 32).astype(np.int64)
    return bool(np.all(q.sum(0) == 0))


def _region_bound(m, prec):
    """A bound on |p| over every cell-block region of the map (coordinates
    relative to the cell-block centre): half the widest cell block plus two
    blocks per axis."""
    bl = 0.5 * RC
    w = [np.diff(c).max() for c in m.get_cuts(prec)]
    return float(np.sqrt(sum(((0.5 * x + 2.0) * bl) ** 2 for x in w)))


def _mixed_within_bars(f, ref, R):
    """FP32 coordinates relative to a cell-block centre, with |p| <= R, put
    up to 2^-24 R into each coordinate difference; a pair at the 0.8 sigma
    minimum separation, whose force goes as r^-13, carries that as
    13 * 2^-24 R / 0.8 of itself (2e-5 at R = 10 sigma; the wide cell
    blocks of the vapour reach R = 47 sigma).  So each atom's error is
    bounded relative to its own force (or the RMS force, for small ones),
    with a factor 2 over that estimate; the RMS error, dominated by the many
    pairs the random packing puts near that separation, is a few 1e-6 of
    the RMS force (about 1e-6 on the liquid at rc 2.5)."""
    err = np.sqrt(((f - ref) ** 2).sum(1))
    fmag = np.sqrt((ref ** 2).sum(1))
    scale = np.sqrt((fmag ** 2).mean())
    bar = 2.0 * 13.0 * 2.0 ** -24 * R / 0.8
    print(f"mixed err: max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} "
          f"rms|F| {scale:.3e} max err/(|F|+rms) {(err / (fmag + scale)).max():.3e}"
          f" bar {bar:.2e} (R {R:.1f})", flush=True)
    assert (err / (fmag + scale)).max() < bar
    assert np.sqrt((err ** 2).mean()) / scale < 5e-6


def _balanced(m, prec, tune):
    """The tuned launch shape with boundaries at equal atom shares."""
    m.set_config(prec, tune["tiles"], tune["block_threads"], tune["run_groups"],
                 tune["lane_queue"], tune["tile_edge"], tune["workgroups"],
                 balance=1)
    assert m.get_config(prec)["balance"] == 1


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_slab_autotuned_balanced_map(E, oracle, slab, prec):
    pos, vel, box = slab
    print(f"[slab rc3 {prec}] n={len(pos)} box={box.round(2).tolist()}",
          flush=True)
    ref, t, hc = oracle.forces(pos, box, RC, home_counts=True, threads=THREADS)
    with E.MidpointLJ(len(pos), box, RC) as m:
        m.set_state(pos, vel)
        _balanced(m, prec, m.autotune(prec, trials=1))
        f, s, _ = m.forces(prec)
        got = m.home_pair_counts(prec)
        R = _region_bound(m, prec)
    assert s.pairs_cand == t.cand
    if prec == "fp64":
        assert _rel(f, ref) < FP64_TOL
        assert abs(s.pe - t.pe) < FP64_TOL * abs(t.pe)
        assert s.pairs_cut == t.cut
        np.testing.assert_array_equal(got, hc)
    else:
        _mixed_within_bars(f, ref, R)
        assert abs(s.pairs_cut - t.cut) <= 2
        assert np.abs(got - hc).sum() <= 4
        assert abs(s.pe - t.pe) < 1e-6 * abs(t.pe)
        assert _fixed_sum_zero(f)


def test_slab_named_map(E, oracle, slab):
    """32 x 12 x 12 cell blocks with boundaries at equal atom shares (the
    workload's named map) computes every pair once."""
    pos, vel, box = slab
    print(f"[slab rc3 32x12x12] n={len(pos)}", flush=True)
    ref, t, hc = oracle.forces(pos, box, RC, home_counts=True, threads=THREADS)
    with E.MidpointLJ(len(pos), box, RC) as m:
        m.set_state(pos, vel)
        m.set_config("fp64", (32, 12, 12), 256, balance=1)
        m.set_config("mixed", (32, 12, 12), 256, balance=1)
        cuts = m.get_cuts("fp64")
        assert tuple(len(c) - 1 for c in cuts) == (32, 12, 12)
        f64, s64, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
        fmx, smx, _ = m.forces("mixed")
        R = _region_bound(m, "mixed")
    assert _rel(f64, ref) < FP64_TOL
    assert (s64.pairs_cand, s64.pairs_cut) == (t.cand, t.cut)
    np.testing.assert_array_equal(got, hc)
    _mixed_within_bars(fmx, ref, R)
    assert smx.pairs_cand == t.cand
    assert _fixed_sum_zero(fmx)
    # the boundaries follow the atoms: cell blocks are narrower in the slab
    w = np.diff(cuts[0])
    lo, hi = len(w) // 3, 2 * len(w) // 3
    assert w[lo:hi].mean() < w[:lo].mean()


def test_slab_full_size(E, oracle):
    """The full 4,000,000-atom slab: FP64 against the threaded oracle,
    P_cand against the occupancy formula, the fixed-point sum exactly zero,
    under 32 x 16 x 16 cell blocks at equal atom shares.  The named
    32 x 12 x 12 map does not fit one workgroup's LDS at this size, split
    evenly (a dense region of 5,594 atoms) or balanced (a vapour cell block
    56 blocks wide at the slab face needs about 190 KB with its block table
    and 15% slack), and is refused with a status rather than run."""
    print("[slab rc3 full size] generating 4,000,000 atoms", flush=True)
    pos, vel, box = H.slab_system(E, 4000000, seed=4)
    print(f"[slab rc3 full size] n={len(pos)} oracle on {THREADS} threads",
          flush=True)
    ref, t, hc = oracle.forces(pos, box, RC, home_counts=True, threads=THREADS)
    cand = H.candidate_pairs(pos, box, RC)
    assert t.cand == cand
    with E.MidpointLJ(len(pos), box, RC) as m:
        m.set_state(pos, vel)
        with pytest.raises(E.ExamdError) as ei:
            m.set_config("mixed", (32, 12, 12), 1024, balance=1)
        assert ei.value.status == E.ERR_TILE_OVERFLOW
        for prec in ("fp64", "mixed"):
            m.set_config(prec, (32, 16, 16), 1024, balance=1)
        print("[slab rc3 full size] fp64", flush=True)
        f64, s64, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
        print("[slab rc3 full size] mixed", flush=True)
        fmx, smx, _ = m.forces("mixed")
        R = _region_bound(m, "mixed")
    assert _rel(f64, ref) < FP64_TOL
    assert abs(s64.pe - t.pe) < FP64_TOL * abs(t.pe)
    assert (s64.pairs_cand, s64.pairs_cut) == (t.cand, t.cut)
    np.testing.assert_array_equal(got, hc)
    assert smx.pairs_cand == cand
    assert _fixed_sum_zero(fmx)
    _mixed_within_bars(fmx, ref, R)
