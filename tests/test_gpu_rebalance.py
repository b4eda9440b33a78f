"""Periodic re-partition of the load-balanced cell-block map during a run
(examd_tune.rebalance_every), on a slab that translates along x so the
balanced boundaries have to follow it.

A run that re-partitions every R steps must be exactly a run that stops
every R steps, applies the balanced map of the current positions with
set_config and starts again: mixed results bit for bit (its sums are
deterministic), FP64 to the rounding of its summation order.  The map after
the run is the balanced map of the final positions (the device's block
counts against the host's binning of the same positions), and the forces
under the re-partitioned map match the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILES = (6, 2, 2)
DT = 0.005
R = 50
STEPS = 200


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _moving_slab(E, seed):
    """Slab ρ 0.8 in the middle third of a 36x12x12 box, vapour ρ 0.05,
    everything drifting at +3 along x (3σ over the 200 steps)."""
    box = np.array([36.0, 12.0, 12.0])
    pos, vel = E.generate_slab(box, 12.0, 24.0, 0.8, 0.05, min_sep=0.85,
                               temperature=0.85, seed=seed)
    vel[:, 0] += 3.0
    return pos, vel, box


def _same(a, b, prec):
    """Bit for bit (mixed) or within 1e-10 relative (FP64)."""
    a, b = np.asarray(a), np.asarray(b)
    if prec == "mixed":
        return np.array_equal(a, b)
    return np.abs(a - b).max() <= 1e-10 * max(1.0, np.abs(b).max())


@pytest.mark.parametrize("prec,chunk", [("mixed", 25), ("mixed", 0),
                                        ("fp64", 25)])
def test_rebalance_equals_stop_and_reapply(E, oracle, prec, chunk):
    print(f"[rebalance] {prec} chunk={chunk}", flush=True)
    pos, vel, box = _moving_slab(E, 7)
    n = len(pos)
    # A: one run, the map re-partitioned every R steps inside it
    with E.MidpointLJ(n, box, 2.5, graph_chunk=chunk) as m:
        m.set_state(pos, vel)
        m.set_config(prec, TILES, 256, balance=1, rebalance_every=R)
        assert m.get_config(prec)["rebalance_every"] == R
        cuts0 = m.get_cuts(prec)
        st_a = m.run(STEPS, DT, prec, energy_every=R)
        cuts_a = m.get_cuts(prec)
        p_a, v_a = m.get_state()
        f_a, s_a, _ = m.forces(prec)
    # B: stop every R steps and apply the balanced map by hand
    st_b = []
    with E.MidpointLJ(n, box, 2.5, graph_chunk=chunk) as m:
        m.set_state(pos, vel)
        m.set_config(prec, TILES, 256, balance=1)
        for _ in range(STEPS // R):
            st_b += m.run(R, DT, prec, energy_every=R)
            m.set_config(prec, TILES, 256, balance=1)
        cuts_b = m.get_cuts(prec)
        p_b, v_b = m.get_state()
    # the slab moved, and the boundaries along x with it
    assert not np.array_equal(cuts0[0], cuts_a[0])
    for a in range(3):
        np.testing.assert_array_equal(cuts_a[a], cuts_b[a])
    assert [x.step for x in st_a] == list(range(R, STEPS + 1, R))
    assert len(st_b) == STEPS // R
    for x, y in zip(st_a, st_b):
        assert (x.pairs_cand, x.pairs_cut) == (y.pairs_cand, y.pairs_cut)
        assert _same(x.pe, y.pe, prec)
        # a tally inside a run sums v^2 in other groups than the closing
        # kick of a run that ends there
        assert abs(x.ke - y.ke) <= 1e-12 * abs(y.ke)
    assert _same(p_a, p_b, prec) and _same(v_a, v_b, prec)
    # forces under the re-partitioned map against the oracle
    ref, t, _ = oracle.forces(p_a, box, 2.5)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert s_a.pairs_cand == t.cand
    if prec == "fp64":
        assert np.abs(f_a - ref).max() / scale < 1e-10
        assert s_a.pairs_cut == t.cut
    else:
        assert np.abs(f_a - ref).max() / scale < 2e-5
        assert abs(s_a.pairs_cut - t.cut) <= 2


def test_rebalance_batch_equals_stop_and_reapply(E):
    """The same in a batch of two moving slabs with maps from a tile edge:
    every system's map re-partitioned, mixed results bit for bit."""
    print("[rebalance batch]", flush=True)
    sy = [_moving_slab(E, s) for s in (8, 9)]
    systems = [(len(p), box) for p, _, box in sy]
    pos = np.vstack([p for p, _, _ in sy])
    vel = np.vstack([v for _, v, _ in sy])
    out = []
    for manual in (False, True):
        with E.MidpointLJBatch(systems, 2.5, graph_chunk=25) as b:
            b.set_state(pos, vel)
            b.set_config("mixed", block_threads=256, tile_edge=5, balance=1,
                         rebalance_every=0 if manual else R)
            c0 = [b.get_cuts("mixed", k) for k in range(2)]
            if manual:
                for _ in range(STEPS // R):
                    b.run(R, DT, "mixed")
                    b.set_config("mixed", block_threads=256, tile_edge=5,
                                 balance=1)
            else:
                b.run(STEPS, DT, "mixed")
            c1 = [b.get_cuts("mixed", k) for k in range(2)]
            p, v = b.get_state()
        out.append((c0, c1, p, v))
    (c0a, c1a, pa, va), (c0b, c1b, pb, vb) = out
    for k in range(2):
        assert not np.array_equal(c0a[k][0], c1a[k][0])
        for a in range(3):
            np.testing.assert_array_equal(c1a[k][a], c1b[k][a])
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_rebalance_config_checks(E):
    print("[rebalance config]", flush=True)
    pos, vel, box = _moving_slab(E, 10)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        with pytest.raises(E.ExamdError):
            m.set_config("mixed", TILES, 256, balance=1, rebalance_every=-1)
        m.set_config("mixed", TILES, 256, balance=0, rebalance_every=10)
        cuts = m.get_cuts("mixed")
        # an even map stays even: re-partition only renews the capacity
        m.run(30, DT, "mixed")
        after = m.get_cuts("mixed")
    assert [len(c) for c in cuts] == [t + 1 for t in TILES]
    for a in range(3):
        np.testing.assert_array_equal(cuts[a], after[a])
        assert cuts[a][0] == 0


def test_rebalance_with_species_in_a_batch(E):
    """Periodic re-partition with two species (Kob-Andersen rows) in a batch
    of two moving slabs: equal to stopping and re-applying the balanced map,
    mixed results bit for bit."""
    from tests import helpers as H
    print("[rebalance species batch]", flush=True)
    sy = [_moving_slab(E, s) for s in (11, 12)]
    systems = [(len(p), box) for p, _, box in sy]
    pos = np.vstack([p for p, _, _ in sy])
    vel = np.vstack([v for _, v, _ in sy])
    types = H.binary_types(len(pos), 0.2, 13)
    out = []
    for manual in (False, True):
        with E.MidpointLJBatch(systems, 2.5, graph_chunk=25) as b:
            b.set_state(pos, vel)
            b.set_species(types, H.KA_EPS, H.KA_SIG, H.KA_RC)
            b.set_config("mixed", block_threads=256, tile_edge=5, balance=1,
                         rebalance_every=0 if manual else R)
            if manual:
                for _ in range(STEPS // R):
                    b.run(R, DT, "mixed")
                    b.set_config("mixed", block_threads=256, tile_edge=5,
                                 balance=1)
            else:
                b.run(STEPS, DT, "mixed")
            cuts = [b.get_cuts("mixed", k) for k in range(2)]
            p, v = b.get_state()
        out.append((cuts, p, v))
    (ca, pa, va), (cb, pb, vb) = out
    for k in range(2):
        for a in range(3):
            np.testing.assert_array_equal(ca[k][a], cb[k][a])
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
