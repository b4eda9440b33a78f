"""The headline system at its full size: 2,048,000 atoms (80^3 FCC at
ρ 0.8442, positions perturbed by up to ±0.05σ), 107^3 midpoint blocks, under
the launch shapes the autotuner picks for the liquid (mixed: 12^3 balanced
cell blocks, 1024 threads, 2 run groups, queues of 4, 256 persistent
workgroups; FP64: 16^3 balanced, 1024 threads, 2 run groups, queues of 8).

Against the oracle on the same positions: cell assignment of every atom,
candidate and in-cutoff pair counts and per-home-block ownership exact, FP64
forces within 1e-10 relative and energy within 1e-10; the mixed path within
its FP32 bars, with Newton's third law exact in fixed point and a
bit-identical repeat of a 50-step run.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

RHO = 0.8442
MIXED = dict(tiles=(12, 12, 12), block_threads=1024, run_groups=2,
             lane_queue=4, workgroups=256, balance=1)
FP64 = dict(tiles=(16, 16, 16), block_threads=1024, run_groups=2,
            lane_queue=8, balance=1)


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


@pytest.fixture(scope="module")
def system(oracle):
    pos, L = H.perturbed_fcc(80, RHO, 0.05, 7)
    box = np.array([L, L, L])
    print("[full size] oracle on 2,048,000 atoms", flush=True)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    return pos, box, ref, t, hc


def test_fullsize_fp64_matches_oracle(E, oracle, system):
    pos, box, ref, t, hc = system
    print("[full size] fp64", flush=True)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        assert m.nblocks == 107 ** 3
        m.set_state(pos)
        m.set_config("fp64", **FP64)
        f, s, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
        num, _ = m.block_ids()
    _, _, bid = oracle.bin_atoms(oracle.domain_create(box, 1.25), pos)
    np.testing.assert_array_equal(num, bid + 1)   # bin_atoms counts from 0
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert np.abs(f - ref).max() / scale < 1e-10
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    assert abs(s.pe - t.pe) < 1e-10 * abs(t.pe)
    np.testing.assert_array_equal(got, hc)


def test_fullsize_mixed_near_oracle(E, oracle, system):
    """On the liquid (the lattice melted by 300 mixed steps under the default
    shape; the balanced 12^3 map is the autotuner's choice for the liquid and
    does not fit the lattice's regions): forces, pairs and energy against the
    oracle on the same positions, where close pairs make the largest FP32
    error (relative to the RMS force) about 3e-3."""
    pos, box, _, _, _ = system
    print("[full size] mixed", flush=True)
    rng = np.random.default_rng(8)
    vel = rng.normal(0.0, np.sqrt(0.72), pos.shape)
    vel -= vel.mean(0)
    with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=25) as m:
        m.set_state(pos, vel)
        m.run(300, 0.005, "mixed")
        p0, v0 = m.get_state()
        m.set_config("mixed", **MIXED)
        f, s, _ = m.forces("mixed")
        got = m.home_pair_counts("mixed")
        # a 50-step run twice from the same state: bit for bit
        ends = []
        for _ in range(2):
            m.set_state(p0, v0)
            m.set_config("mixed", **MIXED)
            st = m.run(50, 0.005, "mixed", energy_every=25)
            ends.append((m.get_state(), [(x.pe, x.ke, x.pairs_cut) for x in st]))
    print("[full size] oracle on the liquid", flush=True)
    ref, t, hc = oracle.forces(p0, box, 2.5, home_counts=True)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    err = np.sqrt(((f - ref) ** 2).sum(1))
    assert err.max() / scale < 1e-2
    assert np.sqrt((err ** 2).mean()) / scale < 2e-5
    assert s.pairs_cand == t.cand
    assert abs(s.pairs_cut - t.cut) <= max(2, int(1e-5 * t.cut))
    assert abs(s.pe - t.pe) < 1e-6 * abs(t.pe)
    assert np.abs(got - hc).sum() <= max(2, int(1e-5 * t.cut))
    q = np.round(f * 2.0 ** 32).astype(np.int64)
    assert np.all(q.sum(0) == 0)
    (s1, t1), (s2, t2) = ends
    assert np.array_equal(s1[0], s2[0]) and np.array_equal(s1[1], s2[1])
    assert t1 == t2
