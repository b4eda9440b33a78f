"""Batching form on the GPU: independent periodic systems of uneven sizes in
one handle, every kernel of a step covering all systems in one launch.

Bars: each system's FP64 forces within 1e-10 relative of the oracle run on
that system alone, pair counts and cell assignment bit-exact; the mixed path
bit-identical to the same system run in its own handle (fixed-point sums are
order-independent and the cell-block maps are the same), over single
evaluations and over MD runs replayed from captured graphs.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10

# (FCC cells per side, density, perturbation, seed): 500 to 4000 atoms,
# 6 to 13 blocks per side (one cell block per axis for the smallest)
SYSTEMS = [(5, 0.90, 0.10, 1), (7, 0.60, 0.05, 2), (6, 0.75, 0.20, 3),
           (10, 0.8442, 0.05, 4), (5, 0.65, 0.10, 5)]


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _systems(E, temperature=0.0):
    out = []
    for nc, rho, amp, seed in SYSTEMS:
        pos, vel, box = E.generate_fcc(nc, rho, temperature, perturb=amp,
                                       seed=seed)
        out.append((pos, vel, box))
    return out


def _batch(E, sy, **kw):
    b = E.MidpointLJBatch([(len(p), box) for p, _, box in sy], 2.5, **kw)
    b.set_state(np.vstack([p for p, _, _ in sy]),
                np.vstack([v for _, v, _ in sy]))
    return b


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def test_batch_fp64_forces_match_oracle(E, oracle):
    print("[batch fp64 vs oracle]", flush=True)
    sy = _systems(E)
    with _batch(E, sy) as b:
        f, stats, _ = b.forces("fp64")
        off = b.offsets
    for k, (pos, _, box) in enumerate(sy):
        ref, t, _ = oracle.forces(pos, box, 2.5)
        assert _rel(f[off[k]:off[k + 1]], ref) < FP64_TOL, k
        assert (stats[k].pairs_cand, stats[k].pairs_cut) == (t.cand, t.cut), k
        assert abs(stats[k].pe - t.pe) < FP64_TOL * abs(t.pe), k


def test_single_cell_block_per_axis_matches_oracle(E, oracle):
    """A system whose map has one cell block per axis: the region's ghost
    shell holds periodic images of the block's own atoms."""
    print("[one cell block]", flush=True)
    pos, _, box = E.generate_fcc(5, 0.9, 0.0, perturb=0.1, seed=9)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    with E.MidpointLJ(len(pos), box, 2.5, tiles=(1, 1, 1)) as m:
        m.set_state(pos)
        f, s, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    np.testing.assert_array_equal(got, hc)


def test_batch_mixed_equals_standalone(E):
    print("[batch mixed vs standalone]", flush=True)
    sy = _systems(E)
    with _batch(E, sy) as b:
        fb, sb, _ = b.forces("mixed")
        off = b.offsets
    for k, (pos, vel, box) in enumerate(sy):
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos, vel)
            f1, s1, _ = m.forces("mixed")
        assert np.array_equal(fb[off[k]:off[k + 1]], f1), k
        assert (sb[k].pe, sb[k].pairs_cand, sb[k].pairs_cut) == \
            (s1.pe, s1.pairs_cand, s1.pairs_cut), k
        assert abs(sb[k].ke - s1.ke) <= 1e-12 * s1.ke, k


def test_batch_run_equals_standalone(E):
    """45 mixed MD steps, tallies every 10 (two replays of a captured
    20-step chunk of the whole batch, then 5 eager steps), then 10 eager
    steps, against each system run on its own with the same chunking."""
    print("[batch run vs standalone]", flush=True)
    sy = _systems(E, temperature=1.0)
    with _batch(E, sy, graph_chunk=20) as b:
        rows = b.run(45, 0.005, "mixed", energy_every=10)
        rows += b.run(10, 0.005, "mixed", energy_every=10)
        pb, vb = b.get_state()
        off = b.offsets
    assert [r[0].step for r in rows] == [10, 20, 30, 40, 45, 10]
    for k, (pos, vel, box) in enumerate(sy):
        with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=20) as m:
            m.set_state(pos, vel)
            st = m.run(45, 0.005, "mixed", energy_every=10)
            st += m.run(10, 0.005, "mixed", energy_every=10)
            p1, v1 = m.get_state()
        assert np.array_equal(pb[off[k]:off[k + 1]], p1), k
        assert np.array_equal(vb[off[k]:off[k + 1]], v1), k
        for r, s in zip(rows, st):
            assert (r[k].step, r[k].pe, r[k].pairs_cut) == \
                (s.step, s.pe, s.pairs_cut), k
            assert abs(r[k].ke - s.ke) <= 1e-12 * s.ke, k


def test_batch_block_ids_and_ownership(E, oracle):
    print("[batch ids + ownership]", flush=True)
    sy = _systems(E)
    with _batch(E, sy) as b:
        num, xyz = b.block_ids()
        hc = b.home_pair_counts("fp64")
        off = b.offsets
    boff = 0
    for k, (pos, _, box) in enumerate(sy):
        d = oracle.domain_create(box, 1.25)
        want = np.array([oracle.identify_block(d, p)[2] for p in pos])
        np.testing.assert_array_equal(num[off[k]:off[k + 1]], want)
        _, _, ref_hc = oracle.forces(pos, box, 2.5, home_counts=True)
        np.testing.assert_array_equal(hc[boff:boff + d.nblocks], ref_hc)
        boff += d.nblocks
    assert boff == len(hc)


@pytest.mark.parametrize("batch", [False, True])
def test_persistent_grid_matches(E, batch):
    """A persistent force grid (workgroups looping over the cell blocks)
    gives the same mixed forces bit for bit and the same FP64 forces to
    rounding.  Three workgroups run on three of the eight XCDs, so the cell
    blocks of the other five XCDs' ranges are all taken from other ranges."""
    print(f"[persistent grid] batch={batch}", flush=True)
    if batch:
        sy = _systems(E)
        h = _batch(E, sy)
        cfg = {"tile_edge": 5}
    else:
        pos, vel, box = E.generate_fcc(12, 0.8442, 0.0, perturb=0.1, seed=8)
        h = E.MidpointLJ(len(pos), box, 2.5)
        h.set_state(pos)
        cfg = {"tiles": (4, 4, 3)}
    with h:
        out = {}
        for wgs in (0, 3, 16, 24):
            for prec in ("mixed", "fp64"):
                h.set_config(prec, block_threads=256, workgroups=wgs, **cfg)
                assert h.get_config(prec)["workgroups"] == wgs
                out[wgs, prec] = h.forces(prec)[0]
    for wgs in (3, 16, 24):
        assert np.array_equal(out[wgs, "mixed"], out[0, "mixed"])
        assert _rel(out[wgs, "fp64"], out[0, "fp64"]) < 1e-13


def test_batch_autotune(E, oracle):
    print("[batch autotune]", flush=True)
    sy = _systems(E)
    with _batch(E, sy) as b:
        best = b.autotune("mixed", trials=1)
        assert best["tile_edge"] > 0 and best["force_ms"] > 0
        f, _, _ = b.forces("mixed")
        off = b.offsets
        # the tuned shape goes back in as it came out
        cuts = [b.get_cuts("mixed", k) for k in range(len(sy))]
        b.set_config("mixed", **best)
        again = b.get_config("mixed")
        assert {k: v for k, v in again.items() if k != "force_ms"} == \
            {k: v for k, v in best.items() if k != "force_ms"}
        for k in range(len(sy)):
            for a in range(3):
                np.testing.assert_array_equal(b.get_cuts("mixed", k)[a], cuts[k][a])
        f2, _, _ = b.forces("mixed")
    assert np.array_equal(f, f2)
    pos, _, box = sy[3]
    ref, _, _ = oracle.forces(pos, box, 2.5)
    assert _rel(f[off[3]:off[4]], ref) < 2e-5


def _small_slab(E, seed=3):
    box = np.array([24.0, 12.0, 12.0])
    pos, vel = E.generate_slab(box, 8.0, 16.0, 0.8, 0.05, min_sep=0.85,
                               temperature=0.85, seed=seed)
    return pos, vel, box


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_balanced_map_matches_oracle(E, oracle, prec):
    """A map whose cell-block boundaries sit at equal atom shares (a slab:
    dense middle, vapour outside) computes every pair exactly once: FP64
    forces, pair counts and per-home-block ownership against the oracle,
    the mixed path within its FP32 bars."""
    print(f"[balanced map] {prec}", flush=True)
    pos, vel, box = _small_slab(E)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        m.set_config(prec, (6, 2, 2), 256, balance=1)
        assert m.get_config(prec)["balance"] == 1
        f, s, _ = m.forces(prec)
        got = m.home_pair_counts(prec)
        m.set_config(prec, (6, 2, 2), 256, balance=0)
        f0, _, _ = m.forces(prec)
    if prec == "fp64":
        assert _rel(f, ref) < FP64_TOL
        assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
        np.testing.assert_array_equal(got, hc)
    else:
        assert _rel(f, ref) < 2e-5
        assert s.pairs_cand == t.cand and abs(s.pairs_cut - t.cut) <= 2
        assert np.abs(got - hc).sum() <= 4
        assert _rel(f, f0) < 2e-5


def test_balanced_batch_equals_standalone(E, oracle):
    """Balanced maps from a tile edge in a batch: each system's map is
    coarsened as far as the LDS budget of one workgroup allows (some systems
    get fewer cell blocks than from the edge alone), its mixed forces are
    bit-identical to the system alone under the same map, and its FP64
    forces under the coarsened map match the oracle (1e-10, pair counts
    exact)."""
    print("[balanced batch]", flush=True)
    sy = [_small_slab(E, seed=s) for s in (4, 5)] + _systems(E)[:2]
    with _batch(E, sy) as b:
        b.set_config("mixed", block_threads=256, tile_edge=4, balance=1)
        b.set_config("fp64", block_threads=256, tile_edge=4, balance=1)
        fb, _, _ = b.forces("mixed")
        f64, s64, _ = b.forces("fp64")
        off = b.offsets
        cuts = [b.get_cuts("mixed", k) for k in range(len(sy))]
    for k, (pos, _, box) in enumerate(sy):
        ref, t, _ = oracle.forces(pos, box, 2.5)
        assert _rel(f64[off[k]:off[k + 1]], ref) < 1e-10, k
        assert (s64[k].pairs_cand, s64[k].pairs_cut) == (t.cand, t.cut), k
    coarser = 0
    for k, (pos, vel, box) in enumerate(sy):
        T = tuple(len(c) - 1 for c in cuts[k])
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos, vel)
            m.set_config("mixed", block_threads=256, tile_edge=4, balance=1)
            T_edge = tuple(len(c) - 1 for c in m.get_cuts("mixed"))
            m.set_config("mixed", T, 256, balance=1)
            for a in range(3):
                np.testing.assert_array_equal(m.get_cuts("mixed")[a],
                                              cuts[k][a])
            f1, _, _ = m.forces("mixed")
        assert all(t <= te for t, te in zip(T, T_edge)), (k, T, T_edge)
        coarser += T != T_edge
        assert np.array_equal(fb[off[k]:off[k + 1]], f1), k
    assert coarser > 0


def test_batch_log_uniform_32_systems_match_oracle(E, oracle):
    """The batching workload's shape at 32 systems: atom counts log-uniform
    in [500, 20000], densities uniform in [0.60, 0.90], temperatures uniform
    in [0.70, 1.50] (randomly depleted FCC lattices), under maps from a tile
    edge coarsened within the LDS budget of a workgroup, boundaries at equal
    atom shares: every system's FP64 forces within 1e-10 of the oracle on
    that system alone, pair counts exact; the mixed forces within the FP32
    bars and summing to zero per system in fixed point."""
    rng = np.random.default_rng(32)
    n = np.rint(np.exp(rng.uniform(np.log(500), np.log(20000), 32))).astype(int)
    rho = rng.uniform(0.60, 0.90, 32)
    temp = rng.uniform(0.70, 1.50, 32)
    sy = [E.generate_vacancy_fcc(int(n[k]), rho[k], temp[k], perturb=0.05,
                                 seed=3200 + k) for k in range(32)]
    print(f"[batch 32] atoms {sum(len(p) for p, _, _ in sy)}", flush=True)
    with _batch(E, sy) as b:
        for prec in ("fp64", "mixed"):
            b.set_config(prec, block_threads=512, tile_edge=5, balance=1)
        f64, s64, _ = b.forces("fp64")
        fmx, smx, _ = b.forces("mixed")
        off = b.offsets
        R = [H.region_bound(b.get_cuts("mixed", k), 1.25) for k in range(32)]
    for k, (pos, _, box) in enumerate(sy):
        ref, t, _ = oracle.forces(pos, box, 2.5)
        a, z = off[k], off[k + 1]
        assert _rel(f64[a:z], ref) < FP64_TOL, k
        assert (s64[k].pairs_cand, s64[k].pairs_cut) == (t.cand, t.cut), k
        # FP32 rounding at the scale of the system's (coarsened) regions,
        # relative to each atom's pair forces (a perturbed lattice's net
        # forces are small beside them), with a factor 2 of margin
        ncross = abs(smx[k].pairs_cut - t.cut)
        assert ncross <= 2, k
        assert H.mixed_error_bound(fmx[a:z], ref, pos, box, 2.5, R[k],
                                   ncross) < 2.0, k
        assert smx[k].pairs_cand == t.cand, k
        q = np.round(fmx[a:z] * 2.0 ** 32).astype(np.int64)
        assert np.all(q.sum(0) == 0), k
