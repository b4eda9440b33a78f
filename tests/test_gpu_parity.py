"""HIP path against the oracle, through the C ABI, on the GPU.

Bars: cell assignment, candidate/in-cutoff pair counts and per-home-block
pair ownership bit-exact; FP64 per-atom forces within 1e-10 relative
(normwise: max |dF_i| / RMS |F|) and energy within 1e-10 relative; the mixed
path within the FP32 bars stated per test, with bit-exact fixed-point
invariants (sum F = 0, order and tile-size independence).
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

RHO = 0.8442
FP64_TOL = 1e-10


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _system(E, pos, box, rc=2.5, **kw):
    m = E.MidpointLJ(len(pos), box, rc, **kw)
    m.set_state(pos)
    return m


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def _case(nc=10, amp=0.05, seed=1):
    pos, L = H.perturbed_fcc(nc, RHO, amp, seed)
    return pos, np.array([L, L, L])


@pytest.mark.parametrize("nc,amp,seed", [(10, 0.05, 1), (12, 0.2, 2), (7, 0.1, 3)])
def test_fp64_forces_match_oracle(E, oracle, nc, amp, seed):
    print(f"[fp64 parity] nc={nc} amp={amp} seed={seed}", flush=True)
    pos, box = _case(nc, amp, seed)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    with _system(E, pos, box) as m:
        f, s, _ = m.forces("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert s.pairs_cand == t.cand
    assert s.pairs_cut == t.cut
    assert abs(s.pe - t.pe) < FP64_TOL * abs(t.pe)


def test_fp64_config1_shape(E, oracle):
    """Config-1 shape: 32,000 atoms, 20^3 FCC, 26^3 blocks, 4x4x4 cell blocks."""
    print("[fp64 parity] config-1 shape 32000 atoms", flush=True)
    pos, vel, box = E.generate_fcc(20, RHO, 0.72, perturb=0.05, seed=7)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    with _system(E, pos, box, tiles=(4, 4, 4)) as m:
        f, s, _ = m.forces("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)


def test_block_ids_bit_exact(E, oracle):
    print("[block ids]", flush=True)
    pos, box = _case(10, 0.3, 4)
    bl = 1.25
    # add atoms exactly on faces, one ulp below faces, in the enlarged block
    nb = int(box[0] / bl)
    extra = []
    for k in range(nb):
        x = k * bl
        extra.append([x, np.nextafter(x, 0) if k else 0.0, box[0] - 1e-12])
    pos = np.vstack([pos, np.array(extra)])
    d = oracle.domain_create(box, bl)
    want = np.array([oracle.identify_block(d, p)[2] for p in pos])
    with _system(E, pos, box) as m:
        num, xyz = m.block_ids()
    np.testing.assert_array_equal(num, want)
    np.testing.assert_array_equal(
        num, xyz[:, 0] + (xyz[:, 1] - 1) * nb + (xyz[:, 2] - 1) * nb * nb)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_home_pair_ownership_bit_exact(E, oracle, prec):
    print(f"[ownership] {prec}", flush=True)
    pos, box = _case(9, 0.1, 5)
    _, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    with _system(E, pos, box, tiles=(3, 2, 4)) as m:
        got = m.home_pair_counts(prec)
    if prec == "fp64":
        np.testing.assert_array_equal(got, hc)
    else:
        # FP32 r^2 may move a pair within ~1e-7 of rc across the cutoff
        assert np.abs(got - hc).sum() <= max(2, int(1e-5 * t.cut))


@pytest.mark.parametrize("tiles,threads", [((2, 2, 2), 256), ((2, 3, 5), 512),
                                           ((13, 13, 13), 256), ((4, 4, 4), 1024)])
def test_fp64_tile_shapes(E, oracle, tiles, threads):
    print(f"[tiles] {tiles} threads={threads}", flush=True)
    pos, box = _case(10, 0.1, 6)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    with _system(E, pos, box, tiles=tiles, block_threads=threads) as m:
        f, s, _ = m.forces("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)


def test_mixed_forces_near_oracle(E, oracle):
    print("[mixed parity]", flush=True)
    pos, box = _case(12, 0.05, 8)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    with _system(E, pos, box) as m:
        f, s, _ = m.forces("mixed")
    err = np.sqrt(((f - ref) ** 2).sum(1))
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert err.max() / scale < 2e-5
    assert np.sqrt((err ** 2).mean()) / scale < 2e-6
    assert s.pairs_cand == t.cand
    assert abs(s.pairs_cut - t.cut) <= 2
    assert abs(s.pe - t.pe) < 1e-6 * abs(t.pe)


def test_mixed_fixed_point_invariants(E):
    """Sum F = 0 exactly; bit-identical forces under atom permutation and
    under a different workgroup size (order-independent fixed-point
    accumulation).  A different cell-block map moves the FP32 coordinate
    origin, so it agrees to FP32 rounding only."""
    print("[mixed invariants]", flush=True)
    pos, box = _case(11, 0.1, 9)
    with _system(E, pos, box, tiles=(3, 5, 2)) as m:
        f1, _, _ = m.forces("mixed")
        f1b, _, _ = m.forces("mixed")
    assert np.array_equal(f1, f1b)
    q = np.round(f1 * 2.0 ** 32).astype(np.int64)
    assert np.all(q.sum(0) == 0)
    perm = np.random.default_rng(1).permutation(len(pos))
    with _system(E, pos[perm], box, tiles=(3, 5, 2), block_threads=512) as m:
        f2, _, _ = m.forces("mixed")
    assert np.array_equal(f2, f1[perm])
    with _system(E, pos, box, tiles=(4, 4, 4), block_threads=1024) as m:
        f3, _, _ = m.forces("mixed")
    assert _rel(f3, f1) < 2e-5


def test_fp64_run_tracks_oracle(E, oracle):
    print("[fp64 run]", flush=True)
    pos, vel, box = E.generate_fcc(8, RHO, 0.72, perturb=0.02, seed=3)
    p_ref, v_ref, f_ref, ke, pe, cut = oracle.run_vv(pos, vel, box, 2.5,
                                                     0.005, 10)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        st = m.run(10, 0.005, "fp64", energy_every=1)
        p, v = m.get_state()
    dp = p - p_ref
    dp -= np.round(dp / box) * box
    assert np.abs(dp).max() < 1e-10
    assert np.abs(v - v_ref).max() < 1e-10
    assert len(st) == 10
    for k in range(10):
        assert abs(st[k].ke - ke[k]) < 1e-10 * ke[k]
        assert abs(st[k].pe - pe[k]) < 1e-10 * abs(pe[k])
        assert st[k].pairs_cut == cut[k]


def test_mixed_energy_conservation(E):
    print("[mixed nve]", flush=True)
    pos, vel, box = E.generate_fcc(10, RHO, 0.72, seed=11)
    out = {}
    for prec in ("fp64", "mixed"):
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos, vel)
            st = m.run(1000, 0.005, prec, energy_every=10)
        e = np.array([s.ke + s.pe_shifted for s in st])
        out[prec] = e
    for prec, e in out.items():
        assert np.abs(e - e[0]).max() < 5e-4 * abs(e[0]), prec


def test_out_of_box_status(E):
    print("[errors]", flush=True)
    pos, box = _case(6, 0.05, 1)
    pos = pos.copy()
    pos[5] = [-3.0, -3.0, -3.0]
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        with pytest.raises(E.ExamdError) as ei:
            m.forces("fp64")
        assert ei.value.status == E.ERR_OUT_OF_BOX


def test_fixed_tiles_too_large_rejected(E):
    """A fixed cell-block map whose region cannot fit in LDS is refused
    with a status, not run."""
    pos, box = _case(10, 0.1, 6)
    with E.MidpointLJ(len(pos), box, 2.5, tiles=(1, 1, 1)) as m:
        with pytest.raises(E.ExamdError) as ei:
            m.set_state(pos)
        assert ei.value.status == E.ERR_TILE_OVERFLOW


def test_small_grid_rejected(E):
    with pytest.raises(E.ExamdError) as ei:
        E.MidpointLJ(10, [11.0, 11.0, 11.0], 2.5)     # 8 blocks: fine
        E.MidpointLJ(10, [6.0, 6.0, 6.0], 2.5)        # 4 blocks: rejected
    assert ei.value.status == E.ERR_GRID


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_lane_queue_variant_matches(E, oracle, prec):
    """The deferred (lane-queue) pair evaluation gives the oracle's forces
    and pair counts, and the mixed path stays bit-identical to the direct
    evaluation at the same cell-block map."""
    print(f"[lane queue] {prec}", flush=True)
    pos, box = _case(10, 0.1, 12)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    with _system(E, pos, box) as m:
        cfg = m.get_config(prec)
        m.set_config(prec, cfg["tiles"], cfg["block_threads"], 0, 0)
        f0, s0, _ = m.forces(prec)
        m.set_config(prec, cfg["tiles"], cfg["block_threads"], 0, 8)
        f1, s1, _ = m.forces(prec)
        got = m.home_pair_counts(prec)
    assert (s1.pairs_cand, s1.pairs_cut) == (s0.pairs_cand, s0.pairs_cut)
    if prec == "fp64":
        assert _rel(f1, ref) < FP64_TOL
        np.testing.assert_array_equal(got, hc)
    else:
        assert np.array_equal(f1, f0)


def _stats_tuple(s):
    return (s.step, s.ke, s.pe, s.pe_shifted, s.pairs_cand, s.pairs_cut)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_graph_chunks_match_eager(E, prec):
    """run() replaying a captured chunk of 10 steps equals the kernel-by-kernel
    run: every replay's tallies, the state after each run() call (eager
    copies and synchronisations in between), and a trailing partial chunk
    run eagerly.  Mixed precision is bit-identical (order-independent
    fixed-point sums); FP64 sums forces with floating-point atomics, so it
    is compared within a tolerance."""
    print(f"[graph vs eager {prec}]", flush=True)
    pos, vel, box = E.generate_fcc(10, RHO, 0.72, seed=21)
    states = {}
    tallies = {}
    for chunk in (0, 10):
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos, vel)
            m.set_graph_chunk(chunk)
            seq, st = [], []
            for nsteps in (30, 30, 25):
                st += [_stats_tuple(s) for s in
                       m.run(nsteps, 0.005, prec, energy_every=10)]
                seq.append(m.get_state())
            states[chunk], tallies[chunk] = seq, st
    assert [t[0] for t in tallies[10]] == [t[0] for t in tallies[0]]
    assert len(tallies[10]) == 9
    if prec == "mixed":
        assert tallies[10] == tallies[0]
        for (p0, v0), (p1, v1) in zip(states[0], states[10]):
            assert np.array_equal(p0, p1) and np.array_equal(v0, v1)
    else:
        a, b = np.array(tallies[0]), np.array(tallies[10])
        assert np.allclose(a[:, 1:4], b[:, 1:4], rtol=1e-9, atol=0)
        assert np.array_equal(a[:, 4:], b[:, 4:])
        for (p0, v0), (p1, v1) in zip(states[0], states[10]):
            assert np.abs(p0 - p1).max() < 1e-9
            assert np.abs(v0 - v1).max() < 1e-8


def test_graph_chunk_timing_counts_launches(E):
    """Force-kernel timing events captured in a chunk are read back after
    every replay: one timed launch per step."""
    pos, vel, box = E.generate_fcc(10, RHO, 0.72, seed=22)
    with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=20) as m:
        m.set_state(pos, vel)
        m.run(5, 0.005, "mixed")
        m.timing(enable=True, reset=True)
        m.run(60, 0.005, "mixed", energy_every=20)
        ms, launches = m.timing(enable=False, reset=True)
    assert launches == 60
    assert ms > 0.0


def test_sampled_timing_counts_launches(E):
    """Timing every 10th step: steps 5, 15, ... of each captured chunk or
    eager piece; the sampled steps' trajectory is the untimed one, bit for
    bit (the events only bracket the launches)."""
    pos, vel, box = E.generate_fcc(10, RHO, 0.72, seed=23)
    out = []
    for every in (0, 10):
        with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=20) as m:
            m.set_state(pos, vel)
            m.timing(enable=every > 0, reset=True, every=max(every, 1))
            m.run(65, 0.005, "mixed", energy_every=20)  # 3 chunks + 5 eager
            ms, launches = m.timing(enable=False, reset=True)
            out.append((m.get_state(), launches, ms))
    (s0, n0, _), (s1, n1, ms1) = out
    assert n0 == 0 and n1 == 3 * 2 + 1 and ms1 > 0.0
    assert np.array_equal(s0[0], s1[0]) and np.array_equal(s0[1], s1[1])


def test_checkpoint_restart_continues_exactly(E, tmp_path):
    """Mixed precision: a run continued from a checkpoint in a new handle
    matches the uninterrupted run bit for bit."""
    pos, vel, box = E.generate_fcc(10, RHO, 1.0, seed=31)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        m.run(40, 0.005, "mixed")
        p_ref, v_ref = m.get_state()
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        m.run(20, 0.005, "mixed")
        p20, v20 = m.get_state()
    E.write_checkpoint(tmp_path / "c.chk", p20, v20, box, step=20)
    p, v, b, step = E.read_checkpoint(tmp_path / "c.chk")
    assert step == 20
    with E.MidpointLJ(len(p), b, 2.5) as m:
        m.set_state(p, v)
        m.run(20, 0.005, "mixed")
        p40, v40 = m.get_state()
    assert np.array_equal(p40, p_ref) and np.array_equal(v40, v_ref)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_get_state_returns_forces_of_the_last_evaluation(E, prec):
    """examd_hip_get_state with a force array: the forces of the last
    evaluation (after a run, the forces at its final positions), in the host
    layout; before any evaluation the call reports NOT_READY."""
    pos, vel, box = E.generate_fcc(9, RHO, 0.9, seed=41)
    with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=10) as m:
        m.set_state(pos, vel)
        with pytest.raises(E.ExamdError):
            m.get_state(forces=True)
        m.run(25, 0.005, prec)
        p, v, f_run = m.get_state(forces=True)
        f_now, _, _ = m.forces(prec)
        p2, v2, f_again = m.get_state(forces=True)
    assert np.array_equal(p, p2) and np.array_equal(v, v2)
    assert np.array_equal(f_again, f_now)
    if prec == "mixed":
        assert np.array_equal(f_run, f_now)
    else:
        assert _rel(f_run, f_now) < 1e-13


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_persistent_lists_match_fresh_binning(E, prec):
    """A hot liquid (256,000 atoms started at T* 5) in which over 1% of the
    atoms change block every step: 100 steps on block lists kept from step to
    step, whose binning pass moves the atoms that change block (atomic slot
    frees and compare-and-swap takes by threads all over the chip), against
    the same 100 steps with the lists built from scratch before every step
    (a new state each step).  The map is fixed, so both runs see the same
    cell-block origins.  Mixed sums do not depend on the order of the atoms
    in the lists, so the trajectories are bit-identical; the FP64 lists are
    put back in input order and its sums differ only in the order of their
    atomic additions (1e-10 bar on positions and velocities)."""
    pos, vel, box = E.generate_fcc(40, RHO, 5.0, seed=77)
    nb = np.floor(box / 1.25).astype(int)
    shape = dict(tiles=(8, 8, 8), block_threads=1024, run_groups=1,
                 lane_queue=4, workgroups=64, balance=0)
    steps = 100
    print(f"[lists {prec}] n={len(pos)} steps={steps}", flush=True)

    def blocks(p):
        b = np.minimum(np.floor(p / (box / nb)).astype(int), nb - 1)
        return b[:, 0] + nb[0] * (b[:, 1] + nb[1] * b[:, 2])

    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos, vel)
        m.set_config(prec, **shape)
        cuts = m.get_cuts(prec)
        m.run(steps, 0.005, prec)
        p_run, v_run = m.get_state()[:2]
        p, v = pos, vel
        moved = []
        for _ in range(steps):
            m.set_state(p, v)
            b0 = blocks(p)
            m.run(1, 0.005, prec)
            p, v = m.get_state()[:2]
            moved.append(np.mean(blocks(p) != b0))
        assert all(np.array_equal(a, b) for a, b in zip(cuts, m.get_cuts(prec)))
    # the stress the test relies on: thousands of list moves per step
    assert np.mean(moved) > 0.01, np.mean(moved)
    if prec == "mixed":
        assert np.array_equal(p_run, p) and np.array_equal(v_run, v)
    else:
        dp = np.abs(p_run - p)
        dp = np.minimum(dp, box - dp).max()
        assert dp < 1e-10 * box.max(), dp
        assert np.abs(v_run - v).max() < 1e-10 * np.abs(v).max()


def test_fp64_dense_block_of_twelve_matches_oracle(E, oracle):
    """A block that holds 12 atoms (a 3 x 2 x 2 cluster 0.4-0.45 sigma apart
    in a lattice with a hole around it): its list has 16 slots with ids in
    the second eight, which the FP64 force kernel's gather sorts into input
    order together with the first eight.  The cluster atoms sit at random
    places of the input order.  Every atom's force is held to the oracle at
    1e-10 of its own magnitude (cluster) or of the lattice's RMS force
    (the rest); pair counts and the energy as everywhere."""
    pos, L = H.perturbed_fcc(7, RHO, 0.05, 31)
    box = np.array([L, L, L])
    nb = E.domain(box, 1.25)[0][0]
    bl = L / nb
    g = np.stack(np.meshgrid(4 * bl + 0.25 + 0.4 * np.arange(3),
                             4 * bl + 0.35 + 0.45 * np.arange(2),
                             4 * bl + 0.35 + 0.45 * np.arange(2), indexing="ij"),
                 -1).reshape(-1, 3)
    lo, hi = g.min(0) - 0.9, g.max(0) + 0.9
    keep = ~np.all((pos > lo) & (pos < hi), axis=1)
    lat = pos[keep]
    rng = np.random.default_rng(32)
    at = np.sort(rng.choice(len(lat) + len(g), len(g), replace=False))
    is_cl = np.zeros(len(lat) + len(g), bool)
    is_cl[at] = True
    allp = np.empty((len(is_cl), 3))
    allp[is_cl] = g
    allp[~is_cl] = lat
    b = np.floor(allp / bl).astype(int)
    occ = np.bincount(b[:, 0] + nb * (b[:, 1] + nb * b[:, 2]), minlength=nb ** 3)
    assert occ.max() == 12
    print(f"[dense block fp64] n={len(allp)} nb={nb}", flush=True)
    with _system(E, allp, box) as m:
        f, s, _ = m.forces("fp64")
    ref, t, _ = oracle.forces(allp, box, 2.5)
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    assert abs(s.pe - t.pe) < FP64_TOL * abs(t.pe)
    df = np.linalg.norm(f - ref, axis=1)
    fc = np.linalg.norm(ref[is_cl], axis=1)
    assert np.all(df[is_cl] < FP64_TOL * fc)
    rms = np.sqrt((ref[~is_cl] ** 2).sum(1).mean())
    assert df[~is_cl].max() < FP64_TOL * rms
