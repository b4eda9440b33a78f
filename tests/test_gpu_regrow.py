"""A run whose cell-block regions outgrow their LDS capacity mid-run: the
chunk that overflowed is run again from its start with a larger capacity,
and the run's result is the one a run started with that capacity gives.
Likewise a block whose atom list outgrows its slots: the chunk is run again
with twice the slots.

The system is made to compress (velocities pointing to the box centre), so
the central cell blocks' regions grow past a 1% capacity slack.  These are
data conditions the kernels flag and skip (a region beyond the capacity is
not evaluated), never GPU faults."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _compressing(E, seed=5):
    L = 20.0
    box = np.array([L, L, L])
    pos, _ = E.generate_slab(box, 0.0, L, 0.5, 0.5, min_sep=0.9, seed=seed)
    vel = -0.15 * (pos - 0.5 * L)
    vel -= vel.mean(0)
    return pos, vel, box


def _run(E, pos, vel, box, prec, slack, chunk, steps=200):
    with E.MidpointLJ(len(pos), box, 2.5, tiles=(4, 4, 4), cap_slack_pct=slack,
                      graph_chunk=chunk) as m:
        m.set_state(pos, vel)
        st = m.run(steps, 0.005, prec, energy_every=50)
        p, v = m.get_state()
        return p, v, [(s.step, s.ke, s.pe, s.pairs_cut) for s in st], m.regrow_count()


@pytest.mark.parametrize("chunk", [50, 0])
def test_overflow_mid_run_is_recovered_mixed(E, chunk):
    pos, vel, box = _compressing(E)
    print(f"[regrow mixed chunk={chunk}] n={len(pos)}", flush=True)
    pa, va, sa, ra = _run(E, pos, vel, box, "mixed", 1, chunk)
    pb, vb, sb, rb = _run(E, pos, vel, box, "mixed", 300, chunk)
    assert ra >= 1 and rb == 0
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    assert sa == sb


def test_overflow_mid_run_is_recovered_fp64(E):
    pos, vel, box = _compressing(E, seed=6)
    print("[regrow fp64]", flush=True)
    pa, va, sa, ra = _run(E, pos, vel, box, "fp64", 1, 50)
    pb, vb, sb, rb = _run(E, pos, vel, box, "fp64", 300, 50)
    assert ra >= 1 and rb == 0
    d = pa - pb
    d -= np.round(d / box) * box
    assert np.abs(d).max() < 1e-8
    assert [s[3] for s in sa] == [s[3] for s in sb]


def _converging(steps=100, dt=0.05):
    """27 atoms that fly (epsilon = 0: no forces) radially inwards over 5
    sigma onto a 3 x 3 x 3 grid, 0.7 sigma apart, inside the enlarged last
    block of a 13.7-sigma box (rc 2.5: 10 blocks per side, the last one
    2.45 sigma wide), so that the block's list outgrows its slots."""
    L = 13.7
    box = np.array([L, L, L])
    g = 11.5 + 0.7 * np.arange(3)
    tgt = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    u = tgt - tgt.mean(0)
    nrm = np.linalg.norm(u, axis=1)
    u[nrm == 0] = (1.0, 0.0, 0.0)
    u /= np.linalg.norm(u, axis=1)[:, None]
    vel = -u * 5.0 / (steps * dt)
    pos = np.mod(tgt - vel * steps * dt, L)
    # no two atoms ever come closer than 0.1 sigma (minimum image)
    for s in range(steps + 1):
        p = np.mod(pos + vel * s * dt, L)
        r = p[:, None, :] - p[None, :, :]
        r -= np.round(r / L) * L
        r2 = (r ** 2).sum(-1) + np.eye(len(p)) * 1e9
        assert r2.min() > 0.01
    return pos, vel, box


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_block_slots_grow_mid_run(E, oracle, prec):
    """A block that receives more atoms than its list slots during a run:
    the chunk is run again from its start with twice the slots (as often as
    needed), and every tally of the run counts the pairs the oracle counts
    at the same positions, so no atom was left out of the binning."""
    O = oracle
    steps, dt, every = 100, 0.05, 25
    pos, vel, box = _converging()
    print(f"[bin slots {prec}] n={len(pos)}", flush=True)
    with E.MidpointLJ(len(pos), box, 2.5, epsilon=0.0, graph_chunk=every) as m:
        m.set_state(pos, vel)
        st = m.run(steps, dt, prec, energy_every=every)
        grown = m.regrow_count()
        p, _ = m.get_state()
    assert grown >= 1
    expect = np.mod(pos + vel * steps * dt, box)
    d = p - expect
    d -= np.round(d / box) * box
    assert np.abs(d).max() < 1e-12
    for s in st:
        ps = np.mod(pos + vel * s.step * dt, box)
        _, t, _ = O.forces(ps, box, 2.5)
        assert s.pairs_cand == t.cand, s.step
        # FP32 r^2 at the cutoff may move a pair across it on the mixed path
        assert abs(s.pairs_cut - t.cut) <= (2 if prec == "mixed" else 0), s.step
