"""A run whose cell-block regions outgrow their LDS capacity mid-run: the
chunk that overflowed is run again from its start with a larger capacity,
and the run's result is the one a run started with that capacity gives.

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
