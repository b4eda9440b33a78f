"""Failure statuses of the force path, forced by crafted input data (data
conditions the kernels flag, never GPU faults), and their mapping in the
Fortran layer.

Reference behaviour: a pair with |d2| < 0.001 is printed as 'ZERO at' and
evaluated anyway (SRC/exaMD_forces.f90:525-527); the driver scans the forces
for NaN (exaMD.f90:269-286); an atom outside the block grid aborts with 109
(SRC/exaMD_module.f90:115-121).  The mixed path adds one status of its own:
a pair closer than r_fix = 0.638 sigma, below which an atom's 2^-32 fixed-
point force sum could leave its range.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

RHO = 0.8442


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _lattice(nc=8, seed=1):
    pos, L = H.perturbed_fcc(nc, RHO, 0.05, seed)
    return pos.copy(), np.array([L, L, L])


def _move_next_to(pos, i, j, d):
    """Put atom i at distance d from atom j along x (inside the box)."""
    pos[i] = pos[j] + np.array([d, 0.0, 0.0])


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def test_zero_distance_reported_and_evaluated(E, oracle):
    """A pair at 0.02 sigma (d2 = 4e-4 < 0.001): the FP64 path reports
    ERR_ZERO_DISTANCE and still returns the forces of every pair, the
    close one included, as the reference prints and goes on."""
    print("[status zero]", flush=True)
    pos, box = _lattice()
    _move_next_to(pos, 100, 7, 0.02)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    assert t.zero == 1
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        f, s, st = m.forces("fp64", allow=(E.ERR_ZERO_DISTANCE,))
        assert st == E.ERR_ZERO_DISTANCE
        assert np.all(np.isfinite(f))
        assert _rel(f, ref) < 1e-10
        assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
        # the mixed path cannot hold this force in its fixed-point range
        with pytest.raises(E.ExamdError) as ei:
            m.forces("mixed")
        assert ei.value.status == E.ERR_FIXED_RANGE


def test_fixed_range_mixed_only(E, oracle):
    """A pair at 0.6 sigma: the mixed path refuses it (ERR_FIXED_RANGE, no
    forces), the FP64 path evaluates it against the oracle."""
    print("[status fixed range]", flush=True)
    pos, box = _lattice(seed=2)
    _move_next_to(pos, 200, 31, 0.6)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    assert t.zero == 0
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        with pytest.raises(E.ExamdError) as ei:
            m.forces("mixed")
        assert ei.value.status == E.ERR_FIXED_RANGE
        f, s, st = m.forces("fp64")
        assert st == E.OK
        assert _rel(f, ref) < 1e-10
        assert s.pairs_cut == t.cut
        # at 0.66 sigma (beyond r_fix) the mixed path evaluates it
        pos2 = pos.copy()
        _move_next_to(pos2, 200, 31, 0.66)
        ref2, t2, _ = oracle.forces(pos2, box, 2.5)
        m.set_state(pos2)
        f2, s2, st2 = m.forces("mixed")
        assert st2 == E.OK
        err = np.sqrt(((f2 - ref2) ** 2).sum(1)).max()
        assert err < 2e-5 * np.abs(ref2).max()


def test_nonfinite_positions(E):
    """A non-finite position is refused at set_state; one that becomes
    non-finite during a run (an infinite velocity) stops the run with
    ERR_NONFINITE, in both precisions."""
    print("[status nonfinite]", flush=True)
    pos, box = _lattice(seed=3)
    bad = pos.copy()
    bad[17, 1] = np.nan
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        with pytest.raises(E.ExamdError) as ei:
            m.set_state(bad)
        assert ei.value.status == E.ERR_NONFINITE
        vel = np.zeros_like(pos)
        vel[17, 0] = np.inf
        for prec in ("fp64", "mixed"):
            m.set_state(pos, vel)
            with pytest.raises(E.ExamdError) as ei:
                m.run(3, 0.005, prec)
            assert ei.value.status == E.ERR_NONFINITE
        # the handle stays usable with a finite state
        m.set_state(pos)
        _, _, st = m.forces("fp64")
        assert st == E.OK
