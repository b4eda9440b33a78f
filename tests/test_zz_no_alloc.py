"""examd_hip_forces_static, the evaluation that allocates nothing (every
buffer beyond its inputs and outputs in one static device array per GPU),
against examd_hip_forces through a handle on the same input: two calls in a
row, the second on a smaller system, in both precisions.

Bars: FP64 within 1e-12 relative (normwise) of the handle's forces and
within 1e-10 of the oracle, pair counts equal; mixed bit-identical to the
handle's under the same cell-block map (fixed-point sums do not depend on
the order)."""
import ctypes as C

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


class _Dev:
    """A device array of the test (the caller's memory, allocated by the
    caller through the HIP runtime)."""
    hip = None

    def __init__(self, nbytes):
        if _Dev.hip is None:
            _Dev.hip = C.CDLL("libamdhip64.so")
        self.p = C.c_void_p()
        assert _Dev.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert _Dev.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p),
                                  C.c_size_t(a.nbytes), 1) == 0

    def get(self, shape):
        a = np.zeros(shape)
        assert _Dev.hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), self.p,
                                  C.c_size_t(a.nbytes), 2) == 0
        return a

    def free(self):
        _Dev.hip.hipFree(self.p)


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_static_matches_handle_two_calls(E, oracle, prec):
    cases = [(11, 0.1, 21), (7, 0.2, 22)]        # 5324, then 1372 atoms
    nmax = 4 * cases[0][0] ** 3
    pos_d, f_d = _Dev(nmax * 24), _Dev(nmax * 24)
    try:
        for nc, amp, seed in cases:
            pos, L = H.perturbed_fcc(nc, 0.8442, amp, seed)
            pos = np.mod(pos, L)
            box = np.array([L, L, L])
            n = len(pos)
            print(f"[no-alloc {prec}] n={n}", flush=True)
            pos_d.put(pos)
            s, st = E.forces_static(n, box, 2.5, pos_d.p.value, f_d.p.value,
                                    prec)
            assert st == E.OK
            f = f_d.get((n, 3))
            with E.MidpointLJ(n, box, 2.5) as m:
                m.set_state(pos)
                fh, sh, _ = m.forces(prec)
            assert (s.pairs_cand, s.pairs_cut) == (sh.pairs_cand, sh.pairs_cut)
            ref, t, _ = oracle.forces(pos, box, 2.5)
            assert s.pairs_cand == t.cand
            if prec == "fp64":
                assert _rel(f, fh) < 1e-12
                assert _rel(f, ref) < 1e-10
                assert s.pairs_cut == t.cut
                assert abs(s.pe - t.pe) < 1e-10 * abs(t.pe)
            else:
                assert np.array_equal(f, fh)
                assert s.pe == sh.pe
    finally:
        pos_d.free()
        f_d.free()


def test_static_limits(E):
    """Beyond the static limits the call returns a status and runs nothing."""
    d = _Dev(64)
    try:
        L = 60.0
        s, st = E.forces_static(70000, [L, L, L], 2.5, d.p.value, d.p.value,
                                allow=(E.ERR_STATIC_LIMIT,))
        assert st == E.ERR_STATIC_LIMIT
        s, st = E.forces_static(1000, [120.0, 120.0, 120.0], 2.5, d.p.value,
                                d.p.value, allow=(E.ERR_STATIC_LIMIT,))
        assert st == E.ERR_STATIC_LIMIT          # 96^3 blocks
        with pytest.raises(E.ExamdError):
            E.forces_static(1000, [8.0, 8.0, 8.0], 2.5, 0, 0)   # null arrays
    finally:
        d.free()
