"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product package
(examd_amd) never imports it.

Arrays follow the reference's Fortran layout posn(1:3, 1:n): a numpy array of
shape (n, 3), C-contiguous, float64.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class _Domain(C.Structure):
    _fields_ = [("box", C.c_double * 3), ("boxi", C.c_double * 3),
                ("bl", C.c_double), ("nb", C.c_int32 * 3),
                ("nblocks", C.c_int64)]


class _LJ(C.Structure):
    _fields_ = [("eps", C.c_double), ("sig", C.c_double), ("rc", C.c_double)]


class _Species(C.Structure):
    _fields_ = [("ntypes", C.c_int32), ("type", C.POINTER(C.c_int32)),
                ("pair", C.POINTER(_LJ))]


class _Table(C.Structure):
    _fields_ = [("n", C.c_int32), ("r2_lo", C.c_double), ("inv_h", C.c_double),
                ("gc", C.POINTER(C.c_double)), ("vc", C.POINTER(C.c_double))]


class _Tally(C.Structure):
    _fields_ = [("cand", C.c_int64), ("cut", C.c_int64), ("zero", C.c_int64),
                ("pe", C.c_double), ("below", C.c_int64)]


_lib = None


def build() -> None:
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.POINTER
        L.or_domain_create.argtypes = [P(C.c_double), C.c_double, P(_Domain)]
        L.or_identify_block.argtypes = [P(_Domain), P(C.c_double),
                                        P(C.c_int32), P(C.c_int64)]
        L.or_interactions.argtypes = [P(C.c_int32)]
        L.or_stencil_offset.argtypes = [C.c_int32, P(C.c_int32)]
        L.or_bin.argtypes = [P(_Domain), C.c_int64, P(C.c_double),
                             P(C.c_int64), P(C.c_int64), P(C.c_int64)]
        L.or_forces_midpoint_all.argtypes = [P(_Domain), P(_LJ), C.c_int64,
                                             P(C.c_double), P(C.c_double),
                                             P(C.c_int64), P(_Tally)]
        L.or_forces_midpoint_species.argtypes = [
            P(_Domain), P(_LJ), P(_Species), C.c_int64, P(C.c_double),
            P(C.c_double), P(C.c_int64), P(_Tally)]
        L.or_forces_midpoint_table.argtypes = [
            P(_Domain), P(_LJ), P(_Table), C.c_int64, P(C.c_double),
            P(C.c_double), P(C.c_int64), P(_Tally)]
        L.or_run_vv_species.argtypes = [
            P(_Domain), P(_LJ), P(_Species), C.c_int64, C.c_double,
            C.c_double, C.c_int64, P(C.c_double), P(C.c_double),
            P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_int64)]
        L.or_lennardjones_12.argtypes = [P(_LJ), C.c_double, C.c_double,
                                         C.c_double, C.c_double,
                                         P(C.c_double), P(C.c_double)]
        L.or_run_vv.argtypes = [P(_Domain), P(_LJ), C.c_int64, C.c_double,
                                C.c_double, C.c_int64, P(C.c_double),
                                P(C.c_double), P(C.c_double), P(C.c_double),
                                P(C.c_double), P(C.c_int64)]
        L.or_forces_midpoint_mt.argtypes = [P(_Domain), P(_LJ), C.c_int64,
                                            P(C.c_double), P(C.c_double),
                                            P(C.c_int64), P(_Tally), C.c_int32]
        L.or_run_vv_mt.argtypes = [P(_Domain), P(_LJ), C.c_int64, C.c_double,
                                   C.c_double, C.c_int64, P(C.c_double),
                                   P(C.c_double), P(C.c_double), P(C.c_double),
                                   P(C.c_double), P(C.c_int64), C.c_int32]
        L.or_wrap.argtypes = [C.c_double, C.c_double]
        L.or_wrap.restype = C.c_double
        L.or_write_particles.argtypes = [C.c_char_p, C.c_int64,
                                         P(C.c_double), P(C.c_double)]
        L.or_read_params.argtypes = [C.c_char_p, P(C.c_int64), P(C.c_double)]
        L.or_read_particles.argtypes = [C.c_char_p, C.c_int64, P(C.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


@dataclass
class Domain:
    box: np.ndarray
    bl: float
    nb: tuple
    nblocks: int
    _c: _Domain


def domain_create(box, block_length) -> Domain:
    box = np.ascontiguousarray(box, dtype=np.float64).reshape(3)
    d = _Domain()
    rc = lib().or_domain_create(_dp(box), float(block_length), C.byref(d))
    if rc:
        raise ValueError(f"or_domain_create failed ({rc})")
    return Domain(box.copy(), d.bl, tuple(d.nb), d.nblocks, d)


def identify_block(dom: Domain, pos):
    p = np.ascontiguousarray(pos, dtype=np.float64).reshape(3)
    bxyz = (C.c_int32 * 3)()
    num = C.c_int64()
    rc = lib().or_identify_block(C.byref(dom._c), _dp(p), bxyz, C.byref(num))
    return rc, tuple(bxyz), num.value


def interactions() -> np.ndarray:
    out = (C.c_int32 * 126)()
    lib().or_interactions(out)
    return np.array(out, dtype=np.int32).reshape(63, 2)


def stencil_offset(n: int):
    out = (C.c_int32 * 3)()
    lib().or_stencil_offset(int(n), out)
    return tuple(out)


def bin_atoms(dom: Domain, pos):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    start = np.zeros(dom.nblocks + 1, dtype=np.int64)
    order = np.zeros(max(n, 1), dtype=np.int64)
    bid = np.zeros(max(n, 1), dtype=np.int64)
    rc = lib().or_bin(C.byref(dom._c), n, _dp(pos), _ip(start), _ip(order),
                      _ip(bid))
    if rc:
        raise ValueError(f"or_bin failed ({rc})")
    return start, order[:n], bid[:n]


@dataclass
class Tally:
    cand: int
    cut: int
    zero: int
    pe: float
    below: int = 0


def _species(species, n):
    """(types (n,), eps, sig, rc (k, k) each) -> (_Species, keep-alive)."""
    types, eps, sig, rc = species
    types = np.ascontiguousarray(types, dtype=np.int32).reshape(n)
    k = np.asarray(eps).shape[0]
    table = (_LJ * (k * k))()
    for a in range(k):
        for b in range(k):
            table[a * k + b] = _LJ(float(eps[a][b]), float(sig[a][b]),
                                   float(rc[a][b]))
    sp = _Species(k, types.ctypes.data_as(C.POINTER(C.c_int32)), table)
    return sp, (types, table)


def hermite_coefficients(y):
    """Per interval of the samples y (uniform spacing), the coefficients
    (c0, c1, c2, c3) of the cubic in t in [0, 1] through the two points with
    the points' slopes per spacing: fourth-order differences (centred
    inside, one-sided at the two ends), exact for quartics, so that the
    interpolant is exact for cubics.  Fewer than five points: the linear
    interpolant.  Shape (len(y) - 1, 4)."""
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    c = np.zeros((n - 1, 4))
    c[:, 0] = y[:-1]
    if n < 5:
        c[:, 1] = y[1:] - y[:-1]
        return c
    m = np.empty(n)
    m[2:n - 2] = (y[:n - 4] - 8 * y[1:n - 3] + 8 * y[3:n - 1] - y[4:]) / 12
    m[0] = (-25 * y[0] + 48 * y[1] - 36 * y[2] + 16 * y[3] - 3 * y[4]) / 12
    m[1] = (-3 * y[0] - 10 * y[1] + 18 * y[2] - 6 * y[3] + y[4]) / 12
    m[n - 2] = (-y[n - 5] + 6 * y[n - 4] - 18 * y[n - 3] + 10 * y[n - 2]
                + 3 * y[n - 1]) / 12
    m[n - 1] = (3 * y[n - 5] - 16 * y[n - 4] + 36 * y[n - 3] - 48 * y[n - 2]
                + 25 * y[n - 1]) / 12
    c[:, 1] = m[:-1]
    c[:, 2] = 3 * (y[1:] - y[:-1]) - 2 * m[:-1] - m[1:]
    c[:, 3] = 2 * (y[:-1] - y[1:]) + m[:-1] + m[1:]
    return c


def table_arrays(table, cutoff):
    """(r2_lo, g, v) sampled at len(g) points uniform in r^2 up to cutoff^2
    -> (n, r2_lo, inv_h, gc, vc): the cubic coefficients of g and V per
    interval, as the interpolation uses them."""
    r2_lo, g, v = table
    n = len(g)
    inv_h = float(n - 1) / (cutoff * cutoff - float(r2_lo))
    gc = np.ascontiguousarray(hermite_coefficients(g).reshape(-1))
    vc = np.ascontiguousarray(hermite_coefficients(v).reshape(-1))
    return n, float(r2_lo), inv_h, gc, vc


def forces(pos, box, cutoff, eps=1.0, sig=1.0, home_counts=False,
           species=None, table=None, threads=1):
    """One exaMD_forces_midpoint_all evaluation from zeroed forces.  With
    species = (types, eps, sig, rc), every pair uses the parameters of its
    species pair (cutoff sizes the blocks).  threads > 1 (one species, no
    table): the slab-threaded form (same pairs and arithmetic; an atom's
    sums near slab faces in another order).

    Returns (forces (n,3), Tally, home_counts or None)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    dom = domain_create(box, 0.5 * cutoff)
    f = np.zeros((n, 3), dtype=np.float64)
    hc = np.zeros(dom.nblocks, dtype=np.int64) if home_counts else None
    t = _Tally()
    p = _LJ(eps, sig, cutoff)
    if table is not None:
        tn, lo, ih, gc, vc = table_arrays(table, cutoff)
        tab = _Table(tn, lo, ih, _dp(gc), _dp(vc))
        rc = lib().or_forces_midpoint_table(
            C.byref(dom._c), C.byref(p), C.byref(tab), n, _dp(pos), _dp(f),
            _ip(hc) if hc is not None else None, C.byref(t))
    elif species is None and threads > 1:
        rc = lib().or_forces_midpoint_mt(
            C.byref(dom._c), C.byref(p), n, _dp(pos), _dp(f),
            _ip(hc) if hc is not None else None, C.byref(t), int(threads))
    elif species is None:
        rc = lib().or_forces_midpoint_all(
            C.byref(dom._c), C.byref(p), n, _dp(pos), _dp(f),
            _ip(hc) if hc is not None else None, C.byref(t))
    else:
        sp, keep = _species(species, n)
        rc = lib().or_forces_midpoint_species(
            C.byref(dom._c), C.byref(p), C.byref(sp), n, _dp(pos), _dp(f),
            _ip(hc) if hc is not None else None, C.byref(t))
    if rc:
        raise ValueError(f"or_forces_midpoint_all failed ({rc})")
    return f, Tally(t.cand, t.cut, t.zero, t.pe, t.below), hc


def lennardjones_12(dx, dy, dz, d2, cutoff, eps=1.0, sig=1.0):
    f = (C.c_double * 3)()
    e = C.c_double()
    p = _LJ(eps, sig, cutoff)
    rv = lib().or_lennardjones_12(C.byref(p), dx, dy, dz, d2, f, C.byref(e))
    return rv, np.array(f), e.value


def run_vv(pos, vel, box, cutoff, dt, nsteps, mass=1.0, eps=1.0, sig=1.0,
           forces0=None, species=None, threads=1):
    """Velocity Verlet from (pos, vel); returns (pos, vel, forces, ke, pe, cut).
    threads > 1 (one species): the slab-threaded form."""
    pos = np.array(pos, dtype=np.float64, order="C")
    vel = np.array(vel, dtype=np.float64, order="C")
    n = pos.shape[0]
    if forces0 is None:
        forces0, _, _ = forces(pos, box, cutoff, eps, sig, species=species,
                               threads=threads)
    f = np.array(forces0, dtype=np.float64, order="C")
    dom = domain_create(box, 0.5 * cutoff)
    ke = np.zeros(nsteps)
    pe = np.zeros(nsteps)
    cut = np.zeros(nsteps, dtype=np.int64)
    p = _LJ(eps, sig, cutoff)
    if species is None and threads > 1:
        rc = lib().or_run_vv_mt(C.byref(dom._c), C.byref(p), n, mass, dt,
                                nsteps, _dp(pos), _dp(vel), _dp(f), _dp(ke),
                                _dp(pe), _ip(cut), int(threads))
    elif species is None:
        rc = lib().or_run_vv(C.byref(dom._c), C.byref(p), n, mass, dt, nsteps,
                             _dp(pos), _dp(vel), _dp(f), _dp(ke), _dp(pe),
                             _ip(cut))
    else:
        sp, keep = _species(species, n)
        rc = lib().or_run_vv_species(C.byref(dom._c), C.byref(p), C.byref(sp),
                                     n, mass, dt, nsteps, _dp(pos), _dp(vel),
                                     _dp(f), _dp(ke), _dp(pe), _ip(cut))
    if rc:
        raise ValueError(f"or_run_vv failed ({rc})")
    return pos, vel, f, ke, pe, cut


def wrap(x: float, L: float) -> float:
    return lib().or_wrap(float(x), float(L))


def write_particles(path, pos, box):
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    rc = lib().or_write_particles(os.fsencode(path), pos.shape[0], _dp(box),
                                  _dp(pos))
    if rc:
        raise OSError(f"cannot write {path}")


def read_particles(path):
    n = C.c_int64()
    box = np.zeros(3)
    if lib().or_read_params(os.fsencode(path), C.byref(n), _dp(box)):
        raise OSError(f"cannot read {path}")
    pos = np.zeros((n.value, 3))
    if lib().or_read_particles(os.fsencode(path), n.value, _dp(pos)):
        raise OSError(f"cannot read {path}")
    return pos, box
