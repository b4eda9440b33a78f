"""examd_amd -- MI355X (gfx950) midpoint-method Lennard-Jones force path.

Python view of the C ABI in include/examd_hip.h (libexamd_hip.so, built
in-tree by ``python -m examd_amd.build``).  The Fortran drop-in layer in
examd_amd/fortran binds the same ABI for the reference's own call surface.

Loading fails loudly when the HIP library is missing: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexamd_hip.so")

PREC_FP64 = 0
PREC_MIXED = 1
_PREC = {"fp64": PREC_FP64, "mixed": PREC_MIXED, PREC_FP64: PREC_FP64,
         PREC_MIXED: PREC_MIXED}

# status codes (include/examd_hip.h)
OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_NOMEM = 3
ERR_GRID = 4
ERR_TILE_OVERFLOW = 5
ERR_NO_DEVICE = 6
ERR_NONFINITE = 7
ERR_NOT_READY = 8
ERR_FIXED_RANGE = 9
ERR_ZERO_DISTANCE = 10
ERR_TABLE_RANGE = 11
ERR_STATIC_LIMIT = 12
ERR_BIN_OVERFLOW = 13
ERR_OUT_OF_BOX = 109

# every symbol include/examd_hip.h declares
EXPORTS = (
    "examd_hip_domain", "examd_hip_create", "examd_hip_destroy",
    "examd_hip_set_state", "examd_hip_get_state", "examd_hip_forces",
    "examd_hip_run", "examd_hip_block_ids", "examd_hip_home_pair_counts",
    "examd_hip_timing", "examd_hip_generate_fcc", "examd_hip_strerror",
    "examd_hip_phase_profile", "examd_hip_autotune", "examd_hip_get_config",
    "examd_hip_set_config", "examd_hip_set_graph_chunk",
    "examd_hip_batch_create", "examd_hip_batch_info",
    "examd_hip_generate_vacancy_fcc", "examd_hip_generate_slab",
    "examd_hip_set_species", "examd_hip_set_pair_table", "examd_hip_get_cuts",
    "examd_hip_forces_static", "examd_hip_regrow_count",
)

MAX_SPECIES = 8


class ExamdError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {strerror(status)} (status {status})")


class Params(C.Structure):
    _fields_ = [("natoms", C.c_int64), ("box", C.c_double * 3),
                ("cutoff", C.c_double), ("epsilon", C.c_double),
                ("sigma", C.c_double), ("mass", C.c_double),
                ("tiles", C.c_int32 * 3), ("device", C.c_int32),
                ("block_threads", C.c_int32), ("cap_slack_pct", C.c_int32),
                ("run_groups", C.c_int32), ("graph_chunk", C.c_int32),
                ("reserved", C.c_int32 * 2)]


class Tune(C.Structure):
    _fields_ = [("tiles", C.c_int32 * 3), ("block_threads", C.c_int32),
                ("run_groups", C.c_int32), ("lane_queue", C.c_int32),
                ("tile_edge", C.c_int32), ("workgroups", C.c_int32),
                ("balance", C.c_int32), ("rebalance_every", C.c_int32),
                ("force_ms", C.c_double)]

    def as_dict(self):
        return {"tiles": tuple(self.tiles), "block_threads": self.block_threads,
                "run_groups": self.run_groups, "lane_queue": self.lane_queue,
                "tile_edge": self.tile_edge, "workgroups": self.workgroups,
                "balance": self.balance,
                "rebalance_every": self.rebalance_every,
                "force_ms": self.force_ms}


class System(C.Structure):
    _fields_ = [("natoms", C.c_int64), ("box", C.c_double * 3)]


class Stats(C.Structure):
    _fields_ = [("step", C.c_int64), ("ke", C.c_double), ("pe", C.c_double),
                ("pe_shifted", C.c_double), ("etot", C.c_double),
                ("pairs_cand", C.c_int64), ("pairs_cut", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """The loaded libexamd_hip.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m examd_amd.build` (there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp, ip64, ip32 = P(C.c_double), P(C.c_int64), P(C.c_int32)
        L.examd_hip_domain.argtypes = [dp, C.c_double, ip32, ip64]
        L.examd_hip_create.argtypes = [P(Params), P(C.c_void_p)]
        L.examd_hip_batch_create.argtypes = [P(Params), P(System), C.c_int32,
                                             P(C.c_void_p)]
        L.examd_hip_batch_info.argtypes = [C.c_void_p, ip32, ip64, ip64]
        L.examd_hip_set_graph_chunk.argtypes = [C.c_void_p, C.c_int64]
        L.examd_hip_destroy.argtypes = [C.c_void_p]
        L.examd_hip_set_state.argtypes = [C.c_void_p, dp, dp, C.c_int64]
        L.examd_hip_get_state.argtypes = [C.c_void_p, dp, dp, dp, C.c_int64]
        L.examd_hip_forces.argtypes = [C.c_void_p, C.c_int, dp, P(Stats)]
        L.examd_hip_regrow_count.argtypes = [C.c_void_p, ip64]
        L.examd_hip_forces_static.argtypes = [P(Params), C.c_int, C.c_void_p,
                                              C.c_void_p, P(Stats)]
        L.examd_hip_run.argtypes = [C.c_void_p, C.c_int, C.c_int64,
                                    C.c_double, C.c_int64, P(Stats),
                                    C.c_int64, ip64]
        L.examd_hip_block_ids.argtypes = [C.c_void_p, ip64, ip32, C.c_int64]
        L.examd_hip_home_pair_counts.argtypes = [C.c_void_p, C.c_int, ip64,
                                                 C.c_int64]
        L.examd_hip_timing.argtypes = [C.c_void_p, C.c_int, dp, ip64, C.c_int]
        L.examd_hip_phase_profile.argtypes = [C.c_void_p, C.c_int, dp]
        L.examd_hip_autotune.argtypes = [C.c_void_p, C.c_int, C.c_int, P(Tune)]
        L.examd_hip_get_config.argtypes = [C.c_void_p, C.c_int, P(Tune)]
        L.examd_hip_set_config.argtypes = [C.c_void_p, C.c_int, P(Tune)]
        L.examd_hip_generate_fcc.argtypes = [ip32, C.c_double, C.c_double,
                                             C.c_double, C.c_double,
                                             C.c_uint64, dp, dp, dp]
        L.examd_hip_generate_vacancy_fcc.argtypes = [
            C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
            C.c_uint64, dp, dp, dp]
        L.examd_hip_generate_slab.argtypes = [
            dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
            C.c_double, C.c_double, C.c_uint64, dp, dp, C.c_int64, ip64]
        L.examd_hip_set_species.argtypes = [C.c_void_p, C.c_int32, ip32,
                                            C.c_int64, dp, dp, dp]
        L.examd_hip_set_pair_table.argtypes = [C.c_void_p, C.c_int32,
                                               C.c_double, dp, dp]
        L.examd_hip_strerror.argtypes = [C.c_int]
        L.examd_hip_strerror.restype = C.c_char_p
        for name in EXPORTS:
            if name != "examd_hip_strerror":
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def strerror(status: int) -> str:
    return lib().examd_hip_strerror(int(status)).decode()


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _check(st, where, allow=()):
    if st != OK and st not in allow:
        raise ExamdError(st, where)
    return st


def domain(box, block_length):
    """exaMD_domain_create: (nblocks_xyz, nblocks)."""
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(3)
    nb = (C.c_int32 * 3)()
    n = C.c_int64()
    _check(lib().examd_hip_domain(_dp(b), float(block_length), nb,
                                  C.byref(n)), "examd_hip_domain")
    return tuple(nb), n.value


def generate_fcc(nc, rho, temperature=0.0, mass=1.0, perturb=0.0, seed=0):
    """Seeded FCC lattice + Maxwell-Boltzmann velocities: (pos, vel, box)."""
    nc = (C.c_int32 * 3)(*[int(v) for v in (nc if np.ndim(nc) else (nc,) * 3)])
    n = 4 * nc[0] * nc[1] * nc[2]
    pos = np.zeros((n, 3))
    vel = np.zeros((n, 3))
    box = np.zeros(3)
    _check(lib().examd_hip_generate_fcc(nc, float(rho), float(temperature),
                                        float(mass), float(perturb),
                                        int(seed) & (2**64 - 1), _dp(pos),
                                        _dp(vel), _dp(box)),
           "examd_hip_generate_fcc")
    return pos, vel, box


def generate_vacancy_fcc(n, rho, temperature=0.0, mass=1.0, perturb=0.0,
                         seed=0):
    """n atoms at density rho on a randomly depleted FCC lattice in a cubic
    box, Maxwell-Boltzmann velocities: (pos, vel, box)."""
    pos = np.zeros((int(n), 3))
    vel = np.zeros((int(n), 3))
    box = np.zeros(3)
    _check(lib().examd_hip_generate_vacancy_fcc(
        int(n), float(rho), float(temperature), float(mass), float(perturb),
        int(seed) & (2**64 - 1), _dp(pos), _dp(vel), _dp(box)),
        "examd_hip_generate_vacancy_fcc")
    return pos, vel, box


def generate_slab(box, slab_lo, slab_hi, rho_slab, rho_vapour, min_sep=0.8,
                  temperature=0.0, mass=1.0, seed=0):
    """Liquid-vapour slab (random sequential addition with a minimum
    separation), Maxwell-Boltzmann velocities: (pos, vel)."""
    b = np.ascontiguousarray(box, dtype=np.float64).reshape(3)
    n = C.c_int64()
    args = (_dp(b), float(slab_lo), float(slab_hi), float(rho_slab),
            float(rho_vapour), float(min_sep), float(temperature), float(mass),
            int(seed) & (2**64 - 1))
    _check(lib().examd_hip_generate_slab(*args, None, None, 0, C.byref(n)),
           "examd_hip_generate_slab")
    pos = np.zeros((n.value, 3))
    vel = np.zeros((n.value, 3))
    _check(lib().examd_hip_generate_slab(*args, _dp(pos), _dp(vel), n.value,
                                         C.byref(n)),
           "examd_hip_generate_slab")
    return pos, vel


def write_checkpoint(path, pos, vel, box, step=0):
    """MD checkpoint in the Fortran layer's format (exaMD_write_checkpoint):
    'n step', the box, then x y z vx vy vz per atom with 17 significant
    digits, so that a restart continues bit for bit."""
    pos = np.asarray(pos, dtype=np.float64)
    vel = np.asarray(vel, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(f"{len(pos)} {int(step)}\n")
        fh.write("".join(f"{v:25.16E}" for v in np.reshape(box, 3)) + "\n")
        for p, v in zip(pos, vel):
            fh.write("".join(f"{x:25.16E}" for x in (*p, *v)) + "\n")


def read_checkpoint(path):
    """(pos, vel, box, step) from a checkpoint file."""
    with open(path) as fh:
        n, step = (int(x) for x in fh.readline().split())
        box = np.array([float(x) for x in fh.readline().split()])
        data = np.loadtxt(fh, dtype=np.float64, ndmin=2)
    if data.shape != (n, 6):
        raise ValueError(f"{path}: expected {n} atoms with 6 values")
    return data[:, :3].copy(), data[:, 3:].copy(), box, step


class _Handle:
    """Common methods of a handle of the C ABI (one system or a batch)."""

    n: int
    nsys: int

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().examd_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _info(self):
        ns, na, nb = C.c_int32(), C.c_int64(), C.c_int64()
        _check(lib().examd_hip_batch_info(self._h, C.byref(ns), C.byref(na),
                                          C.byref(nb)), "examd_hip_batch_info")
        self.nsys, self.n, self.nblocks = ns.value, na.value, nb.value

    def set_state(self, pos, vel=None):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        vel = None if vel is None else np.ascontiguousarray(vel, np.float64)
        assert pos.shape == (self.n, 3)
        _check(lib().examd_hip_set_state(self._h, _dp(pos), _dp(vel), self.n),
               "examd_hip_set_state")

    def set_species(self, types=None, eps=None, sigma=None, cutoff=None):
        """Several species: types (n,) in [0, k), symmetric (k, k) tables of
        eps, sigma and cutoff (each cutoff at most the handle's).  With no
        arguments, back to the handle's one species."""
        if types is None:
            _check(lib().examd_hip_set_species(self._h, 0, None, self.n, None,
                                               None, None),
                   "examd_hip_set_species")
            return
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(self.n)
        tab = [np.ascontiguousarray(x, dtype=np.float64) for x in
               (eps, sigma, cutoff)]
        k = tab[0].shape[0]
        assert all(t.shape == (k, k) for t in tab)
        _check(lib().examd_hip_set_species(
            self._h, k, types.ctypes.data_as(C.POINTER(C.c_int32)), self.n,
            _dp(tab[0]), _dp(tab[1]), _dp(tab[2])), "examd_hip_set_species")

    def set_pair_table(self, r2_lo=None, g=None, v=None):
        """A tabulated pair function: g (force factor, f12 = g d) and v at
        len(g) points uniform in r^2 from r2_lo to the cutoff squared.  With
        no arguments, back to the Lennard-Jones pair."""
        if g is None:
            _check(lib().examd_hip_set_pair_table(self._h, 0, 0.0, None, None),
                   "examd_hip_set_pair_table")
            return
        g = np.ascontiguousarray(g, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        assert g.shape == v.shape and g.ndim == 1
        _check(lib().examd_hip_set_pair_table(self._h, len(g), float(r2_lo),
                                              _dp(g), _dp(v)),
               "examd_hip_set_pair_table")

    def get_state(self, forces=False):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        f = np.zeros((self.n, 3)) if forces else None
        _check(lib().examd_hip_get_state(self._h, _dp(pos), _dp(vel), _dp(f),
                                         self.n), "examd_hip_get_state")
        return (pos, vel, f) if forces else (pos, vel)

    def _forces(self, precision, allow, out=None):
        if out is None:
            f = np.zeros((self.n, 3))
        else:
            f = out
            assert (f.shape == (self.n, 3) and f.dtype == np.float64
                    and f.flags["C_CONTIGUOUS"])
        s = (Stats * self.nsys)()
        st = _check(lib().examd_hip_forces(self._h, _PREC[precision], _dp(f),
                                           s), "examd_hip_forces", allow)
        return f, [s[i] for i in range(self.nsys)], st

    def _run(self, nsteps, dt, precision, energy_every):
        rows = max(1, int(nsteps) if energy_every else 1)
        if energy_every:
            rows = int(nsteps) // int(energy_every) + 1
        cap = rows * self.nsys
        out = (Stats * cap)()
        nout = C.c_int64()
        _check(lib().examd_hip_run(self._h, _PREC[precision], int(nsteps),
                                   float(dt), int(energy_every), out, cap,
                                   C.byref(nout)), "examd_hip_run")
        return [out[i] for i in range(nout.value)]

    def regrow_count(self):
        """Chunks of run() that ran again from their start with a larger
        LDS capacity after a cell-block region outgrew it."""
        n = C.c_int64()
        _check(lib().examd_hip_regrow_count(self._h, C.byref(n)),
               "examd_hip_regrow_count")
        return n.value

    def set_graph_chunk(self, chunk):
        """Steps per captured hipGraph chunk of run() (0 = launch eagerly)."""
        _check(lib().examd_hip_set_graph_chunk(self._h, int(chunk)),
               "examd_hip_set_graph_chunk")

    def block_ids(self):
        num = np.zeros(self.n, dtype=np.int64)
        xyz = np.zeros((self.n, 3), dtype=np.int32)
        _check(lib().examd_hip_block_ids(
            self._h, num.ctypes.data_as(C.POINTER(C.c_int64)),
            xyz.ctypes.data_as(C.POINTER(C.c_int32)), self.n),
            "examd_hip_block_ids")
        return num, xyz

    def home_pair_counts(self, precision="fp64"):
        c = np.zeros(self.nblocks, dtype=np.int64)
        _check(lib().examd_hip_home_pair_counts(
            self._h, _PREC[precision], c.ctypes.data_as(C.POINTER(C.c_int64)),
            self.nblocks), "examd_hip_home_pair_counts")
        return c

    def autotune(self, precision="mixed", trials=3):
        """Pick the fastest force-kernel launch shape for this state."""
        t = Tune()
        _check(lib().examd_hip_autotune(self._h, _PREC[precision], int(trials),
                                        C.byref(t)), "examd_hip_autotune")
        return t.as_dict()

    def get_config(self, precision="mixed"):
        t = Tune()
        _check(lib().examd_hip_get_config(self._h, _PREC[precision],
                                          C.byref(t)), "examd_hip_get_config")
        return t.as_dict()

    def set_config(self, precision, tiles=(0, 0, 0), block_threads=512,
                   run_groups=0, lane_queue=0, tile_edge=0, workgroups=0,
                   balance=0, rebalance_every=0, force_ms=None):
        """Apply a launch shape; rebalance_every > 0 makes run() re-partition
        the map from the current block counts that often (in steps, at
        captured-chunk boundaries).  Takes the dict get_config() and
        autotune() return: force_ms is ignored, and a map that came from a
        tile edge is derived from the edge again (in a batch, each system's
        map is then coarsened while its largest region fits the LDS budget
        of one workgroup, in atoms and in region blocks)."""
        if int(tile_edge) > 0:
            tiles = (0, 0, 0)
        t = Tune()
        t.tiles = (C.c_int32 * 3)(*[int(v) for v in tiles])
        t.block_threads = int(block_threads)
        t.run_groups = int(run_groups)
        t.lane_queue = int(lane_queue)
        t.tile_edge = int(tile_edge)
        t.workgroups = int(workgroups)
        t.balance = int(balance)
        t.rebalance_every = int(rebalance_every)
        _check(lib().examd_hip_set_config(self._h, _PREC[precision],
                                          C.byref(t)), "examd_hip_set_config")

    def get_cuts(self, precision="mixed", system=0):
        """Cell-block boundaries of one system's map, in blocks: a tuple of
        three arrays (x, y, z), each from 0 to the axis's block count."""
        buf = np.zeros(3 * 1024 + 3, dtype=np.int32)
        n = C.c_int32(0)
        _check(lib().examd_hip_get_cuts(
            self._h, _PREC[precision], int(system),
            buf.ctypes.data_as(C.POINTER(C.c_int32)), len(buf), C.byref(n)),
            "examd_hip_get_cuts")
        c = buf[:n.value].copy()
        # each axis's boundaries start at 0 and increase
        return tuple(np.split(c, np.flatnonzero(c == 0)[1:]))

    def phase_profile(self, precision="mixed"):
        """Mean cycles per cell block: scan, gather, pairs, return, total."""
        out = np.zeros(5)
        _check(lib().examd_hip_phase_profile(self._h, _PREC[precision],
                                             _dp(out)),
               "examd_hip_phase_profile")
        return out

    def timing(self, enable=True, reset=False, every=1):
        """Force-kernel time from HIP events: (total ms, launches) so far.
        every=k > 1 times only every k-th step's force launch in run()."""
        ms = C.c_double()
        n = C.c_int64()
        _check(lib().examd_hip_timing(self._h, int(every) if enable else 0,
                                      C.byref(ms), C.byref(n), int(reset)),
               "examd_hip_timing")
        return ms.value, n.value


def _params(cutoff, epsilon, sigma, mass, device, block_threads,
            cap_slack_pct, run_groups, graph_chunk):
    p = Params()
    p.cutoff, p.epsilon, p.sigma, p.mass = (float(cutoff), float(epsilon),
                                            float(sigma), float(mass))
    p.device = int(device)
    p.block_threads = int(block_threads)
    p.cap_slack_pct = int(cap_slack_pct)
    p.run_groups = int(run_groups)
    p.graph_chunk = int(graph_chunk)
    return p


def forces_static(natoms, box, cutoff, pos_dev, forces_dev, precision="fp64",
                  tiles=(0, 0, 0), epsilon=1.0, sigma=1.0, device=0,
                  block_threads=0, run_groups=0, allow=()):
    """examd_hip_forces_static: one evaluation that allocates nothing.
    pos_dev and forces_dev are device addresses (ints) of (natoms, 3)
    float64 arrays.  Returns (Stats, status)."""
    p = _params(cutoff, epsilon, sigma, 1.0, device, block_threads, 0,
                run_groups, 0)
    p.natoms = int(natoms)
    p.box = (C.c_double * 3)(*[float(v) for v in np.reshape(box, 3)])
    p.tiles = (C.c_int32 * 3)(*[int(t) for t in tiles])
    s = Stats()
    st = _check(lib().examd_hip_forces_static(
        C.byref(p), _PREC[precision], C.c_void_p(int(pos_dev)),
        C.c_void_p(int(forces_dev)), C.byref(s)), "examd_hip_forces_static",
        allow)
    return s, st


class MidpointLJ(_Handle):
    """One LJ system on one GPU (a handle of the C ABI)."""

    def __init__(self, natoms, box, cutoff, epsilon=1.0, sigma=1.0, mass=1.0,
                 tiles=(0, 0, 0), device=0, block_threads=0,
                 cap_slack_pct=0, run_groups=0, graph_chunk=0):
        p = _params(cutoff, epsilon, sigma, mass, device, block_threads,
                    cap_slack_pct, run_groups, graph_chunk)
        p.natoms = int(natoms)
        p.box = (C.c_double * 3)(*[float(v) for v in np.reshape(box, 3)])
        p.tiles = (C.c_int32 * 3)(*[int(t) for t in tiles])
        self.params = p
        self.box = np.array(np.reshape(box, 3), dtype=np.float64)
        self.cutoff = float(cutoff)
        self.nb, _ = domain(self.box, 0.5 * self.cutoff)
        h = C.c_void_p()
        _check(lib().examd_hip_create(C.byref(p), C.byref(h)),
               "examd_hip_create")
        self._h = h
        self._info()

    def forces(self, precision="fp64", allow=(), out=None):
        """(forces (n, 3), Stats, status) at the current positions; out: a
        C-contiguous float64 (n, 3) array to write the forces into (reused
        by callers that evaluate again and again)."""
        f, s, st = self._forces(precision, allow, out)
        return f, s[0], st

    def run(self, nsteps, dt, precision="fp64", energy_every=0):
        """Velocity-Verlet steps; Stats of every tallied step."""
        return self._run(nsteps, dt, precision, energy_every)


class MidpointLJBatch(_Handle):
    """Independent periodic LJ systems of uneven sizes in one handle: every
    kernel of a step covers all systems in one launch.  systems: sequence of
    (natoms, box).  Positions and velocities are concatenated in system
    order."""

    def __init__(self, systems, cutoff, epsilon=1.0, sigma=1.0, mass=1.0,
                 device=0, block_threads=0, cap_slack_pct=0, run_groups=0,
                 graph_chunk=0):
        p = _params(cutoff, epsilon, sigma, mass, device, block_threads,
                    cap_slack_pct, run_groups, graph_chunk)
        arr = (System * len(systems))()
        for k, (n, box) in enumerate(systems):
            arr[k].natoms = int(n)
            arr[k].box = (C.c_double * 3)(*[float(v) for v in np.reshape(box, 3)])
        self.params = p
        self.sizes = np.array([int(n) for n, _ in systems], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)])
        self.cutoff = float(cutoff)
        h = C.c_void_p()
        _check(lib().examd_hip_batch_create(C.byref(p), arr, len(systems),
                                            C.byref(h)),
               "examd_hip_batch_create")
        self._h = h
        self._info()

    def forces(self, precision="fp64", allow=(), out=None):
        """(forces (n, 3), [Stats per system], status); out: a C-contiguous
        float64 (n, 3) array to write the forces into."""
        return self._forces(precision, allow, out)

    def run(self, nsteps, dt, precision="mixed", energy_every=0):
        """Velocity-Verlet steps of every system; one list of per-system
        Stats per tallied step."""
        flat = self._run(nsteps, dt, precision, energy_every)
        k = self.nsys
        return [flat[i:i + k] for i in range(0, len(flat), k)]
