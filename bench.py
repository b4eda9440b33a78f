"""Headline benchmark: atom-steps/s of the 2,048,000-atom LJ liquid on one GPU.

Workload (one GPU's share at real shape): 80^3 FCC lattice, 2,048,000 atoms,
density 0.8442, Maxwell-Boltzmann velocities at T* = 0.72 (seeded, zero net
momentum), sigma = epsilon = mass = 1, cutoff 2.5 (unshifted), cubic periodic
box of 107^3 rc/2 blocks, velocity Verlet dt = 0.005, 10,000 steps per
precision by default (energies every 10 steps), replayed from captured
100-step hipGraph chunks.  Each step rebins on device and evaluates the
midpoint-method forces once.  The launch shape (cell-block map, workgroup
size, run groups, queue depth, persistent grid) is autotuned per precision
on the liquid, after 300 untimed steps that melt the lattice.

value = mixed-precision atom-steps/s (FP32 pair arithmetic, int64 fixed-point
accumulation, FP64 integration); the FP64 path is timed on the same system
for the mixed-vs-FP64 ratio.  Inputs are resident in HBM before the timed
region; the run call returns after the stream is synchronised.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3      # MI355X FP32 vector (packed)
FP64_PEAK_TFLOPS = 78.6       # MI355X FP64 vector
NC = 80                       # FCC cells per side -> 2,048,000 atoms
RHO, TEMP, RC, DT = 0.8442, 0.72, 2.5, 0.005
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles",
                               "k_forces_mixed_2048000_traffic.txt")
# The force kernel is timed with HIP events around the force launch of every
# TIME_EVERY-th step (steps 2, 7, 12, ... of each chunk, none of them an
# energy-tally step): each pair of event records costs the stream about
# 5.6 us, which timing every launch would add to every step.
TIME_EVERY = 5
# Accuracy leg, whatever the timed length: mixed and FP64 steps continued
# from the timed run's state, energies every 10 steps.
DRIFT_STEPS = {"mixed": 10000, "fp64": 1000}
DRIFT_EVERY = 10


def progress(what, t_start):
    """One line on stderr as each leg of the bench ends (the JSON line on
    stdout stays the only output a caller parses)."""
    print(f"bench: {what} done at {time.perf_counter() - t_start:.1f} s",
          file=sys.stderr, flush=True)


def algorithmic_flops(cand, cut, energy):
    """W = 8 P_cand + 18 P_cut (+3 P_cut when energies are tallied)."""
    return 8.0 * cand + 18.0 * cut + (3.0 * cut if energy else 0.0)


def drift_per_1k(stats):
    """Least-squares slope of the shifted total energy per 1,000 steps,
    relative to |E| (None when the accuracy leg was not run)."""
    if stats is None:
        return None
    if len(stats) < 3:
        raise RuntimeError(f"drift needs 3 or more tallies, got {len(stats)}")
    t = np.array([s.step for s in stats], dtype=float)
    e = np.array([s.ke + s.pe_shifted for s in stats])
    slope = np.polyfit(t, e, 1)[0]
    return float(slope * 1000.0 / abs(e.mean()))


def max_dev(stats):
    """max |E(t) - E(0)| / |E(0)| of the shifted total energy."""
    if not stats:
        return None
    e = np.array([s.ke + s.pe_shifted for s in stats])
    return float(np.abs(e - e[0]).max() / abs(e[0]))


def read_traffic():
    """HBM bytes per force launch from the committed rocprofv3 PMC summary
    (line 'traffic_bytes_per_launch: N'), or None."""
    try:
        with open(PROFILE_TRAFFIC) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def parse_shape(text):
    """'Tx,Ty,Tz,threads[,groups[,lane_queue[,workgroups[,balance]]]]' ->
    dict, or None."""
    if not text:
        return None
    v = [int(x) for x in text.split(",")]
    return {"tiles": tuple(v[:3]), "block_threads": v[3],
            "run_groups": v[4] if len(v) > 4 else 0,
            "lane_queue": v[5] if len(v) > 5 else 0,
            "workgroups": v[6] if len(v) > 6 else 0,
            "balance": v[7] if len(v) > 7 else 0}


def error_stats(f, ref):
    """max and RMS of |f - ref| per atom (epsilon / sigma units), and both
    relative to the RMS reference force."""
    d = np.linalg.norm(f - ref, axis=1)
    frms = float(np.sqrt(np.mean(np.sum(ref * ref, axis=1))))
    return {"max_abs": float(d.max()), "rms_abs": float(np.sqrt(np.mean(d * d))),
            "rms_force": frms, "max_rel_to_rms_force": float(d.max()) / frms,
            "rms_rel_to_rms_force": float(np.sqrt(np.mean(d * d))) / frms}


def force_errors(m, box, threads):
    """Per-atom force errors at the handle's current positions: the mixed
    path against the oracle (the reference's FP64 evaluation restated, run
    on `threads` host threads), the FP64 path against the oracle, and the
    mixed path against the FP64 path.  Outside every timed region."""
    from oracle import oracle as O
    pos, _ = m.get_state()
    f64, _, _ = m.forces("fp64")
    fmx, _, _ = m.forces("mixed")
    ref, _, _ = O.forces(pos, box, RC, threads=threads)
    return {"mixed_vs_oracle": error_stats(fmx, ref),
            "fp64_vs_oracle": error_stats(f64, ref),
            "mixed_vs_fp64": error_stats(fmx, f64),
            "oracle": f"oracle C restatement of exaMD_forces_midpoint_all, "
                      f"FP64, {threads} host threads"}


def boundary_ms(m, prec, reps=3):
    """Milliseconds of one evaluation through the boundary with host
    buffers, as the drop-in force call makes it: positions in (set_state),
    bin + sort + force, forces out; and the same with the positions already
    resident (the drop-in call skips set_state when the positions have not
    changed since the last call).  PCIe-inclusive; never the headline."""
    p, _ = m.get_state()
    # the caller's force array is written again on every call, as the
    # drop-in call writes the caller's force array
    f = np.zeros_like(p)
    m.set_state(p)
    m.forces(prec, out=f)
    t0 = time.perf_counter()
    for _ in range(reps):
        m.set_state(p)
        m.forces(prec, out=f)
    t1 = time.perf_counter()
    for _ in range(reps):
        m.forces(prec, out=f)
    t2 = time.perf_counter()
    return (t1 - t0) / reps * 1e3, (t2 - t1) / reps * 1e3


def chunk_used(graph_chunk, steps, every):
    """Steps per captured chunk that examd_hip_run replays for a run of
    `steps` steps tallied every `every` steps (0 = launched kernel by
    kernel): the chunk is made a multiple of the tally stride, and a run
    shorter than the chunk is not captured."""
    if graph_chunk <= 0 or steps <= 0:
        return 0
    c = graph_chunk
    if every > 0:
        c = max(every, c // every * every)
    return c if c <= steps else 0


def timed_run(E, pos, vel, box, prec, steps, warmup, energy_every,
              shape=None, graph_chunk=100, report_error=False, tune_after=300,
              oracle_threads=1):
    """Run `tune_after` untimed steps (melting the lattice), then autotune
    the launch shape for this precision (or apply the given shape), warm up
    and time `steps` MD steps, replayed from captured hipGraph chunks of up
    to `graph_chunk` steps (a shorter run replays a chunk of its own length).
    Then, outside the timed region: the accuracy leg (DRIFT_STEPS more steps
    with energies every DRIFT_EVERY steps) and, with report_error, the
    per-atom force errors at the final positions and the PCIe-inclusive
    time of one evaluation through the boundary."""
    err = None
    chunk = min(graph_chunk, steps) if graph_chunk > 0 else 0
    with E.MidpointLJ(len(pos), box, RC, graph_chunk=chunk) as m:
        m.set_state(pos, vel)
        if tune_after:
            m.run(tune_after, DT, prec, energy_every=0)
        if shape is None:
            tune = m.autotune(prec, trials=3)
        else:
            m.set_config(prec, **shape)
            tune = m.get_config(prec)
        if warmup:
            m.run(warmup, DT, prec, energy_every=0)
        # one untimed chunk with the timed run's settings captures the graph
        # the timed steps replay (the handle keeps one captured chunk)
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        used = chunk_used(chunk, steps, energy_every)
        if used:
            m.run(used, DT, prec, energy_every=energy_every)
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        t0 = time.perf_counter()
        stats = m.run(steps, DT, prec, energy_every=energy_every)
        t1 = time.perf_counter()
        force_ms, launches = m.timing(enable=False, reset=True)
        # accuracy leg, in 100-step chunks
        m.set_graph_chunk(graph_chunk if graph_chunk > 0 else 0)
        drift = (m.run(DRIFT_STEPS[prec], DT, prec, energy_every=DRIFT_EVERY)
                 if DRIFT_STEPS[prec] > 0 else None)
        if report_error:
            err = force_errors(m, box, oracle_threads)
            ms, ms_res = boundary_ms(m, prec)
            err["boundary_eval_ms"] = ms
            err["boundary_eval_ms_resident"] = ms_res
    if launches < steps // TIME_EVERY:
        raise RuntimeError(f"{prec}: {launches} timed force launches for "
                           f"{steps} steps")
    return (t1 - t0, stats, force_ms / max(launches, 1), launches, tune, err,
            drift, used)


def batch_systems(E, nsys=256, seed=5):
    """Batching workload: nsys periodic LJ systems, atom counts log-uniform
    in [500, 20000], densities uniform in [0.60, 0.90], temperatures uniform
    in [0.70, 1.50] (one master seed), each a randomly depleted FCC lattice
    at its density with Maxwell-Boltzmann velocities."""
    rng = np.random.default_rng(seed)
    n = np.rint(np.exp(rng.uniform(np.log(500), np.log(20000), nsys)))
    rho = rng.uniform(0.60, 0.90, nsys)
    temp = rng.uniform(0.70, 1.50, nsys)
    out = []
    for k in range(nsys):
        pos, vel, box = E.generate_vacancy_fcc(int(n[k]), rho[k], temp[k],
                                               perturb=0.02, seed=seed * 1000 + k)
        out.append((pos, vel, box))
    return out


def batch_bench(E, steps, warmup, loop_steps, nsys=256):
    """Aggregate atom-steps/s of the batch in one handle (every kernel of a
    step covers all systems; 100-step hipGraph chunks; tile edge and launch
    shape autotuned after 300 untimed steps that take the perturbed lattices
    towards their liquids, as for the single system) against a per-system
    launch loop: each system in its own handle with the same shape and its
    own 100-step graphs, run in turn for loop_steps steps."""
    sy = batch_systems(E, nsys)
    total = int(sum(len(p) for p, _, _ in sy))
    systems = [(len(p), box) for p, _, box in sy]
    pos = np.vstack([p for p, _, _ in sy])
    vel = np.vstack([v for _, v, _ in sy])
    with E.MidpointLJBatch(systems, RC, graph_chunk=100) as b:
        b.set_state(pos, vel)
        b.run(300, DT, "mixed", energy_every=0)
        tune = b.autotune("mixed", trials=2)
        if warmup:
            b.run(warmup, DT, "mixed")
        # one untimed chunk with the timed run's settings captures the graph
        # the timed steps replay
        b.timing(enable=True, reset=True, every=TIME_EVERY)
        b.run(100, DT, "mixed", energy_every=100)
        b.timing(enable=True, reset=True, every=TIME_EVERY)
        t0 = time.perf_counter()
        rows = b.run(steps, DT, "mixed", energy_every=100)
        wall = time.perf_counter() - t0
        force_ms, launches = b.timing(enable=False, reset=True)
    e0 = np.array([s.ke + s.pe_shifted for s in rows[0]])
    e1 = np.array([s.ke + s.pe_shifted for s in rows[-1]])
    natoms = np.array([len(p) for p, _, _ in sy], dtype=float)
    handles = []
    try:
        for (p, v, box) in sy:
            m = E.MidpointLJ(len(p), box, RC, graph_chunk=100)
            handles.append(m)
            m.set_state(p, v)
            m.set_config("mixed", block_threads=tune["block_threads"],
                         run_groups=tune["run_groups"],
                         lane_queue=tune["lane_queue"],
                         tile_edge=tune["tile_edge"],
                         balance=tune["balance"])
            m.run(warmup if warmup else 1, DT, "mixed")
        t0 = time.perf_counter()
        for m in handles:
            m.run(loop_steps, DT, "mixed", energy_every=100)
        loop_wall = time.perf_counter() - t0
    finally:
        for m in handles:
            m.close()
    rate = total * steps / wall
    loop_rate = total * loop_steps / loop_wall
    return {
        "workload": f"{nsys} periodic LJ systems, atoms log-uniform in "
                    "[500, 20000], rho U[0.60, 0.90], T* U[0.70, 1.50], "
                    "rc 2.5, mixed precision, dt 0.005, 100-step hipGraph "
                    "chunks of the whole batch",
        "systems": nsys, "atoms": total, "steps": steps,
        "atom_steps_per_s": rate, "ms_per_step": 1e3 * wall / steps,
        "force_ms": force_ms / max(launches, 1),
        "per_system_loop_atom_steps_per_s": loop_rate,
        "per_system_loop_steps": loop_steps,
        "speedup_vs_loop": rate / loop_rate,
        "max_energy_change_per_atom": float(np.max(np.abs(e1 - e0) / natoms)),
        "launch": tune,
    }


def slab_bench(E, steps, warmup):
    """Non-uniform density: 4,000,000-atom liquid-vapour slab (box 3:1:1,
    central third of x at density 0.80, the rest at 0.05, uniform with a
    0.8 sigma minimum separation, T* 0.85), rc 3.0, mixed precision,
    autotuned cell-block map."""
    rc = 3.0
    vol = 4.0e6 / (0.80 / 3 + 0.05 * 2 / 3)
    ly = (vol / 3.0) ** (1.0 / 3.0)
    box = np.array([3 * ly, ly, ly])
    t0 = time.perf_counter()
    pos, vel = E.generate_slab(box, ly, 2 * ly, 0.80, 0.05, min_sep=0.8,
                               temperature=0.85, seed=4)
    gen_s = time.perf_counter() - t0
    n = len(pos)
    with E.MidpointLJ(n, box, rc, graph_chunk=100) as m:
        m.set_state(pos, vel)
        # the workload's named map, 32 x 12 x 12 cell blocks at equal atom
        # shares: refused when a region does not fit one workgroup's LDS
        try:
            m.set_config("mixed", (32, 12, 12), 1024, balance=1)
            named = "fits"
        except E.ExamdError as e:
            named = f"refused: {E.strerror(e.status)}"
        tune = m.autotune("mixed", trials=2)
        cuts = m.get_cuts("mixed")
        if warmup:
            m.run(warmup, DT, "mixed")
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        m.run(100, DT, "mixed", energy_every=100)
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        t0 = time.perf_counter()
        st = m.run(steps, DT, "mixed", energy_every=100)
        wall = time.perf_counter() - t0
        force_ms, launches = m.timing(enable=False, reset=True)
    return {
        "map": f"autotuned: {len(cuts[0]) - 1} x {len(cuts[1]) - 1} x "
               f"{len(cuts[2]) - 1} cell blocks"
               + (", boundaries at equal atom shares" if tune["balance"] else
                  ", even split"),
        "named_map_32x12x12_balanced": named,
        "workload": "liquid-vapour slab, box 3:1:1, central third at rho 0.80, "
                    "vapour 0.05, 0.8 sigma minimum separation, T* 0.85, "
                    "rc 3.0, mixed precision, dt 0.005",
        "atoms": n, "box": box.tolist(), "steps": steps,
        "atom_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "force_ms": force_ms / max(launches, 1),
        "pairs_cut": st[-1].pairs_cut, "generate_s": gen_s, "launch": tune,
    }


KA_EPS = [[1.0, 1.5], [1.5, 0.5]]      # Kob-Andersen 80:20 binary mixture
KA_SIG = [[1.0, 0.8], [0.8, 0.88]]
KA_RC = [[2.5, 2.0], [2.0, 2.2]]       # 2.5 sigma of each pair


def species_bench(E, steps, warmup, tune_after=300):
    """Several species: Kob-Andersen 80:20 binary LJ mixture, 2,048,000 atoms
    (80^3 FCC at rho 1.2, 20% species B at random, seeded), T* 1.0, per-pair
    eps, sigma and cutoff 2.5 sigma_ab, mixed precision, dt 0.003, autotuned
    after `tune_after` untimed steps; mixed atom-steps/s and the FP64 force
    time at the same state."""
    dt = 0.003
    pos, vel, box = E.generate_fcc(NC, 1.2, 1.0, seed=1200)
    rng = np.random.default_rng(1200)
    types = np.zeros(len(pos), dtype=np.int32)
    types[rng.choice(len(pos), len(pos) // 5, replace=False)] = 1
    n = len(pos)
    with E.MidpointLJ(n, box, RC, graph_chunk=100) as m:
        m.set_state(pos, vel)
        m.set_species(types, KA_EPS, KA_SIG, KA_RC)
        m.run(tune_after, dt, "mixed", energy_every=0)
        tune = m.autotune("mixed", trials=2)
        if warmup:
            m.run(warmup, dt, "mixed", energy_every=0)
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        t0 = time.perf_counter()
        st = m.run(steps, dt, "mixed", energy_every=100)
        wall = time.perf_counter() - t0
        force_ms, launches = m.timing(enable=False, reset=True)
        tune64 = m.autotune("fp64", trials=2)
    e0 = st[0].ke + st[0].pe_shifted      # each pair's own row's V(rc)
    e1 = st[-1].ke + st[-1].pe_shifted
    return {
        "workload": "Kob-Andersen 80:20 binary LJ mixture, 2,048,000 atoms "
                    "(80^3 FCC at rho 1.2, 20% B), T* 1.0, eps/sigma/cutoff "
                    "per species pair (cutoff 2.5 sigma_ab), mixed precision, "
                    "dt 0.003, 100-step hipGraph chunks",
        "atoms": n, "steps": steps,
        "atom_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "force_ms": force_ms / max(launches, 1),
        "force_ms_fp64": tune64["force_ms"],
        "pairs_cut": st[-1].pairs_cut,
        "shifted_energy_change_per_atom": (e1 - e0) / n,
        "launch": tune, "launch_fp64": tune64,
    }


def table_bench(E, steps, warmup, tune_after=300, npoints=4096, r_lo=0.6):
    """A tabulated pair function: the headline 2,048,000-atom liquid with
    the LJ pair given as a table (npoints uniform in r^2 from r_lo to rc,
    linear between them), the path a general user function takes; mixed
    atom-steps/s, force time against the analytic pair at the same state,
    and the tabulated forces' error against the analytic ones."""
    pos, vel, box = E.generate_fcc(NC, RHO, TEMP, seed=2048)
    r2 = np.linspace(r_lo * r_lo, RC * RC, npoints)
    s6 = (1.0 / r2) ** 3
    g = -24.0 * (2.0 * s6 * s6 - s6) / r2
    v = 4.0 * (s6 * s6 - s6)
    n = len(pos)
    with E.MidpointLJ(n, box, RC, graph_chunk=100) as m:
        m.set_state(pos, vel)
        m.run(tune_after, DT, "mixed", energy_every=0)
        f_lj, _, _ = m.forces("fp64")
        m.set_pair_table(r2[0], g, v)
        f_tab, _, _ = m.forces("fp64")
        tune = m.autotune("mixed", trials=2)
        if warmup:
            m.run(warmup, DT, "mixed", energy_every=0)
        m.timing(enable=True, reset=True, every=TIME_EVERY)
        t0 = time.perf_counter()
        st = m.run(steps, DT, "mixed", energy_every=100)
        wall = time.perf_counter() - t0
        force_ms, launches = m.timing(enable=False, reset=True)
    rms = np.sqrt((f_lj ** 2).sum(1).mean())
    e0 = st[0].ke + st[0].pe_shifted
    e1 = st[-1].ke + st[-1].pe_shifted
    return {
        "workload": "LJ liquid 2,048,000 atoms (the headline system) with the "
                    f"LJ pair as a table of {npoints} points uniform in r^2 "
                    f"from {r_lo} sigma to rc 2.5, cubic Hermite pieces in "
                    "r^2, mixed "
                    "precision, dt 0.005",
        "atoms": n, "steps": steps,
        "atom_steps_per_s": n * steps / wall, "ms_per_step": 1e3 * wall / steps,
        "force_ms": force_ms / max(launches, 1),
        "fp64_table_vs_analytic_max_rel": float(np.abs(f_tab - f_lj).max() / rms),
        "shifted_energy_change_per_atom": (e1 - e0) / n,
        "launch": tune,
    }


def config1(E, steps=100):
    """The reference's CPU-runnable case: 32,000 atoms (20^3 FCC, rho 0.8442,
    T* 0.72), rc 2.5, 4x4x4 cell blocks, FP64, 100 velocity-Verlet steps;
    the GPU FP64 path against the oracle on the same seeded input
    (atom-steps/s of both and the largest position difference after the
    run)."""
    from oracle import oracle as O
    pos, vel, box = E.generate_fcc(20, RHO, TEMP, seed=32000)
    with E.MidpointLJ(len(pos), box, RC, tiles=(4, 4, 4),
                      graph_chunk=steps) as m:
        m.set_state(pos, vel)
        m.run(steps, DT, "fp64")        # parity run (captures the chunk)
        gpos, gvel = m.get_state()
        t0 = time.perf_counter()        # timed: the same chunk replayed
        m.run(steps, DT, "fp64")
        gpu_s = time.perf_counter() - t0
    f0, _, _ = O.forces(pos, box, RC)
    t0 = time.perf_counter()
    opos, ovel = O.run_vv(pos, vel, box, RC, DT, steps, forces0=f0)[:2]
    cpu_s = time.perf_counter() - t0
    d = gpos - opos
    d -= np.round(d / box) * box
    n = len(pos)
    return {"workload": "32,000 atoms, 20^3 FCC, rho 0.8442, T* 0.72, rc 2.5, "
                        "4x4x4 cell blocks, FP64, dt 0.005",
            "steps": steps, "gpu_fp64_atom_steps_per_s": n * steps / gpu_s,
            "cpu_oracle_atom_steps_per_s": n * steps / cpu_s,
            "max_position_diff_vs_oracle": float(np.abs(d).max()),
            "max_velocity_diff_vs_oracle": float(np.abs(gvel - ovel).max())}


def cpu_baseline(pos, vel, box, nsteps, threads=1):
    """The oracle (the C restatement of the reference path, FP64) on a
    bounded sample of the same workload: nsteps VV steps (rebin + force) of
    the full system, after one untimed force evaluation, on `threads` host
    threads (the slab-threaded form when more than one)."""
    from oracle import oracle as O
    f0, _, _ = O.forces(pos, box, RC, threads=threads)
    t0 = time.perf_counter()
    O.run_vv(pos, vel, box, RC, DT, nsteps, forces0=f0, threads=threads)
    dt = time.perf_counter() - t0
    return len(pos) * nsteps / dt


def host_cpu():
    """Logical CPUs the process sees and the CPU model (/proc/cpuinfo)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def pmc_probe(E, shape, melt=300):
    """Short program for rocprofv3 --pmc runs: `melt` MD steps of the bench
    system (kernels without tallies), then three mixed force evaluations
    with tallies (k_forces<false, true, false, NT>, the dispatches the
    summaries use) at the given or autotuned shape, then three
    device-to-device copies of 256 MiB (a plain copy kernel of known size,
    to calibrate the FETCH_SIZE / WRITE_SIZE counters)."""
    import ctypes
    pos, vel, box = E.generate_fcc(NC, RHO, TEMP, seed=2048)
    with E.MidpointLJ(len(pos), box, RC) as m:
        m.set_state(pos, vel)
        if melt:
            m.run(melt, DT, "mixed", energy_every=0)
        if shape is None:
            print("tune", m.autotune("mixed", trials=2), flush=True)
        else:
            m.set_config("mixed", **shape)
        for _ in range(3):
            m.forces("mixed")
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"pmc probe done: copy bytes {nbytes}", flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="oracle MD steps of the one-core CPU baseline")
    ap.add_argument("--cpu-mt-steps", type=int, default=4,
                    help="oracle MD steps of the all-cores CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=5000)
    ap.add_argument("--config1-steps", type=int, default=100)
    ap.add_argument("--loop-steps", type=int, default=200)
    ap.add_argument("--slab-steps", type=int, default=2000)
    ap.add_argument("--species-steps", type=int, default=2000)
    ap.add_argument("--table-steps", type=int, default=1000)
    ap.add_argument("--drift-steps", default="10000,1000",
                    help="mixed,fp64 steps of the accuracy leg (0,0 skips it, "
                         "for profiler runs)")
    ap.add_argument("--graph-chunk", type=int, default=100,
                    help="MD steps per captured hipGraph chunk (0 = eager)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run the short counter-collection program instead")
    ap.add_argument("--mixed-shape", default=None,
                    help="Tx,Ty,Tz,threads,groups,lane_queue,workgroups,"
                         "balance instead of autotuning")
    ap.add_argument("--fp64-shape", default=None,
                    help="Tx,Ty,Tz,threads,groups,lane_queue,workgroups,"
                         "balance instead of autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and the "
              "path does not use a second (run with --gpus 1)",
              file=sys.stderr)
        return 2

    import examd_amd as E
    E.lib()
    if args.pmc_probe:
        return pmc_probe(E, parse_shape(args.mixed_shape))
    pos, vel, box = E.generate_fcc(NC, RHO, TEMP, seed=2048)
    n = len(pos)
    steps = args.steps
    fsteps = args.fp64_steps if args.fp64_steps is not None else steps
    energy_every = 10

    dm, df = (int(x) for x in args.drift_steps.split(","))
    DRIFT_STEPS["mixed"], DRIFT_STEPS["fp64"] = dm, df
    host_threads = max(1, min(16, os.cpu_count() or 1))
    t_start = time.perf_counter()
    (wall_mx, st_mx, force_ms_mx, launches, tune_mx, ferr, drift_mx,
     chunk_mx) = timed_run(E, pos, vel, box, "mixed", steps, args.warmup,
                           energy_every, parse_shape(args.mixed_shape),
                           args.graph_chunk, report_error=True,
                           oracle_threads=host_threads)
    progress("mixed timed run, accuracy leg and force errors", t_start)
    wall_64, st_64, force_ms_64, _, tune_64, _, drift_64, chunk_64 = timed_run(
        E, pos, vel, box, "fp64", fsteps, args.warmup, energy_every,
        parse_shape(args.fp64_shape), args.graph_chunk)
    progress("fp64 timed run and accuracy leg", t_start)

    rate_mx = n * steps / wall_mx
    rate_64 = n * fsteps / wall_64
    boundary = ferr.pop("boundary_eval_ms", None)
    boundary_res = ferr.pop("boundary_eval_ms_resident", None)
    # algorithmic work of one timed force launch: mean P_cand, P_cut over
    # the energy steps tallied; the timed launches tally no energy
    cand = float(np.mean([s.pairs_cand for s in st_mx]))
    cut = float(np.mean([s.pairs_cut for s in st_mx]))
    w = algorithmic_flops(cand, cut, False)
    achieved = w / (force_ms_mx * 1e-3) / 1e12
    w64 = algorithmic_flops(cand, cut, False)
    achieved64 = w64 / (force_ms_64 * 1e-3) / 1e12

    batch = None
    if args.batch_steps > 0:
        batch = batch_bench(E, args.batch_steps, args.warmup, args.loop_steps)
        progress("batch", t_start)
    slab = None
    if args.slab_steps > 0:
        slab = slab_bench(E, args.slab_steps, args.warmup)
        progress("slab", t_start)
    species = None
    if args.species_steps > 0:
        species = species_bench(E, args.species_steps, args.warmup)
        progress("species", t_start)
    table = None
    if args.table_steps > 0:
        table = table_bench(E, args.table_steps, args.warmup)
        progress("table", t_start)

    c1 = None
    if not args.no_cpu and args.config1_steps > 0:
        c1 = config1(E, args.config1_steps)
        progress("32,000-atom case against the oracle", t_start)

    cpu = None
    if not args.no_cpu and args.cpu_mt_steps > 0:
        v = cpu_baseline(pos, vel, box, args.cpu_mt_steps, host_threads)
        cpu = {"value": v, "unit": "atom-steps/s", "cores": host_threads,
               "kind": "port",
               "sample": f"{args.cpu_mt_steps} velocity-Verlet steps (rebin + "
                         f"force) of the same 2,048,000-atom system, FP64, "
                         f"oracle C restatement of the reference path on "
                         f"{host_threads} threads (z slabs in two phases)",
               "host": host_cpu()}
        progress("cpu baseline on all cores", t_start)
        if args.cpu_steps > 0:
            v1 = cpu_baseline(pos, vel, box, args.cpu_steps, 1)
            cpu["one_core"] = {
                "value": v1, "cores": 1,
                "sample": f"{args.cpu_steps} velocity-Verlet steps of the "
                          f"same system, the serial oracle (the reference's "
                          f"one-rank mode)"}

    traffic = read_traffic()
    line = {
        "metric": "atom-steps/s @2.048M-atom LJ: mixed vs FP64 GPU; "
                  "% of FP32 vector peak",
        "value": rate_mx,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall_mx / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed",
        "data": "synthetic",
        "config": {
            "workload": "LJ liquid 2,048,000 atoms (80^3 FCC, rho 0.8442, "
                        "T* 0.72, rc 2.5, dt 0.005), 107^3 midpoint blocks, "
                        "cell-block map chosen by the autotuner, rebin "
                        "every step",
            "natoms": n,
            "precision": "mixed: fp32 pair arithmetic, int64 fixed-point "
                         "(2^-32) force/energy accumulation, fp64 integration",
            "fp64_atom_steps_per_s": rate_64,
            "fp64_steps": fsteps,
            "mixed_vs_fp64": rate_mx / rate_64,
            "force_ms_mixed": force_ms_mx,
            "force_ms_fp64": force_ms_64,
            "force_launches_timed": launches,
            "pairs_cand": cand,
            "pairs_cut": cut,
            "force_frac_of_step_mixed": force_ms_mx * steps / (1e3 * wall_mx),
            "fp64_force_tflops": achieved64,
            "fp64_force_frac_of_fp64_peak": achieved64 / FP64_PEAK_TFLOPS,
            "energy_drift_per_1k_mixed": drift_per_1k(drift_mx),
            "energy_drift_per_1k_fp64": drift_per_1k(drift_64),
            "energy_drift_leg": {
                "steps": [DRIFT_STEPS["mixed"], DRIFT_STEPS["fp64"]],
                "tally_every": DRIFT_EVERY,
                "what": "shifted total energy KE + sum(V - V(rc)), "
                        "least-squares slope per 1,000 steps relative to "
                        "|E|, steps continued from the timed run (mixed, "
                        "fp64), outside the timed region"},
            "energy_max_dev_rel_mixed": max_dev(drift_mx),
            "energy_max_dev_rel_fp64": max_dev(drift_64),
            "force_error_at_end_of_drift_leg": ferr,
            "host_boundary": {
                "eval_ms_mixed": boundary,
                "eval_ms_mixed_positions_resident": boundary_res,
                "what": "one evaluation through the C ABI with host buffers, "
                        "as the drop-in force call makes it: set_state "
                        "(positions in), bin + sort + force, forces out; "
                        "and the same when the positions are unchanged "
                        "since the last call, which the drop-in call then "
                        "does not send again; PCIe-inclusive, not the "
                        "headline"},
            "launch_mixed": tune_mx,
            "launch_fp64": tune_64,
            "graph_chunk_steps": [chunk_mx, chunk_64],
            "graph_chunk_what": "steps per captured hipGraph chunk the timed "
                                "region replayed (mixed, fp64; 0 = launched "
                                "kernel by kernel)",
            "batch": batch,
            "slab": slab,
            "species": species,
            "table": table,
            "config1": c1,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": "k_forces (mixed)",
            "flops_per_launch": w,
            "launch_ms": force_ms_mx,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
