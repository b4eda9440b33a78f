"""The Fortran drop-in layer: builds with amdflang, links the C ABI, and
drives the GPU path with the reference's call sequence."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "examd_amd", "fortran", "examd_md")
FLANG = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "amdflang")


def _driver():
    if not os.path.exists(DRIVER):
        if not os.path.exists(FLANG):
            pytest.skip("amdflang absent and the driver is not prebuilt")
        from examd_amd import build as B
        B.build_fortran()
    return DRIVER


def test_bindings_cover_the_header():
    """Every C entry point has an ISO_C_BINDING interface."""
    hdr = open(os.path.join(ROOT, "include", "examd_hip.h")).read()
    names = set(re.findall(r"\b(examd_hip_\w+)\s*\(", hdr))
    f90 = open(os.path.join(ROOT, "examd_amd", "fortran",
                            "exaMD_hip_binding.f90")).read()
    bound = set(re.findall(r'name="(examd_hip_\w+)"', f90))
    assert names == bound


CALLER = """
Program caller
  Use exaMD_mod
  Use exaMD_forces_mod
  Use exaMD_utils_mod
  Use exaMD_partitioning
  Use exaMD_comms
  Use exaMD_comms_overlap
  Use exaMD_input_formats
  Use user_func_mod
  Implicit None
  Integer, Allocatable :: arrived(:), justin(:), sendlist(:)
  Integer :: n
  Allocate(arrived(1), justin(1), sendlist(1))
  Call exaMD_start_comms_posn()
  Call exaMD_recv_comms_posn(1)
  Call exaMD_wait_comms_posn(1)
  Call exaMD_start_comms_forces(1)
  Call exaMD_wait_comms_forces(1)
  n = exaMD_testsome_comms_posn(arrived, justin, 1)
  Call exaMD_recvsome_comms_posn(justin, n, 1)
  Call exaMD_waitrest_comms_posn(arrived, justin, 1)
  Call exaMD_start_comms_forcessome(0, sendlist, 1)
End Program caller
"""


def test_reference_module_surface_compiles(tmp_path):
    """A caller that uses every module of the reference's force path
    (exaMD_mod, exaMD_forces_mod, exaMD_utils_mod, exaMD_partitioning,
    exaMD_comms, exaMD_comms_overlap, exaMD_input_formats, user_func_mod)
    and the halo routines with the reference's argument lists compiles
    against the drop-in modules."""
    if not os.path.exists(FLANG):
        pytest.skip("amdflang absent")
    _driver()   # builds the modules when needed
    src = tmp_path / "caller.f90"
    src.write_text(CALLER)
    mods = os.path.join(ROOT, "examd_amd", "fortran", "_build")
    r = subprocess.run([FLANG, "-c", str(src), "-I", mods, "-module-dir",
                        str(tmp_path), "-o", str(tmp_path / "caller.o")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr


def test_driver_reports_missing_device(tmp_path):
    """Without a GPU the driver sets up the grid, then stops with the
    status of examd_hip_create instead of crashing."""
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except Exception:
        pass
    r = subprocess.run([_driver(), "-atoms", "4000"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    out = r.stdout + r.stderr
    assert "Number of blocks in each dimension: 13 13 13" in out
    assert "no HIP device" in out


@pytest.mark.gpu
def test_driver_force_check_and_md(tmp_path, oracle):
    """The Fortran driver's force check on a 4,000-atom FCC lattice (the
    reference's Sum f and Sum f^2 report) and a short MD run."""
    from tests import helpers as H
    r = subprocess.run([_driver(), "-atoms", "4000", "-steps", "200",
                        "-writefile"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    out = r.stdout
    print(out[-3000:], flush=True)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Pairs: candidate (\d+), in cutoff (\d+)", out)
    pos, L = H.fcc(10, 0.8442)
    _, t, _ = oracle.forces(pos, [L, L, L], 2.5)
    assert int(m.group(1)) == t.cand and int(m.group(2)) == t.cut
    drift = float(re.search(r"Relative shifted-energy change:\s*(\S+)",
                            out).group(1))
    assert abs(drift) < 1e-3
    assert os.path.exists(tmp_path / "exaMD_particles.dat")


@pytest.mark.gpu
def test_driver_checkpoint_restart_is_exact(tmp_path):
    """Mixed precision: 40 steps in one run and 20 + 20 steps through a
    checkpoint (positions and velocities written with 17 digits, the box and
    the step) end in the same checkpoint, bit for bit."""
    drv = _driver()
    common = ["-atoms", "4000", "-precision", "mixed", "-temp", "1.0"]

    def run(*extra):
        r = subprocess.run([drv, *common, *extra], cwd=tmp_path,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        return r.stdout

    run("-steps", "40", "-checkpoint", "a.chk")
    run("-steps", "20", "-checkpoint", "b1.chk")
    out = run("-restart", "b1.chk", "-steps", "20", "-checkpoint", "b2.chk")
    assert "Checkpoint written" in out
    a = open(tmp_path / "a.chk").read().split("\n")
    b = open(tmp_path / "b2.chk").read().split("\n")
    assert a[0].split() == ["4000", "40"] and b[0].split() == ["4000", "40"]
    assert a[1:] == b[1:]


@pytest.mark.gpu
def test_driver_binary_mixture_pair_counts(tmp_path, oracle):
    """-binary: species through grid%ltype, kktype, ltpvdw and prmvdw (every
    fifth atom B); the driver's pair counts against the oracle's
    species evaluation of the same lattice."""
    from tests import helpers as H
    r = subprocess.run([_driver(), "-atoms", "4000", "-rho", "1.2", "-binary",
                        "-steps", "50", "-precision", "mixed"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    out = r.stdout
    print(out[-3000:], flush=True)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Pairs: candidate (\d+), in cutoff (\d+)", out)
    pos, L = H.fcc(10, 1.2)
    types = np.array([1 if (i + 1) % 5 == 0 else 0 for i in range(len(pos))],
                     dtype=np.int32)
    _, t, _ = oracle.forces(pos, [L, L, L], 2.5,
                            species=(types, H.KA_EPS, H.KA_SIG, H.KA_RC))
    assert int(m.group(1)) == t.cand
    assert abs(int(m.group(2)) - t.cut) <= 2


@pytest.mark.gpu
def test_driver_tabulates_a_general_user_function(tmp_path, oracle):
    """-morse: the user function is not the LJ pair, so the library samples
    it (4096 points from 0.2 rc) and the GPU interpolates the table; the
    driver's pair counts and energy against the oracle's evaluation of the
    same table."""
    from tests import helpers as H
    r = subprocess.run([_driver(), "-atoms", "4000", "-morse", "-steps", "20"],
                       cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    out = r.stdout
    print(out[-3000:], flush=True)
    assert r.returncode == 0, r.stderr
    m = re.search(r"Pairs: candidate (\d+), in cutoff (\d+)", out)
    pe = float(re.search(r"Potential energy:\s*(\S+)", out).group(1))
    pos, L = H.fcc(10, 0.8442)
    r2 = np.linspace((0.2 * 2.5) ** 2, 2.5 ** 2, 4096)
    rr = np.sqrt(r2)
    x = np.exp(-2.0 * (rr - 1.1))
    g = -2 * 0.8 * 2.0 * (1 - x) * x / rr
    v = 0.8 * (1 - x) ** 2 - 0.8
    _, t, _ = oracle.forces(pos, [L, L, L], 2.5, table=(r2[0], g, v))
    assert (int(m.group(1)), int(m.group(2))) == (t.cand, t.cut)
    assert abs(pe - t.pe) < 1e-6 * abs(t.pe)


@pytest.mark.gpu
def test_driver_rebalance(tmp_path):
    """-rebalance 25: exaMD_md_run applies the load-balanced map and
    re-partitions it every 25 steps of a 100-step mixed run; the run keeps
    its energy and ends within rounding-driven divergence of the run on the
    default map."""
    drv = _driver()
    common = ["-atoms", "4000", "-precision", "mixed", "-steps", "100"]
    states = []
    for extra, chk in ((["-rebalance", "25"], "a.chk"), ([], "b.chk")):
        r = subprocess.run([drv, *common, *extra, "-checkpoint", chk],
                           cwd=tmp_path, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr
        drift = float(re.search(r"Relative shifted-energy change:\s*(\S+)",
                                r.stdout).group(1))
        assert abs(drift) < 1e-3
        states.append(np.loadtxt(tmp_path / chk, skiprows=2))
    d = states[0][:, :3] - states[1][:, :3]
    L = (4000 / 0.8442) ** (1 / 3)
    d -= np.round(d / L) * L
    assert np.abs(d).max() < 1e-4


def _particle_file(tmp_path, oracle):
    """exaMD_particles.dat of a non-cubic box whose sides are not multiples
    of the block length (enlarged last blocks on every axis), uniform random
    atoms with a 0.85 sigma minimum separation at density 0.75."""
    import examd_amd as E
    box = np.array([17.3, 15.1, 19.7])
    pos, _ = E.generate_slab(box, 0.0, box[0], 0.75, 0.75, min_sep=0.85,
                             seed=77)
    oracle.write_particles(tmp_path / "exaMD_particles.dat", pos, box)
    return oracle.read_particles(tmp_path / "exaMD_particles.dat")


def _read_dump(path, pos_ref):
    """The driver's per-atom dump (grid order) matched to the file's atom
    order by position (both read the same decimal file)."""
    d = np.loadtxt(path)
    where = {tuple(p): i for i, p in enumerate(pos_ref)}
    idx = np.array([where[tuple(p)] for p in d[:, :3]])
    assert np.array_equal(np.sort(idx), np.arange(len(pos_ref)))
    f = np.zeros_like(pos_ref)
    f[idx] = d[:, 3:]
    return f


def _block_sums(pos, f, box, rc):
    """Per-block sums of squared force components, keyed by the reference's
    1-based (i, j, k)."""
    from tests import helpers as H
    nb = (np.asarray(box) / (0.5 * rc)).astype(np.int64)
    b = H.block_of(pos, 0.5 * rc, nb)
    out = {}
    for bb, ff in zip(map(tuple, b), f):
        key = (bb[0] + 1, bb[1] + 1, bb[2] + 1)
        out[key] = out.get(key, 0.0) + ff ** 2
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("call", ["all", "overlap"])
def test_driver_readfile_forces_per_atom(tmp_path, oracle, call):
    """-readfile through exaMD_forces_midpoint_all (and -overlap through
    exaMD_forces_midpoint_overlap), evaluated twice on unchanged positions:
    every atom's accumulated grid%forces equals twice the oracle's forces of
    the same file (FP64, 1e-10 normwise), the pair counts are exact, the
    PRINTALLFORCES lines equal the oracle's per-block sums of squares, the
    second call does not send the positions again, and the mixed path stays
    within its FP32 bars."""
    pos, box = _particle_file(tmp_path, oracle)
    ref, t, _ = oracle.forces(pos, box, 2.5)
    print(f"[fortran readfile {call}] n={len(pos)} box={box.tolist()}", flush=True)
    extra = ["-overlap"] if call == "overlap" else []
    r = subprocess.run([_driver(), "-readfile", "-evals", "2", "-printforces",
                        "-dumpforces", "f64.txt", *extra], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    out = r.stdout
    print(out[-1500:], flush=True)
    assert r.returncode == 0, r.stderr
    assert f"Number of atoms: {len(pos)}" in out
    m = re.search(r"Pairs: candidate (\d+), in cutoff (\d+)", out)
    assert (int(m.group(1)), int(m.group(2))) == (t.cand, t.cut)
    assert "positions not resent: 1" in out
    f = _read_dump(tmp_path / "f64.txt", pos)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert np.abs(f - 2.0 * ref).max() < 1e-10 * 2.0 * scale
    want = _block_sums(pos, 2.0 * ref, box, 2.5)
    lines = re.findall(r"^\s*(\d{3,})\s*(\d{3,})\s*(\d{3,})\s+(\S+)\s+(\S+)\s+(\S+)\s*$",
                       out, re.M)
    assert len(lines) == sum(1 for v in want.values() if v.sum() > 0)
    for i, j, k, a, b, c in lines:
        got = np.array([float(a), float(b), float(c)])
        w = want[(int(i), int(j), int(k))]
        assert np.all(np.abs(got - w) <= 1e-10 * max(w.sum(), 1e-300)), (i, j, k)
    full = re.search(r"SUM\(fxx, fyy, fzz\) full:\s*(\S+)\s+(\S+)\s+(\S+)", out)
    assert np.abs([float(x) for x in full.groups()]).max() < 1e-10 * np.abs(ref).sum()
    # the mixed path through the same call
    r = subprocess.run([_driver(), "-readfile", "-precision", "mixed",
                        "-dumpforces", "fmx.txt", *extra], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    fm = _read_dump(tmp_path / "fmx.txt", pos)
    err = np.sqrt(((fm - ref) ** 2).sum(1))
    assert err.max() / scale < 2e-5
    assert np.all(np.round(fm * 2.0 ** 32).astype(np.int64).sum(0) == 0)


@pytest.mark.gpu
def test_driver_status_codes(tmp_path, oracle):
    """The Fortran layer's handling of the device statuses, as the
    reference behaves: a coincident pair (0.02 sigma apart) is reported and
    the forces still come back (the reference prints "ZERO at" and goes on,
    SRC/exaMD_forces.f90:525-527); on the mixed path that pair is beyond
    the fixed-point range and the program stops with code 102; an atom
    outside the box stops it with 109 (SRC/exaMD_module.f90:118-123)."""
    pos, box = _particle_file(tmp_path, oracle)
    pos = pos.copy()
    pos[5] = pos[9] + np.array([0.02, 0.0, 0.0])
    pos[5] = np.where(pos[5] >= box, pos[9] - np.array([0.02, 0.0, 0.0]), pos[5])
    oracle.write_particles(tmp_path / "exaMD_particles.dat", pos, box)
    pos, box = oracle.read_particles(tmp_path / "exaMD_particles.dat")
    ref, t, _ = oracle.forces(pos, box, 2.5)
    assert t.zero == 1
    print(f"[fortran status] zero-distance pair, n={len(pos)}", flush=True)
    r = subprocess.run([_driver(), "-readfile", "-dumpforces", "f64.txt"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    out = r.stdout
    assert r.returncode == 0, out[-2000:] + r.stderr
    assert "(status 10)" in out
    m = re.search(r"Pairs: candidate (\d+), in cutoff (\d+)", out)
    assert (int(m.group(1)), int(m.group(2))) == (t.cand, t.cut)
    f = _read_dump(tmp_path / "f64.txt", pos)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert np.abs(f - ref).max() < 1e-10 * scale
    print("[fortran status] the same pair, mixed", flush=True)
    r = subprocess.run([_driver(), "-readfile", "-precision", "mixed"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "(status 9)" in r.stdout and "aborting with code 102" in r.stdout
    print("[fortran status] an atom outside the box", flush=True)
    bad = pos.copy()
    bad[3, 2] = -2.0 * 0.5 * 2.5   # block -1 along z: block number < 1
    oracle.write_particles(tmp_path / "exaMD_particles.dat", bad, box)
    r = subprocess.run([_driver(), "-readfile"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0
    assert "aborting with code 109" in r.stdout
