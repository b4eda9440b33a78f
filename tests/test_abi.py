"""The C ABI library on the CPU: loads, exports what the header declares,
and its host-only entry points behave (no kernel is launched here)."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "examd_hip.h")


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(examd_hip_\w+)\s*\(", text)))


def test_header_matches_python_exports(E):
    assert declared() == sorted(E.EXPORTS)


def test_every_declared_symbol_is_exported(E):
    lib = E.lib()
    for name in declared():
        assert hasattr(lib, name), name


def test_strerror(E):
    assert E.strerror(0) == "ok"
    assert "out of range" in E.strerror(109)
    assert E.strerror(12345) == "unknown status"


def test_domain_matches_oracle(E, oracle):
    for box in ([33.5918, 33.5918, 33.5918], [134.3677, 134.3677, 134.3677],
                [13.9, 12.6, 16.1]):
        nb, n = E.domain(box, 1.25)
        d = oracle.domain_create(box, 1.25)
        assert nb == d.nb and n == d.nblocks


def test_generate_fcc_invariants(E):
    pos, vel, box = E.generate_fcc(5, 0.8442, 0.72, mass=1.0, seed=3)
    n = len(pos)
    assert n == 500
    a = (4 / 0.8442) ** (1 / 3)
    assert np.allclose(box, 5 * a)
    assert np.all(pos >= 0) and np.all(pos < box)
    assert np.abs(vel.sum(0)).max() < 1e-12
    ke = 0.5 * (vel ** 2).sum()
    assert abs(ke - 1.5 * (n - 1) * 0.72) < 1e-10
    # lattice sites: cell x fastest, then the 4-atom basis
    assert np.allclose(pos[1], [0.5 * a, 0.5 * a, 0.0])
    assert np.allclose(pos[4], [a, 0.0, 0.0])
    p2, v2, _ = E.generate_fcc(5, 0.8442, 0.72, seed=3)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel)
    p3, _, _ = E.generate_fcc(5, 0.8442, 0.72, perturb=0.05, seed=3)
    d = p3 - pos
    d -= np.round(d / box) * box
    assert 0 < np.abs(d).max() <= 0.05


def _min_image_min_dist(pos, box):
    d = pos[:, None, :] - pos[None, :, :]
    d -= np.round(d / box) * box
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return np.sqrt(r2.min())


@pytest.mark.parametrize("n,rho", [(500, 0.9), (501, 0.9), (777, 0.6)])
def test_generate_vacancy_fcc_invariants(E, n, rho):
    pos, vel, box = E.generate_vacancy_fcc(n, rho, 1.1, perturb=0.02, seed=5)
    assert pos.shape == (n, 3)
    L = (n / rho) ** (1 / 3)
    assert np.allclose(box, L, rtol=1e-14)
    assert np.all(pos >= 0) and np.all(pos < box)
    nc = 1
    while 4 * nc ** 3 < n:
        nc += 1
    assert _min_image_min_dist(pos, box) >= L / nc / np.sqrt(2) - 0.04 * np.sqrt(3)
    assert np.abs(vel.sum(0)).max() < 1e-12
    assert abs(0.5 * (vel ** 2).sum() - 1.5 * (n - 1) * 1.1) < 1e-9
    p2, v2, _ = E.generate_vacancy_fcc(n, rho, 1.1, perturb=0.02, seed=5)
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel)
    p3, _, _ = E.generate_vacancy_fcc(n, rho, 1.1, perturb=0.02, seed=6)
    assert not np.array_equal(p3, pos)


def test_generate_slab_invariants(E):
    box = np.array([24.0, 8.0, 8.0])
    lo, hi = 8.0, 16.0
    pos, vel = E.generate_slab(box, lo, hi, 0.8, 0.05, min_sep=0.8,
                               temperature=0.85, seed=11)
    ns = round(0.8 * 8 * 64)
    nv = round(0.05 * 16 * 64)
    assert len(pos) == ns + nv
    assert np.all(pos >= 0) and np.all(pos < box)
    assert np.all((pos[:ns, 0] >= lo) & (pos[:ns, 0] < hi))
    assert np.all((pos[ns:, 0] < lo) | (pos[ns:, 0] >= hi))
    assert _min_image_min_dist(pos, box) >= 0.8
    n = len(pos)
    assert np.abs(vel.sum(0)).max() < 1e-12
    assert abs(0.5 * (vel ** 2).sum() - 1.5 * (n - 1) * 0.85) < 1e-9
    p2, _ = E.generate_slab(box, lo, hi, 0.8, 0.05, min_sep=0.8,
                            temperature=0.85, seed=11)
    assert np.array_equal(p2, pos)
    with pytest.raises(E.ExamdError):   # cannot pack 2.0 at 1.0 separation
        E.generate_slab(box, lo, hi, 2.0, 0.0, min_sep=1.0, seed=1)


def test_create_without_device_reports_status(E):
    """On a host without a GPU the ABI returns a status, never aborts."""
    try:
        import torch
        has_gpu = torch.cuda.device_count() > 0
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(E.ExamdError) as ei:
        E.MidpointLJ(100, [13.4] * 3, 2.5)
    assert ei.value.status == E.ERR_NO_DEVICE


def test_bad_arguments_rejected(E):
    with pytest.raises(E.ExamdError) as ei:
        E.domain([10.0, -1.0, 10.0], 1.25)
    assert ei.value.status == E.ERR_ARG


def test_checkpoint_round_trip_is_exact(E, tmp_path):
    pos, vel, box = E.generate_fcc(3, 0.8442, 1.3, perturb=0.1, seed=4)
    pos[0, 0] = np.nextafter(box[0], 0)     # extreme values survive
    vel[1, 2] = -1e-300
    E.write_checkpoint(tmp_path / "s.chk", pos, vel, box, step=1234)
    p2, v2, b2, st = E.read_checkpoint(tmp_path / "s.chk")
    assert st == 1234
    assert np.array_equal(p2, pos) and np.array_equal(v2, vel)
    assert np.array_equal(b2, box)
