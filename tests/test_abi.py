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
