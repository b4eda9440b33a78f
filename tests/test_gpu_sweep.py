"""Seeded sweep over boxes and launch shapes against the oracle: random
non-cubic boxes (edges 7-16σ, most not multiples of the block length, so the
last block of an axis is enlarged), random densities 0.5-1.0 (uniform random
positions with a 0.85σ minimum separation), and random cell-block maps,
workgroup sizes, queue depths, persistent grids and balanced boundaries.

The cases cycle through the pair functions: the one-species LJ pair, three
species with random symmetric (ε, σ, cutoff) tables, and the LJ pair as a
4096-point table.

Bars as in the other parity tests: FP64 forces within 1e-10 relative,
energy within 1e-10, pair counts and per-home-block ownership exact; mixed
within 2e-5 (max) of the oracle with the candidate count exact and Σf
exactly zero in fixed point.
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

CASES = 12
KINDS = ("lj", "species", "table")


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _case(E, k):
    rng = np.random.default_rng(1000 + k)
    box = rng.uniform(7.0, 16.0, 3)
    rho = rng.uniform(0.5, 1.0)
    pos, _ = E.generate_slab(box, 0.0, box[0], rho, 0.0, min_sep=0.85,
                             seed=2000 + k)
    nb = (box / 1.25).astype(int)
    shape = dict(
        tiles=tuple(int(rng.integers(1, max(2, n // 3) + 1)) for n in nb),
        block_threads=int(rng.choice([256, 512, 1024])),
        lane_queue=int(rng.choice([0, 4, 8])),
        workgroups=4 if k % 3 == 2 else 0,
        balance=int(rng.integers(0, 2)))
    return pos, box, shape


def _species(n, k):
    """Three species at random, symmetric random (eps, sigma, cutoff)."""
    rng = np.random.default_rng(3000 + k)
    types = rng.integers(0, 3, n).astype(np.int32)
    tabs = []
    for lo, hi in ((0.5, 1.5), (0.8, 1.1), (2.0, 2.5)):
        a = rng.uniform(lo, hi, (3, 3))
        tabs.append(np.triu(a) + np.triu(a, 1).T)
    return (types, *tabs)


@pytest.mark.parametrize("k", range(CASES))
def test_sweep_matches_oracle(E, oracle, k):
    pos, box, shape = _case(E, k)
    kind = KINDS[k % 3]
    print(f"[sweep {k}] {kind} n={len(pos)} box={np.round(box, 2).tolist()} "
          f"{shape}", flush=True)
    extra = {}
    if kind == "species":
        extra["species"] = _species(len(pos), k)
    elif kind == "table":
        extra["table"] = H.lj_table(4096, 0.6, 2.5)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True, **extra)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        if kind == "species":
            m.set_species(*extra["species"])
        elif kind == "table":
            m.set_pair_table(*extra["table"])
        for prec in ("fp64", "mixed"):
            m.set_config(prec, **shape)
            f, s, _ = m.forces(prec)
            got = m.home_pair_counts(prec)
            if prec == "fp64":
                assert np.abs(f - ref).max() / scale < 1e-10
                assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
                assert abs(s.pe - t.pe) < 1e-10 * abs(t.pe)
                np.testing.assert_array_equal(got, hc)
            else:
                assert np.abs(f - ref).max() / scale < 2e-5
                assert s.pairs_cand == t.cand
                assert abs(s.pairs_cut - t.cut) <= 2
                q = np.round(f * 2.0 ** 32).astype(np.int64)
                assert np.all(q.sum(0) == 0)
