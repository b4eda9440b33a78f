"""Several species on the GPU: per-species-pair Lennard-Jones rows (eps,
sigma, cutoff), through examd_hip_set_species, against the oracle.

Bars: as for one species.  FP64 forces within 1e-10 relative of the oracle,
pair counts, energies and per-home-block ownership exact; the mixed path
within its FP32 bars; a table whose rows are all equal gives the
one-species results (mixed bit for bit).
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
KA = (H.KA_EPS, H.KA_SIG, H.KA_RC)


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def _mixture(nc, amp, seed, rho=1.2, frac_b=0.2):
    pos, L = H.perturbed_fcc(nc, rho, amp, seed)
    return pos, np.array([L, L, L]), H.binary_types(len(pos), frac_b, seed)


@pytest.mark.parametrize("nc,amp,seed", [(8, 0.05, 1), (10, 0.15, 2)])
def test_species_fp64_matches_oracle(E, oracle, nc, amp, seed):
    print(f"[species fp64] nc={nc} amp={amp} seed={seed}", flush=True)
    pos, box, ty = _mixture(nc, amp, seed)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True,
                               species=(ty,) + KA)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        m.set_species(ty, *KA)
        f, s, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    assert abs(s.pe - t.pe) < FP64_TOL * abs(t.pe)
    np.testing.assert_array_equal(got, hc)
    # shifted energy: each pair's own row's V(rc), rounded once per row to
    # 2^-32 and tallied as integers, so within pairs * 2^-33 of the sum
    shift = H.species_shift_sum(pos, box, ty, *KA)
    assert abs(s.pe_shifted - (t.pe - shift)) <= t.cut * 2.0 ** -33 + 1e-12 * abs(t.pe)


def test_species_mixed_matches_oracle(E, oracle):
    print("[species mixed]", flush=True)
    pos, box, ty = _mixture(10, 0.1, 3)
    ref, t, _ = oracle.forces(pos, box, 2.5, species=(ty,) + KA)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        m.set_species(ty, *KA)
        f, s, _ = m.forces("mixed")
    assert _rel(f, ref) < 2e-5
    assert s.pairs_cand == t.cand and abs(s.pairs_cut - t.cut) <= 2
    assert abs(s.pe - t.pe) < 1e-5 * abs(t.pe)
    # fixed-point sums: Newton's third law holds exactly
    assert np.abs(f.sum(0)).max() < 1e-6


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_species_equal_rows_match_one_species(E, prec):
    """Every species-pair row equal to the handle's (eps, sigma, cutoff):
    the species path gives the one-species results (mixed forces and
    energy bit for bit; FP64 forces and energy to the rounding of their
    summation order, which the global atomics of the ghost forces and the
    in-tile work order leave open; pair counts exactly); clearing the
    species returns to the one-species path."""
    print(f"[species equal rows] {prec}", flush=True)
    pos, L = H.perturbed_fcc(10, 0.8442, 0.1, 4)
    box = np.array([L, L, L])
    ty = H.binary_types(len(pos), 0.4, 4)
    k = 3
    ty[::7] = 2
    eps, sig, rc = (np.full((k, k), v) for v in (1.0, 1.0, 2.5))
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        f1, s1, _ = m.forces(prec)
        m.set_species(ty, eps, sig, rc)
        f2, s2, _ = m.forces(prec)
        m.set_species()
        f3, s3, _ = m.forces(prec)
    for f, s in ((f2, s2), (f3, s3)):
        assert (s.pairs_cand, s.pairs_cut) == (s1.pairs_cand, s1.pairs_cut)
        if prec == "mixed":
            assert np.array_equal(f, f1) and s.pe == s1.pe
        else:
            assert _rel(f, f1) < 1e-13
            assert abs(s.pe - s1.pe) < 1e-12 * abs(s1.pe)


def test_species_md_matches_oracle(E, oracle):
    """20 FP64 velocity-Verlet steps of a hot binary mixture, captured in
    10-step graph chunks, against the oracle's integration."""
    print("[species md]", flush=True)
    pos, box, ty = _mixture(7, 0.05, 5)
    rng = np.random.default_rng(5)
    vel = rng.normal(0.0, 1.0, pos.shape)
    vel -= vel.mean(0)
    p_ref, v_ref, _, ke, pe, cut = oracle.run_vv(pos, vel, box, 2.5, 0.002, 20,
                                                 species=(ty,) + KA)
    with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=10) as m:
        m.set_state(pos, vel)
        m.set_species(ty, *KA)
        st = m.run(20, 0.002, "fp64", energy_every=10)
        p, v = m.get_state()
    d = p - p_ref
    d -= np.round(d / box) * box
    assert np.abs(d).max() < 1e-11
    assert np.abs(v - v_ref).max() < 1e-10
    assert st[-1].pairs_cut == cut[-1]
    assert abs(st[-1].pe - pe[-1]) < 1e-10 * abs(pe[-1])


def test_species_batch_equals_standalone(E):
    """Species in a batch: each system's mixed forces bit-identical to the
    same system alone."""
    print("[species batch]", flush=True)
    sy = [_mixture(nc, 0.1, s) for nc, s in ((6, 11), (8, 12), (7, 13))]
    b = E.MidpointLJBatch([(len(p), box) for p, box, _ in sy], 2.5)
    with b:
        b.set_state(np.vstack([p for p, _, _ in sy]))
        b.set_species(np.concatenate([t for _, _, t in sy]), *KA)
        fb, sb, _ = b.forces("mixed")
        off = b.offsets
    for k, (pos, box, ty) in enumerate(sy):
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos)
            m.set_species(ty, *KA)
            f1, s1, _ = m.forces("mixed")
        assert np.array_equal(fb[off[k]:off[k + 1]], f1), k
        assert (sb[k].pe, sb[k].pairs_cut) == (s1.pe, s1.pairs_cut), k


def test_species_rejects_bad_tables(E):
    print("[species errors]", flush=True)
    pos, box, ty = _mixture(6, 0.05, 6)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        eps = np.array(H.KA_EPS)
        bad = eps.copy()
        bad[0, 1] = 1.4                                   # not symmetric
        for args in ((ty, bad, H.KA_SIG, H.KA_RC),
                     (ty, eps, H.KA_SIG, [[2.5, 2.6], [2.6, 2.2]]),  # > cutoff
                     (np.where(ty == 1, 2, 0), eps, H.KA_SIG, H.KA_RC)):
            with pytest.raises(E.ExamdError):
                m.set_species(*args)
        # nothing changed: still the one-species path
        m.set_species(ty, *KA)
        m.set_species()
        f, s, _ = m.forces("fp64")
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        f1, s1, _ = m.forces("fp64")
    assert _rel(f, f1) < 1e-13
    assert (s.pairs_cand, s.pairs_cut) == (s1.pairs_cand, s1.pairs_cut)
    assert abs(s.pe - s1.pe) < 1e-12 * abs(s1.pe)
