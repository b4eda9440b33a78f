"""The CPU oracle against what the mathematics fixes.

The reference ships no golden vectors and cannot be built here without
stand-ins (it needs MPI and METIS), so every oracle function is pinned by
closed forms, independent brute force and invariants chosen so that a
plausible mistake (dropped term, wrong sign or index, transposed operand,
multiply-by-reciprocal instead of divide, floor instead of truncation) fails
one of them.
"""
import json
import os

import numpy as np
import pytest

from tests import helpers as H

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------ domain ----
def test_domain_create_counts(oracle):
    # config-1 and config-2 shapes: L = 20 a and 80 a, a = (4/0.8442)^(1/3)
    a = (4 / 0.8442) ** (1 / 3)
    for nc, nb in ((20, 26), (80, 107), (10, 13)):
        d = oracle.domain_create([nc * a] * 3, 1.25)
        assert d.nb == (nb, nb, nb)
        assert d.nblocks == nb ** 3
    d = oracle.domain_create([12.0, 7.6, 100.0], 1.25)   # Int truncates
    assert d.nb == (9, 6, 80)


# ----------------------------------------------------- identify_block ----
def test_identify_block_faces_and_clamp(oracle):
    d = oracle.domain_create([13.4, 13.4, 13.4], 1.25)   # 10 blocks, last enlarged
    assert d.nb == (10, 10, 10)
    cases = [
        ((0.0, 0.0, 0.0), (1, 1, 1)),
        ((1.25, 0.0, 0.0), (2, 1, 1)),          # exactly on a face
        ((2.5, 3.75, 5.0), (3, 4, 5)),
        ((12.5, 0.0, 0.0), (10, 1, 1)),         # start of the last block
        ((13.3999, 0.0, 0.0), (10, 1, 1)),      # clamp: Int(11.7) -> 10
        ((np.nextafter(1.25, 0), 0.0, 0.0), (1, 1, 1)),
    ]
    for pos, want in cases:
        rc, bxyz, num = oracle.identify_block(d, pos)
        assert rc == 0 and bxyz == want, (pos, bxyz, want)
        # blocknum: x fastest (Fortran order ravel)
        assert num == 1 + np.ravel_multi_index(np.array(want) - 1, d.nb,
                                               order="F")


def test_identify_block_divides_not_multiplies(oracle):
    """Int(x/bl + 1) with an IEEE division: find x where multiplying by
    1/bl rounds to the other side of a face and check the division wins."""
    bl = 1.25                      # 1/bl = 0.8 is inexact in binary
    d = oracle.domain_create([60.0, 60.0, 60.0], bl)
    inv = 1.0 / bl
    found = 0
    for k in range(1, 48):
        x = k * bl
        for step in range(-3, 4):
            xx = x
            for _ in range(abs(step)):
                xx = np.nextafter(xx, np.inf if step > 0 else -np.inf)
            by_div = int(xx / bl + 1.0)
            by_mul = int(xx * inv + 1.0)
            if by_div != by_mul:
                rc, bxyz, _ = oracle.identify_block(d, (xx, 0.0, 0.0))
                assert rc == 0 and bxyz[0] == min(by_div, d.nb[0])
                found += 1
    assert found > 0


def test_identify_block_matches_exact_floor_away_from_faces(oracle):
    from fractions import Fraction
    rng = np.random.default_rng(5)
    L = 21.7
    bl = 1.25
    d = oracle.domain_create([L, L, L], bl)
    for x in rng.uniform(0, L, 300):
        q = Fraction(float(x)) / Fraction(bl)
        frac = q - (q.numerator // q.denominator)
        if frac < Fraction(1, 10 ** 12) or frac > 1 - Fraction(1, 10 ** 12):
            continue
        want = min(q.numerator // q.denominator + 1, d.nb[0])
        _, bxyz, _ = oracle.identify_block(d, (x, x, x))
        assert bxyz == (want,) * 3


def test_identify_block_out_of_range_is_109(oracle):
    d = oracle.domain_create([13.4, 13.4, 13.4], 1.25)
    rc, _, _ = oracle.identify_block(d, (-2.0, -2.0, -2.0))
    assert rc == 109


# ------------------------------------------------------------- table ----
def test_table_covers_every_offset_class_once(oracle):
    t = oracle.interactions()
    assert t.shape == (63, 2)
    assert [tuple(r) for r in t] == H.TABLE
    seen = set()
    for n1, n2 in t:
        o1 = np.array(oracle.stencil_offset(n1))
        o2 = np.array(oracle.stencil_offset(n2))
        assert tuple(o1) == H.stencil(n1) and tuple(o2) == H.stencil(n2)
        d = tuple(o2 - o1)
        key = min(d, tuple(-v for v in d))
        assert key not in seen
        seen.add(key)
    assert len(seen) == 63      # (5^3 - 1)/2 + self
    assert tuple(t[0]) == (14, 14)   # row 1: self


# ---------------------------------------------------------- LJ pair ----
def test_lj_closed_forms(oracle):
    rc = 2.5
    r0 = 2 ** (1 / 6)
    rv, f, e = oracle.lennardjones_12(r0, 0, 0, r0 * r0, rc)
    assert rv == 0 and abs(f[0]) < 1e-13 and abs(e + 1.0) < 1e-14
    rv, f, e = oracle.lennardjones_12(1.0, 0, 0, 1.0, rc)
    # r = sigma: V = 0; force on atom 1 points away from atom 2 (-x)
    assert rv == 0 and e == 0.0 and f[0] == -24.0
    for eps, sig, r in ((0.7, 1.3, 1.6), (2.0, 0.9, 1.05)):
        rv, f, e = oracle.lennardjones_12(0, r, 0, r * r, rc, eps, sig)
        s6 = (sig / r) ** 6
        assert abs(e - 4 * eps * (s6 * s6 - s6)) < 1e-12
        assert abs(f[1] - (-24 * eps * (2 * s6 * s6 - s6) / r)) < 1e-12
    # d2 == rc^2 is inside (test is d2 > rc^2), one ulp above is not
    rv, _, _ = oracle.lennardjones_12(rc, 0, 0, rc * rc, rc)
    assert rv == 0
    rv, _, _ = oracle.lennardjones_12(rc, 0, 0, np.nextafter(rc * rc, 9), rc)
    assert rv == 1


# ------------------------------------------------------------ forces ----
@pytest.mark.parametrize("seed,nc,amp", [(1, 5, 0.05), (2, 6, 0.2), (3, 7, 0.1)])
def test_forces_match_all_pairs(oracle, seed, nc, amp):
    pos, L = H.perturbed_fcc(nc, 0.8442, amp, seed)
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5)
    fb, pe, ncut = H.brute_force(pos, [L] * 3, 2.5)
    scale = np.sqrt((fb ** 2).sum(1).mean())
    assert np.abs(f - fb).max() / scale < 1e-12
    assert t.cut == ncut
    assert abs(t.pe - pe) < 1e-10 * abs(pe)
    assert t.zero == 0


def test_forces_noncubic_uneven_box(oracle):
    rng = np.random.default_rng(7)
    box = np.array([13.9, 12.6, 16.1])           # last blocks enlarged
    pos = rng.uniform(0, 1, (600, 3)) * box
    # push apart coincident atoms so forces stay finite
    fb, pe, ncut = H.brute_force(pos, box, 2.5)
    f, t, _ = oracle.forces(pos, box, 2.5)
    scale = np.abs(fb).max()
    assert np.abs(f - fb).max() / scale < 1e-12
    assert t.cut == ncut


def test_newton_third_law(oracle):
    pos, L = H.perturbed_fcc(8, 0.8442, 0.05, 11)
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5)
    assert np.abs(f.sum(0)).max() < 1e-11 * np.abs(f).sum()


def test_candidate_pairs_closed_form(oracle):
    """P_cand = sum_b n_b(n_b-1)/2 + sum over the 62 offset classes of
    sum_b n_b n_{b+d}, counted from block occupancies alone."""
    pos, L = H.perturbed_fcc(6, 0.8442, 0.1, 4)
    nb = int(L / 1.25)
    b = H.block_of(pos, 1.25, (nb,) * 3)
    occ = np.zeros((nb,) * 3, dtype=np.int64)
    np.add.at(occ, (b[:, 2], b[:, 1], b[:, 0]), 1)
    cand = int((occ * (occ - 1) // 2).sum())
    for n1, n2 in H.TABLE[1:]:
        o1, o2 = np.array(H.stencil(n1)), np.array(H.stencil(n2))
        d = o2 - o1
        cand += int((occ * np.roll(occ, shift=(-d[2], -d[1], -d[0]),
                                   axis=(0, 1, 2))).sum())
    _, t, _ = oracle.forces(pos, [L] * 3, 2.5)
    assert t.cand == cand


def test_ownership_pair_centric(oracle):
    pos, L = H.perturbed_fcc(5, 0.8442, 0.15, 9)
    _, t, hc = oracle.forces(pos, [L] * 3, 2.5, home_counts=True)
    want = H.ownership_counts(pos, [L] * 3, 2.5)
    assert hc.sum() == t.cut
    np.testing.assert_array_equal(hc, want)


def test_fcc_lattice_energy(oracle):
    pos, L = H.fcc(6, 0.8442)
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5)
    want = H.fcc_lattice_pe_per_atom(0.8442, 2.5)
    assert abs(t.pe / len(pos) - want) < 1e-12 * abs(want)
    assert np.abs(f).max() < 1e-11       # symmetric lattice: no net force
    assert t.cut == len(pos) * 54 // 2   # shells 12+6+24+12 within 2.5


def test_two_atoms_across_periodic_face(oracle):
    L = 13.0
    pos = np.array([[0.2, 6.0, 6.0], [12.5, 6.0, 6.0]])    # r = 0.7 via image
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5)
    r = 0.7
    s6 = r ** -6
    mag = 24 * (2 * s6 * s6 - s6) / r
    assert t.cut == 1
    assert abs(f[0, 0] - mag) < 1e-9 * mag       # pushed to +x, away
    assert abs(f[1, 0] + mag) < 1e-9 * mag


def test_grid_too_small_rejected(oracle):
    pos = np.zeros((2, 3)) + 1.0
    with pytest.raises(ValueError):
        oracle.forces(pos, [5.0, 5.0, 5.0], 2.4)     # 4 blocks per side


# -------------------------------------------------------- integrator ----
def _state(nc=5, seed=3):
    pos, L = H.perturbed_fcc(nc, 0.8442, 0.05, seed)
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, np.sqrt(0.72), pos.shape)
    vel -= vel.mean(0)
    return pos, vel, L


def test_vv_time_reversible(oracle):
    pos0, vel0, L = _state()
    p, v, f, *_ = oracle.run_vv(pos0, vel0, [L] * 3, 2.5, 0.005, 40)
    p2, v2, *_ = oracle.run_vv(p, -v, [L] * 3, 2.5, 0.005, 40, forces0=f)
    dp = p2 - pos0
    dp -= np.round(dp / L) * L
    assert np.abs(dp).max() < 1e-9
    assert np.abs(v2 + vel0).max() < 1e-9


def test_vv_conserves_shifted_energy_and_momentum(oracle):
    pos0, vel0, L = _state(6, 8)
    p, v, f, ke, pe, cut = oracle.run_vv(pos0, vel0, [L] * 3, 2.5, 0.005, 200)
    vrc = 4 * (2.5 ** -12 - 2.5 ** -6)
    e = ke + pe - cut * vrc
    assert np.abs(e - e[0]).max() < 2e-4 * abs(e[0])
    assert np.abs(v.sum(0)).max() < 1e-10


def test_vv_first_step_closed_form(oracle):
    """One step by hand: v' = v + h F, x' = x + dt v', then v'' = v' + h F'."""
    pos0, vel0, L = _state(5, 2)
    f0, _, _ = oracle.forces(pos0, [L] * 3, 2.5)
    dt = 0.005
    vh = vel0 + 0.5 * dt * f0
    x1 = np.mod(pos0 + dt * vh, L)
    f1, _, _ = oracle.forces(x1, [L] * 3, 2.5)
    v1 = vh + 0.5 * dt * f1
    p, v, f, ke, pe, cut = oracle.run_vv(pos0, vel0, [L] * 3, 2.5, dt, 1)
    assert np.abs(p - x1).max() < 1e-12
    assert np.abs(v - v1).max() < 1e-12
    assert abs(ke[0] - 0.5 * (v1 ** 2).sum()) < 1e-10 * ke[0]


def test_wrap_rule(oracle):
    L = 10.0
    assert oracle.wrap(-1e-18, L) == 0.0          # x + L rounds to L -> 0
    assert oracle.wrap(10.0, L) == 0.0
    assert oracle.wrap(-0.5, L) == 9.5
    assert oracle.wrap(10.25, L) == 0.25


# ------------------------------------------------------ particle file ----
def test_particle_file_round_trip(oracle, tmp_path):
    pos, L = H.perturbed_fcc(3, 0.8442, 0.05, 1)
    p = tmp_path / "exaMD_particles.dat"
    oracle.write_particles(str(p), pos, [L, L, L])
    lines = p.read_text().split("\n")
    assert int(lines[0]) == len(pos)            # N, then box, then x y z
    assert len(lines[1].split()) == 3
    back, box = oracle.read_particles(str(p))
    np.testing.assert_array_equal(back, pos)
    np.testing.assert_array_equal(box, [L, L, L])


# ------------------------------------------------------ golden files ----
# ----------------------------------------------------------- species ----
def _ka(pos):
    return (H.binary_types(len(pos), 0.2, 5), H.KA_EPS, H.KA_SIG, H.KA_RC)


def test_species_forces_match_all_pairs(oracle):
    """Binary mixture with per-pair eps, sigma and cutoff against an
    all-pairs sum that looks the parameters up pair by pair."""
    pos, L = H.perturbed_fcc(6, 1.2, 0.1, 21)
    sp = _ka(pos)
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5, species=sp)
    fb, pe, ncut = H.species_brute_force(pos, [L] * 3, *sp)
    scale = np.sqrt((fb ** 2).sum(1).mean())
    assert np.abs(f - fb).max() / scale < 1e-12
    assert t.cut == ncut
    assert abs(t.pe - pe) < 1e-10 * abs(pe)
    assert np.abs(f.sum(0)).max() < 1e-11 * np.abs(f).sum()


def test_species_identical_rows_reduce_to_one_species(oracle):
    """Every row equal to (eps, sig, rc): bit-identical to the one-species
    evaluation, tallies included."""
    pos, L = H.perturbed_fcc(6, 0.8442, 0.1, 22)
    types = H.binary_types(len(pos), 0.5, 6)
    one = [[1.3, 1.3], [1.3, 1.3]], [[0.9, 0.9], [0.9, 0.9]], [[2.5, 2.5], [2.5, 2.5]]
    f1, t1, h1 = oracle.forces(pos, [L] * 3, 2.5, eps=1.3, sig=0.9, home_counts=True)
    f2, t2, h2 = oracle.forces(pos, [L] * 3, 2.5, home_counts=True,
                               species=(types,) + one)
    assert np.array_equal(f1, f2)
    assert (t1.cand, t1.cut, t1.pe) == (t2.cand, t2.cut, t2.pe)
    np.testing.assert_array_equal(h1, h2)


def test_species_pair_closed_form_and_own_cutoff(oracle):
    """One A and one B atom: the force is the AB row's LJ at r, and it
    vanishes between the AB cutoff (2.0) and the block cutoff (2.5)."""
    L = 13.0
    types = np.array([0, 1], dtype=np.int32)
    for r, inside in ((0.9, True), (1.95, True), (2.05, False), (2.4, False)):
        pos = np.array([[4.0, 6.0, 6.0], [4.0 + r, 6.0, 6.0]])
        f, t, _ = oracle.forces(pos, [L] * 3, 2.5,
                                species=(types, H.KA_EPS, H.KA_SIG, H.KA_RC))
        s6 = (0.8 / r) ** 6
        mag = 24 * 1.5 * (2 * s6 * s6 - s6) / r
        if inside:
            assert t.cut == 1
            assert abs(f[1, 0] - mag) < 1e-12 * abs(mag)   # on B, along +x
            assert abs(f[0, 0] + mag) < 1e-12 * abs(mag)
            assert abs(t.pe - 4 * 1.5 * (s6 * s6 - s6)) < 1e-12
        else:
            assert t.cut == 0 and not f.any()
    # the same two atoms as A-A interact out to 2.5
    pos = np.array([[4.0, 6.0, 6.0], [6.4, 6.0, 6.0]])
    _, t, _ = oracle.forces(pos, [L] * 3, 2.5, species=(
        np.array([0, 0], dtype=np.int32), H.KA_EPS, H.KA_SIG, H.KA_RC))
    assert t.cut == 1


def test_species_bad_table_rejected(oracle):
    pos, L = H.perturbed_fcc(5, 0.8442, 0.1, 23)
    types = H.binary_types(len(pos), 0.2, 7)
    rc = [[2.5, 2.6], [2.6, 2.5]]                  # beyond the block cutoff
    with pytest.raises(ValueError):
        oracle.forces(pos, [L] * 3, 2.5, species=(types, H.KA_EPS, H.KA_SIG, rc))
    bad = types.copy()
    bad[3] = 2                                     # no such species
    with pytest.raises(ValueError):
        oracle.forces(pos, [L] * 3, 2.5, species=(bad, H.KA_EPS, H.KA_SIG, H.KA_RC))


# ------------------------------------------------------------- table ----
def test_table_forces_match_all_pairs(oracle):
    """A tabulated pair function against an all-pairs sum that interpolates
    the same table independently."""
    pos, L = H.perturbed_fcc(6, 0.8442, 0.1, 31)
    tab = H.lj_table(2001, 0.8, 2.5)
    f, t, _ = oracle.forces(pos, [L] * 3, 2.5, table=tab)
    fb, pe, ncut = H.table_brute_force(pos, [L] * 3, 2.5, tab)
    scale = np.sqrt((fb ** 2).sum(1).mean())
    assert np.abs(f - fb).max() / scale < 1e-12
    assert t.cut == ncut and t.below == 0
    assert abs(t.pe - pe) < 1e-10 * abs(pe)


def test_table_of_lj_approaches_lj(oracle):
    """The LJ pair tabulated on a fine grid reproduces the analytic
    evaluation within the cubic interpolation's bound, which falls as h^4."""
    pos, L = H.perturbed_fcc(6, 0.8442, 0.1, 32)
    ref, t0, _ = oracle.forces(pos, [L] * 3, 2.5)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    errs = []
    for npts in (1001, 4001):
        f, t, _ = oracle.forces(pos, [L] * 3, 2.5, table=H.lj_table(npts, 0.8, 2.5))
        errs.append(np.abs(f - ref).max() / scale)
        assert t.cut == t0.cut
    assert errs[1] < 1e-9
    assert 180 < errs[0] / errs[1] < 360          # fourth order: about 256x


def test_table_cubic_in_r2_is_exact(oracle):
    """g and V cubic in r^2 are interpolated exactly (the slopes of the
    fourth-order differences are exact for them): two atoms at r give the
    closed form, and a pair closer than the table is counted."""
    L = 13.0
    r2 = np.linspace(0.25, 6.25, 7)
    gp = lambda x: 3.0 - 0.5 * x + 0.25 * x * x - 0.03 * x ** 3
    vp = lambda x: 1.0 + 2.0 * x - 0.4 * x * x + 0.02 * x ** 3
    g, v = gp(r2), vp(r2)
    for r in (0.7, 1.3, 2.2, 2.5):
        pos = np.array([[4.0, 6.0, 6.0], [4.0 + r, 6.0, 6.0]])
        f, t, _ = oracle.forces(pos, [L] * 3, 2.5, table=(0.25, g, v))
        want = gp(r * r) * r
        assert abs(f[0, 0] - want) < 1e-12 and abs(f[1, 0] + want) < 1e-12
        assert abs(t.pe - vp(r * r)) < 1e-12
    pos = np.array([[4.0, 6.0, 6.0], [4.3, 6.0, 6.0]])    # r^2 = 0.09 < 0.25
    _, t, _ = oracle.forces(pos, [L] * 3, 2.5, table=(0.25, g, v))
    assert t.below == 1


def test_short_table_is_linear(oracle):
    """Below five points a table is interpolated linearly: g and V linear in
    r^2 on three points come out exactly."""
    L = 13.0
    r2 = np.linspace(0.25, 6.25, 3)
    g, v = 3.0 - 0.5 * r2, 1.0 + 2.0 * r2
    for r in (0.7, 1.9):
        pos = np.array([[4.0, 6.0, 6.0], [4.0 + r, 6.0, 6.0]])
        f, t, _ = oracle.forces(pos, [L] * 3, 2.5, table=(0.25, g, v))
        assert abs(f[0, 0] - (3.0 - 0.5 * r * r) * r) < 1e-13
        assert abs(t.pe - (1.0 + 2.0 * r * r)) < 1e-13


def test_golden_kats(oracle):
    """Committed known-answer vectors (tests/golden/make_golden.py)."""
    path = os.path.join(GOLDEN, "kats.json")
    kats = json.load(open(path))
    for k in kats["identify_block"]:
        d = oracle.domain_create(k["box"], k["bl"])
        rc, bxyz, num = oracle.identify_block(d, k["pos"])
        assert (rc, list(bxyz), num) == (k["rc"], k["blockxyz"], k["blocknum"])
    for k in kats["lj"]:
        rv, f, e = oracle.lennardjones_12(*k["args"])
        assert rv == k["retval"]
        if rv == 0:
            # relative to the size of the cancelling terms 24 eps s^6 / r^2
            scale = 24 * k["args"][5] * (k["args"][6] ** 2 / k["args"][3]) ** 3
            assert np.allclose(f, k["f12"], rtol=1e-13, atol=1e-14 * scale)
            assert np.isclose(e, k["e12"], rtol=1e-13, atol=1e-14 * scale)


@pytest.mark.parametrize("nc,threads", [(8, 4), (13, 8), (16, 16)])
def test_threaded_form_matches_serial(oracle, nc, threads):
    """The slab-threaded evaluation (the all-cores CPU baseline) visits the
    same pairs with the same arithmetic: pair counts and per-home-block
    ownership equal, forces and energy within FP64 summation order."""
    pos, L = H.perturbed_fcc(nc, 0.8442, 0.1, nc)
    stretch = np.array([1.0, 1.1, 0.95])
    pos = np.mod(pos, L) * stretch
    box = L * stretch
    f1, t1, h1 = oracle.forces(pos, box, 2.5, home_counts=True)
    fm, tm, hm = oracle.forces(pos, box, 2.5, home_counts=True, threads=threads)
    assert (t1.cand, t1.cut, t1.zero) == (tm.cand, tm.cut, tm.zero)
    np.testing.assert_array_equal(h1, hm)
    scale = np.sqrt((f1 ** 2).sum(1).mean())
    assert np.abs(f1 - fm).max() < 1e-13 * scale
    assert abs(t1.pe - tm.pe) < 1e-13 * abs(t1.pe)
    vel = np.random.default_rng(nc).normal(0.0, 0.8, pos.shape)
    r1 = oracle.run_vv(pos, vel, box, 2.5, 0.005, 3)
    rm = oracle.run_vv(pos, vel, box, 2.5, 0.005, 3, threads=threads)
    assert np.abs(r1[0] - rm[0]).max() < 1e-12
    assert np.abs(r1[1] - rm[1]).max() < 1e-12
    np.testing.assert_array_equal(r1[5], rm[5])
