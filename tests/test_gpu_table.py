"""A tabulated pair function on the GPU (examd_hip_set_pair_table): the
general user function of exaMD_forces_midpoint_all sampled on a grid in r^2
and interpolated by cubic Hermite pieces in r^2, against the oracle's
evaluation of the same table.

Bars: FP64 forces within 1e-10 relative of the oracle's table evaluation
(the same interpolation), pair counts exact, energy within 1e-10; the mixed
path within its FP32 bars; the table of the LJ pair within the
interpolation error of the analytic LJ evaluation; a pair function cubic in
r^2 reproduced exactly (FP64 to rounding).
"""
import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def _rel(f, ref):
    return np.abs(f - ref).max() / np.sqrt((ref ** 2).sum(1).mean())


def _morse_table(n, r_lo, rc, de=0.8, a=2.0, r0=1.1):
    """A Morse pair V = De (1 - exp(-a (r - r0)))^2 - De, g = -V'(r) / r."""
    r2 = np.linspace(r_lo * r_lo, rc * rc, n)
    r = np.sqrt(r2)
    x = np.exp(-a * (r - r0))
    v = de * (1 - x) ** 2 - de
    g = -2 * de * a * (1 - x) * x / r
    return r_lo * r_lo, g, v


@pytest.mark.parametrize("name", ["lj", "morse"])
def test_table_fp64_matches_oracle(E, oracle, name):
    print(f"[table fp64] {name}", flush=True)
    pos, L = H.perturbed_fcc(10, 0.8442, 0.1, 41)
    box = np.array([L, L, L])
    tab = H.lj_table(3001, 0.75, 2.5) if name == "lj" else _morse_table(2049, 0.6, 2.5)
    ref, t, hc = oracle.forces(pos, box, 2.5, table=tab, home_counts=True)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        m.set_pair_table(*tab)
        f, s, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
    assert _rel(f, ref) < FP64_TOL
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    assert abs(s.pe - t.pe) < FP64_TOL * abs(t.pe)
    np.testing.assert_array_equal(got, hc)
    assert abs(s.pe_shifted - (t.pe - t.cut * tab[2][-1])) < 1e-9 * abs(t.pe)


def test_table_mixed_matches_oracle(E, oracle):
    print("[table mixed]", flush=True)
    pos, L = H.perturbed_fcc(10, 0.8442, 0.1, 42)
    box = np.array([L, L, L])
    tab = _morse_table(2049, 0.6, 2.5)
    ref, t, _ = oracle.forces(pos, box, 2.5, table=tab)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        m.set_pair_table(*tab)
        f, s, _ = m.forces("mixed")
    assert _rel(f, ref) < 2e-5
    assert s.pairs_cand == t.cand and abs(s.pairs_cut - t.cut) <= 2
    assert abs(s.pe - t.pe) < 1e-5 * abs(t.pe)
    assert np.abs(f.sum(0)).max() < 1e-6


def test_table_of_lj_close_to_lj_and_cleared(E, oracle):
    """The LJ pair tabulated at 4001 points gives the analytic LJ forces
    within the interpolation error (the oracle's table error bounds it);
    clearing the table returns to the analytic pair bit for bit (mixed)."""
    print("[table of lj]", flush=True)
    pos, L = H.perturbed_fcc(10, 0.8442, 0.1, 43)
    box = np.array([L, L, L])
    tab = H.lj_table(4001, 0.75, 2.5)
    ref_t, _, _ = oracle.forces(pos, box, 2.5, table=tab)
    ref, _, _ = oracle.forces(pos, box, 2.5)
    bound = _rel(ref_t, ref)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        f0, s0, _ = m.forces("mixed")
        m.set_pair_table(*tab)
        f64, _, _ = m.forces("fp64")
        m.set_pair_table()
        f1, s1, _ = m.forces("mixed")
    assert _rel(f64, ref) < 1.01 * bound + 1e-12
    assert bound < 1e-8
    assert np.array_equal(f0, f1) and s0.pe == s1.pe


def test_table_cubic_in_r2_is_exact(E):
    """g and V cubic in r^2 come out exactly from seven points: two atoms at
    several distances give the closed form, FP64 to rounding and mixed
    within FP32 rounding of the coefficients."""
    print("[table cubic]", flush=True)
    r2 = np.linspace(0.25, 6.25, 7)
    gp = lambda x: 3.0 - 0.5 * x + 0.25 * x * x - 0.03 * x ** 3
    vp = lambda x: 1.0 + 2.0 * x - 0.4 * x * x + 0.02 * x ** 3
    box = np.array([13.0, 13.0, 13.0])
    for r in (0.7, 1.3, 2.2, 2.5):
        pos = np.array([[4.0, 6.0, 6.0], [4.0 + r, 6.0, 6.0]])
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos)
            m.set_pair_table(0.25, gp(r2), vp(r2))
            f64, s64, _ = m.forces("fp64")
            f32, s32, _ = m.forces("mixed")
        want = gp(r * r) * r
        assert abs(f64[0, 0] - want) < 1e-12 and abs(f64[1, 0] + want) < 1e-12, r
        assert abs(s64.pe - vp(r * r)) < 1e-12, r
        assert abs(f32[0, 0] - want) < 1e-6 * max(1.0, abs(want)), r
        assert abs(s32.pe - vp(r * r)) < 1e-6 * max(1.0, abs(vp(r * r))), r


def test_table_range_and_arguments(E):
    """A pair closer than the table's range is reported; a table does not
    combine with several species; bad arguments change nothing."""
    print("[table range]", flush=True)
    pos, L = H.perturbed_fcc(6, 0.8442, 0.05, 44)
    box = np.array([L, L, L])
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        m.set_state(pos)
        m.set_pair_table(*H.lj_table(1001, 1.2, 2.5))    # nearest pairs ~1.1
        for prec in ("fp64", "mixed"):
            with pytest.raises(E.ExamdError) as ei:
                m.forces(prec)
            assert ei.value.status == E.ERR_TABLE_RANGE
        with pytest.raises(E.ExamdError):
            m.set_species(np.zeros(len(pos), np.int32), [[1.0]], [[1.0]], [[2.5]])
        lo, g, v = H.lj_table(1001, 0.8, 2.5)
        with pytest.raises(E.ExamdError):
            m.set_pair_table(7.0, g, v)                  # beyond the cutoff
        with pytest.raises(E.ExamdError):
            m.set_pair_table(lo, g[:1], v[:1])           # one point
        m.set_pair_table(lo, g, v)
        f, s, _ = m.forces("fp64")
        assert s.pairs_cut > 0


def test_table_md_conserves_energy(E):
    """200 mixed steps with the tabulated LJ pair (a hard core, so no pair
    leaves the table's range), replayed from graph chunks: the shifted
    energy (V(rc) = the table's last point) drifts by less than 1e-3
    relative."""
    print("[table md]", flush=True)
    pos, vel, box = E.generate_fcc(8, 0.8442, 0.7, perturb=0.02, seed=45)
    with E.MidpointLJ(len(pos), box, 2.5, graph_chunk=50) as m:
        m.set_state(pos, vel)
        m.set_pair_table(*H.lj_table(4097, 0.6, 2.5))
        st = m.run(200, 0.004, "mixed", energy_every=50)
    e = np.array([r.ke + r.pe_shifted for r in st])
    assert np.all(np.isfinite(e))
    assert np.abs(e - e[0]).max() < 1e-3 * np.abs(e).max()


def test_table_batch_equals_standalone(E):
    """A table in a batch handle: each system's mixed forces bit-identical
    to the same system alone with the same table."""
    print("[table batch]", flush=True)
    tab = _morse_table(2049, 0.6, 2.5)
    sy = [H.perturbed_fcc(nc, 0.8442, 0.1, s) for nc, s in ((6, 51), (8, 52))]
    sy = [(p, np.array([L, L, L])) for p, L in sy]
    with E.MidpointLJBatch([(len(p), b) for p, b in sy], 2.5) as bt:
        bt.set_state(np.vstack([p for p, _ in sy]))
        bt.set_pair_table(*tab)
        fb, sb, _ = bt.forces("mixed")
        off = bt.offsets
    for k, (pos, box) in enumerate(sy):
        with E.MidpointLJ(len(pos), box, 2.5) as m:
            m.set_state(pos)
            m.set_pair_table(*tab)
            f1, s1, _ = m.forces("mixed")
        assert np.array_equal(fb[off[k]:off[k + 1]], f1), k
        assert (sb[k].pe, sb[k].pairs_cut) == (s1.pe, s1.pairs_cut), k
