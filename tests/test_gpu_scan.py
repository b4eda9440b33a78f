"""A grid of 164^3 = 4,410,944 blocks, most of them empty: a thin liquid
slab (ρ 0.8, 4σ thick) in a 206σ box, so that the block lists (8 slots per
block) span 141 MB and block ids reach 2^22.  FP64 forces within 1e-10 of
the oracle, pair counts and per-home-block ownership exact; the mixed path
within its FP32 bars."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    import examd_amd
    examd_amd.lib()
    return examd_amd


def test_two_level_scan_large_box(E, oracle):
    print("[two-level scan] 164^3 blocks", flush=True)
    box = np.array([206.0, 206.0, 206.0])
    pos, vel = E.generate_slab(box, 100.0, 104.0, 0.8, 0.0, min_sep=0.85,
                               temperature=0.85, seed=21)
    ref, t, hc = oracle.forces(pos, box, 2.5, home_counts=True)
    with E.MidpointLJ(len(pos), box, 2.5) as m:
        assert m.nblocks == 164 ** 3
        m.set_state(pos, vel)
        f, s, _ = m.forces("fp64")
        got = m.home_pair_counts("fp64")
        fm, sm, _ = m.forces("mixed")
        # a few MD steps through the same sort
        st = m.run(20, 0.005, "mixed", energy_every=10)
    scale = np.sqrt((ref ** 2).sum(1).mean())
    assert np.abs(f - ref).max() / scale < 1e-10
    assert (s.pairs_cand, s.pairs_cut) == (t.cand, t.cut)
    np.testing.assert_array_equal(got, hc)
    assert np.abs(fm - ref).max() / scale < 2e-5
    assert sm.pairs_cand == t.cand
    assert len(st) == 2 and st[-1].pairs_cand > 0
