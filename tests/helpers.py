"""Independent mathematical references used by the tests.

Nothing here follows the reference's block/table loop: forces come from an
all-pairs minimum-image sum, ownership from a pair-centric application of
the 63-row table, lattice energies from a direct lattice-vector sum.  They
pin the oracle; the oracle pins the HIP library.
"""
from __future__ import annotations

import itertools

import numpy as np

# interactions(1:2, 1:63) of SRC/exaMD_module.f90:282-344, as data
TABLE = [(14, 14), (14, 15), (14, 16), (14, 17), (14, 18), (14, 19), (14, 20),
         (14, 21), (14, 22), (14, 23), (14, 24), (14, 25), (14, 26), (14, 27),
         (10, 18), (10, 27), (11, 17), (11, 18), (11, 25), (11, 26), (11, 27),
         (12, 16), (12, 17), (12, 25), (13, 15), (13, 18), (13, 24), (13, 27),
         (15, 16), (15, 19), (15, 22), (15, 25), (16, 21), (16, 24), (17, 19),
         (17, 20), (17, 21), (18, 19), (19, 5), (19, 9), (20, 5), (20, 7),
         (20, 8), (20, 9), (21, 5), (21, 7), (22, 5), (22, 6), (22, 9),
         (23, 5), (24, 5), (24, 4), (24, 7), (25, 5), (25, 3), (25, 6),
         (26, 5), (26, 2), (26, 3), (27, 1), (27, 2), (27, 4), (27, 5)]


def stencil(n):
    """neighrefs index n (1..27) -> offset, ii fastest (numpy unravel)."""
    kk, jj, ii = np.unravel_index(n - 1, (3, 3, 3))
    return (int(ii) - 1, int(jj) - 1, int(kk) - 1)


def fcc(nc, rho):
    """Perfect FCC lattice, cells x fastest; returns (pos, L)."""
    a = (4.0 / rho) ** (1.0 / 3.0)
    basis = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]])
    k, j, i = np.meshgrid(np.arange(nc), np.arange(nc), np.arange(nc),
                          indexing="ij")
    cells = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    pos = (cells[:, None, :] + basis[None]).reshape(-1, 3) * a
    return pos, nc * a


def brute_force(pos, box, rc, eps=1.0, sig=1.0):
    """All-pairs minimum-image LJ: (forces, pe, n_cut)."""
    pos = np.asarray(pos, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    n = len(pos)
    f = np.zeros((n, 3))
    pe = 0.0
    ncut = 0
    blk = 512
    for s in range(0, n, blk):
        d = pos[None, :, :] - pos[s:s + blk, None, :]      # x_j - x_i
        d -= np.round(d / box) * box
        r2 = (d ** 2).sum(-1)
        idx = np.arange(s, min(n, s + blk))
        r2[np.arange(len(idx)), idx] = np.inf
        m = r2 <= rc * rc
        sr6 = np.where(m, (sig * sig / np.where(m, r2, 1.0)) ** 3, 0.0)
        dg = np.where(m, -24 * eps * (2 * sr6 * sr6 - sr6) / np.where(m, r2, 1.0), 0.0)
        f[idx] += (dg[:, :, None] * d).sum(1)
        pe += 0.5 * np.where(m, 4 * eps * (sr6 * sr6 - sr6), 0.0).sum()
        ncut += int(m.sum())
    return f, pe, ncut // 2


def block_of(pos, bl, nb):
    """exaMD_identify_block as written: Int(x/bl + 1) clamped (0-based)."""
    q = np.floor(np.asarray(pos) / bl + 1.0).astype(np.int64)
    q = np.minimum(q, np.asarray(nb)) - 1
    return q


def ownership_counts(pos, box, rc):
    """In-cutoff pairs per home block, pair by pair.

    For a pair with block offset d (minimum periodic offset, |d_a| <= 2),
    the row of the table whose o2 - o1 is d or -d fixes the home block
    (block of the o1-side atom minus o1)."""
    box = np.asarray(box, dtype=np.float64)
    bl = rc / 2
    nb = (box / bl).astype(np.int64)
    b = block_of(pos, bl, nb)
    rows = {}
    for (n1, n2) in TABLE:
        o1, o2 = np.array(stencil(n1)), np.array(stencil(n2))
        rows[tuple(o2 - o1)] = o1
    counts = np.zeros(int(np.prod(nb)), dtype=np.int64)
    n = len(pos)
    for i in range(n):
        d = pos[i + 1:] - pos[i]
        d -= np.round(d / box) * box
        r2 = (d ** 2).sum(1)
        for jj in np.nonzero(r2 <= rc * rc)[0]:
            j = i + 1 + jj
            db = b[j] - b[i]
            db = (db + nb // 2) % nb - nb // 2      # periodic, |d| <= 2
            key = tuple(int(v) for v in db)
            if key in rows:
                home = b[i] - rows[key]
            else:
                home = b[j] - rows[tuple(-v for v in key)]
            home %= nb
            counts[home[0] + nb[0] * (home[1] + nb[1] * home[2])] += 1
    return counts


def fcc_lattice_pe_per_atom(rho, rc, eps=1.0, sig=1.0):
    """Half the sum of 4 eps((s/R)^12 - (s/R)^6) over FCC lattice vectors."""
    a = (4.0 / rho) ** (1.0 / 3.0)
    m = int(np.ceil(2 * rc / a)) + 1
    e = 0.0
    for i, j, k in itertools.product(range(-2 * m, 2 * m + 1), repeat=3):
        if (i + j + k) % 2 or (i == j == k == 0):
            continue
        r2 = (a / 2) ** 2 * (i * i + j * j + k * k)
        if r2 <= rc * rc:
            s6 = (sig * sig / r2) ** 3
            e += 4 * eps * (s6 * s6 - s6)
    return 0.5 * e


def perturbed_fcc(nc, rho, amp, seed):
    pos, L = fcc(nc, rho)
    rng = np.random.default_rng(seed)
    pos = np.mod(pos + rng.uniform(-amp, amp, pos.shape), L)
    pos[pos >= L] = 0.0
    return pos, L


# Kob-Andersen binary mixture (A 80%, B 20%): eps, sigma and cutoffs 2.5 sigma
KA_EPS = [[1.0, 1.5], [1.5, 0.5]]
KA_SIG = [[1.0, 0.8], [0.8, 0.88]]
KA_RC = [[2.5, 2.0], [2.0, 2.2]]


def species_brute_force(pos, box, types, eps, sig, rc):
    """All-pairs minimum-image LJ with per-species-pair parameters:
    (forces, pe, n_cut).  Pair (i, j) uses eps[ti][tj], sig[ti][tj] and
    applies the force when r^2 <= rc[ti][tj]^2."""
    pos = np.asarray(pos, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    types = np.asarray(types)
    E, S, R = (np.asarray(x, dtype=np.float64) for x in (eps, sig, rc))
    n = len(pos)
    f = np.zeros((n, 3))
    pe = 0.0
    ncut = 0
    for i in range(n):
        d = pos - pos[i]
        d -= np.round(d / box) * box
        r2 = (d ** 2).sum(1)
        r2[i] = np.inf
        e, s, c = E[types[i], types], S[types[i], types], R[types[i], types]
        m = r2 <= c * c
        r2m = np.where(m, r2, 1.0)
        sr6 = np.where(m, (s * s / r2m) ** 3, 0.0)
        dg = np.where(m, -24 * e * (2 * sr6 * sr6 - sr6) / r2m, 0.0)
        f[i] += (dg[:, None] * d).sum(0)
        pe += 0.5 * np.where(m, 4 * e * (sr6 * sr6 - sr6), 0.0).sum()
        ncut += int(m.sum())
    return f, pe, ncut // 2


def binary_types(n, frac_b, seed):
    """Species 0/1 with round(frac_b n) atoms of species 1 at random."""
    rng = np.random.default_rng(seed)
    t = np.zeros(n, dtype=np.int32)
    t[rng.choice(n, int(round(frac_b * n)), replace=False)] = 1
    return t


def species_shift_sum(pos, box, types, eps, sig, rc):
    """Sum over in-cutoff pairs of V(rc) of the pair's own species row,
    V(rc) = 4 eps ((sig/rc)^12 - (sig/rc)^6)."""
    pos = np.asarray(pos, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    types = np.asarray(types)
    E, S, R = (np.asarray(x, dtype=np.float64) for x in (eps, sig, rc))
    V = 4 * E * ((S / R) ** 12 - (S / R) ** 6)
    tot = 0.0
    for i in range(len(pos) - 1):
        d = pos[i + 1:] - pos[i]
        d -= np.round(d / box) * box
        r2 = (d ** 2).sum(1)
        c = R[types[i], types[i + 1:]]
        m = r2 <= c * c
        tot += V[types[i], types[i + 1:]][m].sum()
    return tot


def lj_table(npoints, r_lo, rc, eps=1.0, sig=1.0):
    """The LJ pair sampled for a tabulated pair function: (r2_lo, g, v) at
    npoints uniform in r^2 from r_lo^2 to rc^2, g = f12 / d."""
    r2 = np.linspace(r_lo * r_lo, rc * rc, npoints)
    s6 = (sig * sig / r2) ** 3
    g = -24.0 * eps * (2.0 * s6 * s6 - s6) / r2
    v = 4.0 * eps * (s6 * s6 - s6)
    return r_lo * r_lo, g, v


def _quartic_slopes(y):
    """Slope per spacing at every sample: the derivative of the quartic
    through the five nearest samples (centred, or the first or last five at
    the ends)."""
    n = len(y)
    m = np.zeros(n)
    for k in range(n):
        a = min(max(k - 2, 0), n - 5)
        c = np.polyfit(np.arange(5.0) - (k - a), y[a:a + 5], 4)
        m[k] = np.polyval(np.polyder(c), 0.0)
    return m


def table_interpolate(y, u, m=None):
    """A table's value at fractional index u: on each interval the cubic
    Hermite interpolant of the two samples and their slopes m (the Hermite
    basis functions; m from _quartic_slopes when not given), or the linear
    one below five samples."""
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    k = np.minimum(np.asarray(u).astype(np.int64), n - 2)
    t = u - k
    if n < 5:
        return y[k] + t * (y[k + 1] - y[k])
    if m is None:
        m = _quartic_slopes(y)
    h00, h10 = 2 * t ** 3 - 3 * t ** 2 + 1, t ** 3 - 2 * t ** 2 + t
    h01, h11 = -2 * t ** 3 + 3 * t ** 2, t ** 3 - t ** 2
    return h00 * y[k] + h10 * m[k] + h01 * y[k + 1] + h11 * m[k + 1]


def table_brute_force(pos, box, rc, table):
    """All-pairs minimum-image sum of a tabulated pair function, cubic
    Hermite in r^2 between the points: (forces, pe, n_cut)."""
    r2_lo, g, v = table
    g, v = np.asarray(g, dtype=np.float64), np.asarray(v, dtype=np.float64)
    n_pts = len(g)
    inv_h = (n_pts - 1) / (rc * rc - r2_lo)
    mg = _quartic_slopes(g) if n_pts >= 5 else None
    mv = _quartic_slopes(v) if n_pts >= 5 else None
    pos = np.asarray(pos, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64).reshape(3)
    n = len(pos)
    f = np.zeros((n, 3))
    pe = 0.0
    ncut = 0
    for i in range(n):
        d = pos - pos[i]
        d -= np.round(d / box) * box
        r2 = (d ** 2).sum(1)
        r2[i] = np.inf
        m = r2 <= rc * rc
        u = np.clip((r2[m] - r2_lo) * inv_h, 0.0, None)
        gg = table_interpolate(g, u, mg)
        f[i] += (gg[:, None] * d[m]).sum(0)
        pe += 0.5 * table_interpolate(v, u, mv).sum()
        ncut += int(m.sum())
    return f, pe, ncut // 2


def candidate_pairs(pos, box, rc):
    """P_cand of the midpoint stencil from block occupancies alone:
    sum_b n_b (n_b - 1) / 2 + sum over the 62 offset classes d of the table
    of sum_b n_b n_{b+d} (periodic)."""
    box = np.asarray(box, dtype=np.float64)
    bl = 0.5 * rc
    nb = (box / bl).astype(np.int64)
    b = block_of(pos, bl, nb)
    occ = np.zeros((nb[2], nb[1], nb[0]), dtype=np.int64)
    np.add.at(occ, (b[:, 2], b[:, 1], b[:, 0]), 1)
    cand = int((occ * (occ - 1) // 2).sum())
    for n1, n2 in TABLE[1:]:
        d = np.array(stencil(n2)) - np.array(stencil(n1))
        cand += int((occ * np.roll(occ, shift=(-d[2], -d[1], -d[0]),
                                   axis=(0, 1, 2))).sum())
    return cand


def slab_system(E, natoms, seed, min_sep=0.8, temperature=0.85):
    """The non-uniform density workload at a given size: box 3:1:1, the
    central third of x at density 0.80, the rest at 0.05, uniform with a
    minimum separation (random sequential addition), Maxwell-Boltzmann
    velocities at T*: (pos, vel, box)."""
    vol = natoms / (0.80 / 3 + 0.05 * 2 / 3)
    ly = (vol / 3.0) ** (1.0 / 3.0)
    box = np.array([3 * ly, ly, ly])
    pos, vel = E.generate_slab(box, ly, 2 * ly, 0.80, 0.05, min_sep=min_sep,
                               temperature=temperature, seed=seed)
    return pos, vel, box


def pair_force_sums(pos, box, rc, eps=1.0, sig=1.0):
    """Per atom, the sum of the magnitudes of its LJ pair forces within rc
    (the scale an FP32 evaluation's rounding is relative to), and the
    smallest pair distance.  Periodic minimum image (rc < L/2)."""
    from scipy.spatial import cKDTree
    box = np.asarray(box, dtype=np.float64)
    p = np.mod(np.asarray(pos, dtype=np.float64), box)
    p = np.where(p >= box, 0.0, p)
    tree = cKDTree(p, boxsize=box)
    pairs = tree.query_pairs(rc, output_type="ndarray")
    d = p[pairs[:, 1]] - p[pairs[:, 0]]
    d -= np.round(d / box) * box
    r2 = (d * d).sum(1)
    s6 = (sig * sig / r2) ** 3
    fmag = np.abs(24.0 * eps * (2.0 * s6 * s6 - s6)) / np.sqrt(r2)
    out = np.zeros(len(p))
    np.add.at(out, pairs[:, 0], fmag)
    np.add.at(out, pairs[:, 1], fmag)
    return out, float(np.sqrt(r2.min()))


def region_bound(cuts, bl):
    """A bound on |p| over every cell-block region of a map (coordinates
    relative to the cell-block centre): half the widest cell block plus two
    blocks per axis."""
    return float(np.sqrt(sum(((0.5 * np.diff(c).max() + 2.0) * bl) ** 2
                             for c in cuts)))


def mixed_error_bound(f, ref, pos, box, rc, R, ncross=0, eps=1.0, sig=1.0):
    """Largest per-atom error of FP32 pair forces relative to the bound the
    arithmetic implies: FP32 coordinates relative to a cell-block centre,
    with |p| <= R, put up to 2^-24 R into each difference, which moves a
    pair force that goes as r^-13 by about 13 * 2^-24 R / r of itself, so
    an atom's error is at most that times the sum of its pair forces'
    magnitudes; and each of the ncross pairs that FP32 r^2 puts on the
    other side of the cutoff than FP64 adds or drops the force at the
    cutoff, |F(rc)| (0.039 at rc = 2.5), on its two atoms.  Returns
    max_i |dF_i| / bound_i."""
    s, rmin = pair_force_sums(pos, box, rc, eps, sig)
    err = np.sqrt(((np.asarray(f) - ref) ** 2).sum(1))
    s6 = (sig / rc) ** 6
    fc = abs(24.0 * eps * (2.0 * s6 * s6 - s6) / rc)
    bound = (13.0 * 2.0 ** -24 * R / rmin * np.maximum(s, 1e-30) +
             ncross * 1.01 * fc)
    return float((err / bound).max())
