"""Writes tests/golden/kats.json: known-answer vectors from exact arithmetic.

The reference ships no test vectors and cannot be compiled here, so these
KATs are computed from the reference's formulas in exact rational
arithmetic (fractions.Fraction, correctly rounded once per IEEE operation),
independently of the floating-point hardware and of the oracle:

* exaMD_identify_block (SRC/exaMD_module.f90:109-122):
  Int(fl(fl(x / bl) + 1)), clamped to nblocks_xyz, blocknum x fastest;
  cases on block faces, one ulp either side, in the enlarged last block,
  and out of range (109).
* lennardjones_12 (user_func.f90:40-75) plus the pair energy
  4 eps (d12 - d6): exact values rounded once to double.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
import os
from fractions import Fraction as F

HERE = os.path.dirname(os.path.abspath(__file__))


def rn(q: F) -> float:
    """Round an exact rational to the nearest double (ties to even)."""
    return float(q)


def identify(box, bl, pos):
    nb = [int(rn(F(b) / F(bl))) for b in box]
    bxyz = []
    for x, n in zip(pos, nb):
        q = rn(F(x) / F(bl))          # fl(x / bl)
        q = rn(F(q) + 1)              # fl(. + 1)
        b = math.trunc(q)             # Int()
        bxyz.append(min(b, n))
    num = bxyz[0] + (bxyz[1] - 1) * nb[0] + (bxyz[2] - 1) * nb[0] * nb[1]
    rc = 0 if 1 <= num <= nb[0] * nb[1] * nb[2] else 109
    return rc, bxyz, num


def lj_exact(dx, dy, dz, d2, rc, eps, sig):
    if F(d2) > F(rc) * F(rc):
        return 1, None, None
    s6 = (F(sig) ** 2 / F(d2)) ** 3
    dg = -24 * F(eps) * (2 * s6 * s6 - s6) / F(d2)
    f = [rn(dg * F(v)) for v in (dx, dy, dz)]
    e = rn(4 * F(eps) * (s6 * s6 - s6))
    return 0, f, e


def main():
    kats = {"identify_block": [], "lj": []}
    boxes = [([13.4, 13.4, 13.4], 1.25), ([33.591871, 33.591871, 33.591871], 1.25),
             ([7.3, 9.1, 8.8], 0.3)]
    for box, bl in boxes:
        nb = int(box[0] / bl)
        xs = []
        for k in (0, 1, 2, nb - 1, nb):
            x = k * bl
            xs += [x, math.nextafter(x, math.inf), math.nextafter(x, -math.inf)]
        xs += [box[0] - 1e-9, box[0] * 0.5, -0.1 * bl, -1.5 * bl]
        for i, x in enumerate(xs):
            if not (x > -2 * bl):
                continue
            pos = [x, xs[(i + 3) % len(xs)] % box[1], abs(xs[(i + 5) % len(xs)]) % box[2]]
            rc, bxyz, num = identify(box, bl, pos)
            kats["identify_block"].append(
                {"box": box, "bl": bl, "pos": pos, "rc": rc,
                 "blockxyz": bxyz, "blocknum": num})
    for (dx, dy, dz, rc, eps, sig) in [
            (1.0, 0.0, 0.0, 2.5, 1.0, 1.0), (0.9, 0.3, -0.2, 2.5, 1.0, 1.0),
            (2 ** (1 / 6), 0.0, 0.0, 2.5, 1.0, 1.0), (-1.1, 1.4, 0.6, 2.5, 0.8, 1.1),
            (2.5, 0.0, 0.0, 2.5, 1.0, 1.0), (1.7, -1.8, 0.2, 2.5, 1.0, 1.0),
            (0.5, 0.5, 0.5, 3.0, 2.0, 0.9)]:
        d2 = rn(F(dx) ** 2)
        d2 = rn(F(d2) + F(rn(F(dy) ** 2)))
        d2 = rn(F(d2) + F(rn(F(dz) ** 2)))
        rv, f, e = lj_exact(dx, dy, dz, d2, rc, eps, sig)
        kats["lj"].append({"args": [dx, dy, dz, d2, rc, eps, sig],
                           "retval": rv, "f12": f, "e12": e})
    with open(os.path.join(HERE, "kats.json"), "w") as fh:
        json.dump(kats, fh, indent=1)


if __name__ == "__main__":
    main()
