"""Writes the golden vectors under tests/golden/ (run from the repo root).

entropy_vectors.json   -I(F;C) for count tuples (w, x, y, z), evaluated with
                       mpmath at 60 digits from the definition
                       H(F|C) - H(F) (entropy_inline.h:58-79 semantics: 0 when
                       y+z == 0, guarded terms vanish).
fisher_vectors.json    log10 of the two-sided Fisher p for the table
                       [[w, y-w], [x, z-x]], from exact rationals: the sum of
                       P(j) over every j with P(j) <= P(obs) * (1 + 1e-7).
fisher_ref_vectors.json  the reference's own fisher_pval (src/misc/fisher_p.c),
                       compiled unmodified by oracle/Makefile into
                       oracle/_ref/libfisher_ref.so, evaluated on a grid of
                       tables; float32 results stored as bit patterns.

None of these depend on the oracle's or the product's code.
"""
import ctypes
import json
import os
import random
import struct
import sys
from fractions import Fraction
from math import comb

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
mp.mp.dps = 60


def L(t):
    t = mp.mpf(t)
    return t * mp.log(t, 2) if t > 0 else mp.mpf(0)


def entropy_exact(w, x, y, z):
    m = y + z
    if m == 0:
        return mp.mpf(0)
    xx = w + x
    num = (L(y) + L(z) + L(xx) + L(m - xx)
           - L(w) - L(y - w) - L(x) - L(z - x) - L(m))
    return num / m


def fisher_exact_log10(w, x, y, z):
    a, b, c, d = w, y - w, x, z - x
    r0, c0, c1 = a + b, a + c, b + d
    m = a + b + c + d
    if m == 0:
        return mp.mpf(0)
    lo, hi = max(0, r0 - c1), min(r0, c0)
    den = comb(m, r0)
    probs = {j: Fraction(comb(c0, j) * comb(c1, r0 - j), den) for j in range(lo, hi + 1)}
    thr = probs[a] * (Fraction(1) + Fraction(1, 10**7))
    p = sum((v for v in probs.values() if v <= thr), Fraction(0))
    if p == 1:
        return mp.mpf(0)
    return mp.log10(mp.mpf(p.numerator) / mp.mpf(p.denominator))


def tuples(rng, big, nrand):
    out = []
    for y in range(7):
        for z in range(7):
            for w in range(y + 1):
                for x in range(z + 1):
                    out.append((w, x, y, z))
    for _ in range(nrand):
        y = rng.randint(0, big)
        z = rng.randint(0, big)
        out.append((rng.randint(0, y), rng.randint(0, z), y, z))
    # near-independence and extreme tables at realistic sizes
    for y, z in ((5734, 2458), (1024, 1024), (128, 128), (45, 44)):
        for fw, fx in ((0.5, 0.5), (0.3, 0.31), (0.01, 0.02), (0.0, 1.0), (1.0, 0.0), (0.2, 0.2)):
            out.append((int(fw * y), int(fx * z), y, z))
    return out


def main():
    rng = random.Random(20121220)
    ent = [(w, x, y, z, mp.nstr(entropy_exact(w, x, y, z), 30))
           for (w, x, y, z) in tuples(rng, 8192, 400)]
    with open(os.path.join(HERE, "entropy_vectors.json"), "w") as f:
        json.dump({"fields": ["w", "x", "y", "z", "neg_mi_bits"], "rows": ent}, f)
    fis = [(w, x, y, z, mp.nstr(fisher_exact_log10(w, x, y, z), 30))
           for (w, x, y, z) in tuples(rng, 600, 150)]
    with open(os.path.join(HERE, "fisher_vectors.json"), "w") as f:
        json.dump({"fields": ["w", "x", "y", "z", "log10_p"], "rows": fis}, f)

    ref = os.path.join(ROOT, "oracle", "_ref", "libfisher_ref.so")
    if not os.path.exists(ref):
        print("oracle/_ref/libfisher_ref.so missing; run make -C oracle ref", file=sys.stderr)
        return
    lib = ctypes.CDLL(ref)
    lib.fisher_pval.restype = ctypes.c_float
    lib.fisher_pval.argtypes = [ctypes.c_int] * 4
    grid = [(a, b, c, d) for a in range(6) for b in range(6) for c in range(6) for d in range(6)]
    for _ in range(400):
        n = rng.choice((50, 300, 2000, 8192))
        grid.append(tuple(rng.randint(0, n // 2) for _ in range(4)))
    rows = []
    for t in grid:
        v = lib.fisher_pval(*t)
        rows.append(list(t) + [struct.unpack("<I", struct.pack("<f", v))[0]])
    with open(os.path.join(HERE, "fisher_ref_vectors.json"), "w") as f:
        json.dump({"fields": ["n11", "n12", "n21", "n22", "f32_bits"], "rows": rows}, f)


if __name__ == "__main__":
    main()
