"""The inequalities behind the count kernels' fused filters (CPU, exact math).

The device filters reject an evaluation only when a bound proves it cannot
reach the top-k; a wrong bound would silently drop results.  These checks
evaluate each inequality exactly (lgamma / log in double, far from the
margins involved) over exhaustive small tables and random large ones:

  I_nats <= ln(1 + chi2/m)                                  (Jensen on KL)
  e^{n H(k/n)} / sqrt(2n) <= C(n, k) <= e^{n H(k/n)}        (H in nats)
  C(n, k) <= e^{n H(k/n)} sqrt(n / (2 pi k (n-k)))          (1 <= k <= n-1)
  ln P(obs) >= -m I_nats - ln(4 y z) / 2 + g(m)             (Fisher bound)
  g(m) = ln(2 pi (m-1) / m) / 2 >= 0.9189 - 1/m             (1 <= xx <= m-1)
  expm1(T) >= T (1 + T/2)                                   (T >= 0)

with P(obs) = C(y,w) C(z,x) / C(m,xx) the hypergeometric probability of the
observed 2x2 table (w of y controls and x of z cases carry the function).
"""
import math

import numpy as np


def H(p):
    return 0.0 if p <= 0.0 or p >= 1.0 else -(p * math.log(p) + (1 - p) * math.log(1 - p))


def lchoose(n, k):
    return math.lgamma(n + 1) - math.lgamma(k + 1) - math.lgamma(n - k + 1)


def mi_nats(w, x, y, z):
    m = y + z
    xx = w + x
    return H(xx / m) - (y / m) * H(w / y if y else 0.0) - (z / m) * H(x / z if z else 0.0)


def chi2_over_m(w, x, y, z):
    m = y + z
    xx = w + x
    q = xx * (m - xx)
    if q == 0 or y == 0 or z == 0:
        return 0.0
    d = w * z - x * y
    return d * d / (y * z * q)


def tables(rng, count, nmax):
    for _ in range(count):
        y = int(rng.integers(0, nmax))
        z = int(rng.integers(0, nmax))
        if y + z == 0:
            continue
        yield int(rng.integers(0, y + 1)), int(rng.integers(0, z + 1)), y, z


def all_small_tables(nmax):
    for y in range(nmax + 1):
        for z in range(nmax + 1):
            if y + z == 0:
                continue
            for w in range(y + 1):
                for x in range(z + 1):
                    yield w, x, y, z


def test_binomial_entropy_bounds_exhaustive():
    for n in list(range(1, 300)) + [1000, 4097, 8192, 46340]:
        ks = range(n + 1) if n < 300 else np.linspace(0, n, 97).astype(int)
        for k in ks:
            lc = lchoose(n, int(k))
            nh = n * H(int(k) / n)
            assert lc <= nh + 1e-9 * max(1.0, nh), (n, k)
            assert lc >= nh - 0.5 * math.log(2 * n) - 1e-9 * max(1.0, nh), (n, k)
            if 0 < k < n:
                k = int(k)
                ub = nh + 0.5 * math.log(n / (2 * math.pi * k * (n - k)))
                assert lc <= ub + 1e-9 * max(1.0, nh), (n, k)


def test_jensen_chi2_bound():
    rng = np.random.default_rng(7)
    checked = 0
    for w, x, y, z in list(all_small_tables(14)) + list(tables(rng, 20000, 5000)):
        i = mi_nats(w, x, y, z)
        b = math.log1p(chi2_over_m(w, x, y, z))
        assert i <= b + 1e-12 * max(1.0, b), (w, x, y, z, i, b)
        checked += 1
    assert checked > 30000


def test_fisher_pobs_lower_bound():
    rng = np.random.default_rng(11)
    for w, x, y, z in list(all_small_tables(14)) + list(tables(rng, 20000, 5000)):
        m = y + z
        xx = w + x
        lp = lchoose(y, w) + lchoose(z, x) - lchoose(m, xx)
        bound = -m * mi_nats(w, x, y, z) - 0.5 * math.log(4 * max(y, 1) * max(z, 1))
        assert lp >= bound - 1e-9 * max(1.0, abs(lp)), (w, x, y, z, lp, bound)
        if 0 < xx < m:
            g = 0.5 * math.log(2 * math.pi * (m - 1) / m)
            assert g >= 0.9189 - 1.0 / m
            assert lp >= bound + g - 1e-9 * max(1.0, abs(lp)), (w, x, y, z, lp, bound + g)


def test_fisher_chain_rejects_only_unreachable_tables():
    """The chain the kernel applies: for a threshold thr on log10 p and class
    sizes n0 >= y, n1 >= z, a table with D^2 < T (1 + T/2) y z q and
    T = (-(thr + 1e-6) ln 10 - ln(4 n0 n1)/2 + 0.9189 - 1/m) / m has
    log10 P(obs) - 1e-6 > thr, so its p-value (>= P(obs)) misses thr.
    (Tables with xx = 0 or m have P(obs) = 1 and q = 0.)"""
    rng = np.random.default_rng(13)
    rejected = 0
    for t in rng.random(1000) * 3:
        assert math.expm1(t) >= t * (1 + t / 2)
    for thr in (-2.0, -5.0, -12.0):
        for w, x, y, z in tables(rng, 20000, 3000):
            m = y + z
            xx = w + x
            n0 = y + int(rng.integers(0, 3)) * int(rng.integers(0, 50))
            n1 = z + int(rng.integers(0, 3)) * int(rng.integers(0, 50))
            T = (-(thr + 1e-6) * math.log(10) - 0.5 * math.log(4 * max(n0, 1) * max(n1, 1))
                 + 0.9189 - 1.0 / m) / m
            if T <= 0:
                continue
            d = w * z - x * y
            q = xx * (m - xx)
            if q == 0:
                continue
            if d * d < T * (1 + T / 2) * y * z * q:
                rejected += 1
                lp10 = (lchoose(y, w) + lchoose(z, x) - lchoose(m, xx)) / math.log(10)
                assert lp10 - 1e-6 > thr, (thr, w, x, y, z, lp10)
    assert rejected > 1000


def test_entropy_chain_rejects_only_unreachable_tables():
    """Entropy objective -I_bits <= thr: rejection when D^2 < (2^-thr - 1) y z q
    implies I_bits < -thr."""
    rng = np.random.default_rng(17)
    rejected = 0
    for thr in (-0.001, -0.01, -0.1):
        k = 2.0 ** (-thr) - 1.0
        for w, x, y, z in tables(rng, 20000, 3000):
            m = y + z
            xx = w + x
            d = w * z - x * y
            if d * d < k * y * z * xx * (m - xx):
                rejected += 1
                assert mi_nats(w, x, y, z) / math.log(2) < -thr, (thr, w, x, y, z)
    assert rejected > 1000
