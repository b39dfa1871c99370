"""The scan at the full large size (100,000 variants x 8,192 samples, 1e11
evaluations): the entropy top-10,000 against the committed result of the
oracle's gold scan of the whole panel, and size-independent properties:

- every returned record's counts, recomputed with numpy popcounts from the
  panel, and its objective, recomputed by the oracle, are bit-identical;
- the records are in strictly increasing canonical key order;
- completeness on a variant subset holding both variants of many top pairs:
  every evaluation of the subset that the oracle ranks at or below the k-th
  returned key is among the returned records, and every returned record
  inside the subset is in the oracle's top-k of the subset;
- the three count paths return identical lists.
"""
import hashlib
import os

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                    "large_entropy_top10000.npz")


def order_key(obj):
    """The canonical score bits (ascending order of doubles as u64)."""
    b = np.asarray(obj, np.float64).view(np.uint64)
    return np.where(b >> np.uint64(63), ~b, b | np.uint64(1 << 63))


def keys(res):
    idx = (res["sA"].astype(np.uint64) << np.uint64(32)) | (res["sB"].astype(np.uint64) << np.uint64(3)) \
        | res["fun"].astype(np.uint64)
    return order_key(res["obj"]), idx


@pytest.fixture(scope="module")
def large():
    print("case large panel generation", flush=True)
    p = synth.make_panel("large")
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    panel = ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"])
    bits, valid = panel.download()
    bed_sha = hashlib.sha256(np.ascontiguousarray(p["payload"]).tobytes()).hexdigest()
    pheno_sha = hashlib.sha256(np.asarray(p["pheno"], np.int64).tobytes()).hexdigest()
    yield dict(panel=panel, m0=m0, m1=m1, bits=bits, valid=valid, n=p["n_samples"],
               bed_sha=bed_sha, pheno_sha=pheno_sha)
    panel.close()


def recount(bits, valid, m0, m1, res):
    """(w, x, y, z) of each record from the planes (reference semantics)."""
    a, b = res["sA"], res["sB"]
    out = np.zeros((len(res), 4), np.int64)
    for lo in range(0, len(res), 2048):
        sl = slice(lo, lo + 2048)
        A, B = bits[a[sl]], bits[b[sl]]
        v = valid[a[sl]] & valid[b[sl]]
        f = res["fun"][sl][:, None]
        F = np.where(f == 0, A & B, np.where(f == 1, ~A & B, np.where(f == 2, A & ~B,
                     np.where(f == 3, ~(A | B), A ^ B)))) & v
        for c, m in ((0, m0), (1, m1)):
            out[sl, c] = np.bitwise_count(F & m).sum(axis=1)
            out[sl, 2 + c] = np.bitwise_count(v & m).sum(axis=1)
    return out


def check_records(d, res, k, ffun):
    assert len(res) == k
    hi, lo = keys(res)
    assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))), "key order"
    cnt = recount(d["bits"], d["valid"], d["m0"], d["m1"], res)
    np.testing.assert_array_equal(cnt, np.stack([res["w"], res["x"], res["y"], res["z"]], axis=1))
    score = orc.entropy_exact if ffun == 0 else orc.fisher_rank
    exact = np.array([score(int(w), int(x), int(y), int(z)) for w, x, y, z in cnt])
    np.testing.assert_array_equal(exact.view(np.uint64), res["obj"].view(np.uint64))


def check_subset_completeness(d, res, k, ffun, n_top_vars=600, n_rand=600, seed=5):
    """Oracle gold scan of a variant subset (both variants of many top pairs
    plus random ones), mapped back to panel rows."""
    rng = np.random.default_rng(seed)
    m = n_top_vars // 2  # both variants of the first m records
    top_vars = np.unique(np.concatenate([res["sA"][:m] // 2, res["sB"][:m] // 2]))
    n_var = d["bits"].shape[0] // 2
    rand_vars = rng.choice(n_var, size=n_rand, replace=False)
    vars_ = np.unique(np.concatenate([top_vars, rand_vars]))
    rows = np.stack([2 * vars_, 2 * vars_ + 1], axis=1).ravel()
    sub = orc.Panel(d["bits"][rows], d["valid"][rows], d["m0"], d["m1"], d["n"])
    ref = orc.gold_scan(sub, k, ffun=ffun)
    # map subset rows to panel rows (row order is preserved, so sA >= sB holds)
    ref_sA, ref_sB = rows[ref["sA"]], rows[ref["sB"]]
    got = {(int(a), int(b), int(f)) for a, b, f in zip(res["sA"], res["sB"], res["fun"])}
    kth_hi, kth_lo = [x[-1] for x in keys(res)]
    rh = order_key(ref["obj"])
    rl = (ref_sA.astype(np.uint64) << np.uint64(32)) | (ref_sB.astype(np.uint64) << np.uint64(3)) \
        | ref["fun"].astype(np.uint64)
    inside = (rh < kth_hi) | ((rh == kth_hi) & (rl <= kth_lo))
    assert inside.sum() > 50, "subset holds too few top records to test"
    for a, b, f in zip(ref_sA[inside], ref_sB[inside], ref["fun"][inside]):
        assert (int(a), int(b), int(f)) in got, ("missed", a, b, f)
    in_sub = set(rows.tolist())
    refset = {(int(a), int(b), int(f)) for a, b, f in zip(ref_sA, ref_sB, ref["fun"])}
    for a, b, f in got:
        if a in in_sub and b in in_sub:
            assert (a, b, f) in refset, ("not in the subset's top-k", a, b, f)


def test_large_entropy_top10000(large):
    print("case large entropy top-10000", flush=True)
    res, st = large["panel"].scan(large["m0"], large["m1"], topmost=10000)
    assert st["count_path"] == 3
    check_records(large, res, 10000, 0)
    check_subset_completeness(large, res, 10000, 0)


def test_large_entropy_top10000_equals_gold(large):
    """The whole large scan against the oracle's gold scan of all 1e11
    evaluations (tests/golden/make_large_gold.py): the same 10,000 records,
    record by record, in the same order, with the same counts and score bits."""
    print("case large entropy top-10000 vs the full gold scan", flush=True)
    g = np.load(GOLD)
    assert str(g["bed_sha256"]) == large["bed_sha"], "regenerated panel differs from the gold's"
    assert str(g["pheno_sha256"]) == large["pheno_sha"], "regenerated phenotypes differ"
    k = int(g["topk"])
    res, st = large["panel"].scan(large["m0"], large["m1"], topmost=k)
    assert len(res) == k
    for f in ("sA", "sB", "fun", "w", "x", "y", "z"):
        if not np.array_equal(res[f], g[f].astype(res[f].dtype)):
            i = int(np.flatnonzero(res[f] != g[f].astype(res[f].dtype))[0])
            raise AssertionError("field %s differs first at rank %d: got %r, gold %r" % (
                f, i, res[i], {h: g[h][i] for h in ("sA", "sB", "fun", "w", "x", "y", "z")}))
    np.testing.assert_array_equal(res["obj"].view(np.uint64), g["obj_bits"])


def test_large_count_paths_agree(large):
    print("case large paths agree", flush=True)
    ref, _ = large["panel"].scan(large["m0"], large["m1"], topmost=10000, count_path=3)
    for path in (1, 2, 4):
        got, st = large["panel"].scan(large["m0"], large["m1"], topmost=10000, count_path=path)
        assert st["count_path"] == path
        for f in ("sA", "sB", "fun", "w", "x", "y", "z"):
            np.testing.assert_array_equal(got[f], ref[f])
        np.testing.assert_array_equal(got["obj"].view(np.uint64), ref["obj"].view(np.uint64))


def test_large_fisher_top1000(large):
    print("case large Fisher top-1000", flush=True)
    res, st = large["panel"].scan(large["m0"], large["m1"], topmost=1000,
                                  ffun=ge.GEPI_FFUNC_FISHERPV)
    assert st["count_path"] == 3
    check_records(large, res, 1000, 1)
    check_subset_completeness(large, res, 1000, 1, n_top_vars=500, n_rand=300)
    # the LDS-staged kernel agrees
    got, st = large["panel"].scan(large["m0"], large["m1"], topmost=1000,
                                  ffun=ge.GEPI_FFUNC_FISHERPV, count_path=4)
    assert st["count_path"] == 4
    for f in ("sA", "sB", "fun", "w", "x", "y", "z"):
        np.testing.assert_array_equal(got[f], res[f])
    np.testing.assert_array_equal(got["obj"].view(np.uint64), res["obj"].view(np.uint64))
