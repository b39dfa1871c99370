"""The many-scan workloads at their real sizes.

- The uneven mix (64 scans, variants logU[500, 50,000], samples logU[200,
  5,000], case fraction U[0.2, 0.8], top-100 each) in one batch of the
  engine: every item equals the same scan run alone; every returned record's
  counts (numpy popcounts of the panel) and score bits (oracle) are exact and
  the keys strictly increase; items small enough for the oracle's gold scan
  in a second or so equal it record by record.
- The serving loop under Poisson arrivals: 80% requests from that mix, 20%
  re-scans of a resident 50,000 x 4,000 panel with new phenotype vectors;
  every response equals the same scan run alone, re-scans pass the same
  record checks, small mix items equal the oracle.
"""
import threading
import time

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
from oracle import oracle as orc
from test_gpu_properties import keys, recount

pytestmark = pytest.mark.gpu

FIELDS = ("sA", "sB", "fun", "w", "x", "y", "z", "obj")
ORACLE_MAX_EVALS = 1e8          # gold-scan items up to this many evaluations (entropy)
ORACLE_MAX_EVALS_FISHER = 4e6   # Fisher's gold scan walks every table's support
BIG = dict(n_var=50000, n_samples=4000, n_case=2000, maf=("uniform", 0.01, 0.5), miss=0.01,
           planted=0, seed=1005)


def assert_same(got, ref, label=""):
    assert len(got) == len(ref), (label, len(got), len(ref))
    for k in FIELDS:
        if not np.array_equal(got[k], ref[k]):
            bad = int(np.flatnonzero(got[k] != ref[k])[0])
            raise AssertionError("%s: field %s differs first at rank %d: got %r ref %r" % (
                label, k, bad, got[bad], ref[bad]))


def check_records(bits, valid, m0, m1, res, k, ffun, label):
    assert len(res) == k, (label, len(res), k)
    hi, lo = keys(res)
    assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1]))), label
    cnt = recount(bits, valid, m0, m1, res)
    np.testing.assert_array_equal(cnt, np.stack([res["w"], res["x"], res["y"], res["z"]], axis=1),
                                  err_msg=label)
    score = orc.entropy_exact if ffun == 0 else orc.fisher_rank
    exact = np.array([score(int(w), int(x), int(y), int(z)) for w, x, y, z in cnt])
    np.testing.assert_array_equal(exact.view(np.uint64), res["obj"].view(np.uint64), err_msg=label)


def evals_of(n_var):
    R = 2 * n_var
    return 5 * R * (R + 1) // 2


@pytest.fixture(scope="module")
def mix():
    print("case mix generation (64 panels)", flush=True)
    out = []
    for sp in synth.mix_specs():
        p = synth.make_panel(spec=sp)
        m0, m1 = synth.masks_from_pheno(p["pheno"])
        h = ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"])
        out.append(dict(spec=sp, p=p, m0=m0, m1=m1, h=h))
    yield out
    for d in out:
        d["h"].close()


def planes_of(d):
    if "bits" not in d:
        d["bits"], d["valid"] = d["h"].download()
    return d["bits"], d["valid"]


@pytest.mark.parametrize("ffun", [0, 1])
def test_mix64_batched_equals_alone_oracle_and_records(mix, ffun):
    print("case mix64 batched ffun=%d" % ffun, flush=True)
    items = [dict(panel=d["h"], mask_ctrl=d["m0"], mask_case=d["m1"], topmost=d["spec"]["topk"],
                  ffun=ffun) for d in mix]
    with ge.Engine() as eng:
        res, bstats = eng.scan_batch(items)
    assert bstats["items_failed"] == 0
    n_oracle = 0
    for i, (d, (got, st, status)) in enumerate(zip(mix, res)):
        assert status == 0, (i, status)
        alone, _ = d["h"].scan(d["m0"], d["m1"], topmost=d["spec"]["topk"], ffun=ffun,
                               count_path=3)
        assert_same(got, alone, "mix item %d vs alone" % i)
        bits, valid = planes_of(d)
        check_records(bits, valid, d["m0"], d["m1"], got, d["spec"]["topk"], ffun,
                      "mix item %d" % i)
        limit = ORACLE_MAX_EVALS if ffun == 0 else ORACLE_MAX_EVALS_FISHER
        if evals_of(d["p"]["n_var"]) <= limit:
            ref = orc.gold_scan(orc.Panel(bits, valid, d["m0"], d["m1"], d["p"]["n_samples"]),
                                d["spec"]["topk"], ffun=ffun)
            assert_same(got, ref, "mix item %d vs oracle" % i)
            n_oracle += 1
    print("mix64 ffun=%d: 64 items equal alone, %d equal the oracle" % (ffun, n_oracle), flush=True)
    assert n_oracle >= (25 if ffun == 0 else 4)


def test_serve_mix_and_resident_rescans(mix):
    """Poisson arrivals at 40 QPS: 80 requests, 20% re-scans of the resident
    50,000 x 4,000 panel with new phenotypes, the rest scans of the mix."""
    print("case serve mix + resident re-scans", flush=True)
    big = synth.make_panel(spec=BIG)
    rng = np.random.Generator(np.random.PCG64(1006))
    n_req, qps = 80, 40.0
    arrivals = np.cumsum(rng.exponential(1.0 / qps, size=n_req))
    with ge.Server() as srv:
        ids = [srv.add_panel_bed(d["p"]["payload"], d["p"]["n_var"], d["p"]["n_samples"])
               for d in mix]
        big_id = srv.add_panel_bed(big["payload"], big["n_var"], big["n_samples"])
        reqs = []
        for i in range(n_req):
            if rng.random() < 0.2:
                n = BIG["n_samples"]
                pheno = synth.pheno_for(n, int(round(rng.uniform(0.2, 0.8) * n)), 2000 + i)
                m0, m1 = synth.masks_from_pheno(pheno)
                reqs.append(dict(pid=big_id, m0=m0, m1=m1, j=None))
            else:
                j = int(rng.integers(0, len(mix)))
                reqs.append(dict(pid=ids[j], m0=mix[j]["m0"], m1=mix[j]["m1"], j=j))
        tickets = [None] * n_req
        t0 = time.perf_counter()

        def submitter():
            for i, r in enumerate(reqs):
                dly = t0 + arrivals[i] - time.perf_counter()
                if dly > 0:
                    time.sleep(dly)
                tickets[i] = srv.submit(r["pid"], r["m0"], r["m1"], topmost=100)

        th = threading.Thread(target=submitter)
        th.start()
        got = [None] * n_req
        for i in range(n_req):
            while tickets[i] is None:
                time.sleep(0.0005)
            got[i] = srv.wait(tickets[i])
        th.join()
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] >= n_req
    assert st["max_active"] > 1, "no two requests were ever on the device together"
    n_rescan = 0
    with ge.Panel.from_bed(big["payload"], big["n_var"], big["n_samples"]) as hb:
        bb, bv = hb.download()
        for i, (r, (res, _, status)) in enumerate(zip(reqs, got)):
            assert status == 0, (i, status)
            if r["j"] is None:
                alone, _ = hb.scan(r["m0"], r["m1"], topmost=100, count_path=3)
                assert_same(res, alone, "re-scan %d vs alone" % i)
                check_records(bb, bv, r["m0"], r["m1"], res, 100, 0, "re-scan %d" % i)
                n_rescan += 1
            else:
                d = mix[r["j"]]
                alone, _ = d["h"].scan(r["m0"], r["m1"], topmost=100, count_path=3)
                assert_same(res, alone, "request %d (mix %d) vs alone" % (i, r["j"]))
                if evals_of(d["p"]["n_var"]) <= ORACLE_MAX_EVALS:
                    bits, valid = planes_of(d)
                    ref = orc.gold_scan(orc.Panel(bits, valid, r["m0"], r["m1"],
                                                  d["p"]["n_samples"]), 100)
                    assert_same(res, ref, "request %d vs oracle" % i)
    assert n_rescan >= 8
    print("serve: %d responses equal alone (%d re-scans), p50 %.1f ms p99 %.1f ms" % (
        n_req, n_rescan, st["p50_ms"], st["p99_ms"]), flush=True)


def test_fisher_wave_tile_parking_item59():
    """Fisher on the fp4 wave tiles parks flagged variant pairs in global
    regions that workgroups claim and reuse within a launch.  Regions are
    claimed per XCD: a region reused across XCDs read stale records when
    another XCD's L2 wrote its dirty lines back over it, which this scan
    (mix item 59: 30,286 x 4,966, 3,089 cases, top-100) showed on every run.
    Three scans on path 3 must equal the LDS-staged kernel's (path 4)."""
    print("case fisher wave-tile parking, mix item 59", flush=True)
    sp = synth.mix_specs()[59]
    p = synth.make_panel(spec=sp)
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    with ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"]) as h:
        ref, st = h.scan(m0, m1, topmost=100, ffun=1, count_path=4)
        assert st["count_path"] == 4
        for _ in range(3):
            got, st = h.scan(m0, m1, topmost=100, ffun=1)
            assert st["count_path"] == 3
            assert_same(got, ref, "item 59 path 3 vs path 4")
