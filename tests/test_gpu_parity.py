"""GPU parity: the HIP scan (through the C ABI) against the CPU oracle.

Counts, objective bits and top-k order must be identical.  Sizes are chosen so
that the oracle finishes in seconds; workloads at full size are covered by
size-independent properties in test_gpu_properties.py.
"""
import os
import subprocess

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")
FIELDS = ("sA", "sB", "fun", "w", "x", "y", "z", "obj")


def _keys(r):
    return [(int(a), int(b), int(f)) for a, b, f in zip(r["sA"], r["sB"], r["fun"])]


def mismatch_report(got, ref):
    """Which evaluations are missing, extra or repeated in got, and got's
    record for the reference's first differing evaluation."""
    gk, rk = _keys(got), _keys(ref)
    gs, rs = set(gk), set(rk)
    dup = sorted({k for k in gk if gk.count(k) > 1})[:5]
    lines = ["missing %s" % sorted(rs - gs)[:5], "extra %s" % sorted(gs - rs)[:5],
             "repeated %s" % dup]
    for i, k in enumerate(rk):
        if i >= len(gk) or gk[i] != k or any(got[f][i] != ref[f][i] for f in FIELDS):
            same = [got[j] for j in range(len(gk)) if gk[j] == k]
            lines.append("ref rank %d %r; got for that evaluation %r" % (i, ref[i], same))
            break
    return "; ".join(lines)


def assert_same(got, ref, label=""):
    assert len(got) == len(ref), (label, len(got), len(ref), mismatch_report(got, ref))
    for k in FIELDS:
        if not np.array_equal(got[k], ref[k]):
            bad = int(np.flatnonzero(got[k] != ref[k])[0])
            raise AssertionError("%s: field %s differs first at rank %d: got %r ref %r; %s" % (
                label, k, bad, got[bad], ref[bad], mismatch_report(got, ref)))


def random_panel(seed, n_var, n, miss, cls_p=(0.5, 0.45, 0.05)):
    rng = np.random.default_rng(seed)
    codes = orc.random_codes(rng, n_var, n, miss)
    cls = rng.choice([0, 1, -1], size=n, p=list(cls_p))
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    return bits, valid, m0, m1


PATHS = [1, 2, 3, 4]  # popcount (VALU), int8 MFMA, fp4 wave tiles, fp4 LDS-staged tiles


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("seed,n_var,n,miss", [
    (1, 1, 7, 0.0), (2, 5, 64, 0.0), (3, 33, 65, 0.05), (4, 40, 200, 0.1), (5, 70, 129, 0.02),
    (6, 32, 300, 0.0), (7, 130, 70, 0.03)])
def test_all_evaluations_entropy(seed, n_var, n, miss, path):
    print("case seed=%d n_var=%d n=%d miss=%g path=%d" % (seed, n_var, n, miss, path), flush=True)
    bits, valid, m0, m1 = random_panel(seed, n_var, n, miss)
    R = 2 * n_var
    total = 5 * R * (R + 1) // 2
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=total, count_path=path)
    assert st["count_path"] == path
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), total)
    assert_same(got, ref, "all-evals")
    assert st["evals"] == total


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("seed,n_var,n,miss,k", [
    (11, 150, 400, 0.02, 50), (12, 200, 700, 0.0, 500), (13, 97, 1000, 0.05, 7),
    (14, 260, 256, 0.01, 3000)])
def test_topk_entropy_threshold_active(seed, n_var, n, miss, k, path):
    print("case seed=%d n_var=%d n=%d k=%d path=%d" % (seed, n_var, n, k, path), flush=True)
    bits, valid, m0, m1 = random_panel(seed, n_var, n, miss)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=k, count_path=path)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), k)
    assert_same(got, ref, "topk")
    # the filter must have pruned most evaluations
    assert st["exact_evals"] < st["evals"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("seed,n_var,n,miss,k", [(21, 12, 90, 0.05, None), (22, 60, 300, 0.02, 40)])
def test_fisher_objective(seed, n_var, n, miss, k, path):
    print("case fisher seed=%d n_var=%d n=%d path=%d" % (seed, n_var, n, path), flush=True)
    bits, valid, m0, m1 = random_panel(seed, n_var, n, miss)
    R = 2 * n_var
    kk = k or 5 * R * (R + 1) // 2
    got, _ = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=kk, ffun=ge.GEPI_FFUNC_FISHERPV,
                              count_path=path)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), kk, ffun=1)
    assert_same(got, ref, "fisher")


@pytest.mark.parametrize("path", [3, 4])
@pytest.mark.parametrize("ffun", [0, 1])
@pytest.mark.parametrize("n0,n1", [(4096, 4096), (4200, 4200), (1100, 15000), (1000, 16000)])
def test_fp4_filter_count_types(n0, n1, ffun, path):
    """The fp4 path filters on its fp32 accumulators when every product of a
    class-0 and a class-1 count is exact in float (n0 n1 <= 2^24; 4096 x 4096
    is the largest balanced case, where products reach 2^24) and the samples
    number at most 16,384 (1100 + 15000 is near both limits), and on int32
    conversions otherwise (4200 x 4200; 1000 + 16000 samples).  Both must
    admit every evaluation the exact top-k needs."""
    print("case fp4 filter n0=%d n1=%d ffun=%d path=%d" % (n0, n1, ffun, path), flush=True)
    rng = np.random.default_rng(83)
    n = n0 + n1
    codes = orc.random_codes(rng, 48, n, 0.02)
    cls = rng.permutation(np.r_[np.zeros(n0, np.int64), np.ones(n1, np.int64)])
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=300, ffun=ffun, count_path=path)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), 300, ffun=ffun)
    assert st["n_class0"] == n0 and st["n_class1"] == n1 and st["count_path"] == path
    assert_same(got, ref, "fp4-filter")


@pytest.mark.parametrize("path", [3, 4])
@pytest.mark.parametrize("n0,n1", [(512, 1024), (640, 384), (768, 1), (1, 896), (1024, 0),
                                   (500, 1020), (130, 4), (3, 2), (384, 1152), (385, 383)])
def test_fp4_super_step_residues(n0, n1, path):
    """A super-step of the LDS-staged fp4 layout (path 4) holds 4 K steps
    of 128 samples, the fourth in the e2m1 sign bit with a per-row
    correction; one of the wave-tile layout (path 3) holds 3, and its
    accumulators run across the class boundary.  Each class starts a new
    super-step.  Class sizes cover 0, 1, 2 and 3 K steps past full
    super-steps of either layout in either class, exact multiples of 384
    and 512 samples, one or no sample in a class, and partial last K steps.
    Every evaluation must equal the oracle's."""
    print("case fp4 super-step n0=%d n1=%d path=%d" % (n0, n1, path), flush=True)
    rng = np.random.default_rng(91 + n0 + 7 * n1)
    n = n0 + n1
    codes = orc.random_codes(rng, 21, n, 0.03)
    cls = rng.permutation(np.r_[np.zeros(n0, np.int64), np.ones(n1, np.int64)])
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    R = bits.shape[0]
    k = 5 * R * (R + 1) // 2
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=k, count_path=path)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), k)
    assert st["n_class0"] == n0 and st["n_class1"] == n1 and st["count_path"] == path
    assert_same(got, ref, "fp4 super-step")


@pytest.mark.parametrize("path", [3, 4])
def test_fisher_parked_pairs_overflow(monkeypatch, path):
    """The fp4 kernels park variant pairs that pass Fisher's chi-square
    bound (the LDS-staged kernel, path 4, in LDS; the wave tiles, path 3, in
    a global region each workgroup claims) and finish them (function bits,
    Fisher's exact check) one per thread after the tile; with a loose
    threshold (large k, two-tile bands) a tile flags more pairs than the
    list holds and the rest are finished in the lane.  Together they must
    keep exactly the oracle's top-k."""
    print("case fisher parked pairs path=%d" % path, flush=True)
    bits, valid, m0, m1 = random_panel(57, 300, 400, 0.02)
    k = 50000
    monkeypatch.setenv("GEPI_BAND_TILES", "2")
    got, st = ge.process_gpu_1(bits, valid, m0, m1, 400, topmost=k,
                               ffun=ge.GEPI_FFUNC_FISHERPV, count_path=path)
    monkeypatch.delenv("GEPI_BAND_TILES")
    assert st["count_path"] == path
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, 400), k, ffun=1)
    assert st["bands"] > 2
    assert_same(got, ref, "fisher-parked")


@pytest.mark.parametrize("seed,n_var,n,k,band,cls_p", [
    (61, 150, 8500, 3000, None, (0.82, 0.14, 0.04)),    # float filter counts
    (62, 97, 8500, 40000, "3", (0.82, 0.14, 0.04)),     # loose threshold, odd 3-tile bands
    (63, 161, 12500, 2000, None, (0.5, 0.45, 0.05))])   # n0 n1 > 2^24: int32 filter counts
def test_entropy_parking_long_k(monkeypatch, seed, n_var, n, k, band, cls_p):
    """Entropy scans of at least 64 K steps of 128 samples run the wave-tile
    kernel built to park flagged variant pairs (a global region per
    workgroup, finished one per thread after the filter; past the region's
    capacity, in the lane).  Small panels (diagonal and edge tiles), a loose
    threshold with odd bands, and class sizes past the float filter's range
    must keep exactly the oracle's top-k."""
    print("case entropy parking seed=%d n=%d" % (seed, n), flush=True)
    bits, valid, m0, m1 = random_panel(seed, n_var, n, 0.02, cls_p)
    n0 = int(sum(bin(int(w)).count("1") for w in m0))
    n1 = int(sum(bin(int(w)).count("1") for w in m1))
    assert (n0 + 127) // 128 + (n1 + 127) // 128 >= 64
    if band:
        monkeypatch.setenv("GEPI_BAND_TILES", band)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=k)
    if band:
        monkeypatch.delenv("GEPI_BAND_TILES")
    assert st["count_path"] == 3
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), k)
    assert_same(got, ref, "entropy parking")


def test_kat_test_fixture_top5():
    d = ge.read_plink(os.path.join(DATA, "test"))
    got, _ = ge.process_gpu_1(d.bit_data, d.valid_data, d.mask_ctrl, d.mask_case, d.n_samples,
                              topmost=5)
    assert [(int(r["sA"]), int(r["sB"]), int(r["fun"])) for r in got] == \
        [(2, 0, 0), (2, 1, 1), (2, 2, 0), (2, 2, 3), (3, 0, 2)]
    assert np.all(got["obj"] == got["obj"][0])
    assert abs(got["obj"][0] + 0.4591479170272448) < 1e-15


@pytest.mark.parametrize("ffun", [0, 1])
def test_thinhap1_top1000(ffun):
    root = os.path.join(DATA, "thinhap1")
    d = ge.read_plink(root)
    with ge.Panel.from_bed(d.bed_payload, d.n_rows // 2, d.n_samples) as p:
        got, _ = p.scan(d.mask_ctrl, d.mask_case, topmost=1000, ffun=ffun)
    ref = orc.gold_scan(orc.read_plink(root), 1000, ffun=ffun)
    assert_same(got, ref, "thinhap1")


def test_small_workload_top100_both_objectives():
    p = synth.make_panel("small")
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    with ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"]) as panel:
        bits, valid = panel.download()
        op = orc.Panel(bits, valid, m0, m1, p["n_samples"])
        for ffun in (0, 1):
            ref = orc.gold_scan(op, 100, ffun=ffun)
            for path in PATHS:
                got, _ = panel.scan(m0, m1, topmost=100, ffun=ffun, count_path=path)
                assert_same(got, ref, "small ffun=%d path=%d" % (ffun, path))


def test_medium_subset_top1000():
    """Exact match vs the gold scan on a 2,000-variant subset of the medium workload."""
    p = synth.make_panel("medium", n_var=2000)
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    with ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"]) as panel:
        bits, valid = panel.download()
        ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, p["n_samples"]), 1000)
        for path in PATHS:
            got, st = panel.scan(m0, m1, topmost=1000, count_path=path)
            assert_same(got, ref, "medium-2000 path=%d" % path)
    # the planted XOR pairs dominate the ranking
    top_pairs = {(int(r["sA"]) // 2, int(r["sB"]) // 2) for r in got[:50]}
    planted = {(max(a, b), min(a, b)) for a, b in p["planted"]}
    assert len(top_pairs & planted) >= 3


def test_bed_decode_equals_host_planes():
    d = ge.read_plink(os.path.join(DATA, "thinhap1"))
    with ge.Panel.from_bed(d.bed_payload, d.n_rows // 2, d.n_samples) as p:
        bits, valid = p.download()
    assert np.array_equal(bits, d.bit_data & d.valid_data)
    assert np.array_equal(valid, d.valid_data)


def test_panel_create_download_round_trip():
    """Panels created from planes (the path of gepi_process_gpu_1) and from
    .bed bytes read back exactly, many tiny panels in a row: the zero fill,
    the upload and the read-back all run on the panel's own stream."""
    print("case panel round trip", flush=True)
    rng = np.random.default_rng(77)
    for i in range(40):
        n_var = int(rng.integers(1, 7))
        n = int(rng.integers(1, 200))
        codes = orc.random_codes(rng, n_var, n, 0.1)
        bits, valid = orc.planes_from_codes(codes)
        with ge.Panel.from_planes(bits, valid, n) as p:
            b2, v2 = p.download()
        assert np.array_equal(b2, bits & valid), ("planes", i, n_var, n)
        assert np.array_equal(v2, valid), ("planes valid", i, n_var, n)
        with ge.Panel.from_bed(synth.pack_codes(codes), n_var, n) as p:
            b3, v3 = p.download()
        assert np.array_equal(b3, bits & valid), ("bed", i, n_var, n)
        assert np.array_equal(v3, valid), ("bed valid", i, n_var, n)


def test_resident_panel_rescans_new_phenotypes():
    bits, valid, m0, m1 = random_panel(31, 90, 500, 0.02)
    with ge.Panel.from_planes(bits, valid, 500) as panel:
        rng = np.random.default_rng(5)
        for i in range(3):
            cls = rng.choice([0, 1, -1], size=500, p=[0.4, 0.5, 0.1])
            a, b = orc.masks_from_classes(cls)
            got, _ = panel.scan(a, b, topmost=200)
            ref = orc.gold_scan(orc.Panel(bits, valid, a, b, 500), 200)
            assert_same(got, ref, "rescan %d" % i)


def test_overlapping_and_empty_masks():
    bits, valid, m0, m1 = random_panel(41, 20, 130, 0.0)
    both = m0 | m1
    for a, b in ((both, m1), (m0, np.zeros_like(m1)), (np.zeros_like(m0), np.zeros_like(m1))):
        got, _ = ge.process_gpu_1(bits, valid, a, b, 130, topmost=400)
        ref = orc.gold_scan(orc.Panel(bits, valid, a, b, 130), 400)
        assert_same(got, ref, "masks")


@pytest.mark.parametrize("path", PATHS)
def test_pending_overflow_recovery(monkeypatch, path):
    """A forced 32-tile first band with no threshold yet overflows the pending
    list; the band is redone in smaller pieces."""
    bits, valid, m0, m1 = random_panel(51, 500, 300, 0.01)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, 300), 100)
    monkeypatch.setenv("GEPI_CAND_CAP", "1")
    monkeypatch.setenv("GEPI_BAND_TILES", "32")
    got, st = ge.process_gpu_1(bits, valid, m0, m1, 300, topmost=100, count_path=path)
    monkeypatch.delenv("GEPI_CAND_CAP")
    monkeypatch.delenv("GEPI_BAND_TILES")
    assert st["pend_overflows"] > 0
    assert_same(got, ref, "pending-overflow")


@pytest.mark.parametrize("path", PATHS)
def test_candidate_overflow_recovery(monkeypatch, path):
    """Forced 8-tile bands with a pending list larger than the candidate
    buffers admit more candidates than fit: the threshold is taken from the
    records kept, the band's records are dropped by workgroup tile (square or
    64 x 32 geometry) and the band is redone.  Duplicated or lost records
    would break the exact match."""
    bits, valid, m0, m1 = random_panel(53, 300, 200, 0.01)
    k = 20000
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, 200), k)
    monkeypatch.setenv("GEPI_CAND_CAP", "1")
    monkeypatch.setenv("GEPI_PEND_CAP", "4000000")
    monkeypatch.setenv("GEPI_BAND_TILES", "8")
    got, st = ge.process_gpu_1(bits, valid, m0, m1, 200, topmost=k, count_path=path)
    for v in ("GEPI_CAND_CAP", "GEPI_PEND_CAP", "GEPI_BAND_TILES"):
        monkeypatch.delenv(v)
    assert st["cand_overflows"] > 0
    assert_same(got, ref, "candidate-overflow")


@pytest.mark.parametrize("path", [3, 4])
@pytest.mark.parametrize("cand_cap", [None, "1"])
def test_fp4_tile_panels_and_band_edges(monkeypatch, cand_cap, path):
    """fp4 tiles are numbered in panels of 8 A rows (a shared rectangle
    column by column, then the rows' remaining tiles); 620 variants give 10 A
    rows, a full panel and a partial one of 2.  Bands of 7 tiles straddle
    panel and row boundaries; with the candidate buffer forced to overflow,
    the records of a band are dropped through the inverse of that numbering
    and the band is redone.  The wave-tile kernel (path 3) takes two tiles
    per workgroup, so a band of 7 ends in a workgroup with one tile.  Lost
    or duplicated tiles would break the exact match."""
    print("case fp4 panels cand_cap=%s path=%d" % (cand_cap, path), flush=True)
    bits, valid, m0, m1 = random_panel(61, 620, 200, 0.02)
    k = 6000
    monkeypatch.setenv("GEPI_BAND_TILES", "7")
    if cand_cap:
        monkeypatch.setenv("GEPI_CAND_CAP", cand_cap)
        monkeypatch.setenv("GEPI_PEND_CAP", "4000000")
    got, st = ge.process_gpu_1(bits, valid, m0, m1, 200, topmost=k, count_path=path)
    for v in ("GEPI_BAND_TILES", "GEPI_CAND_CAP", "GEPI_PEND_CAP"):
        monkeypatch.delenv(v, raising=False)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, 200), k)
    assert st["bands"] > 10
    if cand_cap:
        assert st["cand_overflows"] > 0
    assert_same(got, ref, "fp4-panels")


def test_mfma_class_encoding_limits():
    """The int8 MFMA path keeps one class in the low 14 bits of its
    accumulators (at most 16383 samples there).  A class above that puts the
    other class in the low bits; when both exceed it, an explicit int8 request
    fails with GEPI_ERR_LIMIT, while the fp4 path (separate accumulators, the
    automatic choice) and the popcount path still count exactly."""
    rng = np.random.default_rng(71)
    n = 16600
    codes = orc.random_codes(rng, 3, n, 0.01)
    bits, valid = orc.planes_from_codes(codes)
    cls = np.where(np.arange(n) < 16400, 0, 1)      # 16400 controls, 200 cases
    m0, m1 = orc.masks_from_classes(cls)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), 10**6)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=10**6, count_path=2)
    assert_same(got, ref, "mfma-big-class")
    n = 33000
    codes = orc.random_codes(rng, 3, n, 0.01)
    bits, valid = orc.planes_from_codes(codes)
    cls = np.where(np.arange(n) < 16500, 0, 1)      # both classes > 16383
    m0, m1 = orc.masks_from_classes(cls)
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), 10**6)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=10**6)
    assert st["count_path"] == 3
    assert_same(got, ref, "fp4-two-big-classes")
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=10**6, count_path=1)
    assert_same(got, ref, "popcount-two-big-classes")
    with pytest.raises(ge.GepiError) as e:
        ge.process_gpu_1(bits, valid, m0, m1, n, topmost=10, count_path=2)
    assert e.value.code == -3


def test_topk_zero_and_more_than_all():
    bits, valid, m0, m1 = random_panel(61, 4, 70, 0.0)
    got, _ = ge.process_gpu_1(bits, valid, m0, m1, 70, topmost=0)
    assert len(got) == 0
    got, _ = ge.process_gpu_1(bits, valid, m0, m1, 70, topmost=10**6)
    assert len(got) == 5 * 8 * 9 // 2


def test_cli_report_equals_hand_derived_kat_report():
    """gepicount -i data/test -N 10 prints the hand-written report (every
    column derived from the .bed/.bim/.fam bytes; FPv from the reference's
    fisher_p.c) byte for byte on stderr."""
    from test_oracle import KAT_REPORT_N10
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    r = subprocess.run([exe, "-i", os.path.join(DATA, "test"), "-N", "10"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    start = r.stderr.index(KAT_REPORT_N10[:40])
    assert r.stderr[start:start + len(KAT_REPORT_N10)] == KAT_REPORT_N10


def test_cli_intermediate_progress():
    """-D: after every band the CLI prints the reference's progress lines and
    a strided sample of the current ranking; the final report is unchanged."""
    print("case cli -D", flush=True)
    root = os.path.join(DATA, "thinhap1")
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    env = dict(os.environ, GEPI_BAND_TILES="4")
    r = subprocess.run([exe, "-i", root, "-N", "100", "-D"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stderr.splitlines()
    progress = [l for l in lines if l.startswith("K1: Total time ")]
    assert len(progress) >= 3, len(progress)
    assert progress[-1].startswith("K1: Total time ") and "(100%)" in progress[-1]
    header = ("Rank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t"
              "(c00,c01,c10,c11)")
    starts = [i for i, l in enumerate(lines) if l == header]
    assert len(starts) == len(progress) + 1  # one sample per band, then the report
    report = lines[starts[-1] - 1:starts[-1] + 101]
    gold = subprocess.run([os.path.join(ROOT, "oracle", "gepicount_gold"), "-i", root, "-N", "100"],
                          capture_output=True, text=True, timeout=120)
    assert report == gold.stderr.splitlines()[:102]
    # the last intermediate sample shows the final ranking's first records
    assert lines[starts[-2] + 1] == report[2]


def test_cli_report_equals_oracle_report():
    root = os.path.join(DATA, "thinhap1")
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    r = subprocess.run([exe, "-i", root, "-N", "100"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = r.stderr.splitlines()
    start = lines.index("Rank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t"
                        "(c00,c01,c10,c11)")
    report = lines[start - 1:start + 101]
    gold = subprocess.run([os.path.join(ROOT, "oracle", "gepicount_gold"), "-i", root, "-N", "100"],
                          capture_output=True, text=True, timeout=120)
    assert report == gold.stderr.splitlines()[:102]
    assert "Cleaned" in r.stdout


@pytest.mark.parametrize("n0,n1,ffun,path", [(390, 310, 0, 4), (640, 640, 0, 3), (640, 512, 0, 4),
                                              (640, 512, 1, 4), (1200, 1200, 1, 3), (1200, 1200, 0, 3)])
def test_automatic_count_path(n0, n1, ffun, path):
    """count_path 0 takes the fp4 wave tiles (3) for scans of at least 10 K
    steps of 128 samples (both classes, each rounded up) and the LDS-staged
    fp4 kernel (4) below that; either way the result equals the oracle's."""
    print("case auto path n0=%d n1=%d ffun=%d" % (n0, n1, ffun), flush=True)
    rng = np.random.default_rng(n0 + 3 * n1 + ffun)
    n = n0 + n1
    codes = orc.random_codes(rng, 30, n, 0.02)
    cls = rng.permutation(np.r_[np.zeros(n0, np.int64), np.ones(n1, np.int64)])
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    got, st = ge.process_gpu_1(bits, valid, m0, m1, n, topmost=500, ffun=ffun)
    assert st["count_path"] == path
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, n), 500, ffun=ffun)
    assert_same(got, ref, "auto path")
