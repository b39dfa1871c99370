"""Pins the CPU oracle before anything is checked against it.

- Appendix-A known answers for data/test (hand-derived from the .bed bytes)
- exact golden vectors (mpmath / exact rationals, tests/golden/make_golden.py)
- the reference's own fisher_p.c compiled unmodified (oracle/_ref) as fixture
- per-sample brute force on random tiny panels
"""
import json
import os
import struct

import numpy as np
import pytest

from oracle import oracle as orc

DATA = os.path.join(os.path.dirname(__file__), "golden", "data")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# (sA, sB) -> per function (AND, NAND, ANDN, NANDN, XOR): (c00, c01, c10, c11)
# Hand derivation for data/test: snp1 = [0,2,3,2,3,3], snp2 = [2,2,0,3,2,3],
# ctrl = samples 0-2, case = samples 3-5.
KAT_TEST = {
    (0, 0): [(2, 1, 1, 2), (0, 3, 0, 3), (0, 3, 0, 3), (1, 2, 2, 1), (0, 3, 0, 3)],
    (1, 0): [(1, 2, 1, 2), (1, 2, 0, 3), (1, 2, 2, 1), (0, 3, 0, 3), (2, 1, 2, 1)],
    (1, 1): [(2, 1, 3, 0), (0, 3, 0, 3), (0, 3, 0, 3), (1, 2, 0, 3), (0, 3, 0, 3)],
    (2, 0): [(2, 1, 0, 3), (0, 3, 1, 2), (1, 2, 1, 2), (0, 3, 1, 2), (1, 2, 2, 1)],
    (2, 1): [(2, 1, 1, 2), (0, 3, 2, 1), (1, 2, 0, 3), (0, 3, 0, 3), (1, 2, 2, 1)],
    (2, 2): [(3, 0, 1, 2), (0, 3, 0, 3), (0, 3, 0, 3), (0, 3, 2, 1), (0, 3, 0, 3)],
    (3, 0): [(2, 1, 1, 2), (0, 3, 0, 3), (0, 3, 2, 1), (1, 2, 0, 3), (0, 3, 2, 1)],
    (3, 1): [(1, 2, 3, 0), (1, 2, 0, 3), (1, 2, 0, 3), (0, 3, 0, 3), (2, 1, 0, 3)],
    (3, 2): [(2, 1, 1, 2), (1, 2, 0, 3), (0, 3, 2, 1), (0, 3, 0, 3), (1, 2, 2, 1)],
    (3, 3): [(2, 1, 3, 0), (0, 3, 0, 3), (0, 3, 0, 3), (1, 2, 0, 3), (0, 3, 0, 3)],
}
KAT_ENT = {(2, 1, 1, 2): -0.0817041659, (0, 3, 0, 3): 0.0, (1, 2, 2, 1): -0.0817041659,
           (1, 2, 1, 2): 0.0, (1, 2, 0, 3): -0.1908745046, (2, 1, 2, 1): 0.0,
           (2, 1, 3, 0): -0.1908745046, (2, 1, 0, 3): -0.4591479170,
           (0, 3, 1, 2): -0.1908745046, (0, 3, 2, 1): -0.4591479170,
           (3, 0, 1, 2): -0.4591479170, (1, 2, 3, 0): -0.4591479170}
KAT_TOP5 = [(2, 0, 0), (2, 1, 1), (2, 2, 0), (2, 2, 3), (3, 0, 2)]


def counts_to_cells(w, x, y, z):
    return (w, y - w, x, z - x)


def test_reader_test_fixture_planes():
    p = orc.read_plink(os.path.join(DATA, "test"))
    assert p.n_rows == 4 and p.n_samples == 6 and p.words == 1
    rows = [sorted(s for s in range(6) if int(p.bits[r, 0]) >> s & 1) for r in range(4)]
    assert rows == [[0, 1, 3], [1, 2, 3, 4, 5], [0, 1, 2, 4], [0, 1, 3, 4, 5]]
    assert all(int(p.valid[r, 0]) == 0b111111 for r in range(4))
    assert int(p.mask_ctrl[0]) == 0b000111 and int(p.mask_case[0]) == 0b111000
    assert p.snp_id == ["snp1", "snp1", "snp2", "snp2"]
    assert p.allele == ["A", "C", "G", "T"]
    assert p.pos == [1, 1, 2, 2] and p.chrom == [1, 1, 1, 1]


def test_kat_counts_and_scores():
    p = orc.read_plink(os.path.join(DATA, "test"))
    cnt = orc.all_counts(p)
    e = 0
    for j in range(4):
        for k in range(j + 1):
            for f in range(5):
                w, x, y, z = (int(v) for v in cnt[e, f])
                cells = counts_to_cells(w, x, y, z)
                assert cells == KAT_TEST[(j, k)][f], (j, k, f)
                assert abs(orc.entropy_exact(w, x, y, z) - KAT_ENT[cells]) < 1e-9
            e += 1


def test_kat_top5_order():
    p = orc.read_plink(os.path.join(DATA, "test"))
    tops = orc.gold_scan(p, 5)
    assert [(int(r["sA"]), int(r["sB"]), int(r["fun"])) for r in tops] == KAT_TOP5
    assert all(r["obj"] == tops[0]["obj"] for r in tops)
    # nine evaluations tie at -0.4591479170; all from four distinct count tuples
    allt = orc.gold_scan(p, 1000)
    assert len(allt) == 50
    assert int(np.sum(allt["obj"] == tops[0]["obj"])) == 9


def test_kat_fisher_mirror_tie():
    # masses {1,9,9,1}/20: the mirror term ties exactly; p = 1
    assert orc.fisher_rank(2, 1, 3, 3) == 0.0
    # observed (2,1,0,3): p = 0.4 -> log10 = -0.39794
    assert abs(orc.fisher_rank(2, 0, 3, 3) - np.log10(0.4)) < 1e-14


def test_entropy_golden():
    d = json.load(open(os.path.join(GOLDEN, "entropy_vectors.json")))
    worst = 0.0
    for w, x, y, z, ref in d["rows"]:
        ref = float(ref)
        got = orc.entropy_exact(w, x, y, z)
        err = abs(got - ref)
        assert err <= 4e-16 * abs(ref) + 1e-24, (w, x, y, z, got, ref)
        assert got <= 1e-24 and got >= -1.0 - 1e-15
        worst = max(worst, err / max(abs(ref), 1e-300))
        # the reference's own float32 expression agrees to float tolerance
        f32 = orc.entropy_ref_f32(w, x, y, z)
        assert abs(f32 - ref) <= 2e-6, (w, x, y, z, f32, ref)


def test_entropy_symmetries_bit_identical():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        y, z = (int(v) for v in rng.integers(0, 3000, 2))
        w, x = int(rng.integers(0, y + 1)), int(rng.integers(0, z + 1))
        s = orc.entropy_exact(w, x, y, z)
        assert orc.entropy_exact(x, w, z, y) == s          # class swap
        assert orc.entropy_exact(y - w, z - x, y, z) == s  # outcome swap
        # transpose: classes <-> outcome
        assert orc.entropy_exact(w, y - w, w + x, y + z - w - x) == s


def test_fisher_rank_golden():
    d = json.load(open(os.path.join(GOLDEN, "fisher_vectors.json")))
    for w, x, y, z, ref in d["rows"]:
        ref = float(ref)
        got = orc.fisher_rank(w, x, y, z)
        assert abs(got - ref) <= 1e-12 * max(abs(ref), 1.0), (w, x, y, z, got, ref)


def test_fisher_ref_matches_reference_binary_fixture():
    d = json.load(open(os.path.join(GOLDEN, "fisher_ref_vectors.json")))
    for n11, n12, n21, n22, bits in d["rows"]:
        got = struct.unpack("<I", struct.pack("<f", orc.fisher_ref(n11, n12, n21, n22)))[0]
        assert got == bits, (n11, n12, n21, n22)


def test_fisher_ref_matches_reference_binary_live():
    if not os.path.exists(orc.REF_FISHER_PATH):
        pytest.skip("oracle/_ref/libfisher_ref.so not built")
    import ctypes
    ref = ctypes.CDLL(orc.REF_FISHER_PATH)
    ref.fisher_pval.restype = ctypes.c_float
    ref.fisher_pval.argtypes = [ctypes.c_int] * 4
    rng = np.random.default_rng(11)
    for _ in range(3000):
        t = [int(v) for v in rng.integers(0, 400, 4)]
        a, b = ref.fisher_pval(*t), orc.fisher_ref(*t)
        assert (a == b) or (np.isnan(a) and np.isnan(b)), t


def test_log2_own_accuracy():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(1e-300, 1, 2000), rng.uniform(1, 1e6, 2000),
                         [1.0, 2.0, 0.5, 3.0, 1e-300, 1e300]])
    for v in xs:
        got = orc.lib().orc_log2(float(v))
        ref = np.log2(v)
        assert abs(got - ref) <= 4e-16 * max(abs(ref), 1.0), v


@pytest.mark.parametrize("seed,N,n,miss", [(1, 5, 37, 0.0), (2, 7, 70, 0.1), (3, 4, 130, 0.3)])
def test_counts_brute_force(seed, N, n, miss):
    rng = np.random.default_rng(seed)
    codes = orc.random_codes(rng, N, n, miss)
    cls = rng.choice([0, 1, -1], size=n, p=[0.45, 0.45, 0.1])
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    panel = orc.Panel(bits, valid, m0, m1, n)
    got = orc.all_counts(panel)
    ref = orc.brute_counts(codes, cls)
    assert np.array_equal(got, ref)


def test_gold_ranking_matches_full_sort():
    rng = np.random.default_rng(7)
    codes = orc.random_codes(rng, 12, 90, 0.05)
    cls = rng.choice([0, 1], size=90)
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    panel = orc.Panel(bits, valid, m0, m1, 90)
    cnt = orc.brute_counts(codes, cls)
    R = 24
    keys = []
    e = 0
    for j in range(R):
        for k in range(j + 1):
            for f in range(5):
                w, x, y, z = (int(v) for v in cnt[e, f])
                keys.append((orc.entropy_exact(w, x, y, z), j, k, f, w, x, y, z))
            e += 1
    keys.sort()
    for ffun, nt in ((0, 1), (0, 3)):
        tops = orc.gold_scan(panel, 200, ffun=ffun, nthreads=nt)
        got = [(float(r["obj"]), int(r["sA"]), int(r["sB"]), int(r["fun"]), int(r["w"]),
                int(r["x"]), int(r["y"]), int(r["z"])) for r in tops]
        assert got == keys[:200]
    allt = orc.gold_scan(panel, 10**9)
    assert len(allt) == 5 * R * (R + 1) // 2


def test_reader_rejects_bad_inputs(tmp_path):
    src = os.path.join(DATA, "test")
    fam = open(src + ".fam").read()
    bim = open(src + ".bim").read()
    bed = open(src + ".bed", "rb").read()

    def write(name, fam_t, bim_t, bed_b):
        root = str(tmp_path / name)
        open(root + ".fam", "w").write(fam_t)
        open(root + ".bim", "w").write(bim_t)
        open(root + ".bed", "wb").write(bed_b)
        return root

    with pytest.raises(ValueError, match="magic"):
        orc.read_plink(write("magic", fam, bim, b"\x6c\x1c" + bed[2:]))
    with pytest.raises(ValueError, match="size"):
        orc.read_plink(write("size", fam, bim, bed + b"\x00"))
    with pytest.raises(ValueError, match="sample-major"):
        orc.read_plink(write("mode", fam, bim, bed[:2] + b"\x00" + bed[3:]))
    with pytest.raises(ValueError, match="malformed"):
        orc.read_plink(write("bim", fam, bim.replace("\t0\t1\t", "\tzz\t1\t"), bed))
    with pytest.raises(ValueError, match="cannot open"):
        orc.read_plink(str(tmp_path / "absent"))


# The data/test report for -N 10, written out by hand: ranks and (sA, sB, fun)
# from the Appendix-A scores under the canonical key (nine evaluations tie at
# -I = -0.4591479170, then (1, 0, NAND) leads the -0.1908745046 group), row
# names from data/test.bim (row 2i = column 5 allele, row 2i+1 = column 6),
# cells (w, y-w, x, z-x) from the Appendix-A count table, Obj as float %g.
# The FPv column is the reference's own fisher_pval (src/misc/fisher_p.c,
# compiled unmodified into oracle/_ref) of those cells: -0.39794 for the
# 3-1/1-3 split tables (two-sided p = 0.4), 0 for (1,2,0,3).
KAT_REPORT_N10 = """
Rank:\tidxA[id allele chr:pos]\tFunc\tidxB[id allele chr:pos]\tObj\tFPv\t(c00,c01,c10,c11)
0:\t2[snp2 G 1:2]\tAND\t0[snp1 A 1:1]\t-0.459148\t-0.39794\t(2,1,0,3)
1:\t2[snp2 G 1:2]\tNAND\t1[snp1 C 1:1]\t-0.459148\t-0.39794\t(0,3,2,1)
2:\t2[snp2 G 1:2]\tAND\t2[snp2 G 1:2]\t-0.459148\t-0.39794\t(3,0,1,2)
3:\t2[snp2 G 1:2]\tNANDN\t2[snp2 G 1:2]\t-0.459148\t-0.39794\t(0,3,2,1)
4:\t3[snp2 T 1:2]\tANDN\t0[snp1 A 1:1]\t-0.459148\t-0.39794\t(0,3,2,1)
5:\t3[snp2 T 1:2]\tXOR\t0[snp1 A 1:1]\t-0.459148\t-0.39794\t(0,3,2,1)
6:\t3[snp2 T 1:2]\tAND\t1[snp1 C 1:1]\t-0.459148\t-0.39794\t(1,2,3,0)
7:\t3[snp2 T 1:2]\tXOR\t1[snp1 C 1:1]\t-0.459148\t-0.39794\t(2,1,0,3)
8:\t3[snp2 T 1:2]\tANDN\t2[snp2 G 1:2]\t-0.459148\t-0.39794\t(0,3,2,1)
9:\t1[snp1 C 1:1]\tNAND\t0[snp1 A 1:1]\t-0.190875\t0\t(1,2,0,3)
"""


def test_kat_report_text_hand_derived():
    """The oracle's gold command line and its report_lines reproduce the
    hand-written data/test report byte for byte."""
    import subprocess
    exe = os.path.join(os.path.dirname(orc.HERE), "oracle", "gepicount_gold")
    r = subprocess.run([exe, "-i", os.path.join(DATA, "test"), "-N", "10"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stderr == KAT_REPORT_N10
    p = orc.read_plink(os.path.join(DATA, "test"))
    tops = orc.gold_scan(p, 10)
    assert "\n".join(orc.report_lines(p, tops)) + "\n" == KAT_REPORT_N10


def test_reader_thinhap1_equals_planes_from_codes():
    """data/thinhap1 (517 missing calls): the oracle reader's planes against
    the per-sample decoding of the raw .bed codes (BinPlinkReader.h:171-201)."""
    root = os.path.join(DATA, "thinhap1")
    p = orc.read_plink(root)
    raw = np.fromfile(root + ".bed", np.uint8)
    assert raw[:3].tolist() == [0x6C, 0x1B, 0x01]
    n, N = p.n_samples, p.n_rows // 2
    payload = raw[3:].reshape(N, (n + 3) // 4)
    q = np.stack([(payload >> (2 * i)) & 3 for i in range(4)], axis=-1).reshape(N, -1)[:, :n]
    assert int((q == 1).sum()) == 517
    bits, valid = orc.planes_from_codes(q)
    assert np.array_equal(p.valid, valid)
    assert np.array_equal(p.bits, bits)
    assert p.total_missing == 517


def test_entropy_approx_within_bound():
    """The gold scan's skip test: the plain-double sum stays within 1e-12 of
    the exact score over the golden tuples and random large tables."""
    with open(os.path.join(GOLDEN, "entropy_vectors.json")) as f:
        vec = json.load(f)
    tuples = [tuple(v[:4]) for v in vec["rows"]]
    rng = np.random.default_rng(3)
    for _ in range(20000):
        y, z = int(rng.integers(0, 9000)), int(rng.integers(0, 9000))
        tuples.append((int(rng.integers(0, y + 1)), int(rng.integers(0, z + 1)), y, z))
    for w, x, y, z in tuples:
        a, e = orc.lib().orc_entropy_approx(w, x, y, z), orc.entropy_exact(w, x, y, z)
        assert abs(a - e) < 1e-12, (w, x, y, z, a, e)


def test_gold_scan_rows_partition():
    """Row ranges partition the gold scan: the top-k of the union of the
    ranges' top-k lists is the whole scan's top-k."""
    rng = np.random.default_rng(11)
    codes = orc.random_codes(rng, 40, 300, 0.02)
    cls = rng.choice([0, 1], size=300)
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    panel = orc.Panel(bits, valid, m0, m1, 300)
    whole = orc.gold_scan(panel, 500)
    parts = [orc.gold_scan_rows(panel, a, b, 500) for a, b in ((0, 30), (30, 57), (57, 80))]
    allp = np.concatenate(parts)
    b = allp["obj"].view(np.uint64)
    hi = np.where(b >> np.uint64(63), ~b, b | np.uint64(1 << 63))
    lo = (allp["sA"].astype(np.uint64) << np.uint64(32)) | \
        (allp["sB"].astype(np.uint64) << np.uint64(3)) | allp["fun"].astype(np.uint64)
    merged = allp[np.lexsort((lo, hi))][:500]
    assert merged.tobytes() == whole.tobytes()
    # every range's records lie in its rows
    for (a, bnd), r in zip(((0, 30), (30, 57), (57, 80)), parts):
        assert np.all((r["sA"] >= a) & (r["sA"] < bnd))


def test_gold_scan_ref_f32_is_the_stable_float_ranking():
    """The reference-style single-thread gold (16-bit table popcounts, float32
    scores, 2k buffer with a strict limit) returns the k smallest float32
    scores in scan order among ties, with brute-force counts, also when the
    scan stops early at j_end."""
    rng = np.random.default_rng(17)
    codes = orc.random_codes(rng, 9, 150, 0.05)
    cls = rng.choice([0, 1, -1], size=150, p=[0.5, 0.45, 0.05])
    bits, valid = orc.planes_from_codes(codes)
    m0, m1 = orc.masks_from_classes(cls)
    panel = orc.Panel(bits, valid, m0, m1, 150)
    cnt = orc.brute_counts(codes, cls)
    R = bits.shape[0]
    recs = []
    e = 0
    for j in range(R):
        for k in range(j + 1):
            for f in range(5):
                w, x, y, z = (int(v) for v in cnt[e, f])
                recs.append((np.float32(orc.entropy_ref_f32(w, x, y, z)), len(recs), j, k, f,
                             w, x, y, z))
            e += 1
    for topk, j_end in ((7, R), (60, R), (60, 11), (5000, R)):
        sub = [r for r in recs if r[2] < j_end]
        ref = sorted(sub, key=lambda r: (r[0], r[1]))[:topk]
        got, sec = orc.gold_scan_ref_f32(panel, topk, j_end)
        assert sec >= 0.0
        assert [(np.float32(g["obj"]), int(g["sA"]), int(g["sB"]), int(g["fun"]), int(g["w"]),
                 int(g["x"]), int(g["y"]), int(g["z"])) for g in got] == \
            [(r[0],) + r[2:] for r in ref]


def test_large_gold_fixture_self_consistent():
    """The committed gold top-10,000 of the large workload: keys strictly
    increase, every score is the oracle's exact score of its counts, and the
    counts fit the class sizes (5,734 controls, 2,458 cases)."""
    g = np.load(os.path.join(GOLDEN, "large_entropy_top10000.npz"))
    assert int(g["topk"]) == 10000 and len(g["sA"]) == 10000
    assert len(str(g["bed_sha256"])) == 64 and len(str(g["pheno_sha256"])) == 64
    b = g["obj_bits"]
    hi = np.where(b >> np.uint64(63), ~b, b | np.uint64(1 << 63))
    lo = (g["sA"].astype(np.uint64) << np.uint64(32)) | \
        (g["sB"].astype(np.uint64) << np.uint64(3)) | g["fun"].astype(np.uint64)
    assert np.all((hi[1:] > hi[:-1]) | ((hi[1:] == hi[:-1]) & (lo[1:] > lo[:-1])))
    assert np.all(g["sA"] >= g["sB"]) and np.all(g["sA"] < 200000) and np.all(g["fun"] < 5)
    w, x, y, z = (g[f].astype(np.int64) for f in "wxyz")
    assert np.all((0 <= w) & (w <= y) & (y <= 5734) & (0 <= x) & (x <= z) & (z <= 2458))
    exact = np.array([orc.entropy_exact(int(a), int(c), int(d), int(e))
                      for a, c, d, e in zip(w, x, y, z)])
    assert np.array_equal(exact.view(np.uint64), b)
