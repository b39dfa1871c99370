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
