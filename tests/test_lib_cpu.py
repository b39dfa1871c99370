"""CPU-only checks of the product library: it loads, exports the C ABI, its host
reader/report/Fisher restatements equal the oracle's, and the scan entry points
fail loudly (no CPU fallback) when no GPU is present."""
import ctypes
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import api, synth
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")
HEADER = os.path.join(ROOT, "include", "gepicount.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gepi_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(api.EXPORTS)


def test_binding_declares_every_signature():
    """Every entry point with parameters has ctypes argtypes: without them
    ctypes passes Python ints as 32-bit C ints and truncates addresses."""
    lib = ge.lib()
    for name in api.EXPORTS:
        fn = getattr(lib, name)
        if name in ("gepi_last_error", "gepi_version"):
            continue
        assert fn.argtypes is not None, name
        if name not in ("gepi_params_init", "gepi_plink_close"):  # void
            assert fn.restype is not None, name


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(api.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert ge.lib().gepi_version().decode().startswith("gepicount")


def test_fisher_pval_matches_reference_binary_fixture():
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "fisher_ref_vectors.json")))
    for n11, n12, n21, n22, bits in d["rows"]:
        got = struct.unpack("<I", struct.pack("<f", ge.fisher_pval(n11, n12, n21, n22)))[0]
        assert got == bits, (n11, n12, n21, n22)


def _bits(v):
    return struct.pack("<d", v)


def test_score_table_equals_oracle():
    """gepi_score_table runs the scan's exact stage (gepi_score.h, built with
    the device code's compiler and flags) on the host.  Every table scores as
    the oracle's restatement, bit for bit: all tables up to 12 + 12 samples,
    then random ones up to 9,000 + 9,000 (near-proportional ones cross the
    mode of the Fisher walks, whose early exits must not change a bit)."""
    rng = np.random.default_rng(91)
    tables = [(w, x, y, z) for y in range(13) for z in range(13)
              for w in range(y + 1) for x in range(z + 1)]
    for i in range(3000):
        y, z = (int(v) for v in rng.integers(1, 9001, size=2))
        if i % 3 == 0:
            w, x = int(rng.integers(0, y + 1)), int(rng.integers(0, z + 1))
        else:
            f = rng.random()
            w = int(f * y)
            x = int(min(max((f + rng.uniform(-0.1, 0.1)) * z, 0), z))
        tables.append((w, x, y, z))
    for w, x, y, z in tables:
        assert _bits(ge.score_table(w, x, y, z, ge.GEPI_FFUNC_ENTROPY)) == \
            _bits(orc.entropy_exact(w, x, y, z)), (w, x, y, z)
        assert _bits(ge.score_table(w, x, y, z, ge.GEPI_FFUNC_FISHERPV)) == \
            _bits(orc.fisher_rank(w, x, y, z)), (w, x, y, z)


def test_score_table_rejects_bad_tables():
    for args in [(-1, 0, 3, 3), (4, 0, 3, 3), (0, 4, 3, 3), (0, 0, -1, 3)]:
        with pytest.raises(ge.GepiError):
            ge.score_table(*args)
    with pytest.raises(ge.GepiError):
        ge.score_table(0, 0, 30000, 30000)
    with pytest.raises(ge.GepiError):
        ge.score_table(1, 1, 3, 3, ffun=7)


@pytest.mark.parametrize("name", ["test", "thinhap1"])
def test_reader_equals_oracle_reader(name):
    root = os.path.join(DATA, name)
    a = ge.read_plink(root)
    b = orc.read_plink(root)
    assert (a.n_rows, a.n_samples, a.words) == (b.n_rows, b.n_samples, b.words)
    assert np.array_equal(a.bit_data, b.bits) and np.array_equal(a.valid_data, b.valid)
    assert np.array_equal(a.mask_ctrl, b.mask_ctrl) and np.array_equal(a.mask_case, b.mask_case)
    assert a.snp_id == b.snp_id and a.allele == b.allele
    assert a.chrom == b.chrom and a.pos == b.pos and a.pheno == b.pheno
    assert a.sample_name == b.sample_name and a.total_missing == b.total_missing


def test_reader_thinhap1_shape():
    a = ge.read_plink(os.path.join(DATA, "thinhap1"))
    assert a.n_rows == 2040 and a.n_samples == 89 and a.words == 2
    assert int(sum(int(w).bit_count() for w in a.mask_ctrl)) == 45
    assert int(sum(int(w).bit_count() for w in a.mask_case)) == 44


def test_reader_rejects_bad_files(tmp_path):
    src = os.path.join(DATA, "test")
    fam, bim = open(src + ".fam").read(), open(src + ".bim").read()
    bed = open(src + ".bed", "rb").read()

    def write(name, bed_b, bim_t=bim):
        root = str(tmp_path / name)
        open(root + ".fam", "w").write(fam)
        open(root + ".bim", "w").write(bim_t)
        open(root + ".bed", "wb").write(bed_b)
        return root

    for root, code in ((write("magic", b"\x6c\x1c" + bed[2:]), -7),
                       (write("size", bed[:-1]), -7),
                       (write("mode", bed[:2] + b"\x00" + bed[3:]), -7),
                       (write("bim", bed, bim.replace("snp2\t0\t2", "snp2\t0\tq")), -7),
                       (str(tmp_path / "none"), -6)):
        with pytest.raises(ge.GepiError) as e:
            ge.read_plink(root)
        assert e.value.code == code


def test_check_quality_counts(tmp_path):
    p = synth.make_panel("small", n_var=64)
    codes = synth.unpack_codes(p["payload"], 256)
    codes[3, :40] = 1       # variant 3 misses 40/256 samples -> both its rows flagged
    codes[:, 7] = 1         # sample 7 misses everything
    p["payload"] = synth.pack_codes(codes)
    root = str(tmp_path / "qa")
    synth.write_plink(root, p)
    d = ge.read_plink(root)
    bad_s, bad_r = d.check_quality(0.05, 0.05)
    assert bad_s == 1 and bad_r == 2


def test_check_quality_counts_ragged_samples(tmp_path):
    """Missing-call counts per sample and per row against numpy, with a sample
    count that leaves a partial last word (150) and thresholds around the
    observed fractions, so that the float comparisons decide."""
    spec = dict(n_var=90, n_samples=150, n_case=70, maf=("uniform", 0.05, 0.5), miss=0.06,
                planted=0, seed=71)
    p = synth.make_panel(spec=spec)
    root = str(tmp_path / "qa_ragged")
    synth.write_plink(root, p)
    d = ge.read_plink(root)
    R, n = d.n_rows, d.n_samples
    bits = np.unpackbits(d.valid_data.view(np.uint8).reshape(R, -1), axis=1,
                         bitorder="little")[:, :n].astype(np.int64)
    f_s = (R - bits.sum(axis=0)).astype(np.float32) / np.float32(R)
    f_r = (n - bits.sum(axis=1)).astype(np.float32) / np.float32(n)
    for thr in (0.03, 0.05, 0.06, 0.07, 0.09):
        bad_s, bad_r = d.check_quality(thr, thr)
        assert bad_s == int((f_s > np.float32(thr)).sum()), thr
        assert bad_r == int((f_r > np.float32(thr)).sum()), thr


def test_report_format_matches_oracle(tmp_path):
    """printInteractions output of the library == the oracle's print_misc.h restatement."""
    root = os.path.join(DATA, "thinhap1")
    d = ge.read_plink(root)
    op = orc.read_plink(root)
    tops = orc.gold_scan(op, 40)
    res = np.zeros(len(tops), dtype=api.RESULT_DTYPE)
    for k in ("sA", "sB", "w", "x", "y", "z", "obj", "fun"):
        res[k] = tops[k]
    script = ("import sys, numpy as np; sys.path.insert(0, %r)\n"
              "import gepicount_amd as ge\nfrom gepicount_amd import api\n"
              "d = ge.read_plink(%r)\nres = np.load(%r)\n"
              "api.lib().gepi_print_interactions(d._p, res.ctypes.data, len(res), 1, 0)\n"
              % (ROOT, root, str(tmp_path / "r.npy")))
    np.save(str(tmp_path / "r.npy"), res)
    out = subprocess.run(["python3", "-c", script], capture_output=True, text=True, check=True)
    assert out.stderr.splitlines() == orc.report_lines(op, tops)


def test_scan_fails_loudly_without_gpu():
    if ge.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = synth.make_panel("small", n_var=8)
    with pytest.raises(ge.GepiError) as e:
        ge.Panel.from_bed(p["payload"], 8, 256)
    assert e.value.code == -5
    bits = np.zeros((16, 4), np.uint64)
    with pytest.raises(ge.GepiError) as e:
        ge.process_gpu_1(bits, bits, None, None, 256, topmost=10)
    assert e.value.code == -5
    with pytest.raises(ge.GepiError) as e:
        ge.Engine()
    assert e.value.code == -5
    with pytest.raises(ge.GepiError) as e:
        ge.Server()
    assert e.value.code == -5


def test_layout_and_argument_errors_before_device():
    bits = np.zeros((3, 1), np.uint64)
    with pytest.raises(ge.GepiError) as e:
        ge.Panel.from_planes(bits, bits, 10)
    assert e.value.code == -2
    valid = np.zeros((4, 1), np.uint64)
    valid[1, 0] = 1
    with pytest.raises(ge.GepiError) as e:
        ge.Panel.from_planes(np.zeros((4, 1), np.uint64), valid, 10)
    assert e.value.code == -2
    with pytest.raises(ge.GepiError) as e:
        ge.process_gpu_1(np.zeros((4, 1), np.uint64), np.zeros((4, 1), np.uint64), None, None, 10,
                         topmost=-1)
    assert e.value.code == -1


def test_cli_refuses_gold_and_bad_args():
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    r = subprocess.run([exe, "-g", "-i", os.path.join(DATA, "test")], capture_output=True, text=True)
    assert r.returncode == 3 and "oracle/gepicount_gold" in r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 3 and "Insufficient arguments" in r.stderr
    r = subprocess.run([exe, "-i", "x", "--block-sizes", "3", "5"], capture_output=True, text=True)
    assert r.returncode == 3 and "multiple of war size" in r.stderr


def test_synth_pack_roundtrip_and_masks():
    rng = np.random.default_rng(0)
    codes = rng.integers(0, 4, size=(7, 131)).astype(np.uint8)
    assert np.array_equal(synth.unpack_codes(synth.pack_codes(codes), 131), codes)
    p = synth.make_panel("small", n_var=16)
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    for s in range(256):
        assert ((int(m0[s // 64]) >> (s % 64)) & 1) == (p["pheno"][s] == 1)
        assert ((int(m1[s // 64]) >> (s % 64)) & 1) == (p["pheno"][s] == 2)


def test_synth_written_panel_reads_back(tmp_path):
    p = synth.make_panel("medium", n_var=300)
    root = str(tmp_path / "m")
    synth.write_plink(root, p)
    d = ge.read_plink(root)
    o = orc.read_plink(root)
    assert np.array_equal(d.bit_data, o.bits)
    assert d.n_rows == 600 and d.n_samples == 2048
    assert np.array_equal(d.bed_payload.reshape(300, -1), p["payload"])
    # missing rate ~1%
    assert 0.005 < d.total_missing / (300 * 2048) < 0.02
