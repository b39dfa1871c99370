"""Missing-call QA on the device (gepi_panel_qa) and checkQuality(remove=true)
(gepi_panel_filter, --qa-remove): counts against numpy on the genotype codes
and the host reader, the filtered panel against the original rows, and a
scan of the filtered data against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def qa_panel(seed, n_var=300, n=500):
    """Codes with uneven missingness: most variants ~1%, some 4-12%; a few
    samples missing 20% of their calls."""
    rng = np.random.default_rng(seed)
    codes = orc.random_codes(rng, n_var, n, 0.0)
    vrate = np.where(rng.random(n_var) < 0.2, rng.uniform(0.04, 0.12, n_var), 0.01)
    srate = np.where(rng.random(n) < 0.05, 0.2, 0.0)
    miss = rng.random((n_var, n)) < np.maximum(vrate[:, None], srate[None, :])
    codes = np.where(miss, 1, codes)
    pheno = np.where(rng.random(n) < 0.45, 2, 1)
    return codes, pheno


def expected_qa(codes, thr_snp, thr_samp):
    n_var, n = codes.shape
    vm = (codes == 1).sum(axis=1).astype(np.int32)
    sm = (codes == 1).sum(axis=0).astype(np.int32)
    bad_v = (vm.astype(np.float32) / np.float32(n)) > np.float32(thr_snp)
    bad_s = ((2 * sm).astype(np.float32) / np.float32(2 * n_var)) > np.float32(thr_samp)
    return vm, sm, bad_v, bad_s


def test_qa_counts_match_codes_and_host_reader(tmp_path):
    codes, pheno = qa_panel(1)
    n_var, n = codes.shape
    vm_e, sm_e, bad_v, bad_s = expected_qa(codes, 0.05, 0.05)
    with ge.Panel.from_bed(synth.pack_codes(codes), n_var, n) as p:
        vm, sm, nbv, nbs = p.qa(0.05, 0.05)
    np.testing.assert_array_equal(vm, vm_e)
    np.testing.assert_array_equal(sm, sm_e)
    assert nbv == bad_v.sum() and nbs == bad_s.sum()
    assert 0 < nbv < n_var and 0 < nbs < n
    # the host reader's checkQuality counts: rows (two per variant) and samples
    root = str(tmp_path / "qa")
    synth.write_plink(root, dict(payload=synth.pack_codes(codes), pheno=pheno, n_var=n_var,
                                 n_samples=n))
    d = ge.read_plink(root)
    bad_samples, bad_rows = d.check_quality(0.05, 0.05)
    assert bad_rows == 2 * nbv and bad_samples == nbs


def test_filter_panel_and_scan_equal_oracle():
    codes, pheno = qa_panel(2)
    n_var, n = codes.shape
    _, _, bad_v, bad_s = expected_qa(codes, 0.05, 0.05)
    m0, m1 = synth.masks_from_pheno(pheno)
    with ge.Panel.from_bed(synth.pack_codes(codes), n_var, n) as p:
        bits, valid = p.download()
        q, kept, keep = p.filter(0.05, 0.05)
        with q:
            np.testing.assert_array_equal(kept, np.flatnonzero(~bad_v))
            rows = np.stack([2 * kept, 2 * kept + 1], axis=1).ravel()
            qb, qv = q.download()
            np.testing.assert_array_equal(qb, bits[rows])
            np.testing.assert_array_equal(qv, valid[rows])
            want = np.zeros(len(keep) * 64, np.uint8)
            want[:n] = ~bad_s
            exp_keep = np.packbits(want.reshape(-1, 64)[:, ::-1], axis=1).view(">u8").ravel()
            np.testing.assert_array_equal(keep, exp_keep.astype(np.uint64))
            got, _ = q.scan(m0 & keep, m1 & keep, topmost=500)
    ref = orc.gold_scan(orc.Panel(bits[rows], valid[rows], m0 & keep, m1 & keep, n), 500)
    for f in ("sA", "sB", "fun", "w", "x", "y", "z"):
        np.testing.assert_array_equal(got[f], ref[f])
    np.testing.assert_array_equal(got["obj"].view(np.uint64), ref["obj"].view(np.uint64))


def test_cli_qa_remove_report(tmp_path):
    """--qa-remove: the report equals the oracle's on the filtered data, with
    the original SNP rows."""
    codes, pheno = qa_panel(3, n_var=120, n=300)
    n_var, n = codes.shape
    _, _, bad_v, bad_s = expected_qa(codes, 0.05, 0.05)
    root = str(tmp_path / "qa")
    synth.write_plink(root, dict(payload=synth.pack_codes(codes), pheno=pheno, n_var=n_var,
                                 n_samples=n))
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    r = subprocess.run([exe, "-i", root, "-N", "50", "--qa-remove"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "QA removal: %d of %d SNPs kept" % ((~bad_v).sum(), n_var) in r.stdout
    full = orc.read_plink(root)
    kept = np.flatnonzero(~bad_v)
    rows = np.stack([2 * kept, 2 * kept + 1], axis=1).ravel()
    keep = np.zeros(((n + 63) // 64) * 64, np.uint8)
    keep[:n] = ~bad_s
    kmask = np.packbits(keep.reshape(-1, 64)[:, ::-1], axis=1).view(">u8").ravel().astype(np.uint64)
    sub = orc.Panel(full.bits[rows], full.valid[rows], full.mask_ctrl & kmask,
                    full.mask_case & kmask, n)
    tops = orc.gold_scan(sub, 50)
    tops["sA"] = rows[tops["sA"]]
    tops["sB"] = rows[tops["sB"]]
    expect = orc.report_lines(full, tops)
    lines = r.stderr.splitlines()
    start = lines.index(expect[1])
    assert lines[start - 1:start + 51] == expect
