"""gepi_scan_noalloc: the scan of a resident panel with every buffer in the
library's fixed per-GPU scratch array (no device allocation).  Results must
equal gepi_scan's bit for bit, on consecutive calls that reuse the array
with a smaller input second, and the call must not change the device's free
memory."""
import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("sA", "sB", "fun", "w", "x", "y", "z", "obj")


def assert_same(got, ref, label):
    assert len(got) == len(ref), (label, len(got), len(ref))
    for k in FIELDS:
        if not np.array_equal(got[k], ref[k]):
            bad = int(np.flatnonzero(got[k] != ref[k])[0])
            raise AssertionError("%s: field %s differs first at rank %d: got %r ref %r" % (
                label, k, bad, got[bad], ref[bad]))


def free_bytes():
    """Free device memory from the HIP runtime the library itself uses."""
    import ctypes
    ge.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")  # already loaded by the library
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_equals_scan_on_consecutive_calls():
    print("case noalloc: 6000 x 2048 then 300 x 700", flush=True)
    big = synth.make_panel("medium", n_var=6000)
    small = synth.make_panel(spec=dict(n_var=300, n_samples=700, n_case=310,
                                       maf=("uniform", 0.01, 0.5), miss=0.02, planted=0, seed=77))
    # a scan on another panel first, so that the runtime's own lazily created
    # resources (staging buffers, module state) exist before free memory is read
    with ge.Panel.from_bed(small["payload"], small["n_var"], small["n_samples"]) as warm:
        m0, m1 = synth.masks_from_pheno(small["pheno"])
        warm.scan(m0, m1, topmost=10)
        warm.scan_noalloc(m0, m1, topmost=10)
    cases = [(big, 1000, 0), (big, 300, 1), (small, 200, 0), (small, 5000, 1)]
    for i, (p, k, ffun) in enumerate(cases):
        m0, m1 = synth.masks_from_pheno(p["pheno"])
        with ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"]) as h:
            got, st = h.scan_noalloc(m0, m1, topmost=k, ffun=ffun)
            # the fp4 kernel chosen by objective and size: wave tiles for
            # entropy scans of at least 10 K steps of 128 samples
            ks = -(-st["n_class0"] // 128) - (-st["n_class1"] // 128)
            assert st["count_path"] == (3 if ffun == 0 and ks >= 10 else 4), st
            assert st["bytes_allocated"] == 0, st
            ref, st_ref = h.scan(m0, m1, topmost=k, ffun=ffun)
            assert st_ref["bytes_allocated"] > 0, "a fresh panel's first scan grows its workspace"
            assert_same(got, ref, "case %d vs gepi_scan" % i)
            # a repeated call: the runtime's own lazily created resources
            # exist by now, so free device memory must not move at all
            before = free_bytes()
            again, st2 = h.scan_noalloc(m0, m1, topmost=k, ffun=ffun)
            after = free_bytes()
            assert before == after, ("free memory changed", i, before, after)
            assert st2["bytes_allocated"] == 0
            assert_same(again, ref, "case %d repeated" % i)
    # the smaller panel also against the oracle
    m0, m1 = synth.masks_from_pheno(small["pheno"])
    with ge.Panel.from_bed(small["payload"], small["n_var"], small["n_samples"]) as h:
        got, _ = h.scan_noalloc(m0, m1, topmost=500)
        bits, valid = h.download()
    ref = orc.gold_scan(orc.Panel(bits, valid, m0, m1, small["n_samples"]), 500)
    assert_same(got, ref, "small vs oracle")


def test_noalloc_rejects_what_does_not_fit():
    print("case noalloc limits", flush=True)
    rng = np.random.default_rng(5)
    n_var, n = 100000, 2048
    payload = rng.integers(0, 256, size=(n_var, n // 4), dtype=np.uint8)
    pheno = synth.pheno_for(n, 1024, 9)
    m0, m1 = synth.masks_from_pheno(pheno)
    with ge.Panel.from_bed(payload, n_var, n) as h:
        with pytest.raises(ge.GepiError) as e:
            h.scan_noalloc(m0, m1, topmost=100)
        assert e.value.code == -3  # GEPI_ERR_LIMIT
        assert "scratch" in str(e.value)
    small = synth.make_panel("small", n_var=40)
    m0, m1 = synth.masks_from_pheno(small["pheno"])
    with ge.Panel.from_bed(small["payload"], small["n_var"], small["n_samples"]) as h:
        with pytest.raises(ge.GepiError) as e:
            h.scan_noalloc(m0, m1, topmost=100, count_path=2)
        assert e.value.code == -1 and "fp4" in str(e.value)  # GEPI_ERR_ARG
        # the array is usable after a rejected call
        got, _ = h.scan_noalloc(m0, m1, topmost=100)
        ref, _ = h.scan(m0, m1, topmost=100)
        assert_same(got, ref, "after rejection")
