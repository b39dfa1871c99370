"""Writes tests/golden/large_entropy_top10000.npz: the oracle's gold scan of
the whole large workload (100,000 variants x 8,192 samples, 2,458 cases /
5,734 controls, 2% missing; 2e10 row pairs, 1e11 evaluations), entropy
objective, top-10,000, in canonical key order.

The panel is regenerated from its seed by gepicount_amd.synth.make_panel,
written as a PLINK fileset and read back by the oracle's reader; the fixture
stores the SHA-256 of the .bed payload and of the phenotype vector so that a
test can check that its regenerated panel is the same one.

The oracle's sources are compiled here for the local CPU (-O3 -march=native,
same -ffp-contract=off; the counts are integers and the score uses only IEEE
+, -, / and a binary128-built table, so the bits do not depend on the
build) into a temporary directory.  The scan runs in row ranges of equal
work; each range's top-10,000 is kept in a checkpoint directory, so an
interrupted run resumes.  The top-10,000 of the whole scan is the top-10,000
of the union of the ranges' lists because the key is total.

Usage (from the repository root):
    python tests/golden/make_large_gold.py [--work DIR] [--chunks 48] [--threads 0]
About 15 minutes on 8 cores with AVX-512.
"""
import argparse
import hashlib
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "large_entropy_top10000.npz")
TOPK = 10000


def build_native_oracle(dst):
    src = os.path.join(ROOT, "oracle")
    lib = os.path.join(dst, "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared",
           "-o", lib] + [os.path.join(src, f) for f in
                         ("orc_plink.c", "orc_score.c", "orc_gold.c", "orc_report.c")] + \
        ["-lquadmath", "-lm"]
    subprocess.run(cmd, check=True)
    return lib


def order_key(obj):
    b = np.asarray(obj, np.float64).view(np.uint64)
    return np.where(b >> np.uint64(63), ~b, b | np.uint64(1 << 63))


def canonical_sort(res):
    idx = (res["sA"].astype(np.uint64) << np.uint64(32)) | \
        (res["sB"].astype(np.uint64) << np.uint64(3)) | res["fun"].astype(np.uint64)
    o = np.lexsort((idx, order_key(res["obj"])))
    return res[o]


def panel_hashes(p):
    return (hashlib.sha256(np.ascontiguousarray(p["payload"]).tobytes()).hexdigest(),
            hashlib.sha256(np.asarray(p["pheno"], np.int64).tobytes()).hexdigest())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--work", default=os.path.join(tempfile.gettempdir(), "gepi_large_gold"))
    ap.add_argument("--chunks", type=int, default=48)
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    os.makedirs(a.work, exist_ok=True)
    os.environ["GEPI_ORACLE_LIB"] = build_native_oracle(a.work)
    sys.path.insert(0, ROOT)
    from gepicount_amd import synth
    from oracle import oracle as orc

    t0 = time.time()
    p = synth.make_panel("large")
    bed_sha, pheno_sha = panel_hashes(p)
    print("panel generated (%.1f s), bed sha256 %s" % (time.time() - t0, bed_sha), flush=True)
    root = os.path.join(a.work, "large")
    synth.write_plink(root, p)
    pan = orc.read_plink(root)
    R = pan.n_rows
    # row ranges of equal pair counts: rows [0, j) hold j(j+1)/2 pairs
    bounds = sorted({int(round(R * np.sqrt(i / a.chunks))) for i in range(a.chunks + 1)})
    parts = []
    for c in range(len(bounds) - 1):
        j0, j1 = bounds[c], bounds[c + 1]
        ck = os.path.join(a.work, "rows_%d_%d_%s.npy" % (j0, j1, bed_sha[:12]))
        if not os.path.exists(ck):
            t = time.time()
            r = orc.gold_scan_rows(pan, j0, j1, TOPK, ffun=0, nthreads=a.threads)
            np.save(ck + ".tmp.npy", r)
            os.replace(ck + ".tmp.npy", ck)
            print("rows [%d, %d): %.1f s" % (j0, j1, time.time() - t), flush=True)
        parts.append(np.load(ck))
    allr = canonical_sort(np.concatenate(parts))[:TOPK]
    np.savez_compressed(
        OUT, sA=allr["sA"].astype(np.int32), sB=allr["sB"].astype(np.int32),
        fun=allr["fun"].astype(np.uint8), w=allr["w"].astype(np.int16),
        x=allr["x"].astype(np.int16), y=allr["y"].astype(np.int16),
        z=allr["z"].astype(np.int16), obj_bits=allr["obj"].view(np.uint64),
        bed_sha256=np.array(bed_sha), pheno_sha256=np.array(pheno_sha),
        workload=np.array("large"), topk=np.array(TOPK))
    print("wrote %s (%d records, %.0f s)" % (OUT, len(allr), time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
