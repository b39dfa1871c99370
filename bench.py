"""Benchmark: all-pairs SNP x Boolean-function scan throughput on one MI355X.

A step is one complete scan of a panel already resident in HBM: class layout,
every count-kernel band, top-k maintenance and the final extraction of the
top-k to the host.  The workload is the 100,000-variant x 8,192-sample panel
(2,458 cases, 2% missing, top-10,000, entropy objective).

Prints one JSON line (rank 0, N = 1):
  value       evals/s = 5 * R(R+1)/2 / mean step seconds, R = 2 x variants
  roofline    the count kernel: algorithmic ops per launch / mean launch time
              (HIP events on the scan stream) against the MI355X dense fp4
              MFMA peak (fp4 path, the default), the dense int8 MFMA peak
              (int8 path) or the gfx950 VALU peak (popcount path)
  cpu_baseline  the oracle's OpenMP gold scan on a bounded leading subset of
              the same panel, on this host's cores
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SNP-pair x Boolean-fn evals/s at 100k SNPs x 8k samples; p50/p99 scan latency"
# gfx950: 256 CUs x 64 lanes x 2.4 GHz of 32-bit VALU lane operations
VALU_PEAK_OPS = 256 * 64 * 2.4e9
# MI355X dense int8 matrix peak
INT8_MFMA_PEAK_OPS = 5.0e15
# MI355X dense fp4 (e2m1) matrix peak
FP4_MFMA_PEAK_FLOPS = 10.1e15


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="large",
                    choices=["small", "medium", "large", "mix64", "serve", "cpuref"],
                    help="mix64: 64 uneven scans, batched (gepi_engine) against sequential; "
                         "serve: Poisson arrivals on the serving loop (gepi_server); "
                         "cpuref: the reference-style one-thread CPU gold on the small "
                         "workload and the medium workload's first 2,000 variants")
    ap.add_argument("--qps", type=float, default=40.0, help="serve: arrival rate")
    ap.add_argument("--requests", type=int, default=200, help="serve: requests in the run")
    ap.add_argument("--n-var", type=int, default=None, help="override variant count")
    ap.add_argument("--topk", type=int, default=10000)
    ap.add_argument("--ffun", type=int, default=0)
    ap.add_argument("--autotune", type=int, default=1,
                    help="mix64 / serve: run the batch autotuner before the timed region")
    ap.add_argument("--count-path", type=int, default=0)
    ap.add_argument("--cli", type=int, default=1,
                    help="large/medium/small: also time the command line end to end")
    ap.add_argument("--cpu-seconds", type=float, default=25.0,
                    help="target CPU time of the cpu_baseline sample (0 = skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (counter scale)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size, so that HBM byte counters
    (FETCH_SIZE / WRITE_SIZE) can be scaled: it reads and writes mib MiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_mib": mib, "bytes_each_way": mib << 20}))


def class_words(n):
    return ((n + 127) // 128) * 4


def cpu_baseline(panel, m0, m1, n_samples, target_s, topk, ffun):
    from oracle import oracle as orc
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    bits, valid = panel.download()

    def run(v):
        sub = orc.Panel(bits[:2 * v], valid[:2 * v], m0, m1, n_samples)
        t = time.perf_counter()
        orc.gold_scan(sub, topk, ffun=ffun, nthreads=threads)
        dt = time.perf_counter() - t
        R = 2 * v
        return 5 * R * (R + 1) // 2, dt

    v = 400
    ev, dt = run(v)
    # work grows with v^2: pick the subset that takes about target_s
    v2 = int(v * (target_s / max(dt, 1e-3)) ** 0.5)
    v2 = max(v, min(v2, bits.shape[0] // 2))
    ev, dt = run(v2)
    out = {"value": ev / dt, "unit": "evals/s", "cores": threads, "kind": "port",
           "sample": "oracle gold scan (OpenMP, exact scores, top-%d) of the first %d variants "
                     "(%d evals) of the same panel, %.1f s" % (topk, v2, ev, dt)}
    if ffun == 0:
        out["reference_style"] = reference_style(orc, bits, valid, m0, m1, n_samples, topk,
                                                 target_s / 3.0)
    return out


def reference_style(orc, bits, valid, m0, m1, n_samples, topk, target_s):
    """The reference's own CPU gold as it runs there (one thread, 16-bit
    table popcounts, float32 scores, src/proc_gold_1.h:68-178) over the
    leading rows of the panel, extrapolated to the whole scan the way its
    progress estimate does (src/proc_gold_1.h:140-153): the rows j < J hold
    J(J+1)/2 of the R(R+1)/2 row pairs."""
    R = bits.shape[0]
    panel = orc.Panel(bits, valid, m0, m1, n_samples)
    J = 64
    _, sec = orc.gold_scan_ref_f32(panel, topk, J)
    J = int(min(R, J * (target_s / max(sec, 1e-3)) ** 0.5))
    _, sec = orc.gold_scan_ref_f32(panel, topk, J)
    done = J * (J + 1) // 2
    total_s = sec * (R * (R + 1) / 2) / done
    return {"value": 5 * R * (R + 1) / 2 / total_s, "unit": "evals/s", "cores": 1,
            "kind": "port",
            "sample": "single-thread gold scan as the reference runs it (16-bit table popcounts, "
                      "float32 scores) over the first %d of %d rows (%d row pairs) in %.1f s; "
                      "whole scan extrapolated to %.0f s by the pair count" % (
                          J, R, done, sec, total_s)}


def run_mix(args):
    """64 uneven scans (variants logU[500, 50,000], samples logU[200, 5,000],
    case fraction U[0.2, 0.8], top-100 each), panels resident: one batch on
    the engine against the same scans run one after another.  A step is one
    pass over all 64 scans."""
    import numpy as np

    import gepicount_amd as ge
    from gepicount_amd import synth

    t = time.perf_counter()
    specs = synth.mix_specs()
    panels, items, evals = [], [], 0
    for sp in specs:
        p = synth.make_panel(spec=sp)
        m0, m1 = synth.masks_from_pheno(p["pheno"])
        h = ge.Panel.from_bed(p["payload"], p["n_var"], p["n_samples"])
        panels.append(h)
        items.append(dict(panel=h, mask_ctrl=m0, mask_case=m1, topmost=sp["topk"], ffun=args.ffun))
        R = 2 * p["n_var"]
        evals += 5 * R * (R + 1) // 2
    gen_s = time.perf_counter() - t
    eng = ge.Engine()
    tune, tune_s = [], 0.0
    if args.autotune:
        t = time.perf_counter()
        tune = eng.autotune(repeats=2)
        tune_s = time.perf_counter() - t

    def sequential():
        return [it["panel"].scan(it["mask_ctrl"], it["mask_case"], topmost=it["topmost"],
                                 ffun=it["ffun"])[0] for it in items]

    for _ in range(args.warmup):
        sequential()
        eng.scan_batch(items)
    seq_t, bat_t, bstats, lat = [], [], [], []
    for _ in range(args.steps):
        ts = time.perf_counter()
        seq = sequential()
        seq_t.append(time.perf_counter() - ts)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        res, bs = eng.scan_batch(items)
        bat_t.append(time.perf_counter() - ts)
        bstats.append(bs)
        lat.extend(st["ms_scan"] for _, st, _ in res)
    total = time.perf_counter() - t0
    identical = all(np.array_equal(r[0], q) for r, q in zip(res, seq))
    k_ms = statistics.mean(b["ms_count_kernels"] for b in bstats)
    var_pairs_samples = sum((sp["n_var"] * (sp["n_var"] + 1) // 2) * sp["n_samples"] for sp in specs)
    ops = var_pairs_samples * 9 * 2.0
    lat = sorted(lat)
    line = {
        "metric": "SNP-pair x Boolean-fn evals/s, 64 uneven scans batched", "value": evals / statistics.mean(bat_t),
        "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * total / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp4 (e2m1, fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "mix64: 64 scans, variants logU[500, 50000], samples logU[200, 5000], "
                               "case fraction U[0.2, 0.8], top-100 each, %s" %
                               ("Fisher" if args.ffun else "entropy"),
                   "scans": len(items), "evals": evals},
        "sequential_ms": 1000.0 * statistics.mean(seq_t),
        "batched_ms": 1000.0 * statistics.mean(bat_t),
        "speedup_vs_sequential": statistics.mean(seq_t) / statistics.mean(bat_t),
        "results_identical_to_sequential": identical,
        "rounds": bstats[-1]["rounds"], "graph_rounds": bstats[-1]["graph_rounds"],
        "autotune": {"s": tune_s, "classes": [{k: c[k] for k in ("max_ksteps", "tile", "wg_per_cu",
                                                                 "band_evals", "ms_best", "ms_default")}
                                              for c in tune]} if args.autotune else None,
        "scan_latency_in_batch_ms": {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1,
                                                                         int(0.99 * len(lat)))]},
        "roofline": {"bound": "mfma", "kernel": " + ".join(
                         {3: "k_count_f4w_batch", 4: "k_count_f4_batch"}[cp]
                         for cp in sorted(set(st["count_path"] for _, st, _ in res))),
                     "achieved": ops / (k_ms / 1000.0) / 1e12, "peak": FP4_MFMA_PEAK_FLOPS / 1e12,
                     "unit": "TFLOP/s", "frac": ops / (k_ms / 1000.0) / FP4_MFMA_PEAK_FLOPS,
                     "traffic": None, "count_kernel_ms": k_ms},
        "cpu_baseline": None, "gen_s": gen_s,
    }
    eng.close()
    for h in panels:
        h.close()
    print(json.dumps(line))


def run_serve(args):
    """Serving loop under Poisson arrivals at a fixed rate: 80% of the requests
    are scans of the uneven mix (its panels and phenotypes, top-100), 20%
    re-scan a resident 50,000 x 4,000 panel with a new phenotype vector
    (case fraction U[0.2, 0.8], top-100).  Panels are resident before the
    run; latency is submit -> completion as the server measures it."""
    import threading

    import numpy as np

    import gepicount_amd as ge
    from gepicount_amd import synth

    t = time.perf_counter()
    specs = synth.mix_specs()
    big_spec = dict(n_var=50000, n_samples=4000, n_case=2000, maf=("uniform", 0.01, 0.5),
                    miss=0.01, planted=0, seed=1005)
    t = time.perf_counter()
    srv = ge.Server(autotune=bool(args.autotune))
    tune_s = time.perf_counter() - t
    mix = []
    for sp in specs:
        p = synth.make_panel(spec=sp)
        pid = srv.add_panel_bed(p["payload"], p["n_var"], p["n_samples"])
        m0, m1 = synth.masks_from_pheno(p["pheno"])
        mix.append((pid, m0, m1, p["n_var"], p))
    big = synth.make_panel(spec=big_spec)
    big_id = srv.add_panel_bed(big["payload"], big["n_var"], big["n_samples"])
    gen_s = time.perf_counter() - t - tune_s
    rng = np.random.Generator(np.random.PCG64(1006))
    arrivals = np.cumsum(rng.exponential(1.0 / args.qps, size=args.requests))
    reqs = []
    for i in range(args.requests):
        if rng.random() < 0.2:
            n = big_spec["n_samples"]
            pheno = synth.pheno_for(n, int(round(rng.uniform(0.2, 0.8) * n)), 2000 + i)
            m0, m1 = synth.masks_from_pheno(pheno)
            reqs.append((big_id, m0, m1, big_spec["n_var"], "rescan", None))
        else:
            j = int(rng.integers(0, len(mix)))
            pid, m0, m1, nv, p = mix[j]
            reqs.append((pid, m0, m1, nv, "mix", p))
    # warm the server (first launches, allocator) with a few requests
    for pid, m0, m1, _, _, _ in reqs[:4]:
        srv.wait(srv.submit(pid, m0, m1, topmost=100, ffun=args.ffun))
    base = srv.stats()
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def submitter():
        for i, (pid, m0, m1, _, _, _) in enumerate(reqs):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(pid, m0, m1, topmost=100, ffun=args.ffun)

    th = threading.Thread(target=submitter)
    th.start()
    results = [None] * len(reqs)
    for i in range(len(reqs)):
        while tickets[i] is None:
            time.sleep(0.0005)
        results[i] = srv.wait(tickets[i])
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat = sorted(r[1]["ms_scan"] for r in results)
    evals = sum(r[1]["evals"] for r in results)
    # saturated: every request submitted at once (the server's own limit)
    t0 = time.perf_counter()
    sat_tickets = [srv.submit(pid, m0, m1, topmost=100, ffun=args.ffun)
                   for pid, m0, m1, _, _, _ in reqs]
    sat_results = [srv.wait(t) for t in sat_tickets]
    sat_wall = time.perf_counter() - t0
    srv.close()
    # one at a time: the same requests as single scans of resident panels,
    # one after another; every response (scans of the mix and re-scans of the
    # resident panel) must equal the served ones
    handles = {}
    for pid, m0, m1, _, kind, p in reqs:
        if pid not in handles:
            src = big if kind == "rescan" else p
            handles[pid] = ge.Panel.from_bed(src["payload"], src["n_var"], src["n_samples"])
    t0 = time.perf_counter()
    alone = [handles[pid].scan(m0, m1, topmost=100, ffun=args.ffun)[0]
             for pid, m0, m1, _, _, _ in reqs]
    seq_wall = time.perf_counter() - t0
    identical = all(np.array_equal(a, r[0]) and np.array_equal(a, q[0])
                    for a, r, q in zip(alone, results, sat_results))
    checked = len(alone)
    for h in handles.values():
        h.close()
    line = {
        "metric": "SNP-pair x Boolean-fn evals/s served under Poisson arrivals; p50/p99 request latency",
        "value": evals / wall, "unit": "evals/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4,
        "ms_per_step": 1000.0 * wall / len(reqs), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp4 (e2m1, fp32 accumulate)", "data": "synthetic",
        "config": {"workload": "serve: Poisson %.0f QPS, 80%% scans of the uneven 64-scan mix, "
                               "20%% re-scans of a resident 50000 x 4000 panel, top-100, %s"
                               % (args.qps, "Fisher" if args.ffun else "entropy"),
                   "requests": len(reqs), "qps": args.qps},
        "latency_ms": {"p50": lat[len(lat) // 2], "p99": lat[min(len(lat) - 1, int(0.99 * len(lat)))],
                       "mean": sum(lat) / len(lat), "max": lat[-1]},
        "achieved_qps": len(reqs) / wall,
        "gpu_busy_frac": (st["busy_ms"] - base["busy_ms"]) / (1000.0 * wall),
        "rounds": st["rounds"] - base["rounds"], "max_active": st["max_active"],
        "failed": st["failed"], "responses_identical_to_alone": identical,
        "responses_checked": checked,
        "saturated": {"evals_per_s": evals / sat_wall, "wall_s": sat_wall,
                      "note": "all requests submitted at once"},
        "one_at_a_time": {"evals_per_s": evals / seq_wall, "wall_s": seq_wall,
                          "note": "the same requests as single scans, one after another"},
        "served_vs_one_at_a_time": seq_wall / sat_wall,
        "roofline": None, "cpu_baseline": None, "gen_s": gen_s,
        "autotune_s": tune_s if args.autotune else None,
    }
    print(json.dumps(line))


def cli_wall(p, args):
    """End-to-end wall time of the gepicount command line on the same panel
    written as a PLINK fileset: .bed/.bim/.fam reading, upload and decode,
    QA report, scan and the printed top-k report (stderr, discarded)."""
    import shutil
    import subprocess
    import tempfile

    from gepicount_amd import synth
    exe = os.path.join(ROOT, "gepicount_amd", "gepicount")
    if not os.path.exists(exe):
        return None
    d = tempfile.mkdtemp(prefix="gepi_bench_")
    try:
        root = os.path.join(d, "panel")
        synth.write_plink(root, p)
        cmd = [exe, "-i", root, "-N", str(args.topk)] + (["-f"] if args.ffun else [])
        t = time.perf_counter()
        r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
        wall = time.perf_counter() - t
        if r.returncode != 0:
            return {"error": "exit status %d" % r.returncode}
        return {"wall_s": wall, "command": "gepicount -i <prefix> -N %d%s" % (
            args.topk, " -f" if args.ffun else "")}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def run_cpuref(args):
    """The reference's CPU gold as it runs there (one thread, 16-bit table
    popcounts, float32 scores) on the small workload (1,024 variants x 256
    samples, whole scan, top-100) and on the first 2,000 variants of the
    medium workload (whole subset scan, top-1,000), the latter extrapolated
    to all 20,000 variants by the row-pair count.  CPU only."""
    import shutil
    import tempfile

    from gepicount_amd import synth
    from oracle import oracle as orc
    out = {"metric": "reference-style CPU gold scan, one thread", "unit": "evals/s", "cores": 1,
           "kind": "port", "runs": []}
    d = tempfile.mkdtemp(prefix="gepi_cpuref_")
    try:
        for name, n_var, topk, full_var in (("small", None, 100, 1024), ("medium", 2000, 1000, 20000)):
            p = synth.make_panel(name, n_var=n_var)
            root = os.path.join(d, name)
            synth.write_plink(root, p)
            pan = orc.read_plink(root)
            _, sec = orc.gold_scan_ref_f32(pan, topk)
            R = pan.n_rows
            ev = 5 * R * (R + 1) // 2
            Rf = 2 * full_var
            out["runs"].append({"workload": name, "variants": pan.n_rows // 2,
                                "samples": pan.n_samples, "topk": topk, "seconds": sec,
                                "evals_per_s": ev / sec,
                                "extrapolated_full_s": sec * (Rf * (Rf + 1) / 2) / (R * (R + 1) / 2)})
    finally:
        shutil.rmtree(d, ignore_errors=True)
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.workload == "cpuref":
        run_cpuref(args)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine runs on one GPU per process and "
              "has no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "mix64":
        run_mix(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    import numpy as np

    import gepicount_amd as ge
    from gepicount_amd import synth

    t = time.perf_counter()
    p = synth.make_panel(args.workload, n_var=args.n_var)
    gen_s = time.perf_counter() - t
    n_var, n = p["n_var"], p["n_samples"]
    m0, m1 = synth.masks_from_pheno(p["pheno"])
    n0, n1 = int(np.sum(p["pheno"] == 1)), int(np.sum(p["pheno"] == 2))
    R = 2 * n_var
    evals = 5 * R * (R + 1) // 2
    panel = ge.Panel.from_bed(p["payload"], n_var, n)

    for _ in range(args.warmup):
        panel.scan(m0, m1, topmost=args.topk, ffun=args.ffun, count_path=args.count_path)
    times, stats = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        res, st = panel.scan(m0, m1, topmost=args.topk, ffun=args.ffun,
                             count_path=args.count_path)
        times.append(time.perf_counter() - ts)
        stats.append(st)
    total = time.perf_counter() - t0
    ms_per_step = 1000.0 * total / args.steps

    # count kernel roofline.  All paths form 9 plane products per variant pair
    # over the n class samples.  Popcount: 2 VALU ops (and, bcnt) per 32-bit
    # word of each product.  MFMA: 9 x n MACs = 18 n ops per variant pair
    # (int8 operands, or fp4 operands with fp32 accumulation).
    var_pairs = n_var * (n_var + 1) // 2
    path = stats[-1]["count_path"]
    k_ms = statistics.mean(s["ms_count_kernels"] for s in stats)
    launches = statistics.mean(s["bands"] for s in stats)
    if path in (3, 4):
        ops_total = var_pairs * 9 * 2.0 * (n0 + n1)
        peak, bound, unit, kernel = (FP4_MFMA_PEAK_FLOPS, "mfma", "TFLOP/s",
                                     "k_count_f4w" if path == 3 else "k_count_f4")
    elif path == 2:
        ops_total = var_pairs * 9 * 2.0 * (n0 + n1)
        peak, bound, unit, kernel = INT8_MFMA_PEAK_OPS, "mfma", "TOP/s", "k_count_mfma"
    else:
        ops_total = var_pairs * 9 * 2 * ((n0 + n1) / 32.0)
        peak, bound, unit, kernel = VALU_PEAK_OPS, "valu", "Tops/s", "k_count_popc"
    achieved = ops_total / (k_ms / 1000.0)
    roof = {"bound": bound, "kernel": kernel, "achieved": achieved / 1e12, "peak": peak / 1e12,
            "unit": unit, "frac": achieved / peak, "traffic": None,
            "ops_per_launch": ops_total / launches, "launch_ms": k_ms / launches,
            "launches": launches,
            "valu_popcount_roofline_frac": (R * (R + 1) / 2) * (n / 4.0) / (k_ms / 1000.0)
            / VALU_PEAK_OPS}
    prof = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (roof["kernel"], args.workload))
    if os.path.exists(prof):
        roof["traffic"] = json.load(open(prof)).get("bytes_per_launch")

    cpu = None
    if args.cpu_seconds > 0:
        cpu = cpu_baseline(panel, m0, m1, n, args.cpu_seconds, args.topk, args.ffun)
    panel.close()
    cli = cli_wall(p, args) if args.cli else None

    lat = sorted(times)
    line = {
        "metric": METRIC, "value": evals / statistics.mean(times), "unit": "evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {3: "fp4 (e2m1, fp32 accumulate)", 2: "i8"}.get(path, "u32"),
        "data": "synthetic",
        "config": {"workload": "%s: %d variants x %d samples (%d cases / %d controls), top-%d, %s"
                               % (args.workload, n_var, n, n1, n0, args.topk,
                                  "entropy" if args.ffun == 0 else "fisher"),
                   "variants": n_var, "samples": n, "row_pairs": R * (R + 1) // 2,
                   "evals": evals, "topk": args.topk, "count_path": stats[-1]["count_path"]},
        "p50_ms": 1000.0 * lat[len(lat) // 2], "p99_ms": 1000.0 * lat[min(len(lat) - 1,
                                                                           int(0.99 * len(lat)))],
        "median_ms": 1000.0 * statistics.median(times),
        "cli": cli,
        "variant_pairs_per_s": (n_var * (n_var - 1) // 2) / statistics.mean(times),
        "exact_evals": stats[-1]["exact_evals"], "gen_s": gen_s,
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
