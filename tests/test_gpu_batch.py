"""Batched scans (gepi_engine_scan_batch): every item's result must be
identical to the same scan run alone (Panel.scan on the fp4 path) and to the
oracle's gold scan, whatever the other items in the batch."""
import numpy as np
import pytest

import gepicount_amd as ge
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("sA", "sB", "fun", "w", "x", "y", "z", "obj")


def assert_same(got, ref, label=""):
    assert len(got) == len(ref), (label, len(got), len(ref))
    for k in FIELDS:
        if not np.array_equal(got[k], ref[k]):
            bad = int(np.flatnonzero(got[k] != ref[k])[0])
            raise AssertionError("%s: field %s differs first at rank %d: got %r ref %r" % (
                label, k, bad, got[bad], ref[bad]))


def make_mix(seed, n_items, vmax=300, nmax=1200):
    """Uneven scans: log-uniform variants and samples, case fraction U[0.2, 0.8],
    both objectives, several items sharing one panel with new phenotypes."""
    rng = np.random.default_rng(seed)
    panels, items = [], []
    for i in range(n_items):
        if panels and rng.random() < 0.3:
            pi = int(rng.integers(0, len(panels)))       # re-scan of a resident panel
        else:
            n_var = int(np.exp(rng.uniform(np.log(2), np.log(vmax))))
            n = int(np.exp(rng.uniform(np.log(20), np.log(nmax))))
            codes = orc.random_codes(rng, n_var, n, float(rng.choice([0.0, 0.02])))
            bits, valid = orc.planes_from_codes(codes)
            panels.append((bits, valid, n))
            pi = len(panels) - 1
        bits, valid, n = panels[pi]
        frac = rng.uniform(0.2, 0.8)
        cls = np.where(rng.random(n) < frac, 1, 0)
        cls[rng.random(n) < 0.05] = -1                   # excluded samples
        m0, m1 = orc.masks_from_classes(cls)
        R = bits.shape[0]
        total = 5 * R * (R + 1) // 2
        k = int(min(total, rng.choice([1, 17, 100, 1000, 5000])))
        items.append(dict(panel_index=pi, mask_ctrl=m0, mask_case=m1, topmost=k,
                          ffun=int(rng.random() < 0.25)))
    return panels, items


def run_mix(panels, items, engine):
    handles = [ge.Panel.from_planes(b, v, n) for b, v, n in panels]
    try:
        batch = [dict(it, panel=handles[it["panel_index"]]) for it in items]
        res, bstats = engine.scan_batch(batch)
        alone = [handles[it["panel_index"]].scan(it["mask_ctrl"], it["mask_case"],
                                                 topmost=it["topmost"], ffun=it["ffun"],
                                                 count_path=3)[0] for it in items]
    finally:
        for h in handles:
            h.close()
    return res, bstats, alone


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_items_equal_alone_and_oracle(seed):
    print("case batch mix seed=%d" % seed, flush=True)
    panels, items = make_mix(seed, 14)
    with ge.Engine() as eng:
        res, bstats, alone = run_mix(panels, items, eng)
    assert bstats["items_failed"] == 0
    assert bstats["rounds"] >= 1
    for i, (it, (got, st, status)) in enumerate(zip(items, res)):
        assert status == 0, (i, status)
        # the default schedule: Fisher on wave tiles, entropy below 48 K steps on whole tiles
        assert st["count_path"] == (3 if it["ffun"] == 1 else 4), (i, st["count_path"])
        assert_same(got, alone[i], "item %d vs alone" % i)
        bits, valid, n = panels[it["panel_index"]]
        ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                            it["topmost"], ffun=it["ffun"])
        assert_same(got, ref, "item %d vs oracle" % i)


def test_batch_repeated_on_one_engine_is_identical():
    panels, items = make_mix(5, 8)
    with ge.Engine() as eng:
        first, _, _ = run_mix(panels, items, eng)
        second, _, _ = run_mix(panels, items, eng)
    for i in range(len(items)):
        assert_same(second[i][0], first[i][0], "repeat item %d" % i)


def test_batch_overflow_recovery(monkeypatch):
    """Forced 8-tile bands and small candidate buffers make items overflow
    their pending lists and candidate buffers inside the batch."""
    panels, items = make_mix(9, 6, vmax=260, nmax=300)
    for it in items:
        it["topmost"] = 20000 if it["ffun"] == 0 else 2000
        it["ffun"] = 0 if it["ffun"] == 0 else 1
    monkeypatch.setenv("GEPI_CAND_CAP", "1")
    monkeypatch.setenv("GEPI_PEND_CAP", "4000000")
    monkeypatch.setenv("GEPI_BAND_TILES", "8")
    with ge.Engine() as eng:
        handles = [ge.Panel.from_planes(b, v, n) for b, v, n in panels]
        res, _ = eng.scan_batch([dict(it, panel=handles[it["panel_index"]]) for it in items])
        for h in handles:
            h.close()
    for v in ("GEPI_CAND_CAP", "GEPI_PEND_CAP", "GEPI_BAND_TILES"):
        monkeypatch.delenv(v)
    assert sum(st["cand_overflows"] for _, st, _ in res) > 0
    for i, (it, (got, st, status)) in enumerate(zip(items, res)):
        assert status == 0
        bits, valid, n = panels[it["panel_index"]]
        ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                            it["topmost"], ffun=it["ffun"])
        assert_same(got, ref, "overflow item %d" % i)


def test_batch_failing_and_degenerate_items_leave_others_intact():
    rng = np.random.default_rng(21)
    codes = orc.random_codes(rng, 40, 300, 0.01)
    bits, valid = orc.planes_from_codes(codes)
    cls = rng.choice([0, 1], size=300)
    m0, m1 = orc.masks_from_classes(cls)
    z = np.zeros_like(m0)
    with ge.Engine() as eng, ge.Panel.from_planes(bits, valid, 300) as p:
        items = [dict(panel=p, mask_ctrl=m0, mask_case=m1, topmost=500, ffun=0),
                 dict(panel=p, mask_ctrl=m0, mask_case=m1, topmost=500, ffun=7),   # bad objective
                 dict(panel=p, mask_ctrl=z, mask_case=z, topmost=300, ffun=0),     # no samples
                 dict(panel=p, mask_ctrl=m0, mask_case=m1, topmost=0, ffun=0),     # nothing wanted
                 dict(panel=p, mask_ctrl=m1, mask_case=m0, topmost=500, ffun=1)]
        res, bstats = eng.scan_batch(items)
    assert bstats["items_failed"] == 1
    assert res[1][2] == -1 and len(res[1][0]) == 0
    assert res[3][2] == 0 and len(res[3][0]) == 0
    for i in (0, 2, 4):
        it = items[i]
        ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], 300),
                            it["topmost"], ffun=it["ffun"])
        assert res[i][2] == 0
        assert_same(res[i][0], ref, "item %d" % i)


@pytest.mark.parametrize("per_cu", ["0", "1", "4"])
def test_batch_launch_forms_agree(per_cu, monkeypatch):
    """The batched count kernel's launch form (one tile per workgroup, or k
    persistent workgroups per CU taking tiles from the work counters) does
    not change any result."""
    print("case batch launch form per_cu=%s" % per_cu, flush=True)
    panels, items = make_mix(11, 10, vmax=400)
    monkeypatch.setenv("GEPI_F4_WG_PER_CU", per_cu)
    with ge.Engine() as eng:
        res, bstats, alone = run_mix(panels, items, eng)
    assert bstats["items_failed"] == 0
    for i, (got, st, status) in enumerate(res):
        assert status == 0, (i, status)
        assert_same(got, alone[i], "item %d (per_cu %s) vs alone" % (i, per_cu))


def test_batch_everything_survives_early():
    """Large top-k on small panels: the first bands admit every evaluation
    (no threshold yet), so staged survivors overflow the LDS stage and the
    pending list takes them one by one."""
    print("case batch heavy survivors", flush=True)
    rng = np.random.default_rng(21)
    panels, items = [], []
    for j in range(4):
        n_var, n = int(rng.integers(150, 260)), int(rng.integers(300, 900))
        codes = orc.random_codes(rng, n_var, n, 0.01)
        bits, valid = orc.planes_from_codes(codes)
        panels.append((bits, valid, n))
        m0, m1 = orc.masks_from_classes(np.where(rng.random(n) < 0.4, 1, 0))
        R = bits.shape[0]
        items.append(dict(panel_index=j, mask_ctrl=m0, mask_case=m1,
                          topmost=int(min(5 * R * (R + 1) // 2, 60000)), ffun=0))
    with ge.Engine() as eng:
        res, bstats, alone = run_mix(panels, items, eng)
    assert bstats["items_failed"] == 0
    for i, (got, st, status) in enumerate(res):
        assert status == 0, (i, status)
        assert_same(got, alone[i], "item %d vs alone" % i)
        bits, valid, n = panels[items[i]["panel_index"]]
        ref = orc.gold_scan(orc.Panel(bits, valid, items[i]["mask_ctrl"], items[i]["mask_case"], n),
                            items[i]["topmost"])
        assert_same(got, ref, "item %d vs oracle" % i)


def run_batch(panels, items, engine):
    handles = [ge.Panel.from_planes(b, v, n) for b, v, n in panels]
    try:
        return engine.scan_batch([dict(it, panel=handles[it["panel_index"]]) for it in items])
    finally:
        for h in handles:
            h.close()


def test_batch_graph_replays_equal_eager(monkeypatch):
    """Rounds replayed from the captured graphs (GEPI_ENGINE_GRAPH=1) against
    rounds launched one by one (the default): three batches in a row on one
    engine (6 items; then 12, which grows the descriptor tables and captures
    again; then 4, which leaves descriptor slots empty).  Every round is a
    replay, with the host's uploads and report reads between replays; each
    item must equal the eager run's result and the oracle's."""
    print("case batch graph replays", flush=True)
    mixes = [make_mix(s, n, vmax=260, nmax=900) for s, n in ((31, 6), (32, 12), (33, 4))]
    monkeypatch.delenv("GEPI_F4_WG_PER_CU", raising=False)
    monkeypatch.setenv("GEPI_ENGINE_GRAPH", "0")
    eager = []
    with ge.Engine() as eng:
        for panels, items in mixes:
            res, bs = run_batch(panels, items, eng)
            assert bs["graph_rounds"] == 0 and bs["items_failed"] == 0
            eager.append(res)
    monkeypatch.setenv("GEPI_ENGINE_GRAPH", "1")
    with ge.Engine() as eng:
        for j, (panels, items) in enumerate(mixes):
            res, bs = run_batch(panels, items, eng)
            assert bs["items_failed"] == 0
            assert bs["graph_rounds"] == bs["rounds"] >= 2, bs
            for i, it in enumerate(items):
                got, st, status = res[i]
                assert status == 0, (j, i, status)
                assert_same(got, eager[j][i][0], "batch %d item %d graph vs eager" % (j, i))
                bits, valid, n = panels[it["panel_index"]]
                ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                                    it["topmost"], ffun=it["ffun"])
                assert_same(got, ref, "batch %d item %d vs oracle" % (j, i))


def test_batch_autotune_and_tables(monkeypatch):
    """gepi_engine_autotune returns one schedule per size class from its
    candidate set; batches under the tuned table and under a hand-set table
    of extreme schedules (1 workgroup per CU, one tile per workgroup,
    64-tile bands) equal the default engine's results and the oracle's; bad tables
    are rejected and leave the engine usable."""
    print("case batch autotune", flush=True)
    monkeypatch.delenv("GEPI_F4_WG_PER_CU", raising=False)
    panels, items = make_mix(41, 8, vmax=900, nmax=1500)
    with ge.Engine() as eng:
        base, bs0 = run_batch(panels, items, eng)
        table = eng.autotune(repeats=1)
        print("autotune table", table, flush=True)
        assert [c["max_ksteps"] for c in table] == [2, 8, 32, 2**31 - 1]
        for c in table:
            assert c["wg_per_cu"] in (0, 1, 2, 3) and c["band_evals"] in (5e8, 2e9, 8e9), c
            assert c["tile"] in (0, 1, 2), c
            assert c["tile"] != 2 or c["wg_per_cu"] == 2, c
            assert 0 < c["ms_best"] <= c["ms_default"], c
        tuned, _ = run_batch(panels, items, eng)
        eng.set_tuning([dict(max_ksteps=8, wg_per_cu=1, tile=1, band_evals=1e6),
                        dict(max_ksteps=20, wg_per_cu=0, band_evals=3e6),
                        dict(max_ksteps=2**31 - 1, wg_per_cu=2, tile=1, band_evals=3e6)])
        extreme, bs2 = run_batch(panels, items, eng)
        assert bs2["rounds"] >= bs0["rounds"]
        for bad in ([dict(max_ksteps=8, wg_per_cu=2, band_evals=2e9)] * 2,
                    [dict(max_ksteps=8, wg_per_cu=-1, band_evals=2e9)],
                    [dict(max_ksteps=8, wg_per_cu=2, tile=3, band_evals=2e9)],
                    [dict(max_ksteps=8, wg_per_cu=2, band_evals=10.0)]):
            with pytest.raises(ge.GepiError):
                eng.set_tuning(bad)
        eng.set_tuning([])
        again, _ = run_batch(panels, items, eng)
    for i, it in enumerate(items):
        bits, valid, n = panels[it["panel_index"]]
        ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                            it["topmost"], ffun=it["ffun"])
        for label, res in (("default", base), ("tuned", tuned), ("extreme", extreme), ("reset", again)):
            assert res[i][2] == 0, (label, i, res[i][2])
            assert_same(res[i][0], ref, "%s item %d vs oracle" % (label, i))


@pytest.mark.parametrize("per_cu", [0, 2, 3])
def test_batch_half_tile_jobs(per_cu, monkeypatch):
    """Scans whose class runs half-tile jobs (32 x 32 variants, 16 x 16 per
    wave) next to scans on whole tiles, both objectives, so one round has up
    to four count launches; every item equals the oracle.  The last two
    variants also force overflows and replay captured-graph rounds."""
    print("case batch half tiles per_cu=%d" % per_cu, flush=True)
    monkeypatch.setenv("GEPI_F4_WG_PER_CU", str(per_cu))
    panels, items = make_mix(61 + per_cu, 12, vmax=700, nmax=1400)
    table = [dict(max_ksteps=6, wg_per_cu=per_cu, tile=1, band_evals=2e9),
             dict(max_ksteps=2**31 - 1, wg_per_cu=per_cu, tile=0, band_evals=2e9)]
    with ge.Engine() as eng:
        eng.set_tuning(table)
        res, _ = run_batch(panels, items, eng)
        if per_cu == 3:
            big = [dict(it, topmost=20000 if it["ffun"] == 0 else 2000) for it in items]
            monkeypatch.setenv("GEPI_CAND_CAP", "1")
            monkeypatch.setenv("GEPI_PEND_CAP", "4000000")
            monkeypatch.setenv("GEPI_BAND_TILES", "8")
            ovf, _ = run_batch(panels, big, eng)
            for v in ("GEPI_CAND_CAP", "GEPI_PEND_CAP", "GEPI_BAND_TILES"):
                monkeypatch.delenv(v)
            assert sum(st["cand_overflows"] for _, st, _ in ovf) > 0
    runs = [("half", res, items)] + ([("overflow", ovf, big)] if per_cu == 3 else [])
    if per_cu == 2:
        monkeypatch.setenv("GEPI_ENGINE_GRAPH", "1")
        with ge.Engine() as eng:
            eng.set_tuning(table)
            g1, bs = run_batch(panels, items, eng)
            g2, _ = run_batch(panels, items[::-1], eng)
        assert bs["graph_rounds"] > 0
        runs += [("graph", g1, items), ("graph reversed", g2[::-1], items)]
    for label, r, its in runs:
        for i, it in enumerate(its):
            bits, valid, n = panels[it["panel_index"]]
            ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                                it["topmost"], ffun=it["ffun"])
            assert r[i][2] == 0, (label, i, r[i][2])
            assert_same(r[i][0], ref, "%s item %d vs oracle" % (label, i))


def ksteps(m0, m1):
    n0 = int(sum(bin(int(w)).count("1") for w in np.asarray(m0, np.uint64)))
    n1 = int(sum(bin(int(w)).count("1") for w in np.asarray(m1, np.uint64)))
    return (n0 + 127) // 128 + (n1 + 127) // 128, n0, n1


@pytest.mark.parametrize("variant", ["mixed", "overflow"])
def test_batch_wave_tile_jobs(variant, monkeypatch):
    """Scans whose class runs wave-tile jobs (two 64 x 32-variant tiles per
    workgroup as four 32 x 32 wave tiles on the grouped layout, the count
    kernel of count path 3) next to scans on half and whole tiles, both
    objectives, fp32 and int32 filter counts (a class product n0 n1 above
    2^24); every item equals the same scan alone and the oracle.  The
    overflow variant forces 7-tile bands (odd: a unit's second tile lies past
    the band) and a small candidate buffer."""
    print("case batch wave tiles %s" % variant, flush=True)
    monkeypatch.delenv("GEPI_F4_WG_PER_CU", raising=False)
    monkeypatch.delenv("GEPI_ENGINE_GRAPH", raising=False)
    panels, items = make_mix(71 if variant == "mixed" else 72, 8, vmax=500, nmax=3000)
    rng = np.random.default_rng(73)
    for n_var, n, frac in ((300, 1500, 0.3), (90, 8400, 0.5)):
        codes = orc.random_codes(rng, n_var, n, 0.02)
        bits, valid = orc.planes_from_codes(codes)
        panels.append((bits, valid, n))
        for ffun in (0, 1):
            m0, m1 = orc.masks_from_classes(np.where(rng.random(n) < frac, 1, 0))
            items.append(dict(panel_index=len(panels) - 1, mask_ctrl=m0, mask_case=m1,
                              topmost=300 if ffun == 0 else 100, ffun=ffun))
    if variant == "overflow":
        items = [dict(it, topmost=20000 if it["ffun"] == 0 else 2000) for it in items]
    table = [dict(max_ksteps=2, wg_per_cu=2, tile=1, band_evals=2e9),
             dict(max_ksteps=4, wg_per_cu=2, tile=0, band_evals=2e9),
             dict(max_ksteps=2**31 - 1, wg_per_cu=2, tile=2, band_evals=2e9)]
    with ge.Engine() as eng:
        eng.set_tuning(table)
        if variant == "overflow":
            monkeypatch.setenv("GEPI_CAND_CAP", "1")
            monkeypatch.setenv("GEPI_BAND_TILES", "7")
        res, bs, alone = run_mix(panels, items, eng)
        for v in ("GEPI_CAND_CAP", "GEPI_BAND_TILES"):
            monkeypatch.delenv(v, raising=False)
    assert bs["items_failed"] == 0
    n_wave = 0
    for i, it in enumerate(items):
        got, st, status = res[i]
        assert status == 0, (i, status)
        ks, n0, n1 = ksteps(it["mask_ctrl"], it["mask_case"])
        wave = ks > 4 or it["ffun"] == 1          # Fisher runs wave tiles in any class
        assert st["count_path"] == (3 if wave else 4), (i, ks, st["count_path"])
        n_wave += ks > 4 and it["ffun"] == 0
        assert_same(got, alone[i], "%s item %d (K steps %d) vs alone" % (variant, i, ks))
        bits, valid, n = panels[it["panel_index"]]
        ref = orc.gold_scan(orc.Panel(bits, valid, it["mask_ctrl"], it["mask_case"], n),
                            it["topmost"], ffun=it["ffun"])
        assert_same(got, ref, "%s item %d (K steps %d, n0 %d, n1 %d) vs oracle"
                    % (variant, i, ks, n0, n1))
    assert n_wave >= 2
    if variant == "overflow":
        assert sum(st["cand_overflows"] for _, st, _ in res) > 0
