"""Serving loop (gepi_server_*): requests submitted while others run join the
batch at the next round; each result must be identical to the same scan run
alone and to the oracle."""
import threading
import time

import numpy as np
import pytest

import gepicount_amd as ge
from gepicount_amd import synth
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


def make_panels(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n_var = int(rng.integers(3, 200))
        n = int(rng.integers(30, 900))
        codes = orc.random_codes(rng, n_var, n, 0.02)
        bits, valid = orc.planes_from_codes(codes)
        out.append(dict(payload=synth.pack_codes(codes), n_var=n_var, n=n, bits=bits, valid=valid))
    return out


def make_requests(seed, panels, count):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        pi = int(rng.integers(0, len(panels)))
        n = panels[pi]["n"]
        cls = np.where(rng.random(n) < rng.uniform(0.2, 0.8), 1, 0)
        cls[rng.random(n) < 0.05] = -1
        m0, m1 = orc.masks_from_classes(cls)
        reqs.append(dict(pi=pi, m0=m0, m1=m1, k=int(rng.choice([5, 100, 2000])),
                         ffun=int(rng.random() < 0.25)))
    return reqs


def check_against_alone_and_oracle(panels, reqs, got):
    for i, r in enumerate(reqs):
        p = panels[r["pi"]]
        res, st, status = got[i]
        assert status == 0, (i, status)
        with ge.Panel.from_bed(p["payload"], p["n_var"], p["n"]) as h:
            alone, _ = h.scan(r["m0"], r["m1"], topmost=r["k"], ffun=r["ffun"], count_path=3)
        assert_same(res, alone, "request %d vs alone" % i)
        ref = orc.gold_scan(orc.Panel(p["bits"], p["valid"], r["m0"], r["m1"], p["n"]), r["k"],
                            ffun=r["ffun"])
        assert_same(res, ref, "request %d vs oracle" % i)


def test_server_concurrent_clients():
    """Three client threads submit and wait concurrently."""
    print("case server concurrent clients", flush=True)
    panels = make_panels(3, 4)
    reqs = make_requests(4, panels, 24)
    got = [None] * len(reqs)
    with ge.Server() as srv:
        ids = [srv.add_panel_bed(p["payload"], p["n_var"], p["n"]) for p in panels]

        def client(lo, hi):
            tickets = [(i, srv.submit(ids[reqs[i]["pi"]], reqs[i]["m0"], reqs[i]["m1"],
                                      topmost=reqs[i]["k"], ffun=reqs[i]["ffun"]))
                       for i in range(lo, hi)]
            for i, t in tickets:
                got[i] = srv.wait(t)

        th = [threading.Thread(target=client, args=(j * 8, j * 8 + 8)) for j in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    check_against_alone_and_oracle(panels, reqs, got)


def test_server_continuous_batching():
    """Requests arrive in three waves between rounds; later ones join the
    running batch.  Duplicate panel bytes map to one resident panel."""
    print("case server continuous batching", flush=True)
    panels = make_panels(3, 4)
    reqs = make_requests(4, panels, 24)
    got = [None] * len(reqs)
    with ge.Server() as srv:
        ids = [srv.add_panel_bed(p["payload"], p["n_var"], p["n"]) for p in panels]
        assert srv.add_panel_bed(panels[0]["payload"], panels[0]["n_var"], panels[0]["n"]) == ids[0]
        assert srv.stats()["panels"] == len(set(ids))
        tickets = []
        for wave in range(3):
            for i in range(wave * 8, wave * 8 + 8):
                r = reqs[i]
                tickets.append((i, srv.submit(ids[r["pi"]], r["m0"], r["m1"], topmost=r["k"],
                                              ffun=r["ffun"])))
            time.sleep(0.02)
        order = np.random.default_rng(8).permutation(len(tickets))
        for j in order:
            i, t = tickets[j]
            got[i] = srv.wait(t)
        st = srv.stats()
    assert st["submitted"] == len(reqs) and st["completed"] == len(reqs) and st["failed"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    check_against_alone_and_oracle(panels, reqs, got)


def test_server_tight_budget_one_request_at_a_time():
    panels = make_panels(5, 2)
    reqs = make_requests(6, panels, 6)
    with ge.Server(hbm_budget=1) as srv:
        ids = [srv.add_panel_bed(p["payload"], p["n_var"], p["n"]) for p in panels]
        tickets = [srv.submit(ids[r["pi"]], r["m0"], r["m1"], topmost=r["k"], ffun=r["ffun"])
                   for r in reqs]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["max_active"] == 1
    check_against_alone_and_oracle(panels, reqs, got)


def test_server_errors_and_panel_release():
    panels = make_panels(7, 1)
    p = panels[0]
    m0, m1 = orc.masks_from_classes(np.arange(p["n"]) % 2)
    with ge.Server() as srv:
        pid = srv.add_panel_bed(p["payload"], p["n_var"], p["n"])
        with pytest.raises(ge.GepiError) as e:
            srv.submit(pid + 1 if pid + 1 else 2, m0, m1)
        assert e.value.code == -1
        bad = srv.submit(pid, m0, m1, topmost=10, ffun=9)
        good = srv.submit(pid, m0, m1, topmost=10)
        res, _, status = srv.wait(bad)
        assert status == -1 and len(res) == 0
        res, _, status = srv.wait(good)
        assert status == 0 and len(res) == 10
        srv.release_panel(pid)
        assert srv.stats()["panels"] == 0
        with pytest.raises(ge.GepiError):
            srv.release_panel(pid)


def test_server_key_collision_keeps_panels_apart(monkeypatch):
    """With every payload forced onto one content key, distinct payloads of
    one shape still get distinct resident panels (the bytes are compared on
    a key hit), and identical bytes still share one."""
    print("case server key collision", flush=True)
    monkeypatch.setenv("GEPI_SERVE_KEY_COLLIDE", "1")
    rng = np.random.default_rng(31)
    n_var, n = 40, 300
    panels = []
    for _ in range(2):
        codes = orc.random_codes(rng, n_var, n, 0.02)
        bits, valid = orc.planes_from_codes(codes)
        panels.append(dict(payload=synth.pack_codes(codes), n_var=n_var, n=n, bits=bits,
                           valid=valid))
    cls = np.where(rng.random(n) < 0.5, 1, 0)
    m0, m1 = orc.masks_from_classes(cls)
    with ge.Server() as srv:
        a = srv.add_panel_bed(panels[0]["payload"], n_var, n)
        b = srv.add_panel_bed(panels[1]["payload"], n_var, n)
        assert a != b
        assert srv.add_panel_bed(panels[0]["payload"], n_var, n) == a
        assert srv.stats()["panels"] == 2
        got = [srv.wait(srv.submit(pid, m0, m1, topmost=300)) for pid in (a, b)]
        # releasing the keyed panel leaves the other one usable
        srv.release_panel(a)
        srv.release_panel(a)
        again = srv.wait(srv.submit(b, m0, m1, topmost=300))
    for (res, _, status), p in zip(got + [again], panels + [panels[1]]):
        assert status == 0
        ref = orc.gold_scan(orc.Panel(p["bits"], p["valid"], m0, m1, n), 300)
        assert_same(res, ref, "collided panel")


def test_server_autotuned():
    """A server whose engine ran the autotuner at creation serves results
    identical to the oracle's."""
    print("case server autotuned", flush=True)
    panels = make_panels(13, 3)
    reqs = make_requests(14, panels, 9)
    with ge.Server(autotune=True) as srv:
        ids = [srv.add_panel_bed(p["payload"], p["n_var"], p["n"]) for p in panels]
        tickets = [srv.submit(ids[r["pi"]], r["m0"], r["m1"], topmost=r["k"], ffun=r["ffun"])
                   for r in reqs]
        got = [srv.wait(t) for t in tickets]
    check_against_alone_and_oracle(panels, reqs, got)


def test_server_graph_rounds_with_concurrent_clients(monkeypatch):
    """Captured-graph rounds (GEPI_ENGINE_GRAPH=1) in the server: the worker
    thread captures and replays rounds while client threads upload panels
    and submit; every response equals the oracle's."""
    print("case server graph rounds", flush=True)
    monkeypatch.setenv("GEPI_ENGINE_GRAPH", "1")
    panels = make_panels(23, 6)
    reqs = make_requests(24, panels, 24)
    got = [None] * len(reqs)
    with ge.Server() as srv:
        ids = [None] * len(panels)
        ids[0] = srv.add_panel_bed(panels[0]["payload"], panels[0]["n_var"], panels[0]["n"])

        def client(j):
            # each client uploads its own panels while other requests run
            for pi in range(1 + j, len(panels), 3):
                ids[pi] = srv.add_panel_bed(panels[pi]["payload"], panels[pi]["n_var"], panels[pi]["n"])
            mine = [i for i in range(len(reqs)) if i % 3 == j]
            tickets = []
            for i in mine:
                while ids[reqs[i]["pi"]] is None:
                    time.sleep(0.001)
                tickets.append((i, srv.submit(ids[reqs[i]["pi"]], reqs[i]["m0"], reqs[i]["m1"],
                                              topmost=reqs[i]["k"], ffun=reqs[i]["ffun"])))
            for i, t in tickets:
                got[i] = srv.wait(t)

        th = [threading.Thread(target=client, args=(j,)) for j in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    assert st["failed"] == 0 and st["completed"] == len(reqs)
    check_against_alone_and_oracle(panels, reqs, got)
