"""The serving session on the GPU: inserts and extract queries submitted
concurrently from several threads; every extract equals the same extract
served alone on the version it reports; extracts run while an insert builds
the next version; waiting extracts share fused sweeps; the HBM budget admits
or rejects inserts."""
from __future__ import annotations

import threading

import time

import numpy as np
import pytest

from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


def clone(g, x):
    h = chip.Graph(g.n, g.dim, g.k)
    h.set_points(x[:g.n])
    ids, dist, flags, count = g.get_knn()
    h.set_knn(ids, dist, flags, count)
    u, v, w, rc = g.get_bridges()
    h.set_bridges(u, v, w, rc)
    return h


def test_session_mixed_stream_equals_alone():
    """Inserts and extracts submitted concurrently from several threads: every
    extract equals the same extract served alone on the version it reports
    (replayed in a second session that applies the same inserts one by one)."""
    print("[gpu] session: concurrent mixed stream vs the same requests alone", flush=True)
    x, k = datagen.workload("mixture-20k-d16")
    n0 = 12000
    g = chip.Graph(n0, x.shape[1], k)
    g.set_points(x[:n0])
    g.build(chip.build_params())
    g2 = clone(g, x)
    rng = np.random.default_rng(5)
    cuts = [n0, 12500, 14000, 14001, 17000, 20000]
    inserts = [(a, b) for a, b in zip(cuts[:-1], cuts[1:])]
    extracts = []
    for _ in range(24):
        m = rng.choice(np.arange(1, k + 1), size=int(rng.integers(1, 9)), replace=False)
        extracts.append([int(v) for v in m])
    extracts.append(list(range(1, k + 1)))

    a_sess = chip.Session(None, k, graph=g)
    ins_ids, ext_ids = [], []
    lock = threading.Lock()

    def submit_inserts():
        for a, b in inserts:
            with lock:
                ins_ids.append(a_sess.submit_insert(x[a:b]))

    def submit_extracts(items):
        for m in items:
            with lock:
                ext_ids.append((a_sess.submit_extract(m, edges=True, labels=True), m))

    threads = [threading.Thread(target=submit_inserts)]
    threads += [threading.Thread(target=submit_extracts, args=(extracts[i::3],)) for i in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    versions = []
    for rid in ins_ids:
        st = a_sess.wait(rid)
        assert st["state"] == chip.REQ_DONE, st["error"]
        versions.append((st["graph_version"], st["graph_n"]))
    assert versions == [(v + 1, b) for v, (_, b) in enumerate(inserts)]
    by_version, batched = {}, 0
    for rid, m in ext_ids:
        st = a_sess.wait(rid)
        assert st["state"] == chip.REQ_DONE, st["error"]
        batched = max(batched, st["batch_requests"])
        by_version.setdefault(st["graph_version"], []).append(
            (m, st["graph_n"], a_sess.result(rid, edges=True, labels=True)))

    b_sess = chip.Session(None, k, graph=g2)
    checked = 0
    for v in range(len(inserts) + 1):
        if v > 0:
            a, b = inserts[v - 1]
            st = b_sess.wait(b_sess.submit_insert(x[a:b]))
            assert st["state"] == chip.REQ_DONE and st["graph_version"] == v, st["error"]
        for m, n_seen, ra in by_version.get(v, []):
            bid = b_sess.submit_extract(m, edges=True, labels=True)
            st = b_sess.wait(bid)
            assert st["batch_requests"] == 1
            assert st["graph_version"] == v and st["graph_n"] == n_seen
            rb = b_sess.result(bid, edges=True, labels=True)
            for p, q in zip(ra, rb):
                for key in ("mpts", "num_edges", "num_components", "rounds"):
                    assert p[key] == q[key]
                assert np.float64(p["weight"]).view(np.uint64) == \
                    np.float64(q["weight"]).view(np.uint64)
                assert np.array_equal(p["u"], q["u"]) and np.array_equal(p["v"], q["v"])
                assert np.array_equal(p["w"].view(np.uint32), q["w"].view(np.uint32))
                assert np.array_equal(p["labels"], q["labels"])
            checked += 1
    print(f"  {checked} extract requests identical over versions {sorted(by_version)}; largest "
          f"fused batch {batched} requests", flush=True)
    assert checked == len(extracts)
    a_sess.close()
    b_sess.close()


def test_session_extract_overlaps_insert():
    """An extract submitted while an insert is building the next version is
    served from the published one and completes first."""
    print("[gpu] session: extract served while an insert runs", flush=True)
    x = datagen.gaussian_mixture(130_000, 16, 12, 0.3, 2.0, 0.05, seed=9)
    k = 16
    s = chip.Session(x[:100_000], k)
    ins = s.submit_insert(x[100_000:])
    ext = s.submit_extract([4, 8])
    se, si = s.wait(ext), s.wait(ins)
    assert si["state"] == chip.REQ_DONE and se["state"] == chip.REQ_DONE
    assert se["graph_version"] == 0 and se["graph_n"] == 100_000
    assert si["graph_version"] == 1 and si["graph_n"] == 130_000
    assert se["latency_ms"] < si["latency_ms"]
    st = s.wait(s.submit_extract([4]))
    assert st["graph_version"] == 1 and st["graph_n"] == 130_000
    print(f"  extract {se['latency_ms']:.1f} ms (version 0) during insert "
          f"{si['latency_ms']:.1f} ms", flush=True)
    s.close()


def test_session_fuses_waiting_extracts():
    """Extracts submitted while a sweep runs join the next sweep together.
    Which requests were waiting is read from the clock rather than assumed:
    those submitted before the first request's sweep had ended (its submit
    time plus its latency, less a margin) were queued behind it, and two or
    more of them must have been served by one sweep."""
    print("[gpu] session: waiting extracts fused", flush=True)
    x = datagen.gaussian_mixture(150_000, 16, 12, 0.3, 2.0, 0.05, seed=10)
    k = 16
    s = chip.Session(x, k)
    t_first = time.perf_counter()
    first = s.submit_extract(list(range(1, k + 1)))  # keeps the worker busy
    ids, t_sub = [], []
    for m in range(1, k + 1):
        t_sub.append(time.perf_counter())
        ids.append(s.submit_extract([m]))
    sts = [s.wait(i) for i in [first] + ids]
    assert all(st["state"] == chip.REQ_DONE for st in sts)
    first_done = t_first + sts[0]["latency_ms"] * 1e-3
    waited = [st for st, t in zip(sts[1:], t_sub) if t < first_done - 2e-4]
    print(f"  {len(waited)} of {k} extracts waited for the first sweep "
          f"({sts[0]['latency_ms']:.2f} ms); batch sizes "
          f"{sorted(set(st['batch_requests'] for st in sts[1:]))}", flush=True)
    assert len(waited) >= 2, "the first sweep ended before two more extracts were submitted"
    assert max(st["batch_requests"] for st in waited) > 1
    s.close()


def test_session_hbm_budget_and_errors():
    x, k = datagen.workload("mixture-20k-d16")
    g = chip.Graph(10000, x.shape[1], k)
    g.set_points(x[:10000])
    g.build(chip.build_params())
    g.prepare_edges()
    budget = g.device_bytes() + 1
    s = chip.Session(None, k, graph=g, hbm_budget=budget)
    st = s.wait(s.submit_insert(x[10000:15000]))
    assert st["state"] == chip.REQ_FAILED and "budget" in st["error"]
    st = s.wait(s.submit_extract([4]))
    assert st["state"] == chip.REQ_DONE and st["graph_n"] == 10000
    with pytest.raises(chip.CoresgHipError):
        s.submit_extract([k + 1])
    with pytest.raises(chip.CoresgHipError):
        s.submit_extract([])
    s.close()
