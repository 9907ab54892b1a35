"""The serving session on the GPU: a mixed stream of inserts and extract
queries submitted concurrently from several threads gives, request by
request, exactly the results of the same requests served one at a time in the
same order; waiting extracts share fused sweeps; the HBM budget admits or
rejects inserts."""
from __future__ import annotations

import threading

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


def test_session_mixed_stream_equals_sequential():
    print("[gpu] session: concurrent mixed stream vs one-at-a-time replay", flush=True)
    x, k = datagen.workload("mixture-20k-d16")
    n0 = 12000
    g = chip.Graph(n0, x.shape[1], k)
    g.set_points(x[:n0])
    g.build(chip.build_params())
    g2 = clone(g, x)
    rng = np.random.default_rng(5)
    script = []
    cuts = [n0, 12500, 14000, 14001, 17000, 20000]
    for a, b in zip(cuts[:-1], cuts[1:]):
        for _ in range(3):
            m = rng.choice(np.arange(1, k + 1), size=int(rng.integers(1, 9)), replace=False)
            script.append(("extract", [int(v) for v in m]))
        script.append(("insert", (a, b)))
    script.append(("extract", list(range(1, k + 1))))

    a_sess = chip.Session(None, k, graph=g)
    order = []  # (id, request) in the session's order
    lock = threading.Lock()

    def submit(items):
        for kind, arg in items:
            with lock:  # the id order is the session order; keep the script's per-thread order
                if kind == "insert":
                    rid = a_sess.submit_insert(x[arg[0]:arg[1]])
                else:
                    rid = a_sess.submit_extract(arg, edges=True, labels=True)
                order.append((rid, kind, arg))

    # inserts stay in script order (one thread); extracts interleave from others
    inserts = [s for s in script if s[0] == "insert"]
    extracts = [s for s in script if s[0] == "extract"]
    threads = [threading.Thread(target=submit, args=(inserts,))]
    threads += [threading.Thread(target=submit, args=(extracts[i::3],)) for i in range(3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    order.sort()
    res_a, batched = {}, 0
    for rid, kind, arg in order:
        st = a_sess.wait(rid)
        assert st["state"] == chip.REQ_DONE, st["error"]
        batched = max(batched, st["batch_requests"])
        if kind == "extract":
            res_a[rid] = (st["graph_version"], a_sess.result(rid, edges=True, labels=True))

    b_sess = chip.Session(None, k, graph=g2)
    checked = 0
    for rid, kind, arg in order:
        if kind == "insert":
            st = b_sess.wait(b_sess.submit_insert(x[arg[0]:arg[1]]))
            assert st["state"] == chip.REQ_DONE, st["error"]
            continue
        bid = b_sess.submit_extract(arg, edges=True, labels=True)
        st = b_sess.wait(bid)
        assert st["batch_requests"] == 1
        version, ra = res_a[rid]
        assert st["graph_version"] == version
        rb = b_sess.result(bid, edges=True, labels=True)
        for p, q in zip(ra, rb):
            for key in ("mpts", "num_edges", "num_components", "rounds"):
                assert p[key] == q[key]
            assert np.float64(p["weight"]).view(np.uint64) == np.float64(q["weight"]).view(np.uint64)
            assert np.array_equal(p["u"], q["u"]) and np.array_equal(p["v"], q["v"])
            assert np.array_equal(p["w"].view(np.uint32), q["w"].view(np.uint32))
            assert np.array_equal(p["labels"], q["labels"])
        checked += 1
    print(f"  {checked} extract requests identical; largest fused batch {batched} requests",
          flush=True)
    assert checked == len(extracts)
    a_sess.close()
    b_sess.close()


def test_session_fuses_waiting_extracts():
    x, k = datagen.workload("mixture-20k-d16")
    s = chip.Session(x, k)
    ids = [s.submit_extract([m]) for m in range(1, k + 1)]
    sts = [s.wait(i) for i in ids]
    assert all(st["state"] == chip.REQ_DONE for st in sts)
    # the worker takes everything that is waiting when it becomes free
    assert max(st["batch_requests"] for st in sts) > 1
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
