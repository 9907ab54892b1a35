"""MST extraction on the GPU against the reference's own MSTs (golden fixtures)
and the oracle: bit-exact edge sets, weights, labels and component counts."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle
from conftest import golden, golden_cases, load_pkg_module

chip = load_pkg_module("coresg_hip")
coresg_io = load_pkg_module("coresg_io")
pytestmark = pytest.mark.gpu


def graph_from_file(f):
    g = chip.Graph(f.n, f.dim, f.k_max)
    g.set_points(f.points)
    g.set_knn(f.ids, f.dist, f.flags, f.count)
    u, v, w = f.bridge_uvw()
    g.set_bridges(u, v, w, f.repair_count)
    return g


def kruskal_order(u, v, w):
    o = np.lexsort((v, u, w))
    return u[o], v[o], w[o]


@pytest.mark.parametrize("name", golden_cases())
def test_extract_matches_reference_msts(name):
    print(f"[gpu] extract vs reference: {name}", flush=True)
    gd = golden(name)
    f = coresg_io.loads(gd["coresg_bytes"].tobytes())
    g = graph_from_file(f)
    mpts = [int(m) for m in gd["mpts"]]
    res = g.extract(mpts, edges=True, labels=True)
    for m, r in zip(mpts, res):
        ru, rv, rw = gd[f"mst{m}_mst_u"], gd[f"mst{m}_mst_v"], gd[f"mst{m}_mst_w"]
        u, v, w = kruskal_order(r["u"], r["v"], r["w"])
        assert r["num_edges"] == ru.size
        assert np.array_equal(u, ru) and np.array_equal(v, rv)
        assert np.array_equal(w.view(np.uint32), rw.view(np.uint32))
        assert r["num_components"] == int(gd[f"mst{m}_num_components"][0])
        assert np.array_equal(r["labels"], gd[f"mst{m}_component"])
        assert np.float64(r["weight"]).view(np.uint64) == gd[f"mst{m}_weight"].view(np.uint64)[0]
    g.close()


@pytest.mark.parametrize("name", golden_cases())
def test_reweighted_edges_match_oracle(name):
    gd = golden(name)
    f = coresg_io.loads(gd["coresg_bytes"].tobytes())
    g = graph_from_file(f)
    ne, ke = g.prepare_edges()
    ou, ov, ow = oracle.undirected_edges(f.ids, f.dist, f.count, f.n, f.k_max)
    assert ke == ou.size and ne == ou.size + f.bridges.shape[0]
    p = __import__("tempfile").mktemp(suffix=".coresg")
    open(p, "wb").write(gd["coresg_bytes"].tobytes())
    sg = oracle.CoreSG.load(p)
    for m in [1, f.k_max // 2 or 1, f.k_max]:
        u, v, w = g.edges(m)
        eu, ev, ew = sg.edges(m)
        a = np.lexsort((ew.view(np.uint32), ev, eu))
        b = np.lexsort((w.view(np.uint32), v, u))
        assert np.array_equal(u[b], eu[a]) and np.array_equal(v[b], ev[a])
        assert np.array_equal(w[b].view(np.uint32), ew[a].view(np.uint32))
        # (u, v) order of the resident list
        assert np.all((u[1:] > u[:-1]) | ((u[1:] == u[:-1]) & (v[1:] >= v[:-1])))
    g.close()


def test_extract_mixture_20k_all_mpts():
    """A 20k-point CORE-SG built by the oracle: every mpts 1..16 bit-exact."""
    print("[gpu] extract 20k x 16 mpts", flush=True)
    datagen = load_pkg_module("datagen")
    x, k = datagen.workload("mixture-20k-d16")
    sg = oracle.CoreSG.build(x, k_max=k)
    ids, dist, flags, count = sg.graph()
    g = chip.Graph(sg.n, sg.dim, sg.k)
    g.set_points(x)
    g.set_knn(ids, dist, flags, count)
    bu, bv, bw = sg.bridges()
    g.set_bridges(bu, bv, bw, sg.repair_count())
    res = g.extract(list(range(1, k + 1)), edges=True, labels=True)
    for m, r in zip(range(1, k + 1), res):
        ref = sg.mst(m)
        u, v, w = kruskal_order(r["u"], r["v"], r["w"])
        assert np.array_equal(u, ref["u"]) and np.array_equal(v, ref["v"])
        assert np.array_equal(w.view(np.uint32), ref["w"].view(np.uint32))
        assert np.array_equal(r["labels"], ref["labels"])
        assert r["num_components"] == ref["ncomp"]
        assert np.float64(r["weight"]).view(np.uint64) == np.float64(ref["weight"]).view(np.uint64)
    g.close()


def test_extract_tiny_and_errors():
    # n = 2 with one edge; mpts 0 acts as 1; mpts > k_max raises.
    g = chip.Graph(2, 2, 1)
    g.set_points(np.array([[0, 0], [3, 4]], np.float32))
    g.set_knn(np.array([1, 0], np.uint32), np.array([5, 5], np.float32), None,
              np.array([1, 1], np.uint32))
    r = g.extract([0, 1])
    assert r[0]["mpts"] == 1 and r[0]["num_edges"] == 1 and r[0]["weight"] == 5.0
    with pytest.raises(chip.CoresgHipError):
        g.extract([2])
    g.close()
    # isolated nodes (empty lists): a forest with +inf cores, no edges
    g = chip.Graph(3, 2, 2)
    g.set_points(np.zeros((3, 2), np.float32))
    g.set_knn(np.full(6, 0xFFFFFFFF, np.uint32), np.full(6, np.inf, np.float32), None,
              np.zeros(3, np.uint32))
    r = g.extract([1, 2], labels=True)
    assert r[0]["num_edges"] == 0 and r[0]["num_components"] == 3
    assert list(r[0]["labels"]) == [0, 1, 2]
    g.close()


def test_batched_sweep_equals_single_queries():
    """One extract call runs up to 32 mpts values together; the result of every
    slot (edges, labels, weight, rounds) equals the same mpts extracted alone,
    across a batch boundary, with repeats and mpts 0."""
    print("[gpu] batched sweep vs single queries", flush=True)
    gd = golden("mix_d24")
    f = coresg_io.loads(gd["coresg_bytes"].tobytes())
    g = graph_from_file(f)
    k = f.k_max
    mpts = [k, 1, 1, 7, 0, 3] + list(range(1, k + 1)) + [5, k, 2]
    assert len(mpts) > 32
    batched = g.extract(mpts, edges=True, labels=True)
    for m, r in zip(mpts, batched):
        one = g.extract([m], edges=True, labels=True)[0]
        for key in ("mpts", "num_edges", "num_components", "rounds"):
            assert r[key] == one[key], (m, key)
        assert np.float64(r["weight"]).view(np.uint64) == np.float64(one["weight"]).view(np.uint64)
        assert np.array_equal(r["u"], one["u"]) and np.array_equal(r["v"], one["v"])
        assert np.array_equal(r["w"].view(np.uint32), one["w"].view(np.uint32))
        assert np.array_equal(r["labels"], one["labels"])
    g.close()
