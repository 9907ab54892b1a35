"""CoreSG::exact_mst on the GPU (Prim over the complete mutual-reachability
graph, coresg.cpp:419-430): bit-exact against the reference's own Prim output
(tests/golden/exact_mst.npz), against the oracle on a fresh graph, and by
Theorem 1 (test_coresg.cpp:32-61) at a size the oracle is not run at."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import oracle
from conftest import golden, load_pkg_module

chip = load_pkg_module("coresg_hip")
coresg_io = load_pkg_module("coresg_io")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu

EXACT = golden("exact_mst")
RUNS = sorted({(m.group(1), int(m.group(2)))
               for m in (re.match(r"^(.*)_exact(\d+)_", k) for k in EXACT) if m})


def graph_from_file(f):
    g = chip.Graph(f.n, f.dim, f.k_max)
    g.set_points(f.points)
    g.set_knn(f.ids, f.dist, f.flags, f.count)
    u, v, w = f.bridge_uvw()
    g.set_bridges(u, v, w, f.repair_count)
    return g


def assert_same(r, ref):
    assert r["num_edges"] == ref["u"].size
    assert np.array_equal(r["u"], ref["u"]) and np.array_equal(r["v"], ref["v"])
    assert np.array_equal(r["w"].view(np.uint32), ref["w"].view(np.uint32))
    assert np.array_equal(r["labels"], ref["labels"])
    assert r["num_components"] == ref["ncomp"]
    assert np.float64(r["weight"]).view(np.uint64) == np.float64(ref["weight"]).view(np.uint64)


@pytest.mark.parametrize("name,mpts", RUNS)
def test_exact_mst_matches_reference(name, mpts):
    print(f"[gpu] exact_mst vs reference: {name} mpts={mpts}", flush=True)
    f = coresg_io.loads(golden(name)["coresg_bytes"].tobytes())
    g = graph_from_file(f)
    r = g.exact_mst(mpts, labels=True)
    p = f"{name}_exact{mpts}_"
    assert_same(r, dict(u=EXACT[p + "mst_u"], v=EXACT[p + "mst_v"], w=EXACT[p + "mst_w"],
                        labels=EXACT[p + "component"],
                        ncomp=int(EXACT[p + "num_components"][0]),
                        weight=EXACT[p + "weight"][0]))
    assert r["rounds"] == f.n - 1
    g.close()


@pytest.mark.parametrize("name,mpts", RUNS)
def test_exact_mst_stepped_launch_matches_reference(name, mpts):
    """The one-launch-per-step form (coresg_hip_graph_set_exact_launch(g, 1):
    the kernel boundary as the step barrier, batches of launches replayed from
    a captured graph) gives the reference's Prim output bit for bit too."""
    print(f"[gpu] exact_mst stepped vs reference: {name} mpts={mpts}", flush=True)
    f = coresg_io.loads(golden(name)["coresg_bytes"].tobytes())
    g = graph_from_file(f)
    g.set_exact_launch(1)
    r = g.exact_mst(mpts, labels=True)
    p = f"{name}_exact{mpts}_"
    assert_same(r, dict(u=EXACT[p + "mst_u"], v=EXACT[p + "mst_v"], w=EXACT[p + "mst_w"],
                        labels=EXACT[p + "component"],
                        ncomp=int(EXACT[p + "num_components"][0]),
                        weight=EXACT[p + "weight"][0]))
    g.close()


def test_exact_mst_stepped_equals_cooperative_at_5k():
    """Both launch forms on a fresh 5 000-point build (more steps than one
    captured batch of launches, 20 vertices per block), back to back with a
    sweep in between: the same edges, weights and labels."""
    print("[gpu] exact_mst stepped vs cooperative: n=5000", flush=True)
    x = datagen.gaussian_mixture(5000, 24, 12, 0.3, 2.0, 0.05, 17)
    g = chip.Graph(x.shape[0], x.shape[1], 16)
    g.set_points(x)
    g.build(chip.build_params())
    a = g.exact_mst(8, labels=True)
    g.extract([4, 8])
    g.set_exact_launch(1)
    b = g.exact_mst(8, labels=True)
    c = g.exact_mst(3, labels=True)
    g.set_exact_launch(0)
    d = g.exact_mst(3, labels=True)
    g.close()
    for p, q in ((a, b), (c, d)):
        assert np.array_equal(p["u"], q["u"]) and np.array_equal(p["v"], q["v"])
        assert np.array_equal(p["w"].view(np.uint32), q["w"].view(np.uint32))
        assert np.array_equal(p["labels"], q["labels"])
        assert np.float64(p["weight"]).view(np.uint64) == np.float64(q["weight"]).view(np.uint64)


def test_exact_mst_matches_oracle_and_clamps_mpts(tmp_path):
    """A fresh device build of 3000 points (12 blocks of 250 vertices): device
    Prim == oracle Prim on the same lists, edge for edge; mpts 0
    reads the first column like mpts 1, and mpts above k_max is clamped
    (coresg.cpp:68-70); a second call on a smaller graph gives its own answer."""
    x = datagen.gaussian_mixture(3000, 16, 10, 0.3, 2.0, 0.05, 5)
    print("[gpu] exact_mst vs oracle: n=3000 d=16 k=16", flush=True)
    g = chip.Graph(x.shape[0], x.shape[1], 16)
    g.set_points(x)
    g.build(chip.build_params())
    ids, dist, flags, count = g.get_knn()
    u, v, w, rc = g.get_bridges()
    fobj = coresg_io.CoreSGFile(16, x.shape[1], 16, 0, 0, oracle.pad(x), ids, dist, flags, count,
                                coresg_io.bridges_from_uvw(u, v, w), rc)
    path = os.path.join(str(tmp_path), "g.coresg")
    coresg_io.save(path, fobj)
    sg = oracle.CoreSG.load(path)
    for m in (1, 7, 16):
        assert_same(g.exact_mst(m, labels=True), sg.exact_mst(m))
    a = g.exact_mst(0)
    b = g.exact_mst(1)
    assert np.array_equal(a["w"].view(np.uint32), b["w"].view(np.uint32))
    c = g.exact_mst(40)
    d = g.exact_mst(16)
    assert np.array_equal(c["u"], d["u"]) and np.array_equal(c["w"], d["w"])
    g.close()
    small = chip.Graph(700, 16, 16)
    small.set_points(x[:700])
    small.build(chip.build_params(brute_force_limit=1000))
    r = small.exact_mst(4)
    assert r["num_edges"] == 699 and r["num_components"] == 1
    small.close()


def test_exact_mst_theorem1_at_20k():
    """Theorem 1: exact k-NN lists (brute force) plus the exact MST at k_max as
    the bridge set, stored with base distances (coresg.cpp:157-163), make the
    CORE-SG MST of every mpts <= k_max weigh what the complete-graph MST
    weighs, weight for weight: the sorted weights of the device sweep (edge
    list + Boruvka) equal those of the device Prim, two unrelated code paths."""
    x = datagen.gaussian_mixture(20000, 16, 10, 0.3, 2.0, 0.05, 0)
    print("[gpu] exact_mst theorem 1: n=20000 d=16 k=16", flush=True)
    g = chip.Graph(x.shape[0], x.shape[1], 16)
    g.set_points(x)
    g.build_knn(chip.build_params(brute_force_limit=20000))
    b = g.exact_mst(16)
    assert b["num_components"] == 1
    _, bw = oracle.l2sqr_pairs(oracle.pad(x), np.stack([b["u"], b["v"]], 1).astype(np.uint32))
    g.set_bridges(b["u"], b["v"], bw, 0)
    mp = [1, 2, 4, 8, 16]
    res = g.extract(mp, edges=True)
    for m, r in zip(mp, res):
        e = g.exact_mst(m)
        assert r["num_edges"] == e["num_edges"] == 19999
        assert np.array_equal(np.sort(r["w"]).view(np.uint32), np.sort(e["w"]).view(np.uint32))
    g.close()


def test_exact_mst_wide_rows_generic_kernel(tmp_path):
    """Rows of 512 floats do not fit the resident kernel's LDS, so the generic
    kernel (state and rows in global memory, counter barrier) runs; lists from
    the oracle's brute force, checked against the oracle's Prim edge for edge."""
    x = datagen.gaussian_mixture(600, 512, 5, 0.3, 2.0, 0.05, 7)
    print("[gpu] exact_mst generic kernel: n=600 d=512 k=8", flush=True)
    ids, dist, flags, count = oracle.brute_force(oracle.pad(x), 8)
    g = chip.Graph(600, 512, 8)
    g.set_points(x)
    g.set_knn(ids, dist, flags, count)
    fobj = coresg_io.CoreSGFile(8, 512, 512, 0, 0, oracle.pad(x), ids, dist, flags, count,
                                np.zeros((0, 3), np.uint32), 0)
    path = os.path.join(str(tmp_path), "w.coresg")
    coresg_io.save(path, fobj)
    sg = oracle.CoreSG.load(path)
    for m in (1, 8):
        assert_same(g.exact_mst(m, labels=True), sg.exact_mst(m))
    g.close()


@pytest.mark.parametrize("name", ["brute_d8", "islands"])
def test_exactmode_build_matches_reference(name):
    """A device build with missing_mode = 1 (ExactMst) and brute-force lists
    against the reference CLI's `build --missing exact` on the same data: the
    same lists, the same bridges (Prim order, base distances) bit for bit, no
    repair edges; missing_mode = 2 (None) leaves no bridges."""
    print(f"[gpu] exact-mode build vs reference: {name}", flush=True)
    f = coresg_io.loads(EXACT[f"{name}_exactmode_coresg_bytes"].tobytes())
    x = golden(name)["data"]
    g = chip.Graph(x.shape[0], x.shape[1], f.k_max)
    g.set_points(x)
    st = g.build(chip.build_params(brute_force_limit=100000, missing_mode=1))
    ids, dist, _, count = g.get_knn()
    assert np.array_equal(ids, f.ids) and np.array_equal(count, f.count)
    assert np.array_equal(dist.view(np.uint32), f.dist.view(np.uint32))
    u, v, w, rc = g.get_bridges()
    fu, fv, fw = f.bridge_uvw()
    assert np.array_equal(u, fu) and np.array_equal(v, fv)
    assert np.array_equal(w.view(np.uint32), fw.view(np.uint32))
    assert rc == f.repair_count == 0
    assert st["antihubs_selected"] == 0
    st = g.build(chip.build_params(brute_force_limit=100000, missing_mode=2))
    u, _, _, _ = g.get_bridges()
    assert u.size == st["repair_edges"]
    g.close()
