"""Batched independent instances on the GPU: many uneven point sets in one
graph, every CoreSG step per instance (CoreSG::build, coresg.cpp:260-323, and
CoreSG::mst, coresg.cpp:408-417, applied to each point set alone).

- init_random of the union equals each instance's own init bit for bit
  (draws inside the instance with the instance-local node index);
- from the same k-NN lists, the union's selection, bridges and per-instance
  MST summaries (edges, components, exact weight) equal each instance built
  and extracted alone, and the union's MST edges are the instances' edges;
- a full union build keeps each instance's recall at the level of separate
  builds.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu

D, K = 16, 16
SIZES = [700, 2300, 1500, 3100]


def instances():
    xs = [datagen.gaussian_mixture(n, D, 4 + 3 * i, 0.3, 2.0, 0.05, seed=40 + i)
          for i, n in enumerate(SIZES)]
    off = np.concatenate([[0], np.cumsum([x.shape[0] for x in xs])]).astype(np.uint32)
    return xs, off


def union_graph(xs, off):
    g = chip.Graph(int(off[-1]), D, K)
    g.set_points(np.concatenate(xs))
    g.set_instances(off)
    return g


def shift(ids, delta):
    ids = ids.copy()
    live = ids != 0xFFFFFFFF
    ids[live] += np.uint32(delta)
    return ids


def test_instances_init_random_bitexact():
    print("[gpu] instances: init_random", flush=True)
    xs, off = instances()
    g = union_graph(xs, off)
    g.build_knn(chip.build_params(iters=0))
    uids, udist, uflags, ucount = g.get_knn()
    g.close()
    for i, x in enumerate(xs):
        h = chip.Graph(x.shape[0], D, K)
        h.set_points(x)
        h.build_knn(chip.build_params(iters=0))
        ids, dist, flags, count = h.get_knn()
        h.close()
        a, b = int(off[i]) * K, int(off[i + 1]) * K
        assert np.array_equal(uids[a:b], shift(ids, off[i]))
        assert np.array_equal(udist[a:b].view(np.uint32), dist.view(np.uint32))
        assert np.array_equal(uflags[a:b], flags)
        assert np.array_equal(ucount[off[i]:off[i + 1]], count)


def test_instances_bridges_and_msts_equal_separate():
    print("[gpu] instances: bridges and MSTs vs separate builds", flush=True)
    xs, off = instances()
    mpts = list(range(2, K + 1))
    lists, alone = [], []
    for i, x in enumerate(xs):
        h = chip.Graph(x.shape[0], D, K)
        h.set_points(x)
        h.build(chip.build_params())
        ids, dist, flags, count = h.get_knn()
        lists.append((shift(ids, off[i]), dist, flags, count))
        bu, bv, bw, rc = h.get_bridges()
        h.prepare_edges()
        res = h.extract(mpts, edges=True)
        alone.append(dict(sel=h.selected(), bridges=(bu, bv, bw), res=res))
        h.close()
    g = union_graph(xs, off)
    g.set_knn(np.concatenate([l[0] for l in lists]), np.concatenate([l[1] for l in lists]),
              np.concatenate([l[2] for l in lists]), np.concatenate([l[3] for l in lists]))
    g.build_bridges(chip.build_params())
    sel = g.selected()
    bu, bv, bw, _ = g.get_bridges()
    g.prepare_edges()
    per = g.extract_instances(mpts)
    union = g.extract(mpts, edges=True)
    pos = 0
    for i in range(len(xs)):
        s_alone = alone[i]["sel"].astype(np.int64) + int(off[i])
        assert np.array_equal(sel[pos:pos + s_alone.size].astype(np.int64), s_alone)
        pos += s_alone.size
        lo, hi = int(off[i]), int(off[i + 1])
        mine = (bu >= lo) & (bu < hi)
        au, av, aw = alone[i]["bridges"]
        assert np.array_equal(bu[mine], au + lo) and np.array_equal(bv[mine], av + lo)
        assert np.array_equal(bw[mine].view(np.uint32), aw.view(np.uint32))
        for q, m in enumerate(mpts):
            a, u = alone[i]["res"][q], per[i][q]
            assert (u["mpts"], u["num_edges"], u["num_components"]) == (
                m, a["num_edges"], a["num_components"])
            assert np.float64(u["weight"]).view(np.uint64) == np.float64(a["weight"]).view(np.uint64)
    assert pos == sel.size
    # the union's MST edges are the instances' edges, instance by instance
    for q in range(len(mpts)):
        ue = set(zip(union[q]["u"].tolist(), union[q]["v"].tolist()))
        ae = set()
        for i in range(len(xs)):
            r = alone[i]["res"][q]
            ae |= set(zip((r["u"] + off[i]).tolist(), (r["v"] + off[i]).tolist()))
        assert ue == ae
    g.close()


def test_instances_union_build_recall():
    """The union descent (candidates never leave an instance) keeps every
    instance's recall at the level of separate builds."""
    print("[gpu] instances: union build recall", flush=True)
    xs, off = instances()
    g = union_graph(xs, off)
    g.build(chip.build_params())
    uids, _, _, ucount = g.get_knn()
    per = g.extract_instances([2, 8])
    g.close()
    for i, x in enumerate(xs):
        a, b = int(off[i]) * K, int(off[i + 1]) * K
        ids = uids[a:b]
        live = ids != 0xFFFFFFFF
        assert np.all((ids[live] >= off[i]) & (ids[live] < off[i + 1]))  # never leaves
        bids, _, _, _ = oracle.brute_force(oracle.pad(x), K)
        got = np.mean([len(set(ids.reshape(-1, K)[r] - off[i]) & set(bids.reshape(-1, K)[r])) / K
                       for r in range(x.shape[0])])
        assert got >= 0.97, (i, got)
        assert per[i][0]["num_edges"] + per[i][0]["num_components"] == x.shape[0]
