"""Pin the oracle (C restatement of the reference CPU path) against vectors the
reference wrote itself (tests/golden/, from tests/golden/make_golden.py), then
against closed forms and brute force for what the fixtures leave open."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import oracle
from conftest import golden, golden_cases, load_pkg_module, maxreps_cases

coresg_io = load_pkg_module("coresg_io")
CASES = golden_cases()


def _flags_for(name):
    if name == "brute_d8":
        return dict(brute_force_limit=1000)
    if name in ("islands_q0", "islands_repair"):
        return dict(min_per_component=0)
    return {}


_built = {}


def built(name):
    if name not in _built:
        g = golden(name)
        _built[name] = (g, oracle.CoreSG.build(g["data"], k_max=int(g["k"]), **_flags_for(name)))
    return _built[name]


@pytest.mark.parametrize("name", CASES)
def test_pair_distances_bitexact(name):
    g = golden(name)
    r, d = oracle.l2sqr_pairs(oracle.pad(g["data"]), g["pairs"])
    assert np.array_equal(r.view(np.uint32), g["pairs_l2sqr"].view(np.uint32))
    assert np.array_equal(d.view(np.uint32), g["pairs_dist"].view(np.uint32))


@pytest.mark.parametrize("name", CASES)
def test_init_random_bitexact(name):
    g = golden(name)
    ids, dist, flags, count = oracle.init_random(oracle.pad(g["data"]), int(g["k"]), 42)
    assert np.array_equal(ids, g["init_ids"])
    assert np.array_equal(dist.view(np.uint32), g["init_dist"].view(np.uint32))
    assert np.array_equal(flags, g["init_flags"])
    assert np.array_equal(count, g["init_count"])


@pytest.mark.parametrize("name", CASES)
def test_build_matches_reference_file(name):
    """Serial NN-Descent (or brute force), hubness, selection, bridges and
    repair reproduce the reference's .coresg byte for byte."""
    g, sg = built(name)
    ref = coresg_io.loads(g["coresg_bytes"].tobytes())
    ids, dist, flags, count = sg.graph()
    assert np.array_equal(ids, ref.ids)
    assert np.array_equal(dist.view(np.uint32), ref.dist.view(np.uint32))
    assert np.array_equal(flags, ref.flags)
    assert np.array_equal(count, ref.count)
    bu, bv, bw = sg.bridges()
    ru, rv, rw = ref.bridge_uvw()
    assert np.array_equal(bu, ru) and np.array_equal(bv, rv)
    assert np.array_equal(bw.view(np.uint32), rw.view(np.uint32))
    assert sg.repair_count() == ref.repair_count
    assert np.array_equal(sg.points(), ref.points)


@pytest.mark.parametrize("case", maxreps_cases(), ids=lambda c: c[0])
def test_repair_cut_above_max_reps_matches_reference_file(case):
    """More components than max_reps with tied scores: the repair keeps the
    representatives the reference's partial_sort keeps (coresg.cpp:221-227),
    so bridges, repair count and the whole file are the reference's. A stable
    sort in the cut's place fails every case here."""
    tag, x, max_reps, raw = case
    ref = coresg_io.loads(raw.tobytes())
    assert ref.repair_count > 0
    sg = oracle.CoreSG.build(x, k_max=5, brute_force_limit=100000, min_per_component=0,
                             max_reps=max_reps)
    bu, bv, bw = sg.bridges()
    ru, rv, rw = ref.bridge_uvw()
    assert np.array_equal(bu, ru) and np.array_equal(bv, rv)
    assert np.array_equal(bw.view(np.uint32), rw.view(np.uint32))
    assert sg.repair_count() == ref.repair_count


@pytest.mark.parametrize("name", CASES)
def test_save_is_byte_identical(name, tmp_path):
    g, sg = built(name)
    p = str(tmp_path / "x.coresg")
    sg.save(p)
    assert open(p, "rb").read() == g["coresg_bytes"].tobytes()
    ref = coresg_io.loads(g["coresg_bytes"].tobytes())
    assert coresg_io.dumps(ref) == g["coresg_bytes"].tobytes()


@pytest.mark.parametrize("name", CASES)
def test_state_after_load(name, tmp_path):
    g = golden(name)
    p = str(tmp_path / "x.coresg")
    open(p, "wb").write(g["coresg_bytes"].tobytes())
    # ref_dump's `state` loads with the default parameters (quota 2)
    sg = oracle.CoreSG.load(p)
    assert np.array_equal(sg.rdeg(), g["state_rdeg"])
    assert np.array_equal(sg.scores().view(np.uint64), g["state_score"].view(np.uint64))
    assert np.array_equal(sg.selected(), g["state_selected"])
    lab, nc = sg.conn()
    assert np.array_equal(lab, g["state_conn_labels"])
    assert nc == int(g["state_conn_components"][0])
    ids, dist, _, count = sg.graph()
    u, v, w = oracle.undirected_edges(ids, dist, count, sg.n, sg.k)
    assert np.array_equal(u, g["state_undirected_u"])
    assert np.array_equal(v, g["state_undirected_v"])
    assert np.array_equal(w.view(np.uint32), g["state_undirected_w"].view(np.uint32))
    assert sg.knn_edges() == u.size


@pytest.mark.parametrize("name", CASES)
def test_mst_sweep_bitexact(name):
    g, sg = built(name)
    for m in g["mpts"]:
        r = sg.mst(int(m))
        assert np.array_equal(r["u"], g[f"mst{m}_mst_u"])
        assert np.array_equal(r["v"], g[f"mst{m}_mst_v"])
        assert np.array_equal(r["w"].view(np.uint32), g[f"mst{m}_mst_w"].view(np.uint32))
        assert np.array_equal(r["labels"], g[f"mst{m}_component"])
        assert r["ncomp"] == int(g[f"mst{m}_num_components"][0])
        assert np.float64(r["weight"]).view(np.uint64) == g[f"mst{m}_weight"].view(np.uint64)[0]


EXACT = golden("exact_mst")
EXACT_RUNS = sorted({(m.group(1), int(m.group(2)))
                     for m in (re.match(r"^(.*)_exact(\d+)_", k) for k in EXACT) if m})


@pytest.mark.parametrize("name,mpts", EXACT_RUNS)
def test_exact_mst_bitexact(name, mpts):
    """CoreSG::exact_mst (coresg.cpp:419-430): Prim's edges in acceptance order,
    labels, component count and weight, as the reference wrote them."""
    _, sg = built(name)
    r = sg.exact_mst(mpts)
    p = f"{name}_exact{mpts}_"
    assert np.array_equal(r["u"], EXACT[p + "mst_u"])
    assert np.array_equal(r["v"], EXACT[p + "mst_v"])
    assert np.array_equal(r["w"].view(np.uint32), EXACT[p + "mst_w"].view(np.uint32))
    assert np.array_equal(r["labels"], EXACT[p + "component"])
    assert r["ncomp"] == int(EXACT[p + "num_components"][0])
    assert np.float64(r["weight"]).view(np.uint64) == EXACT[p + "weight"].view(np.uint64)[0]


@pytest.mark.parametrize("name", ["brute_d8", "islands"])
def test_exactmode_bridges_from_oracle(name, tmp_path):
    """MissingEdgeMode::ExactMst (coresg.cpp:157-163) as the reference CLI built
    it: the bridges are the exact MST at k_max over the file's lists, in Prim
    order, each re-stored with its base distance."""
    f = coresg_io.loads(EXACT[f"{name}_exactmode_coresg_bytes"].tobytes())
    assert f.mode == 1 and f.repair_count == 0
    path = os.path.join(str(tmp_path), "e.coresg")
    coresg_io.save(path, f)
    sg = oracle.CoreSG.load(path)
    e = sg.exact_mst(f.k_max)
    _, bw = oracle.l2sqr_pairs(f.points, np.stack([e["u"], e["v"]], 1).astype(np.uint32))
    u, v, w = f.bridge_uvw()
    assert np.array_equal(u, e["u"]) and np.array_equal(v, e["v"])
    assert np.array_equal(w.view(np.uint32), bw.view(np.uint32))


def test_exact_mst_theorem1_on_exact_knn():
    """Theorem 1 (test_coresg.cpp:32-61): with exact k-NN lists and the exact
    MST at k_max as the bridge set (MissingEdgeMode::ExactMst, coresg.cpp:157-163,
    bridges stored with their base distance), the CORE-SG holds an MST of the
    complete MRD graph for every mpts <= k_max, so the sorted weights of
    exact_mst equal those of Kruskal over the reweighted CORE-SG. brute_d8's
    lists come from brute force, so they are exact; Prim over every pair
    shares no code with the edge list and Kruskal."""
    g, sg = built("brute_d8")
    ids, dist, _, count = sg.graph()
    n, k = sg.n, sg.k
    padded = oracle.pad(g["data"])
    ku, kv, kw = oracle.undirected_edges(ids, dist, count, n, k)
    b = sg.exact_mst(k)
    _, bw = oracle.l2sqr_pairs(padded, np.stack([b["u"], b["v"]], 1).astype(np.uint32))
    u = np.concatenate([ku, b["u"]])
    v = np.concatenate([kv, b["v"]])
    w = np.concatenate([kw, bw])
    cnt = count.astype(np.int64)
    for m in (1, 2, 5, 12):
        col = np.minimum(m, np.maximum(cnt, 1)) - 1
        core = np.where(cnt > 0, dist.reshape(n, k)[np.arange(n), col], np.inf).astype(np.float32)
        rw = np.maximum(w, np.maximum(core[u], core[v]))
        _, _, mw, _, nc, _ = oracle.kruskal(n, u, v, rw)
        e = sg.exact_mst(m)
        assert e["ncomp"] == nc == 1
        assert np.array_equal(np.sort(e["w"]).view(np.uint32), np.sort(mw).view(np.uint32))
    # A planted mistake in the reweight (the second endpoint's core distance
    # left out) breaks it.
    rw = np.maximum(w, core[u])
    _, _, mw, _, _, _ = oracle.kruskal(n, u, v, rw)
    assert not np.array_equal(np.sort(e["w"]), np.sort(mw))


@pytest.mark.parametrize("name", CASES)
def test_mrd_block_and_prim_bitexact(name):
    g = golden(name)
    pts = oracle.pad(g["data"])[g["mrd_ids"]]
    mat = oracle.mrd_block(pts, g["mrd_core"])
    assert np.array_equal(mat.ravel().view(np.uint32), g["mrd_matrix"].view(np.uint32))
    u, v, w = oracle.prim_dense(mat)
    assert np.array_equal(u, g["mrd_prim_u"]) and np.array_equal(v, g["mrd_prim_v"])
    assert np.array_equal(w.view(np.uint32), g["mrd_prim_w"].view(np.uint32))


@pytest.mark.parametrize("name", CASES)
def test_sampled_recall_matches(name):
    g, sg = built(name)
    r, ratio = sg.sampled_recall(200, 42)
    # The reference is built with -ffast-math, which lets the compiler
    # vectorise its double reductions (eval.cpp:86-93), so the last bits of
    # the means depend on its summation order: compare within 1e-12.
    assert abs(r - g["recall_200_42"][0]) <= 1e-12
    assert abs(ratio - g["recall_200_42"][1]) <= 1e-12


# ---- closed forms and brute force (what the fixtures leave open) -------------

def test_mt19937_64_known_answer():
    # ISO C++ [rand.predef]: the 10000th output of a default-seeded
    # mt19937_64 (seed 5489) is 9981545732273789042.
    assert int(oracle.mt64_first(5489, 10000)[-1]) == 9981545732273789042


def test_core_dist_rules():
    # test_knn_graph.cpp:47-56: 1-based, clamped to the filled length, +inf
    # when empty.
    k = 4
    ids, dist, flags, count = oracle.empty_graph(3, k)
    L = oracle.lib()
    for c, d in ((1, 1.0), (2, 2.0), (3, 3.0)):
        L.orc_try_insert(ids, dist, flags, count, k, 0, c, d, 1, None)
    core = np.empty(3, np.float32)
    L.orc_core_distances(dist, count, 3, k, 1, core)
    assert core[0] == 1.0 and np.isinf(core[1])
    L.orc_core_distances(dist, count, 3, k, 3, core)
    assert core[0] == 3.0
    L.orc_core_distances(dist, count, 3, k, 4, core)
    assert core[0] == 3.0  # clamped to count-1


def test_insert_semantics():
    # knn_graph.cpp:37-81: self and duplicate rejected, equal distance goes
    # after existing equals, a full list rejects d >= last and reports eviction.
    import ctypes as C
    k = 3
    ids, dist, flags, count = oracle.empty_graph(2, k)
    L = oracle.lib()
    ev = C.c_uint32()
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 0, 1.0, 1, C.byref(ev)) == 0
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 5, 1.0, 1, C.byref(ev)) == 1
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 5, 0.5, 1, C.byref(ev)) == 0
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 6, 1.0, 0, C.byref(ev)) == 1
    assert list(ids[:2]) == [5, 6]
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 7, 0.5, 1, C.byref(ev)) == 1
    assert list(ids[:3]) == [7, 5, 6]
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 8, 1.0, 1, C.byref(ev)) == 0
    assert L.orc_try_insert(ids, dist, flags, count, k, 0, 9, 0.7, 1, C.byref(ev)) == 1
    assert ev.value == 6 and list(ids[:3]) == [7, 9, 5]


def test_mutual_pair_is_one_edge():
    # test_knn_graph.cpp:58-75.
    k = 2
    ids, dist, flags, count = oracle.empty_graph(3, k)
    L = oracle.lib()
    L.orc_try_insert(ids, dist, flags, count, k, 0, 1, 1.0, 1, None)
    L.orc_try_insert(ids, dist, flags, count, k, 1, 0, 1.0, 1, None)
    L.orc_try_insert(ids, dist, flags, count, k, 2, 1, 2.0, 1, None)
    u, v, w = oracle.undirected_edges(ids, dist, count, 3, k)
    assert sorted(zip(u.tolist(), v.tolist())) == [(0, 1), (1, 2)]


def test_known_tree_weight():
    # test_mst.cpp:14-25: triangle 0-1 (1), 1-2 (2), 0-2 (5) -> weight 3.
    u, v, w, lab, nc, wt = oracle.kruskal(3, np.array([0, 1, 0]), np.array([1, 2, 2]),
                                          np.array([1, 2, 5], np.float32))
    assert wt == 3.0 and nc == 1 and len(u) == 2


def test_kruskal_weight_matches_scipy():
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import minimum_spanning_tree
    rng = np.random.default_rng(5)
    n, m = 300, 3000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    u, v = u[keep], v[keep]
    key = np.minimum(u, v) * n + np.maximum(u, v)
    _, first = np.unique(key, return_index=True)
    u, v = u[first], v[first]
    w = rng.uniform(0.5, 2.0, u.size).astype(np.float32)
    ou, ov, ow, lab, nc, wt = oracle.kruskal(n, u, v, w)
    t = minimum_spanning_tree(coo_matrix((w.astype(np.float64), (u, v)), shape=(n, n)))
    assert abs(t.sum() - wt) < 1e-9 * max(1.0, wt)
    # forest labels agree with scipy's components
    from scipy.sparse.csgraph import connected_components
    ncs, labs = connected_components(coo_matrix((np.ones(u.size), (u, v)), shape=(n, n)),
                                     directed=False)
    assert nc == ncs
    # same partition, and labels dense by first appearance
    first_seen = {}
    for i, l in enumerate(labs):
        first_seen.setdefault(l, len(first_seen))
    assert np.array_equal(lab, np.array([first_seen[l] for l in labs]))


def test_prim_equals_kruskal_weight():
    # test_mst.cpp:65-84.
    rng = np.random.default_rng(2)
    pts = oracle.pad(rng.standard_normal((120, 5)).astype(np.float32))
    core = np.zeros(120, np.float32)
    mat = oracle.mrd_block(pts, core)
    pu, pv, pw = oracle.prim_dense(mat)
    iu, iv = np.triu_indices(120, 1)
    ku, kv, kw, _, _, wt = oracle.kruskal(120, iu.astype(np.uint32), iv.astype(np.uint32),
                                          mat[iu, iv])
    assert abs(float(np.sum(pw.astype(np.float64))) - wt) < 1e-9 * wt


def test_score_closed_form():
    # hubness.cpp:307-352 against an independent numpy statement of the rule.
    g = golden("mix_d16")
    ref = coresg_io.loads(g["coresg_bytes"].tobytes())
    n, k = ref.n, ref.k_max
    rd = oracle.reverse_degrees(ref.ids, ref.count, n, k)
    ids2 = ref.ids.reshape(n, k)
    expect = np.bincount(ids2[np.arange(k)[None, :] < ref.count[:, None]].astype(np.int64),
                         minlength=n)
    assert np.array_equal(rd, expect)
    sc = oracle.scores(ref.ids, ref.count, n, k, rd)
    for v in range(n):
        deg = int(ref.count[v])
        s = int(sum(rd[q] for q in ids2[v, :deg]))
        r = int(rd[v])
        if r > 0 and s <= 0:
            e = float(k)
        elif r <= 0 and s > 0:
            e = 0.0
        elif r <= 0 and s <= 0:
            e = 1.0
        else:
            e = r / (s / deg)
        assert sc[v] == e


def test_select_order_and_quota():
    rng = np.random.default_rng(3)
    n = 500
    score = rng.choice([0.0, 0.5, 1.0, 2.0], n)
    rdeg = rng.integers(0, 3, n).astype(np.int32)
    comp = (np.arange(n) % 7).astype(np.uint32)
    sel = oracle.select(score, rdeg, 20)
    order = np.lexsort((np.arange(n), rdeg, score))
    assert np.array_equal(sel, order[:20])
    sel2 = oracle.select(score, rdeg, 20, comp, 7, 2)
    want = set(order[:20].tolist())
    for c in range(7):
        members = order[comp[order] == c][:2]
        want |= set(members.tolist())
    assert set(sel2.tolist()) == want
    pos = {v: i for i, v in enumerate(order.tolist())}
    assert [pos[v] for v in sel2.tolist()] == sorted(pos[v] for v in sel2.tolist())


def test_budget_formula():
    # hubness.cpp:16-40 and test_hubness.cpp:165-196: llround(d ln|E|),
    # capped by sqrt(2*max_pair_ops/d), floored at min_selected, <= n.
    import math
    b, unc, cap = oracle.budget(1.0, 24, 26531390, 1_000_000)
    assert b == round(24 * math.log(26531390)) == 410 and not cap
    b, unc, cap = oracle.budget(1.0, 1000, 10**6, 10**6)
    assert b == int(math.sqrt(2 * 2e9 / 1000)) and cap
    assert oracle.budget(1.0, 4, 1, 100)[0] == 2
    assert oracle.budget(1.0, 64, 10**6, 50)[0] == 50


def test_nndescent_recall_floor():
    # test_nndescent.cpp:22-44: single-thread recall >= 0.95 (n=2000, d=8,
    # k=20, sample 12, 12 iterations) against brute force.
    rng = np.random.default_rng(11)
    x = oracle.pad(rng.standard_normal((2000, 8)).astype(np.float32))
    (ids, dist, flags, count), st = oracle.nndescent(x, 20, 12, 12, 42)
    assert oracle.lib().orc_validate(ids, dist, count, 2000, 20) == 1
    bids, _, _, _ = oracle.brute_force(x, 20)
    a = ids.reshape(2000, 20)
    b = bids.reshape(2000, 20)
    hits = sum(len(set(a[i]) & set(b[i])) for i in range(2000))
    assert hits / (2000 * 20) >= 0.95


def test_brute_force_ties_keep_lower_id():
    x = oracle.pad(np.zeros((6, 2), np.float32))
    ids, dist, _, count = oracle.brute_force(x, 3)
    assert list(ids.reshape(6, 3)[0]) == [1, 2, 3]
    assert list(ids.reshape(6, 3)[5]) == [0, 1, 2]


def test_mst_invariants_monotone_and_lower_bound():
    # test_coresg.cpp:116-149: every edge weight >= max(core) and the MST
    # weight does not decrease with mpts.
    g, sg = built("mix_d16")
    prev = -1.0
    ids, dist, _, count = sg.graph()
    for m in range(1, 17):
        u, v, w = sg.edges(m)
        core = np.empty(sg.n, np.float32)
        oracle.lib().orc_core_distances(dist, count, sg.n, sg.k, m, core)
        assert np.all(w >= np.maximum(core[u], core[v]))
        wt = sg.mst(m)["weight"]
        assert wt >= prev
        prev = wt
    with pytest.raises(ValueError):
        sg.edges(17)
