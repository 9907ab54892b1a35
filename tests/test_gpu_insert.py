"""CoreSG::insert on the GPU (coresg.cpp:325-389).

- The batch's lists are the exact k nearest by (distance, id), as the
  reference's brute force builds them (nndescent.cpp:242-263).
- Every existing list receives the batch points that try_insert would accept
  (knn_graph.cpp:37-55), merged after equal distances already present.
- Hubness, selection, bridges and MSTs of the grown graph equal the oracle's
  on the same graph.
- After a stream of inserts, sampled recall is at least that of the reference
  CLI's own `update` on the same data.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

import oracle
from conftest import GOLDEN_DIR, REPO, golden, load_pkg_module

sys.path.insert(0, GOLDEN_DIR)
import refdump  # noqa: E402

chip = load_pkg_module("coresg_hip")
coresg_io = load_pkg_module("coresg_io")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu

REF_CLI = os.path.join(REPO, "oracle", "_ref", "coresg_cli")
REF_DUMP = os.path.join(REPO, "oracle", "_ref", "ref_dump")


def expected_lists(xp, n0, k, ids0, dist0, flags0, count0):
    """The lists after inserting points n0.. of xp into lists (ids0, ...) of
    points 0..n0-1: exact brute-force rows for the batch; for each existing
    point, the accepted batch points ((d, id) order, below its k-th when the
    list is full) merged after existing entries of equal distance."""
    n = xp.shape[0]
    bids, bdist, _, bcount = oracle.brute_force(xp, k)
    ids = np.full((n, k), 0xFFFFFFFF, np.uint32)
    dist = np.full((n, k), np.inf, np.float32)
    flags = np.zeros((n, k), np.uint8)
    count = np.zeros(n, np.uint32)
    bids, bdist = bids.reshape(n, k), bdist.reshape(n, k)
    ids[n0:], dist[n0:], count[n0:] = bids[n0:], bdist[n0:], bcount[n0:]
    flags[n0:, :] = (np.arange(k)[None, :] < bcount[n0:, None]).astype(np.uint8)
    new = np.arange(n0, n, dtype=np.uint32)
    pairs = np.stack([np.repeat(np.arange(n0, dtype=np.uint32), n - n0), np.tile(new, n0)], 1)
    _, dq = oracle.l2sqr_pairs(xp, pairs)
    dq = dq.reshape(n0, n - n0)
    ids0, dist0, flags0 = ids0.reshape(n0, k), dist0.reshape(n0, k), flags0.reshape(n0, k)
    for q in range(n0):
        c = int(count0[q])
        last = dist0[q, k - 1] if c == k else np.float32(np.inf)
        ok = dq[q] < last
        cand = sorted(zip(dq[q][ok].tolist(), new[ok].tolist()))[:k]
        entries = [(float(dist0[q, i]), 0, i, int(ids0[q, i]), int(flags0[q, i])) for i in range(c)]
        entries += [(d, 1, r, p, 1) for r, (d, p) in enumerate(cand)]
        entries.sort(key=lambda e: (e[0], e[1], e[2]))
        entries = entries[:k]
        count[q] = len(entries)
        for i, e in enumerate(entries):
            dist[q, i], ids[q, i], flags[q, i] = e[0], e[3], e[4]
    return ids.ravel(), dist.ravel(), flags.ravel(), count


def check_downstream(g, x, k):
    """Selection, bridges and MSTs of the device graph == the oracle's on the
    same lists."""
    ids, dist, flags, count = g.get_knn()
    u, v, w, rc = g.get_bridges()
    path = tempfile.mktemp(suffix=".coresg")
    fobj = coresg_io.CoreSGFile(k, x.shape[1], (x.shape[1] + 7) // 8 * 8, 0, 0, oracle.pad(x), ids,
                                dist, flags, count,
                                coresg_io.bridges_from_uvw(u, v, w), rc)
    coresg_io.save(path, fobj)
    sg = oracle.CoreSG.load(path)
    os.unlink(path)
    assert np.array_equal(g.selected(), sg.selected())
    sg.rebuild_bridges()
    ou, ov, ow = sg.bridges()
    assert np.array_equal(u, ou) and np.array_equal(v, ov)
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32)) and rc == sg.repair_count()
    mp = [1, k // 2 or 1, k]
    for m, rr in zip(mp, g.extract(mp, edges=True, labels=True)):
        o = sg.mst(m)
        oo = np.lexsort((rr["v"], rr["u"], rr["w"]))
        assert np.array_equal(rr["u"][oo], o["u"]) and np.array_equal(rr["v"][oo], o["v"])
        assert np.array_equal(rr["labels"], o["labels"])
        assert np.float64(rr["weight"]).view(np.uint64) == np.float64(o["weight"]).view(np.uint64)


@pytest.mark.parametrize("gram", [1, 2, 3])
@pytest.mark.parametrize("name", ["mix_d16", "mix_d24", "ties"])
def test_insert_lists_exact(name, gram):
    """gram 1: the exact scan; 2 / 3: the FP16 / BF16 Gram-filter scan, which
    must give the same lists."""
    print(f"[gpu] insert exactness: {name} gram={gram}", flush=True)
    gd = golden(name)
    x = gd["data"]
    k = int(gd["k"])
    n = x.shape[0]
    n0 = (3 * n) // 4
    xp = oracle.pad(x)
    g = chip.Graph(n0, x.shape[1], k)
    g.set_points(x[:n0])
    g.build(chip.build_params())
    ids0, dist0, flags0, count0 = g.get_knn()
    st = g.insert(x[n0:], chip.build_params(gram=gram))
    assert g.n == n
    ids, dist, flags, count = g.get_knn()
    eids, edist, eflags, ecount = expected_lists(xp, n0, k, ids0, dist0, flags0, count0)
    assert np.array_equal(count, ecount)
    assert np.array_equal(ids, eids)
    assert np.array_equal(dist.view(np.uint32), edist.view(np.uint32))
    assert np.array_equal(flags, eflags)
    print(f"  applied {st['updates']} proposals to existing lists", flush=True)
    check_downstream(g, x, k)
    g.close()


def test_insert_stream_grows_and_matches_oracle():
    """Three inserts into a 20k x 16 mixture (capacity growth across calls);
    the final graph's bridges and MSTs equal the oracle's on it."""
    print("[gpu] insert stream 12k + 3 batches", flush=True)
    x, k = datagen.workload("mixture-20k-d16")
    g = chip.Graph(12000, x.shape[1], k)
    g.set_points(x[:12000])
    g.build(chip.build_params())
    for a, b in ((12000, 12001), (12001, 16000), (16000, 20000)):
        g.insert(x[a:b])
        assert g.n == b
    ids, dist, flags, count = g.get_knn()
    assert oracle.lib().orc_validate(ids, dist, count, x.shape[0], k) == 1
    # the last batch's lists are exact
    bids, bdist, _, bcount = oracle.brute_force(oracle.pad(x), k)
    sl = slice(16000 * k, 20000 * k)
    assert np.array_equal(ids[sl], bids[sl]) and np.array_equal(count[16000:], bcount[16000:])
    check_downstream(g, x, k)
    g.close()


@pytest.mark.parametrize("gram", [1, 2])
def test_insert_graph_search_lists_valid(gram):
    """insert_search = 1 (beam search over the existing lists instead of the
    scan; the batch's own pairs by the exact scan, gram 1, or the FP16
    Gram-filter scan, gram 2): every stored entry is its pair's exact
    distance, lists are sorted and distinct, the batch's own pairs are exact,
    the new points' lists reach the exact lists closely, and everything
    downstream equals the oracle on the resulting lists."""
    print(f"[gpu] insert graph search: 15k + 5k x 16, gram={gram}", flush=True)
    x, k = datagen.workload("mixture-20k-d16")
    n0, n = 15000, 20000
    xp = oracle.pad(x)
    g = chip.Graph(n0, x.shape[1], k)
    g.set_points(x[:n0])
    g.build(chip.build_params())
    ids0, dist0, flags0, count0 = g.get_knn()
    st = g.insert(x[n0:], chip.build_params(insert_search=1, gram=gram))
    assert g.n == n
    ids, dist, flags, count = g.get_knn()
    assert oracle.lib().orc_validate(ids, dist, count, n, k) == 1
    live = ids != 0xFFFFFFFF
    rows = np.repeat(np.arange(n, dtype=np.uint32), k)[live]
    _, dq = oracle.l2sqr_pairs(xp, np.stack([rows, ids[live]], 1))
    assert np.array_equal(dq.view(np.uint32), dist[live].view(np.uint32))
    eids, edist, eflags, ecount = expected_lists(xp, n0, k, ids0, dist0, flags0, count0)
    new = slice(n0 * k, n * k)
    a, b = ids[new].reshape(-1, k), eids[new].reshape(-1, k)
    rec_new = np.mean([len(set(a[i]) & set(b[i])) / k for i in range(n - n0)])
    a, b = ids[:n0 * k].reshape(-1, k), eids[:n0 * k].reshape(-1, k)
    rec_old = np.mean([len(set(a[i]) & set(b[i])) / k for i in range(n0)])
    print(f"  new points' lists vs exact: {rec_new:.4f}; existing lists vs exact insert: "
          f"{rec_old:.4f}; {st['updates']} proposals applied", flush=True)
    # measured on this workload: 0.9935 and 0.9910 (the refinement rounds
    # after the search close most of what the beam misses)
    assert rec_new >= 0.98 and rec_old >= 0.98
    # the batch's own pairs are scanned exactly (pruned only by the search's
    # k-th key, which no dropped pair can beat): a batch point clearly closer
    # to a new point than that point's final k-th neighbour is in its list
    # (refinement only lowers the k-th, so this holds after it too)
    xb = x[n0:].astype(np.float64)
    nb = xb.shape[0]
    kth = dist.reshape(n, k)[n0:, k - 1].astype(np.float64)
    assert np.all(count[n0:] == k)
    lists = ids.reshape(n, k)[n0:]
    for r0 in range(0, nb, 1000):
        blk = xb[r0:r0 + 1000]
        d = np.sqrt(np.maximum(
            (blk * blk).sum(1)[:, None] + (xb * xb).sum(1)[None, :] - 2.0 * blk @ xb.T, 0.0))
        for r in range(blk.shape[0]):
            q = r0 + r
            near = np.nonzero(d[r] < kth[q] * (1.0 - 1e-4))[0]
            near = near[near != q] + n0
            assert np.isin(near, lists[q]).all(), f"new point {n0 + q}: batch neighbour missing"
    check_downstream(g, x, k)
    g.close()


def test_insert_brute_force_limit_rebuilds_exactly():
    gd = golden("brute_d8")
    x = gd["data"]
    k = int(gd["k"])
    n = x.shape[0]
    p = chip.build_params(brute_force_limit=n)
    g = chip.Graph(n // 2, x.shape[1], k)
    g.set_points(x[:n // 2])
    g.build(p)
    g.insert(x[n // 2:], p)
    ids, dist, flags, count = g.get_knn()
    bids, bdist, _, bcount = oracle.brute_force(oracle.pad(x), k)
    assert np.array_equal(ids, bids) and np.array_equal(dist.view(np.uint32), bdist.view(np.uint32))
    g.close()


def test_insert_errors():
    g = chip.Graph(10, 3, 4)
    g.set_points(np.zeros((10, 3), np.float32))
    with pytest.raises(chip.CoresgHipError):
        g.insert(np.zeros((2, 3), np.float32))  # not built
    with pytest.raises(ValueError):
        g.insert(np.zeros((2, 5), np.float32))
    g.close()


@pytest.mark.skipif(not (os.path.exists(REF_CLI) and os.path.exists(REF_DUMP)),
                    reason="reference CLI not built (make -C oracle ref)")
@pytest.mark.parametrize("search", [0, 1])
def test_insert_recall_vs_reference_update(search):
    """100k seed points of the 1M x 24 mixture, then three batches of 20k:
    the reference CLI's `build` + `update` (16 threads) against the device
    build + insert (exact scan, and graph search), scored by the same sampled
    recall (3000 queries, seed 7)."""
    print(f"[gpu] insert recall at scale vs reference update (search={search})", flush=True)
    xf, k = datagen.workload("mixture-1m-d24")
    seed_n, batches = 100_000, [(100_000, 120_000), (120_000, 140_000), (140_000, 160_000)]
    x = np.ascontiguousarray(xf[:batches[-1][1]])
    with tempfile.TemporaryDirectory() as tmp:
        x[:seed_n].tofile(os.path.join(tmp, "seed.bin"))
        cur = os.path.join(tmp, "g0.coresg")
        common = ["--dim", str(x.shape[1]), "--threads", "16", "--verbose", "0"]
        subprocess.run([REF_CLI, "build", "--data", os.path.join(tmp, "seed.bin"), "--kmax", str(k),
                        "--instances", "1", "--gpu", "0", "--sample", "16", "--iters", "10",
                        "--out", cur] + common, check=True, capture_output=True, timeout=120)
        for i, (a, b) in enumerate(batches):
            bp = os.path.join(tmp, f"b{i}.bin")
            x[a:b].tofile(bp)
            nxt = os.path.join(tmp, f"g{i + 1}.coresg")
            subprocess.run([REF_CLI, "update", "--in", cur, "--data", bp, "--gpu", "0",
                            "--out", nxt] + common, check=True, capture_output=True, timeout=120)
            cur = nxt
        subprocess.run([REF_DUMP, "recall", cur, "3000", "7", os.path.join(tmp, "r.bin")],
                       check=True, capture_output=True, timeout=120)
        ref = refdump.read_tagged(os.path.join(tmp, "r.bin"))["recall"]
    g = chip.Graph(seed_n, x.shape[1], k)
    g.set_points(x[:seed_n])
    p = chip.build_params(sample=16, iters=10)
    g.build(p)
    for a, b in batches:
        g.insert(x[a:b], chip.build_params(sample=16, iters=10, insert_search=search))
    ours, ratio = g.sampled_recall(3000, 7)
    g.close()
    print(f"  recall gpu={ours:.5f} reference={ref[0]:.5f}", flush=True)
    assert ours >= ref[0]
