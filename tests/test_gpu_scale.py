"""Parity at the benchmark configurations' own sizes.

- mixture-1m-d24 (n = 1M, d = 24, k_max = 32; the headline build): the device
  builds the CORE-SG, and its MSTs for mpts 1, 16 and 32 -- edges in Kruskal
  order, weight bits, labels, component count and the double total -- equal
  the oracle's Kruskal (mst.cpp:19-59) over the same lists and bridges. At this
  size the edge list has hub segments of several hundred entries (the long
  segment sort) and 26.5M edges.
- anisotropic 200k x 128 (64 components, eigenvalues log-uniform in
  [0.01, 4]) built with the BF16 Gram filter: the filter's bound holds on every
  visited pair (FP64 audit), its recall matches the exact join's, and the MSTs
  for mpts 4, 16 and 32 equal the oracle's over the same lists.
- batched uneven instances at k_max = 24 with d in {8, 32, 64}, mpts 2..24:
  per-instance MST summaries of the union sweep equal the oracle's Kruskal over
  each instance's own lists and bridges, for three instances per d.
"""
from __future__ import annotations

import os
import struct
import tempfile

import numpy as np
import pytest

import oracle
from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
coresg_io = load_pkg_module("coresg_io")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


def oracle_from_device(x, k, ids, dist, flags, count, bu, bv, bw, rc):
    """The oracle's CoreSG over the device's lists and bridges (through the
    reference's own file format)."""
    d = x.shape[1]
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "g.coresg")
        coresg_io.save(p, coresg_io.CoreSGFile(k, d, oracle.stride_of(d), 0, 0, oracle.pad(x),
                                               ids, dist, flags, count,
                                               coresg_io.bridges_from_uvw(bu, bv, bw), rc))
        return oracle.CoreSG.load(p)


def assert_mst_equal(r, ref):
    o = np.lexsort((r["v"], r["u"], r["w"]))  # Kruskal (edge_less) order
    assert np.array_equal(r["u"][o], ref["u"]) and np.array_equal(r["v"][o], ref["v"])
    assert np.array_equal(r["w"][o].view(np.uint32), ref["w"].view(np.uint32))
    assert np.array_equal(r["labels"], ref["labels"])
    assert r["num_components"] == ref["ncomp"]
    assert np.float64(r["weight"]).view(np.uint64) == np.float64(ref["weight"]).view(np.uint64)


@pytest.mark.timeout(600)
def test_mixture_1m_msts_bitexact_against_oracle():
    print("[gpu] scale: mixture-1m-d24 build, mpts 1/16/32 vs oracle Kruskal", flush=True)
    x, k = datagen.workload("mixture-1m-d24")
    g = chip.Graph(x.shape[0], x.shape[1], k)
    g.set_points(x)
    g.build(chip.build_params(sample=16, iters=10))
    ne, ke = g.prepare_edges()
    mpts = [1, 16, 32]
    res = g.extract(mpts, edges=True, labels=True)
    ids, dist, flags, count = g.get_knn()
    bu, bv, bw, rc = g.get_bridges()
    g.close()
    sg = oracle_from_device(x, k, ids, dist, flags, count, bu, bv, bw, rc)
    assert sg.knn_edges() == ke and ne == ke + bu.size
    # the hubs' segments exceed one 64-entry chunk of the segment sort
    seg = np.bincount(np.concatenate([ids[ids != 0xFFFFFFFF],
                                      np.repeat(np.arange(x.shape[0]), count)]),
                      minlength=x.shape[0])
    assert seg.max() > 64
    for m, r in zip(mpts, res):
        print(f"  mpts {m}: {r['num_edges']} edges, weight {r['weight']:.6f}", flush=True)
        assert_mst_equal(r, sg.mst(m))


@pytest.mark.timeout(600)
def test_anisotropic_200k_d128_bf16_filter_msts_bitexact():
    print("[gpu] scale: anisotropic 200k x 128, BF16 Gram filter, mpts 4/16/32", flush=True)
    f32 = lambda b: struct.unpack("f", struct.pack("I", b & 0xFFFFFFFF))[0]
    n, d, k = 200_000, 128, 32
    x = datagen.anisotropic_mixture(n, d, 64, 0.01, 4.0, seed=4)
    g = chip.Graph(n, d, k)
    g.set_points(x)
    g.build(chip.build_params(gram=1))
    rec_exact, _ = g.sampled_recall(300, 42)
    g.reset_timers()
    g.set_profiling(True)
    g.set_audit(True)
    g.build(chip.build_params(gram=3))
    g.set_audit(False)
    g.set_profiling(False)
    pairs, exact_pairs, err_bits, ratio_bits = g.kernel_work("nnd_gram")
    rec_bf16, _ = g.sampled_recall(300, 42)
    print(f"  audit: {pairs} pairs, {exact_pairs} exact; max rel error vs FP64 "
          f"{f32(err_bits):.3e}, max bound ratio {f32(ratio_bits):.5f}; recall exact "
          f"{rec_exact:.4f}, bf16 {rec_bf16:.4f}", flush=True)
    assert pairs > 0 and 0 < exact_pairs < pairs
    assert f32(ratio_bits) <= 1.0  # the lower bound never exceeds the FP64 distance
    assert rec_bf16 >= rec_exact - 0.02
    g.prepare_edges()
    mpts = [4, 16, 32]
    res = g.extract(mpts, edges=True, labels=True)
    ids, dist, flags, count = g.get_knn()
    bu, bv, bw, rc = g.get_bridges()
    g.close()
    sg = oracle_from_device(x, k, ids, dist, flags, count, bu, bv, bw, rc)
    for m, r in zip(mpts, res):
        assert_mst_equal(r, sg.mst(m))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("d", [8, 32, 64])
def test_batched_instances_k24_against_oracle(d):
    print(f"[gpu] scale: batched instances k_max=24 d={d}, mpts 2..24 vs oracle", flush=True)
    k = 24
    rng = np.random.default_rng(100 + d)
    sizes = np.exp(rng.uniform(np.log(2_000), np.log(200_000), size=6)).astype(int)
    xs = [datagen.gaussian_mixture(int(m), d, int(rng.integers(4, 41)), 0.2, 3.0,
                                   float(rng.uniform(0.0, 0.10)), seed=7000 + 10 * d + i)
          for i, m in enumerate(sizes)]
    off = np.concatenate([[0], np.cumsum([x.shape[0] for x in xs])]).astype(np.uint32)
    g = chip.Graph(int(off[-1]), d, k)
    g.set_points(np.concatenate(xs))
    g.set_instances(off)
    g.build(chip.build_params(sample=12, iters=10))
    mpts = list(range(2, k + 1))
    per = g.extract_instances(mpts)
    ids, dist, flags, count = g.get_knn()
    bu, bv, bw, rc = g.get_bridges()
    g.close()
    # three instances: the smallest, the largest and one between
    order = np.argsort(sizes)
    for i in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        lo, hi = int(off[i]), int(off[i + 1])
        li = ids[lo * k:hi * k].copy()
        live = li != 0xFFFFFFFF
        assert np.all((li[live] >= lo) & (li[live] < hi))
        li[live] -= np.uint32(lo)
        mine = (bu >= lo) & (bu < hi)
        sg = oracle_from_device(xs[i], k, li, dist[lo * k:hi * k].copy(),
                                flags[lo * k:hi * k].copy(), count[lo:hi].copy(),
                                bu[mine] - lo, bv[mine] - lo, bw[mine], 0)
        for q, m in enumerate(mpts):
            ref = sg.mst(m)
            u = per[i][q]
            assert u["num_edges"] == ref["u"].size and u["num_components"] == ref["ncomp"], (i, m)
            assert np.float64(u["weight"]).view(np.uint64) == \
                np.float64(ref["weight"]).view(np.uint64), (i, m)
