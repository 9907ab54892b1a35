"""The sweep's autotuner on the GPU: every candidate launch shape of the
Boruvka round kernel gives the same MSTs (edges, weights, labels, rounds), and
the tuner keeps a shape at least as fast as the default for the graph and its
size bucket."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


def _graph(x, k):
    g = chip.Graph(x.shape[0], x.shape[1], k)
    g.set_points(x)
    g.build(chip.build_params(brute_force_limit=0))
    return g


def test_every_shape_gives_the_same_msts():
    print("[gpu] autotune: all sweep launch shapes agree (20k x 16, 1 and 16 slots)", flush=True)
    x, k = datagen.workload("mixture-20k-d16")
    g = _graph(x, k)
    lists = g.get_knn()
    bridges = g.get_bridges()
    log2n = int(np.floor(np.log2(x.shape[0])))
    for mpts in ([7], list(range(1, k + 1))):
        ref = g.extract(mpts, edges=True, labels=True)
        bp = 1 << (len(mpts) - 1).bit_length()
        for lanes in (4, 8, 16, 32, 64):
            if lanes < bp:
                continue
            for cap in (4096, 32768):
                chip.tune_table_load(f"{log2n} {k} {bp} {lanes} {cap}")
                h = chip.Graph(x.shape[0], x.shape[1], k)
                h.set_points(x)
                h.set_knn(*lists)
                h.set_bridges(bridges[0], bridges[1], bridges[2], bridges[3])
                got = h.extract(mpts, edges=True, labels=True)
                h.close()
                for p, q in zip(ref, got):
                    assert p["rounds"] == q["rounds"] and p["num_edges"] == q["num_edges"]
                    assert np.float64(p["weight"]).view(np.uint64) == \
                        np.float64(q["weight"]).view(np.uint64)
                    assert np.array_equal(p["u"], q["u"]) and np.array_equal(p["v"], q["v"])
                    assert np.array_equal(p["labels"], q["labels"])
    g.close()


def test_autotune_keeps_a_shape_no_slower_than_the_default():
    print("[gpu] autotune: 200k x 24, 32 slots and 1 slot", flush=True)
    x, k = datagen.workload("mixture-1m-d24", n=200_000)
    g = _graph(x, k)
    for mpts in (list(range(1, k + 1)), [16]):
        r = g.autotune_sweep(mpts)
        print(f"  {len(mpts)} slots: {r['lanes']} lanes, <= {r['max_blocks']} blocks, "
              f"{r['best_ms']:.2f} ms (default {r['default_ms']:.2f} ms, "
              f"{r['candidates']} shapes)", flush=True)
        assert r["slots"] == len(mpts) and r["lanes"] >= len(mpts)
        assert r["best_ms"] <= r["default_ms"]
        bp = 1 << (len(mpts) - 1).bit_length()
        assert f" {k} {bp} {r['lanes']} {r['max_blocks']}" in chip.tune_table()
    g.close()
