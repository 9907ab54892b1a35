"""The plain one-launch-per-step form (coresg_hip_graph_set_plain: a host
synchronisation after every launch, the descent's convergence read after each
round, one mpts value per sweep pass with host-driven Boruvka rounds) gives the
fused form's results bit for bit; it exists as the measured baseline of the
fusion, not as a second algorithm."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


def _same_msts(a, b):
    assert len(a) == len(b)
    for p, q in zip(a, b):
        assert p["num_edges"] == q["num_edges"] and p["num_components"] == q["num_components"]
        assert np.float64(p["weight"]).view(np.uint64) == np.float64(q["weight"]).view(np.uint64)
        assert np.array_equal(p["u"], q["u"]) and np.array_equal(p["v"], q["v"])
        assert np.array_equal(p["w"].view(np.uint32), q["w"].view(np.uint32))
        assert np.array_equal(p["labels"], q["labels"])


def test_plain_build_and_sweep_equal_fused_on_exact_lists():
    """Brute-force lists make the whole build deterministic: plain and fused
    builds give the same lists, bridges and MSTs."""
    print("[gpu] plain vs fused: 6000 x 16 brute force, mpts 1..16", flush=True)
    x = datagen.gaussian_mixture(6000, 16, 8, 0.3, 2.0, 0.05, seed=11)
    k = 16
    res = []
    for plain in (False, True):
        g = chip.Graph(x.shape[0], x.shape[1], k)
        g.set_points(x)
        g.set_plain(plain)
        g.build(chip.build_params(brute_force_limit=10000))
        knn = g.get_knn()
        br = g.get_bridges()
        msts = g.extract(list(range(1, k + 1)), edges=True, labels=True)
        res.append((knn, br, msts))
        g.close()
    (ka, ba, ma), (kb, bb, mb) = res
    for a, b in zip(ka, kb):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    assert np.array_equal(ba[0], bb[0]) and np.array_equal(ba[1], bb[1]) and ba[3] == bb[3]
    _same_msts(ma, mb)


def test_plain_descent_and_sweep_on_shared_lists():
    """NN-Descent in the plain form runs its rounds with host reads in between
    and reaches the fused form's recall band; the sweep over one set of lists is
    identical in both forms."""
    print("[gpu] plain vs fused: 30000 x 24 descent + sweep mpts 1..32", flush=True)
    x = datagen.gaussian_mixture(30000, 24, 16, 0.2, 3.0, 0.02, seed=12)
    k = 32
    g = chip.Graph(x.shape[0], x.shape[1], k)
    g.set_points(x)
    g.set_plain(True)
    st = g.build(chip.build_params())
    assert 1 <= st["rounds"] <= 10
    rec_plain, _ = g.sampled_recall(200, 42)
    plain = g.extract(list(range(1, k + 1)), edges=True, labels=True)
    g.set_plain(False)
    fused = g.extract(list(range(1, k + 1)), edges=True, labels=True)
    _same_msts(plain, fused)
    g.build(chip.build_params())
    rec_fused, _ = g.sampled_recall(200, 42)
    g.close()
    assert rec_plain > 0.9 and abs(rec_plain - rec_fused) < 0.02
