"""The reference's .coresg format through the device-resident graph
(coresg_hip_graph_load / _save, CoreSG::load / save, io.cpp:232-330): a
reference-written file loads onto the device, its derived state equals the
reference's state after load, its MSTs are the reference's, and saving it back
gives the same bytes; a device build saved equals the same arrays written by
the Python format module."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden, golden_cases, load_pkg_module

chip = load_pkg_module("coresg_hip")
coresg_io = load_pkg_module("coresg_io")
datagen = load_pkg_module("datagen")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", golden_cases())
def test_reference_file_roundtrip_state_and_msts(name, tmp_path):
    print(f"[gpu] .coresg load/save: {name}", flush=True)
    gd = golden(name)
    src = str(tmp_path / "ref.coresg")
    gd["coresg_bytes"].tofile(src)
    g = chip.Graph.load(src)
    # the derived state the reference's load() recomputes (default parameters)
    rd, sc, lab, nc = g.hubness()
    assert np.array_equal(rd, gd["state_rdeg"])
    assert np.array_equal(sc.view(np.uint64), gd["state_score"].view(np.uint64))
    assert np.array_equal(lab, gd["state_conn_labels"])
    assert nc == int(gd["state_conn_components"][0])
    assert np.array_equal(g.selected(), gd["state_selected"])
    out = str(tmp_path / "out.coresg")
    g.save(out)
    assert open(out, "rb").read() == gd["coresg_bytes"].tobytes()
    for m in gd["mpts"]:
        r = g.extract([int(m)], edges=True, labels=True)[0]
        o = np.lexsort((r["v"], r["u"], r["w"]))
        assert np.array_equal(r["u"][o], gd[f"mst{m}_mst_u"])
        assert np.array_equal(r["v"][o], gd[f"mst{m}_mst_v"])
        assert np.array_equal(r["labels"], gd[f"mst{m}_component"])
        assert np.float64(r["weight"]).view(np.uint64) == gd[f"mst{m}_weight"].view(np.uint64)[0]
    g.close()


def test_device_build_saves_the_reference_layout(tmp_path):
    print("[gpu] .coresg save of a device build", flush=True)
    x = datagen.gaussian_mixture(4000, 12, 6, 0.3, 2.0, 0.05, seed=21)
    g = chip.Graph(x.shape[0], x.shape[1], 10)
    g.set_points(x)
    g.build(chip.build_params())
    p = str(tmp_path / "b.coresg")
    g.save(p)
    ids, dist, flags, count = g.get_knn()
    u, v, w, rc = g.get_bridges()
    want = coresg_io.dumps(coresg_io.CoreSGFile(10, 12, 16, 0, 0, datagen.pad_rows(x), ids, dist,
                                                flags, count, coresg_io.bridges_from_uvw(u, v, w),
                                                rc))
    assert open(p, "rb").read() == want
    h = chip.Graph.load(p)
    a = g.extract([1, 5, 10], edges=True)
    b = h.extract([1, 5, 10], edges=True)
    for p_, q in zip(a, b):
        assert np.float64(p_["weight"]).view(np.uint64) == np.float64(q["weight"]).view(np.uint64)
        assert np.array_equal(p_["u"], q["u"]) and np.array_equal(p_["v"], q["v"])
    g.close()
    h.close()


def test_load_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.coresg"
    bad.write_bytes(b"NOTCORESG" + bytes(64))
    with pytest.raises(chip.CoresgHipError, match="not a coresg file"):
        chip.Graph.load(str(bad))
    gd = golden("mix_d16")
    cut = tmp_path / "cut.coresg"
    cut.write_bytes(gd["coresg_bytes"].tobytes()[:-20])
    with pytest.raises(chip.CoresgHipError, match="truncated"):
        chip.Graph.load(str(cut))


@pytest.mark.parametrize("name", ["brute_d8", "islands"])
def test_exact_mode_file_roundtrip(name, tmp_path):
    """Files the reference CLI wrote with --missing exact (mode 1) keep their
    mode and bytes through the device."""
    print(f"[gpu] .coresg load/save (ExactMst mode): {name}", flush=True)
    raw = golden("exact_mst")[f"{name}_exactmode_coresg_bytes"].tobytes()
    src = tmp_path / "x.coresg"
    src.write_bytes(raw)
    g = chip.Graph.load(str(src))
    out = str(tmp_path / "y.coresg")
    g.save(out)
    assert open(out, "rb").read() == raw
    g.close()
