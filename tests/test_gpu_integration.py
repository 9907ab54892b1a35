"""The drop-in: the reference's own host code (coresg_cli, IncrementalKnn,
CoreSG) running with HipBackend over libcoresg_hip.so. The k-NN descent and
the bridge MRD matrix run on the GPU through the Backend interface; the output
file is checked like any reference build."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from conftest import REPO, golden, load_pkg_module

coresg_io = load_pkg_module("coresg_io")
pytestmark = pytest.mark.gpu
CLI_HIP = os.path.join(REPO, "oracle", "_ref", "coresg_cli_hip")
CLI_REF = os.path.join(REPO, "oracle", "_ref", "coresg_cli")


@pytest.mark.skipif(not os.path.exists(CLI_HIP), reason="integration build not present")
def test_reference_cli_with_hip_backend(tmp_path):
    print("[gpu] reference coresg_cli with HipBackend", flush=True)
    gd = golden("mix_d16")
    x = gd["data"]
    data = tmp_path / "x.bin"
    x.tofile(data)
    out = tmp_path / "g.coresg"
    log = subprocess.run([CLI_HIP, "build", "--data", str(data), "--dim", str(x.shape[1]),
                          "--kmax", "16", "--threads", "1", "--instances", "1", "--out",
                          str(out)], check=True, capture_output=True, text=True).stdout
    assert re.search(r"instances\s+: 1 \(1 on device, 0 on host", log), log
    f = coresg_io.load(str(out))
    assert oracle.lib().orc_validate(f.ids, f.dist, f.count, f.n, f.k_max) == 1
    # the host part of the build, rerun by the oracle on the device-built graph,
    # gives the same bridges: the device MRD matrix is bit-identical
    sg = oracle.CoreSG.load(str(out))
    sg.rebuild_bridges()
    ou, ov, ow = sg.bridges()
    u, v, w = f.bridge_uvw()
    assert np.array_equal(u, ou) and np.array_equal(v, ov)
    assert np.array_equal(w.view(np.uint32), ow.view(np.uint32))
    # extraction through the reference CLI agrees line for line
    a = subprocess.run([CLI_HIP, "extract", "--in", str(out), "--threads", "1", "--mpts",
                        "1,4,16"], check=True, capture_output=True, text=True).stdout
    b = subprocess.run([CLI_REF, "extract", "--in", str(out), "--threads", "1", "--mpts",
                        "1,4,16"], check=True, capture_output=True, text=True).stdout
    strip = lambda s: [re.sub(r"\(.*s\)", "", l) for l in s.splitlines() if l.startswith("mpts")]
    assert strip(a) == strip(b)


EXTRACT_CHECK = os.path.join(REPO, "oracle", "_ref", "extract_check")


@pytest.mark.skipif(not os.path.exists(EXTRACT_CHECK), reason="integration build not present")
@pytest.mark.parametrize("name,mpts", [("mix_d16", "1 2 4 8 16"), ("ties", "1 3 8"),
                                       ("islands_repair", "1 3 5")])
def test_reference_coresg_routes_extraction_to_device(name, mpts, tmp_path):
    """The reference's CoreSG with HipBackend: all_msts routed through the
    DeviceExtract capability gives what the host backend gives, edge for edge
    (tests/cpp/extract_check.cpp)."""
    print(f"[gpu] DeviceExtract routing: {name} mpts {mpts}", flush=True)
    path = tmp_path / "g.coresg"
    golden(name)["coresg_bytes"].tofile(path)
    r = subprocess.run([EXTRACT_CHECK, str(path)] + mpts.split(), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "equal (backend hip)" in r.stdout


@pytest.mark.skipif(not os.path.exists(EXTRACT_CHECK), reason="integration build not present")
@pytest.mark.parametrize("name,mpts", [("mix_d16", "1 4 16 2 8"), ("ties", "1 3 8 2"),
                                       ("islands_repair", "1 3 5 2")])
def test_reference_coresg_exact_mst_routed_back_to_back(name, mpts, tmp_path):
    """The reference's CoreSG::exact_mst routed through HipBackend's
    DeviceExtract capability (the device Prim in its one-launch-per-step form
    over the resident copy, no cooperative launch) for several mpts in a row,
    after two all_msts calls in the same process: every result equals the host
    backend's bit for bit."""
    print(f"[gpu] exact_mst routed to the device: {name} mpts {mpts}", flush=True)
    path = tmp_path / "g.coresg"
    golden(name)["coresg_bytes"].tofile(path)
    r = subprocess.run([EXTRACT_CHECK, "--exact", str(path)] + mpts.split(),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "exact msts equal (backend hip)" in r.stdout





BACKEND_CHECK = os.path.join(REPO, "oracle", "_ref", "backend_check")


@pytest.mark.skipif(not os.path.exists(BACKEND_CHECK), reason="integration build not present")
@pytest.mark.parametrize("mode", ["small", "oom", "oom-stop", "cosine"])
def test_reference_coresg_backend_semantics(mode):
    """The reference's CoreSG::build through HipBackend keeps the CUDA backend's
    plugin semantics (cuda_backend.cpp:15-133; tests/cpp/backend_check.cpp):
    instances below 512 points on the host, a device out-of-memory (forced by
    a device budget) recorded and finished on the host -- or not run with
    oom_fallback off -- and the cosine metric on the device with the host's
    distances."""
    print(f"[gpu] backend semantics: {mode}", flush=True)
    r = subprocess.run([BACKEND_CHECK, mode], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"backend_check {mode}: passed" in r.stdout


HUBNESS_CHECK = os.path.join(REPO, "oracle", "_ref", "hubness_check")


@pytest.mark.skipif(not os.path.exists(HUBNESS_CHECK), reason="integration build not present")
@pytest.mark.parametrize("name,quota", [("mix_d16", 2), ("mix_d24", 2), ("islands", 2),
                                        ("islands_q0", 0), ("ties", 2)])
def test_reference_coresg_hubness_capability(name, quota, tmp_path):
    """The hubness half of CoreSG::build through HipBackend's DeviceExtract
    capability equals what the reference's HubnessIndex and
    ConnectivityTracker hold after load (tests/cpp/hubness_check.cpp)."""
    print(f"[gpu] hubness capability: {name} quota {quota}", flush=True)
    path = tmp_path / "g.coresg"
    golden(name)["coresg_bytes"].tofile(path)
    r = subprocess.run([HUBNESS_CHECK, str(path), str(quota)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "-> identical (backend hip)" in r.stdout
