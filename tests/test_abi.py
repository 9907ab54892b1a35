"""The C-ABI library loads and exports every entry point include/coresg_hip.h
declares (no compute calls: this runs without a GPU)."""
from __future__ import annotations

import os
import re
import subprocess

from conftest import REPO, load_pkg_module

chip = load_pkg_module("coresg_hip")
HEADER = os.path.join(REPO, "include", "coresg_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(coresg_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(chip.LIB_PATH), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", chip.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    bound = {name for name, _, _ in chip.SIGNATURES}
    assert set(declared_functions()) == bound


def test_library_loads_and_reports_abi():
    L = chip.lib()
    assert L.coresg_hip_abi_version() == 1
    p = chip.build_params()
    assert p.sample == 16 and p.iters == 10 and p.min_per_component == 2
