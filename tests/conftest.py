"""Shared test setup: the `gpu` marker, package loading and golden fixtures."""
from __future__ import annotations

import glob
import importlib.util
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(REPO, "incremental-core-sg_amd")
GOLDEN_DIR = os.path.join(REPO, "tests", "golden")
if REPO not in sys.path:
    sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_pkg_module(name: str):
    """Import a module of the package directory (its name has a hyphen, so it
    is loaded by path rather than by `import`)."""
    full = f"coresg_amd_{name}"
    if full in sys.modules:
        return sys.modules[full]
    spec = importlib.util.spec_from_file_location(full, os.path.join(PKG_DIR, f"{name}.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[full] = mod
    spec.loader.exec_module(mod)
    return mod


def golden_cases():
    """Per-input fixture files; exact_mst.npz holds extra outputs for several of
    them, maxreps.npz its own set of builds (the repair's representative cut) and
    tsv.npz the reference CLI's MST output files."""
    names = (os.path.splitext(os.path.basename(p))[0]
             for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))
    return sorted(x for x in names if x not in ("exact_mst", "maxreps", "tsv"))


def maxreps_cases():
    """(tag, data, max_reps, .coresg bytes) of the reference builds whose
    connectivity repair cuts more than max_reps representatives."""
    g = golden("maxreps")
    out = []
    for seed, jitter, max_reps in g["cases"]:
        tag = f"s{seed}j{jitter}r{max_reps}"
        out.append((tag, g[f"{tag}_data"], int(max_reps), g[f"{tag}_coresg_bytes"]))
    return out


def golden(name: str) -> dict:
    with np.load(os.path.join(GOLDEN_DIR, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
