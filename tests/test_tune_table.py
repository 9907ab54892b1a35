"""The autotuner's table of sweep launch shapes, on the host (no GPU): lines
"log2n k_max slots lanes max_blocks" load, merge and print back; malformed or
impossible entries (lanes not a power of two in 4..64, fewer lanes than slots)
are refused."""
from __future__ import annotations

import pytest

from conftest import load_pkg_module

chip = load_pkg_module("coresg_hip")


def test_table_roundtrip_and_merge():
    chip.tune_table_load("30 32 32 64 8192\n30 16 1 4 4096\n")
    t = chip.tune_table()
    assert "30 32 32 64 8192" in t and "30 16 1 4 4096" in t
    chip.tune_table_load("30 32 32 32 16384")
    assert "30 32 32 32 16384" in chip.tune_table()


@pytest.mark.parametrize("bad", ["1 2 3", "19 32 32 48 8192", "19 32 32 16 8192",
                                 "19 32 3 8 8192", "19 32 8 8 0", "x y z"])
def test_table_rejects_bad_entries(bad):
    with pytest.raises(chip.CoresgHipError):
        chip.tune_table_load(bad)
