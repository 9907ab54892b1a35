"""The library's writers of the reference's output formats, on the host (no
GPU): MST edges as TSV byte for byte as the reference CLI's `extract
--out-prefix` wrote them (write_edges_tsv, io.cpp:131-138; tests/golden/tsv.npz),
labels as TSV by the format's definition (write_labels_tsv, io.cpp:140-145)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import golden, load_pkg_module

chip = load_pkg_module("coresg_hip")

TSV = [("mix_d16", 4), ("ties", 3), ("islands_repair", 5)]


@pytest.mark.parametrize("name,m", TSV)
def test_edges_tsv_matches_reference_cli_bytes(name, m, tmp_path):
    g = golden(name)
    want = golden("tsv")[f"{name}_mpts{m}_tsv"].tobytes()
    p = str(tmp_path / "e.tsv")
    # the reference's MST edges in its Kruskal acceptance order
    chip.write_edges_tsv(p, g[f"mst{m}_mst_u"], g[f"mst{m}_mst_v"], g[f"mst{m}_mst_w"])
    assert open(p, "rb").read() == want


def test_edges_tsv_number_format(tmp_path):
    p = str(tmp_path / "e.tsv")
    w = np.array([0.0, 1.0, 2.4376630783081055, 1e-7, 3.4028234663852886e38, np.inf],
                 np.float32)
    chip.write_edges_tsv(p, np.arange(6), np.arange(6) + 1, w)
    lines = open(p).read().splitlines()
    assert lines[0] == "u\tv\tweight"
    assert lines[1:] == ["0\t1\t0.000000000e+00", "1\t2\t1.000000000e+00",
                         "2\t3\t2.437663078e+00", "3\t4\t1.000000012e-07",
                         "4\t5\t3.402823466e+38", "5\t6\tinf"]


def test_labels_tsv(tmp_path):
    p = str(tmp_path / "l.tsv")
    chip.write_labels_tsv(p, np.array([0, 0, 1, 2, 1], np.uint32))
    assert open(p).read() == "id\tlabel\n0\t0\n1\t0\n2\t1\n3\t2\n4\t1\n"


def test_tsv_write_error_is_reported(tmp_path):
    with pytest.raises(chip.CoresgHipError, match="cannot open"):
        chip.write_edges_tsv(str(tmp_path / "no" / "such" / "dir.tsv"), [0], [1], [1.0])
