"""bench.py's counter-traffic bookkeeping: an entry of profiles/traffic.json is
used only while the source functions it was measured on are unchanged."""
from __future__ import annotations

import importlib.util
import os

from conftest import REPO


def _bench():
    spec = importlib.util.spec_from_file_location("bench_for_test", os.path.join(REPO, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


SRC = """
template <int N>
__device__ float helper(const float (&a)[N], float b) { return a[0] + b; }
__global__ void kern(float* x) {
    if (x) { x[0] = helper<1>({1.0f}, 2.0f); }
}
__global__ void other(float* x) { x[0] = 0.0f; }
"""


def test_fingerprint_follows_only_the_named_functions(tmp_path):
    b = _bench()
    p = tmp_path / "k.hip"
    p.write_text(SRC)
    f0 = b.source_fingerprint(str(p), ("helper", "kern"))
    p.write_text(SRC.replace("x[0] = 0.0f;", "x[0] = 1.0f;"))  # another function
    assert b.source_fingerprint(str(p), ("helper", "kern")) == f0
    p.write_text(SRC.replace("a[0] + b", "a[0] - b"))  # a named function's body
    assert b.source_fingerprint(str(p), ("helper", "kern")) != f0
    p.write_text(SRC.replace("helper", "renamed"))  # a named function gone
    assert b.source_fingerprint(str(p), ("helper", "kern")) != f0


def test_committed_join_traffic_matches_its_source():
    b = _bench()
    traffic, gathered = b.measured_traffic("nnd_join")
    assert traffic is not None and gathered is not None, \
        "profiles/traffic.json no longer describes the join's source: profile it again"
    assert traffic > gathered > 0


def test_committed_merge_traffic_matches_its_source():
    b = _bench()
    traffic, gathered = b.measured_traffic("nnd_merge")
    assert traffic is not None and gathered is not None, \
        "profiles/traffic.json no longer describes the merge's source: profile it again"
    assert traffic > gathered > 0
