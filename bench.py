"""CORE-SG benchmark on one MI355X.

A step is one pass of the hot path over one synthetic batch resident in HBM:
build the CORE-SG (k-NN descent + hubness + anti-hub bridges), lay out its edge
list and adjacency, and extract the full mpts sweep (one MRD reweight + Boruvka
MST per mpts, all mpts in one call). Default workload:
n = 1,000,000 points, d = 24, a 32-component Gaussian mixture (std 0.2-3.0,
2% uniform noise, seed 1), k_max = 32, NN-Descent sample 16 / 10 iterations,
mpts 1..32.

Prints one JSON line. `value` is the build rate in points/s; the sweep rate
(MSTs/s), the same sweep as one call per mpts value, the latency percentiles
of single-mpts extract queries, build + a 16-value sweep in the plain
one-launch-per-step form against the fused form, the dominant kernel's
roofline fraction (device time from HIP events on the library's stream) and
the reference CPU path on the same 1M points, timed on this host, sit beside
it.

    python bench.py [--steps K] [--warmup W] [--n N] [--no-cpu-baseline] [--cpu-threads T]
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "incremental-core-sg_amd")

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def _load(name):
    full = f"coresg_amd_{name}"
    if full in sys.modules:
        return sys.modules[full]
    spec = importlib.util.spec_from_file_location(full, os.path.join(PKG, f"{name}.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[full] = mod
    spec.loader.exec_module(mod)
    return mod


def percentile(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    i = min(len(xs) - 1, max(0, int(round(p / 100.0 * (len(xs) - 1)))))
    return xs[i]


# The source functions each counter entry describes: their text is
# fingerprinted, so a change to one of them marks the entry stale.
TRAFFIC_SOURCES = {
    "nnd_join": ("incremental-core-sg_amd/csrc/nndescent.hip",
                 ("l2sqr_reg", "l2sqr_reg_halves", "l2sqr_join", "join_stage_reg", "join_partner",
                  "lanes_below", "wave_max_u32", "join_reg_kernel", "join_list_cap")),
    "nnd_merge": ("incremental-core-sg_amd/csrc/nndescent.hip",
                  ("merge_kernel", "small_unique_top", "sort_unique_top", "select_top_large",
                   "wave_sort_regs", "count_before", "list_slot")),
}


def source_fingerprint(path, names):
    """sha256 over the definitions (signature and body) of the named
    functions in a source file, each found as `name(` followed by a balanced
    parameter list and a braced body."""
    import hashlib
    with open(path) as f:
        text = f.read()
    h = hashlib.sha256()
    for name in names:
        for m in re.finditer(r"\b" + re.escape(name) + r"\s*\(", text):
            i, depth = m.end() - 1, 0
            while i < len(text):  # the parameter list
                depth += {"(": 1, ")": -1}.get(text[i], 0)
                i += 1
                if depth == 0:
                    break
            j = i
            while j < len(text) and text[j] not in "{;":
                j += 1
            if j >= len(text) or text[j] == ";" or text[i:j].strip() not in ("", "const"):
                continue  # a call or a declaration
            depth, k = 0, j
            while k < len(text):  # the body
                depth += {"{": 1, "}": -1}.get(text[k], 0)
                k += 1
                if depth == 0:
                    break
            h.update(text[m.start():k].encode())
            break
        else:
            h.update(("missing:" + name).encode())
    return h.hexdigest()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the same
    counters on a copy kernel of known size; and scaled by a random-row
    gather probe of known bytes), or (None, None) when there is none or when
    the kernel's source functions have changed since the counters were
    taken."""
    path = os.path.join(REPO, "profiles", "traffic.json")
    try:
        with open(path) as f:
            e = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None, None
    if not e or kernel not in TRAFFIC_SOURCES:
        return None, None
    src, names = TRAFFIC_SOURCES[kernel]
    try:
        if source_fingerprint(os.path.join(REPO, src), names) != e.get("source_fingerprint"):
            return None, None  # stale: the counters describe other code
    except OSError:
        return None, None
    return e.get("bytes_per_launch"), e.get("bytes_per_launch_gather_scaled")


def cpu_baseline(x, k, extract_mpts, threads, recall_queries):
    """The reference CPU path (oracle/_ref/coresg_cli, built from the reference
    sources) on the whole workload: the build with its sampled recall on the
    same queries as the device's (sampled_recall, seed 42), then extract on one
    mpts value and on `extract_mpts`; the difference of the two extract runs is
    the per-mpts cost (edges(mpts) + MST) without the file load."""
    cli = os.path.join(REPO, "oracle", "_ref", "coresg_cli")
    if not os.path.exists(cli):
        return None
    n, d = x.shape
    with tempfile.TemporaryDirectory() as tmp:
        data = os.path.join(tmp, "x.bin")
        x.tofile(data)
        sg = os.path.join(tmp, "x.coresg")
        t0 = time.perf_counter()
        out = subprocess.run([cli, "build", "--data", data, "--dim", str(d), "--kmax", str(k),
                              "--threads", str(threads), "--instances", "1", "--gpu", "0",
                              "--sample", "16", "--iters", "10", "--verbose", "0",
                              "--recall-sample", str(recall_queries), "--out", sg],
                             check=True, capture_output=True, text=True).stdout
        build_wall = time.perf_counter() - t0
        m = re.search(r"total ([0-9.]+)", out)
        build_s = float(m.group(1)) if m else build_wall
        m = re.search(r"recall\s*:\s*([0-9.]+) over", out)
        recall = float(m.group(1)) if m else None
        print(f"[bench] cpu baseline: build {build_s:.2f} s on {threads} threads, recall {recall}",
              file=sys.stderr, flush=True)

        def extract(mp):
            t = time.perf_counter()
            o = subprocess.run([cli, "extract", "--in", sg, "--threads", str(threads), "--mpts",
                                ",".join(str(v) for v in mp), "--verbose", "0"],
                               check=True, capture_output=True, text=True).stdout
            return time.perf_counter() - t, sum(float(v) for v in re.findall(r"\(([0-9.]+)s\)", o))

        one_wall, _ = extract(extract_mpts[-1:])
        many_wall, mst_only = extract(extract_mpts)
    per_mpts = (many_wall - one_wall) / (len(extract_mpts) - 1)
    return dict(build_s=build_s, build_pts_per_s=n / build_s, recall_sampled=recall,
                extract_per_mpts_s=per_mpts, extract_load_s=one_wall - per_mpts,
                mst_only_per_mpts_s=mst_only / len(extract_mpts),
                msts_per_s=1.0 / per_mpts, extract_mpts=list(extract_mpts))


def serving_run(chip, datagen, k, seed_n, batches, insert_rate, extract_rate, budget, seed=3):
    """The incremental serving workload on one device-resident CORE-SG: a seed
    graph, then a stream of insert batches (sizes log-uniform in [1k, 50k],
    every fourth from components the seed does not have) and extract queries
    (a random subset of 1..8 mpts values from 2..32), each kind arriving as a
    Poisson process, submitted by an open-loop client while the session's
    worker serves them. Latency = submission to completion, per request."""
    rng = np.random.default_rng(seed)
    d = 24
    seed_pts = datagen.gaussian_mixture(seed_n, d, 32, 0.2, 3.0, 0.02, seed=seed)
    sizes = np.exp(rng.uniform(np.log(1_000), np.log(50_000), size=batches)).astype(int)
    inserts = []
    for b, m in enumerate(sizes):
        # same centre box; new components (a fresh generator seed) every fourth batch
        src_seed = seed if b % 4 else 1000 + b
        inserts.append(datagen.gaussian_mixture(int(m), d, 32 if b % 4 else 4, 0.2, 3.0, 0.02,
                                                seed=src_seed * 7919 + b))
    t_ins = np.cumsum(rng.exponential(1.0 / insert_rate, size=batches))
    horizon = float(t_ins[-1])
    n_ext = max(1, int(rng.poisson(extract_rate * horizon)))
    t_ext = np.sort(rng.uniform(0.0, horizon, size=n_ext))
    ext_sets = [sorted(rng.choice(np.arange(2, 33), size=int(rng.integers(1, 9)), replace=False))
                for _ in range(n_ext)]
    events = sorted([(float(t), "insert", i) for i, t in enumerate(t_ins)] +
                    [(float(t), "extract", i) for i, t in enumerate(t_ext)])
    # inserts by graph search (beam search over the existing lists + restricted
    # refinement); insert_search=0 would scan every point exactly
    params = chip.build_params(sample=16, iters=10, insert_search=1)
    t0 = time.perf_counter()
    sess = chip.Session(seed_pts, k, params, hbm_budget=budget)
    seed_build_s = time.perf_counter() - t0
    print(f"[bench] serving: seed {seed_n} points built in {seed_build_s * 1e3:.1f} ms; "
          f"{batches} inserts ({int(sizes.sum())} points), {n_ext} extracts over {horizon:.2f} s",
          file=sys.stderr, flush=True)
    ids = []
    start = time.perf_counter()
    for t, kind, i in events:
        delay = start + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        rid = sess.submit_insert(inserts[i]) if kind == "insert" else sess.submit_extract(ext_sets[i])
        ids.append((kind, rid, i))
    lat = {"insert": [], "extract": []}
    run_ms = {"insert": [], "extract": []}
    phases = {"knn_seconds": [], "hubness_seconds": [], "bridge_seconds": [], "total_seconds": []}
    batch_sizes, failed, msts = [], 0, 0
    by_size = {}  # extract latency by the number of mpts values asked for
    ins_sized = []  # (batch points, latency ms, run ms) per insert
    for kind, rid, i in ids:
        st = sess.wait(rid)
        if st["state"] != chip.REQ_DONE:
            failed += 1
            continue
        lat[kind].append(st["latency_ms"])
        run_ms[kind].append(st["run_ms"])
        if kind == "insert":
            for key in phases:
                phases[key].append(st["insert_stats"][key])
            ins_sized.append((len(inserts[i]), st["latency_ms"], st["run_ms"]))
        if kind == "extract":
            batch_sizes.append(st["batch_requests"])
            msts += st["num_msts"]
            by_size.setdefault(st["num_msts"], []).append(st["latency_ms"])
        sess.release(rid)
    wall = time.perf_counter() - start
    info = sess.graph_info()
    final_n, dev_bytes = info["n"], info["device_bytes"]
    sess.close()
    return dict(
        workload=f"serving: seed {seed_n} (d=24) + {batches} log-uniform insert batches "
                 f"[1k,50k] + Poisson extract queries (1..8 mpts from 2..32)",
        insert_mode="graph search (beam over the existing lists, restricted refinement)",
        seed_build_ms=seed_build_s * 1e3, inserted_points=int(sizes.sum()), final_points=final_n,
        inserts=len(lat["insert"]), extracts=len(lat["extract"]), failed=failed,
        insert_rate_per_s=insert_rate, extract_rate_per_s=extract_rate, wall_s=wall,
        insert_p50_ms=percentile(lat["insert"], 50), insert_p99_ms=percentile(lat["insert"], 99),
        extract_p50_ms=percentile(lat["extract"], 50),
        extract_p99_ms=percentile(lat["extract"], 99),
        extract_by_mpts_count={str(c): dict(requests=len(v), p50_ms=percentile(v, 50),
                                            p99_ms=percentile(v, 99))
                               for c, v in sorted(by_size.items())},
        extract_msts=msts, extracts_per_sweep_mean=float(np.mean(batch_sizes)) if batch_sizes else None,
        insert_run_ms_mean=float(np.mean(run_ms["insert"])) if run_ms["insert"] else None,
        insert_by_batch_points={
            f"[{lo},{hi})": dict(requests=len(v), latency_p50_ms=percentile([e[1] for e in v], 50),
                               latency_max_ms=max(e[1] for e in v),
                               run_mean_ms=float(np.mean([e[2] for e in v])))
            for lo, hi in ((1000, 5000), (5000, 20000), (20000, 50001))
            for v in [[e for e in ins_sized if lo <= e[0] < hi]] if v},
        extract_run_ms_mean=float(np.mean(run_ms["extract"])) if run_ms["extract"] else None,
        insert_phase_ms_mean={k.replace("_seconds", ""): float(np.mean(v)) * 1e3
                              for k, v in phases.items() if v},
        hbm_budget_bytes=budget, graph_device_bytes=dev_bytes)


def exact_mst_run(chip, datagen, coresg_io, cpu_n=5000):
    """CoreSG::exact_mst (Prim over the complete MRD graph, one cooperative
    launch, and beside it the one-launch-per-step form) at mpts = k_max on
    cfg1's shape (20k x 16) and on the first 100k
    points of the headline generator (d=24), plus a bounded CPU comparison: the
    oracle's Prim (single thread) and the device on the first `cpu_n` points."""
    out = {}
    cases = (("mixture-20k-d16", datagen.gaussian_mixture(20_000, 16, 10, 0.3, 2.0, 0.05, 0), 16),
             ("mixture-100k-d24", datagen.gaussian_mixture(100_000, 24, 32, 0.2, 3.0, 0.02, 1), 32))
    for name, x, k in cases:
        g = chip.Graph(x.shape[0], x.shape[1], k)
        g.set_points(x)
        g.build(chip.build_params())
        g.exact_mst(k)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = g.exact_mst(k)
            ts.append(time.perf_counter() - t0)
        ms = 1e3 * min(ts)
        # the one-launch-per-step form (kernel boundaries as the step barrier)
        g.set_exact_launch(1)
        t0 = time.perf_counter()
        rs = g.exact_mst(k)
        stepped_ms = 1e3 * (time.perf_counter() - t0)
        g.set_exact_launch(0)
        out[name] = dict(n=int(x.shape[0]), d=int(x.shape[1]), mpts=k, ms=ms,
                         us_per_step=1e3 * ms / (x.shape[0] - 1), num_edges=r["num_edges"],
                         weight=r["weight"], stepped_launch_ms=stepped_ms,
                         stepped_same_weight=rs["weight"] == r["weight"])
        g.close()
        print(f"[bench] exact_mst {name}: {ms:.1f} ms (one launch per step: {stepped_ms:.1f} ms)",
              file=sys.stderr, flush=True)
    # Bounded CPU comparison on the first cpu_n points of the 100k case.
    x = cases[1][1][:cpu_n]
    k = cases[1][2]
    g = chip.Graph(x.shape[0], x.shape[1], k)
    g.set_points(x)
    g.build(chip.build_params())
    g.exact_mst(k)
    t0 = time.perf_counter()
    g.exact_mst(k)
    dev_ms = 1e3 * (time.perf_counter() - t0)
    ids, dist, flags, count = g.get_knn()
    u, v, w, rc = g.get_bridges()
    g.close()
    cpu_ms = None
    try:
        sys.path.insert(0, REPO)
        import oracle
        with tempfile.TemporaryDirectory() as tmp:
            path = os.path.join(tmp, "s.coresg")
            coresg_io.save(path, coresg_io.CoreSGFile(
                k, x.shape[1], oracle.stride_of(x.shape[1]), 0, 0, oracle.pad(x), ids, dist,
                flags, count, coresg_io.bridges_from_uvw(u, v, w), rc))
            sg = oracle.CoreSG.load(path)
            t0 = time.perf_counter()
            sg.exact_mst(k)
            cpu_ms = 1e3 * (time.perf_counter() - t0)
    except Exception as e:  # the CPU comparison is optional
        print(f"[bench] exact_mst CPU comparison skipped: {e}", file=sys.stderr, flush=True)
    out["cpu_baseline"] = dict(kind="port", cores=1, n=cpu_n, d=int(x.shape[1]), cpu_ms=cpu_ms,
                               device_ms=dev_ms,
                               sample=f"first {cpu_n} points of mixture-100k-d24, mpts={k}; "
                                      "oracle Prim (mst.cpp:172-220 restated in C), one thread")
    return out


MFMA_PEAK_TFLOPS = 2500.0  # MI355X dense FP16 / BF16 matrix-core peak


def mixed_precision_run(chip, datagen, n=2_000_000, d=128, k=32, mpts=(4, 16, 32)):
    """The mixed-precision distance path: a 64-component anisotropic mixture
    (eigenvalues log-uniform in [0.01, 4]), built with the exact FP32 join and
    with the FP16 and BF16 Gram filters (gfx950 16x16x32 matrix cores, every
    pair the bound cannot rule out computed in exact FP32); per filter, a
    profiled build gives the matrix-core work (16 x 16 tiles x 32-coordinate
    steps, 2 * 16 * 16 * 32 FLOP each) over the join's device time; the BF16
    build's fast distances are audited against FP64 on every visited pair,
    and mpts = 4, 16, 32 are extracted from each graph."""
    import struct
    f32 = lambda b: struct.unpack("f", struct.pack("I", b & 0xFFFFFFFF))[0]
    t0 = time.perf_counter()
    x = datagen.anisotropic_mixture(n, d, 64, 0.01, 4.0, seed=4)
    gen_s = time.perf_counter() - t0
    g = chip.Graph(n, d, k)
    g.set_points(x)
    del x
    out = dict(workload=f"anisotropic-mixture-{n // 1_000_000}m-d{d}", n=n, d=d, k_max=k,
               data_gen_s=gen_s)
    for name, gram in (("exact_fp32", 1), ("fp16_gram_filter", 2), ("bf16_gram_filter", 3)):
        p = chip.build_params(sample=16, iters=10, gram=gram)
        g.build(p)  # warm
        t0 = time.perf_counter()
        st = g.build(p)
        build_s = time.perf_counter() - t0
        g.prepare_edges()
        t1 = time.perf_counter()
        res = g.extract(list(mpts))
        sweep_s = time.perf_counter() - t1
        rec, _ = g.sampled_recall(300, 42)
        out[name] = dict(build_ms=build_s * 1e3, pts_per_s=n / build_s, sweep_ms=sweep_s * 1e3,
                         recall_sampled=rec, rounds=st["rounds"],
                         mst_weights={int(m): r["weight"] for m, r in zip(mpts, res)})
        # the join's device time and work in a profiled build
        g.reset_timers()
        g.set_profiling(True)
        g.build(p)
        g.set_profiling(False)
        jms, jn = g.kernel_time("nnd_join")
        out[name]["join_ms_per_launch"] = jms
        if gram != 1:
            tiles, ksteps = g.kernel_work("nnd_mfma")[:2]
            pairs, exact = g.kernel_work("nnd_gram")[:2]
            flops = tiles * ksteps * 2 * 16 * 16 * 32 / max(1, jn)
            tflops = flops / (jms * 1e-3) / 1e12 if jms else None
            out[name]["mfma"] = dict(
                tiles_per_launch=tiles / max(1, jn), k_steps=ksteps, flops_per_launch=flops,
                tflops=tflops, peak_tflops=MFMA_PEAK_TFLOPS,
                frac=(tflops / MFMA_PEAK_TFLOPS) if tflops else None,
                pairs_per_launch=pairs / max(1, jn), exact_fraction=exact / max(1, pairs))
        print(f"[bench] mixed precision {name}: build {build_s * 1e3:.1f} ms, "
              f"join {jms:.2f} ms/launch, recall {rec:.4f}", file=sys.stderr, flush=True)
    # audit of the BF16 fast distances (FP64 per visited pair; slow, untimed)
    g.reset_timers()
    g.set_profiling(True)
    g.set_audit(True)
    g.build(chip.build_params(sample=16, iters=10, gram=3))
    g.set_profiling(False)
    g.set_audit(False)
    w = g.kernel_work("nnd_gram")
    out["bf16_audit"] = dict(pairs=w[0], exact_recomputed=w[1],
                             exact_fraction=w[1] / max(1, w[0]),
                             max_rel_distance_error_vs_fp64=f32(w[2]),
                             max_bound_ratio=f32(w[3]))
    g.close()
    return out


def batched_run(chip, datagen, n_sets=256, k=24, seed=5):
    """Batched uneven instances: n_sets independent mixtures (n log-uniform in
    [2k, 200k], d in {8, 16, 32, 64}, 4-40 components, 0-10 % noise), one graph
    per d bucket with the instances as disjoint id ranges; each bucket built
    and swept (mpts 2..k) in one pass. Beside it the plain form: every
    instance as its own graph, built and swept one after another."""
    rng = np.random.default_rng(seed)
    dims = rng.choice([8, 16, 32, 64], size=n_sets)
    sizes = np.exp(rng.uniform(np.log(2_000), np.log(200_000), size=n_sets)).astype(int)
    comps = rng.integers(4, 41, size=n_sets)
    noise = rng.uniform(0.0, 0.10, size=n_sets)
    mpts = list(range(2, k + 1))
    sets = [datagen.gaussian_mixture(int(sizes[i]), int(dims[i]), int(comps[i]), 0.2, 3.0,
                                     float(noise[i]), seed=seed * 1000 + i) for i in range(n_sets)]
    params = chip.build_params(sample=min(16, k // 2), iters=10)
    out = dict(workload=f"{n_sets} mixtures, n log-uniform [2k,200k], d in {{8,16,32,64}}, "
                        f"k_max={k}, mpts 2..{k}", instances=n_sets,
               points=int(sizes.sum()), msts=n_sets * len(mpts), buckets={})
    fused_build = fused_sweep = 0.0
    weights = {}
    for d in (8, 16, 32, 64):
        idx = [i for i in range(n_sets) if dims[i] == d]
        if not idx:
            continue
        x = np.concatenate([sets[i] for i in idx])
        off = np.concatenate([[0], np.cumsum([sets[i].shape[0] for i in idx])]).astype(np.uint32)
        g = chip.Graph(x.shape[0], d, k)
        g.set_points(x)
        g.set_instances(off)
        g.build(params)  # warm (allocations)
        g.prepare_edges()
        tune = g.autotune_sweep(mpts)  # this (n, d) bucket's sweep shape
        t0 = time.perf_counter()
        st = g.build(params)
        t1 = time.perf_counter()
        per = g.extract_instances(mpts)
        t2 = time.perf_counter()
        g.close()
        fused_build += t1 - t0
        fused_sweep += t2 - t1
        for j, i in enumerate(idx):
            weights[i] = [r["weight"] for r in per[j]]
        out["buckets"][str(d)] = dict(instances=len(idx), points=int(off[-1]),
                                      sweep_shape=dict(lanes=tune["lanes"],
                                                       max_blocks=tune["max_blocks"]),
                                      build_ms=(t1 - t0) * 1e3, sweep_ms=(t2 - t1) * 1e3,
                                      knn_ms=st["knn_seconds"] * 1e3,
                                      hubness_select_ms=st["hubness_seconds"] * 1e3,
                                      bridges_ms=st["bridge_seconds"] * 1e3)
        print(f"[bench] batched d={d}: {len(idx)} instances, {int(off[-1])} points, build "
              f"{(t1 - t0) * 1e3:.1f} ms, sweep {(t2 - t1) * 1e3:.1f} ms", file=sys.stderr, flush=True)
    # plain form: one graph per instance
    plain_build = plain_sweep = 0.0
    for i in range(n_sets):
        g = chip.Graph(sets[i].shape[0], int(dims[i]), k)
        g.set_points(sets[i])
        t0 = time.perf_counter()
        g.build(params)
        t1 = time.perf_counter()
        g.extract(mpts)
        t2 = time.perf_counter()
        g.close()
        plain_build += t1 - t0
        plain_sweep += t2 - t1
    total = fused_build + fused_sweep
    out.update(build_ms=fused_build * 1e3, sweep_ms=fused_sweep * 1e3,
               pts_per_s=out["points"] / fused_build, msts_per_s=out["msts"] / fused_sweep,
               plain_build_ms=plain_build * 1e3, plain_sweep_ms=plain_sweep * 1e3,
               fused_over_plain=(plain_build + plain_sweep) / total)
    print(f"[bench] batched: build {fused_build * 1e3:.1f} ms + sweep {fused_sweep * 1e3:.1f} ms; "
          f"plain {plain_build * 1e3:.1f} + {plain_sweep * 1e3:.1f} ms", file=sys.stderr, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None, help="override the point count")
    ap.add_argument("--mpts-max", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="host threads of the reference CPU baseline (the box's CPU share)")
    ap.add_argument("--recall-queries", type=int, default=1000)
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched uneven instances (256 mixtures, d in {8,16,32,64})")
    ap.add_argument("--no-exact", action="store_true",
                    help="skip the exact complete-graph MST (Prim) measurement")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision measurement (n=2M, d=128, BF16 Gram filter)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the incremental serving scenario (seed 500k + 64 insert batches "
                         "+ Poisson extract queries)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 1 GiB device-to-device three times (a plain "
                         "copy kernel of known size, the scale for HBM byte counters) and "
                         "run three random-row gather probes of known bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(one process per GPU box, no collectives)", file=sys.stderr)
        sys.exit(2)

    chip = _load("coresg_hip")
    datagen = _load("datagen")
    workload = "mixture-1m-d24"
    x, k = datagen.workload(workload, n=args.n)
    n, d = x.shape
    mpts = list(range(1, (args.mpts_max or k) + 1))
    print(f"[bench] workload {workload} n={n} d={d} k_max={k} mpts=1..{mpts[-1]}",
          file=sys.stderr, flush=True)

    g = chip.Graph(n, d, k)
    g.set_points(x)  # resident in HBM before any timed work
    params = chip.build_params(sample=16, iters=10)

    def step(record):
        t0 = time.perf_counter()
        st = g.build(params)
        t1 = time.perf_counter()
        g.prepare_edges()
        res = g.extract(mpts)  # the whole sweep in one call (batched slots)
        t2 = time.perf_counter()
        if record is not None:
            record.append(dict(build_s=t1 - t0, sweep_s=t2 - t1, stats=st,
                               weights=[r["weight"] for r in res]))

    def single_queries():
        """extract(mpts) latency of one-mpts queries against the resident
        CORE-SG, and the sweep as one call per mpts value."""
        lat = []
        t0 = time.perf_counter()
        for m in mpts:
            ta = time.perf_counter()
            g.extract([m])
            lat.append((time.perf_counter() - ta) * 1e3)
        return lat, time.perf_counter() - t0

    for w in range(args.warmup):
        step(None)
        print(f"[bench] warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    # the sweep's launch shapes, autotuned on this graph for the batched sweep
    # (32 slots), the 16-value sweep, single-mpts queries and 2, 4 and 8
    # slots (the sizes of the serving scenario's extracts, whose graphs take
    # the shape of the nearest tuned size bucket); an untimed build first when
    # there was no warmup step
    if args.warmup == 0:
        g.build(params)
    g.prepare_edges()
    tuned = {}
    for name, mp in (("sweep32", mpts), ("sweep16", list(range(2, k + 1, 2))[:16]),
                     ("single", [mpts[len(mpts) // 2]]), ("sweep2", [8, 24]),
                     ("sweep4", [4, 12, 20, 28]), ("sweep8", list(range(4, k + 1, 4))[:8])):
        tuned[name] = g.autotune_sweep(mp)
        print(f"[bench] autotune {name}: {tuned[name]['lanes']} lanes, <= "
              f"{tuned[name]['max_blocks']} blocks, {tuned[name]['best_ms']:.2f} ms "
              f"(default {tuned[name]['default_ms']:.2f} ms)", file=sys.stderr, flush=True)
    g.reset_timers()
    g.set_profiling(True)
    rec = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        step(rec)
        print(f"[bench] step {s + 1}/{args.steps}: build {rec[-1]['build_s']*1e3:.1f} ms, "
              f"sweep {rec[-1]['sweep_s']*1e3:.1f} ms", file=sys.stderr, flush=True)
    total = time.perf_counter() - t_start
    g.set_profiling(False)

    build_s = float(np.mean([r["build_s"] for r in rec]))
    sweep_s = float(np.mean([r["sweep_s"] for r in rec]))
    st = rec[-1]["stats"]
    lat, plain_s = [], []
    for _ in range(args.steps):
        l1, p1 = single_queries()
        lat += l1
        plain_s.append(p1)
    plain_s = float(np.mean(plain_s))

    kernels = {}
    for name in ("nnd_init", "nnd_join", "nnd_gather", "nnd_sample", "nnd_merge", "boruvka_round1",
                 "boruvka_rounds_graph", "hub_slot_pass", "hub_score", "bridge_mrd_matrix",
                 "bridge_prim"):
        ms, cnt = g.kernel_time(name)
        if cnt:
            kernels[name] = dict(avg_ms=ms, launches=cnt, total_ms=ms * cnt / args.steps)
    dominant = max(kernels, key=lambda kk: kernels[kk]["total_ms"]) if kernels else None
    ne, ke = g.prepare_edges()
    stride = (d + 7) // 8 * 8
    rooflines = {}
    for name in ("nnd_join", "nnd_merge", "boruvka_round1"):
        if name not in kernels:
            continue
        if name == "nnd_merge" and "nnd_join" not in kernels:
            continue
        w = g.kernel_work("nnd_join" if name == "nnd_merge" else name)
        launches = kernels[name]["launches"]
        avg_s = kernels[name]["avg_ms"] * 1e-3
        if name == "nnd_join":
            # per launch: each active node's gathered ids + thresholds (8 B per
            # member) and member rows (4*stride B), the node's count word, and
            # every proposal written (8 B); FP32 work = 3*stride per distance
            rows, dists, props, nodes = w
            algo = (nodes * 4 + rows * (4 * stride + 8) + props * 8) / launches
            flops = dists * 3 * stride / launches
            extra = dict(flops_per_launch=flops, fp32_tflops=flops / avg_s / 1e12,
                         fp32_frac=flops / avg_s / 1e12 / FP32_PEAK_TFLOPS,
                         members_per_launch=rows / launches, proposals_per_launch=props / launches)
        elif name == "nnd_merge":
            # per launch (one per join launch): every node's list (ids, distances,
            # flags: 9k B), its count and proposal-fill words (8 B) and every key
            # the join offered it (8 B); the lists it rewrites are left out
            rows, dists, props, nodes = w
            algo = (nodes * (9 * k + 8) + props * 8) / launches
            extra = dict(proposals_per_launch=props / launches)
        else:
            # per launch (one launch serves B mpts slots): every adjacency
            # entry (neighbour, rank, base weight: 12 B) plus the neighbour's
            # core distance per slot (4 B each); per vertex its two offsets
            # (8 B) and per slot its core distance, best key and active flag
            # (13 B)
            entries, verts, slots = w[0], w[1], max(1, w[2] // max(1, launches))
            algo = (entries * (12 + 4 * slots) + verts * (8 + 13 * slots)) / launches
            extra = dict(adjacency_entries=entries / launches, slots=slots)
        achieved = algo / avg_s / 1e9
        traffic, traffic_gather = measured_traffic(name)
        rooflines[name] = dict(kernel=name, bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS,
                               unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                               traffic=traffic, traffic_gather_scaled=traffic_gather,
                               algo_bytes_per_launch=algo,
                               avg_launch_ms=kernels[name]["avg_ms"], **extra)
    roofline = rooflines.get(dominant) or (
        dict(kernel=dominant, bound="hbm", achieved=None, peak=HBM_PEAK_GBS, unit="GB/s",
             frac=None, traffic=None) if dominant else None)

    recall, ratio = g.sampled_recall(args.recall_queries, 42)

    # Fused against plain on the north-star unit of work: build + a 16-value
    # mpts sweep (2, 4, ..., 32). Plain = the same kernels one launch per step
    # with a host synchronisation after each, the descent's convergence read
    # between rounds, one mpts value per sweep pass with host-driven Boruvka
    # rounds and no captured graphs.
    sweep16 = [m for m in range(2, k + 1, 2)][:16]

    def build_and_sweep16():
        t0 = time.perf_counter()
        g.build(params)
        t1 = time.perf_counter()
        g.prepare_edges()
        res16 = g.extract(sweep16)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, [r["weight"] for r in res16]

    g.set_plain(True)
    build_and_sweep16()  # warm the plain path
    pb, ps, pw = build_and_sweep16()
    g.set_plain(False)
    fb, fs, fw = build_and_sweep16()
    plain_vs_fused = dict(
        unit="CORE-SG build + 16-value mpts sweep (2, 4, ..., 32)",
        plain_build_ms=pb * 1e3, plain_sweep16_ms=ps * 1e3, plain_total_ms=(pb + ps) * 1e3,
        fused_build_ms=fb * 1e3, fused_sweep16_ms=fs * 1e3, fused_total_ms=(fb + fs) * 1e3,
        fused_over_plain=(pb + ps) / (fb + fs),
        sweep16_fused_over_plain=ps / fs)
    print(f"[bench] plain build+sweep16 {(pb + ps) * 1e3:.1f} ms, fused {(fb + fs) * 1e3:.1f} ms",
          file=sys.stderr, flush=True)

    serving = None
    if not args.no_serving:
        g.close()
        serving = serving_run(chip, datagen, k, seed_n=500_000, batches=64, insert_rate=4.0,
                              extract_rate=16.0, budget=64 << 30)
        print(f"[bench] serving: insert p50/p99 {serving['insert_p50_ms']:.1f}/"
              f"{serving['insert_p99_ms']:.1f} ms, extract p50/p99 {serving['extract_p50_ms']:.1f}/"
              f"{serving['extract_p99_ms']:.1f} ms", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batched:
        batched = batched_run(chip, datagen)

    exact = None
    if not args.no_exact:
        exact = exact_mst_run(chip, datagen, _load("coresg_io"))

    mixed = None
    if not args.no_mixed:
        mixed = mixed_precision_run(chip, datagen)

    cpu = None
    if not args.no_cpu_baseline:
        threads = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        ext = [2, 8, 16, 24, 32] if k >= 32 else sorted({2, max(2, k // 2), k})
        cb = cpu_baseline(x, k, ext, threads, args.recall_queries)
        if cb:
            cpu_total16 = cb["build_s"] + 16 * cb["extract_per_mpts_s"]
            gpu_total16 = plain_vs_fused["fused_total_ms"] / 1e3
            cpu = dict(value=cb["build_pts_per_s"], unit="pts/s", cores=threads,
                       kind="reference",
                       sample=f"the whole {workload} workload (n={n}); reference coresg_cli build "
                              f"--threads {threads} --instances 1 --sample 16 --iters 10 "
                              f"--recall-sample {args.recall_queries}, then extract of mpts "
                              f"{ext[-1]} and of {','.join(map(str, ext))}; per-mpts cost = "
                              f"difference of the two extract runs over {len(ext) - 1} values",
                       build_s=cb["build_s"], recall_sampled=cb["recall_sampled"],
                       extract_per_mpts_s=cb["extract_per_mpts_s"],
                       extract_load_s=cb["extract_load_s"],
                       mst_only_per_mpts_s=cb["mst_only_per_mpts_s"],
                       msts_per_s=cb["msts_per_s"],
                       build_plus_sweep16_s=cpu_total16,
                       fused_gpu_over_cpu_build_plus_sweep16=cpu_total16 / gpu_total16,
                       fused_gpu_over_cpu_build=(n / build_s) / cb["build_pts_per_s"])

    out = {
        "metric": "CORE-SG build pts/s + mpts-sweep MSTs/s at n=1M; p50/p99 extract latency",
        "value": n / build_s,
        "unit": "pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "n": n, "d": d, "k_max": k, "sample": 16, "iters": 10,
                   "mpts": f"1..{mpts[-1]}"},
        "build_ms": build_s * 1e3,
        "sweep_ms": sweep_s * 1e3,
        "sweep_msts_per_s": len(mpts) / sweep_s,
        "sweep_one_call_per_mpts_ms": plain_s * 1e3,
        "sweep_batched_over_one_call_per_mpts": plain_s / sweep_s,
        "plain_vs_fused": plain_vs_fused,
        "autotune": tuned,
        "extract_p50_ms": percentile(lat, 50),
        "extract_p99_ms": percentile(lat, 99),
        "knn_rounds": st["rounds"],
        "edges": ne,
        "knn_edges": ke,
        "antihubs_selected": st["antihubs_selected"],
        "bridge_edges": st["bridge_edges"],
        "recall_sampled": recall,
        "recall_queries": args.recall_queries,
        "kernels_ms_per_step": {kk: round(v["total_ms"], 3) for kk, v in kernels.items()},
        "roofline": roofline,
        "rooflines": rooflines,
        "cpu_baseline": cpu,
        "serving": serving,
        "mixed_precision": mixed,
        "exact_mst": exact,
        "batched_instances": batched,
    }
    print(json.dumps(out))
    if args.copy_probe:
        # Through the HIP runtime the library itself is linked against (a second
        # runtime in the process, such as torch's, must not be loaded).
        import ctypes as C
        hip = C.CDLL("libamdhip64.so.7")
        size = C.c_size_t(1 << 30)
        a_, b_ = C.c_void_p(), C.c_void_p()
        if hip.hipMalloc(C.byref(a_), size) or hip.hipMalloc(C.byref(b_), size):
            raise RuntimeError("copy probe: hipMalloc failed")
        for _ in range(3):
            if hip.hipMemcpy(b_, a_, size, 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("copy probe: hipMemcpy failed")
        hip.hipDeviceSynchronize()
        hip.hipFree(a_)
        hip.hipFree(b_)
        print("[bench] copy probe: 3 x 1 GiB device-to-device", file=sys.stderr, flush=True)
        # random whole-row gathers of a 1M x 24 array (96 MB, the join's point
        # rows): 3 x 32M rows of 96 B = 3.07 GB read, 128 MB written per run
        for _ in range(3):
            ms = chip.probe_gather(1_000_000, 24, 32 << 20)
        print(f"[bench] gather probe: 3 x 32M random 96-byte rows of a 96 MB array, "
              f"{ms:.2f} ms each", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
