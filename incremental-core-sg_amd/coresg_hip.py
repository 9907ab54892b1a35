"""ctypes binding of libcoresg_hip.so (the C ABI in include/coresg_hip.h).

Thin plumbing for Python callers (tests, the benchmark): every call goes
straight to the HIP library. There is no CPU fallback; a missing library or a
failing call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcoresg_hip.so")

HOST, DEVICE = 0, 1


class CoresgHipError(RuntimeError):
    pass


class DeviceProps(C.Structure):
    _fields_ = [("id", C.c_int), ("name", C.c_char * 256), ("arch", C.c_char * 64),
                ("total_mem", C.c_size_t), ("free_mem", C.c_size_t), ("cu_count", C.c_int),
                ("major", C.c_int), ("minor", C.c_int)]


class NndConfig(C.Structure):
    _fields_ = [("n", C.c_uint32), ("dim", C.c_uint32), ("stride", C.c_uint32),
                ("k", C.c_uint32), ("sample", C.c_uint32), ("iters", C.c_uint32),
                ("seed", C.c_uint64), ("metric", C.c_int), ("delta", C.c_float),
                ("random_fill", C.c_int)]


class BuildParams(C.Structure):
    _fields_ = [("sample", C.c_uint32), ("iters", C.c_uint32), ("delta", C.c_double),
                ("seed", C.c_uint64), ("brute_force_limit", C.c_uint32),
                ("antihub_multiplier", C.c_double), ("max_pair_ops", C.c_double),
                ("min_selected", C.c_size_t), ("max_selected", C.c_size_t),
                ("min_per_component", C.c_size_t), ("ensure_connected", C.c_int),
                ("reps_per_component", C.c_size_t), ("max_reps", C.c_size_t),
                ("join_capacity", C.c_uint32), ("gram", C.c_uint32),
                ("missing_mode", C.c_uint32), ("insert_search", C.c_uint32),
                ("score_k", C.c_uint32)]


class BuildStats(C.Structure):
    _fields_ = [("knn_seconds", C.c_double), ("hubness_seconds", C.c_double),
                ("bridge_seconds", C.c_double), ("total_seconds", C.c_double),
                ("rounds", C.c_uint32), ("updates", C.c_uint64), ("knn_edges", C.c_uint64),
                ("antihubs_selected", C.c_uint64), ("bridge_edges", C.c_uint64),
                ("repair_edges", C.c_uint64), ("components_before_repair", C.c_uint32),
                ("components_after_repair", C.c_uint32)]

    def asdict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class MstInfo(C.Structure):
    _fields_ = [("mpts", C.c_uint32), ("num_edges", C.c_uint32),
                ("num_components", C.c_uint32), ("rounds", C.c_uint32), ("weight", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("lanes", C.c_uint32), ("max_blocks", C.c_uint32),
                ("candidates", C.c_uint32), ("best_ms", C.c_double), ("default_ms", C.c_double)]

    def asdict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RequestStatus(C.Structure):
    _fields_ = [("id", C.c_uint64), ("kind", C.c_int), ("state", C.c_int),
                ("graph_n", C.c_uint32), ("graph_version", C.c_uint64), ("num_msts", C.c_uint32),
                ("batch_requests", C.c_uint32), ("queued_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double),
                ("insert_stats", BuildStats), ("error", C.c_char * 256)]

    def asdict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("insert_stats", "error")}
        d["insert_stats"] = self.insert_stats.asdict()
        d["error"] = self.error.decode(errors="replace")
        return d


REQ_INSERT, REQ_EXTRACT = 1, 2
REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3

_vp = C.c_void_p
_lib = None

# (name, restype, argtypes) of every exported symbol; tests check the export
# list against include/coresg_hip.h.
SIGNATURES = [
    ("coresg_hip_abi_version", C.c_int, []),
    ("coresg_hip_last_error", C.c_char_p, []),
    ("coresg_hip_device_count", C.c_int, []),
    ("coresg_hip_get_device_props", C.c_int, [C.c_int, C.POINTER(DeviceProps)]),
    ("coresg_hip_nndescent", C.c_int, [C.c_int, C.POINTER(NndConfig), _vp, _vp, _vp, _vp, _vp,
                                       C.POINTER(C.c_uint64)]),
    ("coresg_hip_set_device_budget", C.c_int, [C.c_int, C.c_size_t]),
    ("coresg_hip_mrd_block", C.c_int, [C.c_int, _vp, C.c_size_t, C.c_uint32, C.c_uint32,
                                       C.c_int, _vp, C.c_size_t, C.c_size_t, _vp]),
    ("coresg_hip_mrd_block_noalloc", C.c_int, [C.c_int, _vp, C.c_size_t, C.c_uint32,
                                               C.c_uint32, C.c_int, _vp, C.c_size_t, C.c_size_t,
                                               _vp]),
    ("coresg_hip_mrd_block_noalloc_capacity", C.c_size_t, []),
    ("coresg_hip_mrd_row_min", C.c_int, [C.c_int, _vp, C.c_size_t, C.c_uint32, C.c_uint32,
                                         C.c_int, _vp, _vp, C.c_size_t, C.c_size_t, _vp, _vp]),
    ("coresg_hip_enable_peer_access", C.c_int, [_vp, C.c_size_t]),
    ("coresg_hip_device_memory", C.c_int, [C.c_int, C.POINTER(C.c_size_t),
                                           C.POINTER(C.c_size_t)]),
    ("coresg_hip_release_thread_cache", None, []),
    ("coresg_hip_default_build_params", None, [C.POINTER(BuildParams)]),
    ("coresg_hip_graph_create", _vp, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("coresg_hip_graph_destroy", None, [_vp]),
    ("coresg_hip_graph_set_points", C.c_int, [_vp, _vp, C.c_int, C.c_int]),
    ("coresg_hip_graph_set_knn", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int]),
    ("coresg_hip_graph_get_knn", C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    ("coresg_hip_graph_set_bridges", C.c_int, [_vp, _vp, _vp, _vp, C.c_size_t, C.c_size_t]),
    ("coresg_hip_graph_get_bridges", C.c_size_t, [_vp, _vp, _vp, _vp, C.c_size_t,
                                                  C.POINTER(C.c_size_t)]),
    ("coresg_hip_graph_device_arrays", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("coresg_hip_graph_device_bytes", C.c_size_t, [_vp]),
    ("coresg_hip_build_knn", C.c_int, [_vp, C.POINTER(BuildParams), C.POINTER(BuildStats)]),
    ("coresg_hip_build_bridges", C.c_int, [_vp, C.POINTER(BuildParams),
                                           C.POINTER(BuildStats)]),
    ("coresg_hip_build", C.c_int, [_vp, C.POINTER(BuildParams), C.POINTER(BuildStats)]),
    ("coresg_hip_graph_get_selected", C.c_size_t, [_vp, _vp, C.c_size_t]),
    ("coresg_hip_graph_get_hubness", C.c_int, [_vp, _vp, _vp, _vp, C.POINTER(C.c_uint32)]),
    ("coresg_hip_prepare_edges", C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("coresg_hip_edges", C.c_int, [_vp, C.c_uint32, _vp, _vp, _vp]),
    ("coresg_hip_extract", C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(MstInfo), _vp, _vp, _vp,
                                     _vp]),
    ("coresg_hip_exact_mst", C.c_int, [_vp, C.c_uint32, C.POINTER(MstInfo), _vp, _vp, _vp,
                                       _vp]),
    ("coresg_hip_sampled_recall", C.c_int, [_vp, C.c_size_t, C.c_uint64,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    ("coresg_hip_graph_set_plain", C.c_int, [_vp, C.c_int]),
    ("coresg_hip_graph_set_exact_launch", C.c_int, [_vp, C.c_int]),
    ("coresg_hip_probe_gather", C.c_int, [C.c_int, C.c_size_t, C.c_uint32, C.c_size_t,
                                          C.POINTER(C.c_double)]),
    ("coresg_hip_probe_mfma", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp]),
    ("coresg_hip_graph_derive_hubness", C.c_int, [_vp, C.POINTER(BuildParams)]),
    ("coresg_hip_autotune_sweep", C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(TuneResult)]),
    ("coresg_hip_tune_table", C.c_size_t, [C.c_char_p, C.c_size_t]),
    ("coresg_hip_tune_table_load", C.c_int, [C.c_char_p]),
    ("coresg_hip_graph_save", C.c_int, [_vp, C.c_char_p]),
    ("coresg_hip_graph_load", _vp, [C.c_int, C.c_char_p]),
    ("coresg_hip_write_edges_tsv", C.c_int, [C.c_char_p, _vp, _vp, _vp, C.c_size_t]),
    ("coresg_hip_write_labels_tsv", C.c_int, [C.c_char_p, _vp, C.c_size_t]),
    ("coresg_hip_graph_set_profiling", C.c_int, [_vp, C.c_int]),
    ("coresg_hip_graph_kernel_times", C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_double),
                                                C.POINTER(C.c_uint64)]),
    ("coresg_hip_graph_kernel_work", C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_uint64)]),
    ("coresg_hip_graph_reset_timers", C.c_int, [_vp]),
    ("coresg_hip_graph_set_audit", C.c_int, [_vp, C.c_int]),
    ("coresg_hip_graph_set_instances", C.c_int, [_vp, _vp, C.c_uint32]),
    ("coresg_hip_extract_instances", C.c_int, [_vp, _vp, C.c_uint32, _vp]),
    ("coresg_hip_insert", C.c_int, [_vp, C.c_void_p, C.c_uint32, C.c_int, C.c_int,
                                    C.POINTER(BuildParams), C.POINTER(BuildStats)]),
    ("coresg_hip_graph_size", C.c_uint32, [_vp]),
    ("coresg_hip_session_create", _vp, [C.c_int, _vp, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.POINTER(BuildParams), C.c_size_t,
                                         C.POINTER(BuildStats)]),
    ("coresg_hip_session_adopt", _vp, [_vp, C.POINTER(BuildParams), C.c_size_t]),
    ("coresg_hip_session_submit_insert", C.c_uint64, [_vp, _vp, C.c_uint32]),
    ("coresg_hip_session_submit_extract", C.c_uint64, [_vp, _vp, C.c_uint32, C.c_int, C.c_int]),
    ("coresg_hip_session_wait", C.c_int, [_vp, C.c_uint64, C.c_double,
                                          C.POINTER(RequestStatus)]),
    ("coresg_hip_session_extract_result", C.c_int, [_vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp]),
    ("coresg_hip_session_release", C.c_int, [_vp, C.c_uint64]),
    ("coresg_hip_session_graph", _vp, [_vp]),
    ("coresg_hip_session_device_bytes", C.c_size_t, [_vp]),
    ("coresg_hip_session_destroy", None, [_vp]),
]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CoresgHipError(
                f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g;"
                f" g.build()' or make -C incremental-core-sg_amd)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().coresg_hip_last_error().decode(errors="replace")


def check(ok) -> None:
    if not ok:
        raise CoresgHipError(last_error())


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    return lib().coresg_hip_device_count()


def device_props(device: int = 0) -> dict:
    p = DeviceProps()
    check(lib().coresg_hip_get_device_props(device, C.byref(p)))
    return dict(id=p.id, name=p.name.decode(), arch=p.arch.decode(), total_mem=p.total_mem,
                free_mem=p.free_mem, cu_count=p.cu_count, major=p.major, minor=p.minor)


def build_params(**kw) -> BuildParams:
    p = BuildParams()
    lib().coresg_hip_default_build_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _u32(a):
    return np.ascontiguousarray(a, np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Graph:
    """Device-resident CORE-SG on one GPU."""

    def __init__(self, n: int, dim: int, k_max: int, device: int = 0):
        self.n, self.dim, self.k = int(n), int(dim), int(k_max)
        self.stride = (self.dim + 7) & ~7
        h = lib().coresg_hip_graph_create(device, self.n, self.dim, self.k)
        if not h:
            raise CoresgHipError(last_error())
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            lib().coresg_hip_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- data in / out ------------------------------------------------------
    def set_points(self, rows: np.ndarray):
        rows = _f32(rows)
        padded = int(rows.shape[1] == self.stride and self.stride != self.dim)
        check(lib().coresg_hip_graph_set_points(self.h, ptr(rows), HOST, padded))

    def set_points_device(self, dev_ptr: int, padded: bool):
        check(lib().coresg_hip_graph_set_points(self.h, C.c_void_p(dev_ptr), DEVICE, int(padded)))

    def set_knn(self, ids, dist, flags, count):
        ids, dist, count = _u32(ids), _f32(dist), _u32(count)
        flags = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        check(lib().coresg_hip_graph_set_knn(self.h, ptr(ids), ptr(dist), ptr(flags), ptr(count),
                                             HOST))

    def get_knn(self):
        nk = self.n * self.k
        ids = np.empty(nk, np.uint32)
        dist = np.empty(nk, np.float32)
        flags = np.empty(nk, np.uint8)
        count = np.empty(self.n, np.uint32)
        check(lib().coresg_hip_graph_get_knn(self.h, ptr(ids), ptr(dist), ptr(flags),
                                             ptr(count)))
        return ids, dist, flags, count

    def set_bridges(self, u, v, w, repair_count=0):
        u, v, w = _u32(u), _u32(v), _f32(w)
        check(lib().coresg_hip_graph_set_bridges(self.h, ptr(u), ptr(v), ptr(w), u.size,
                                                 repair_count))

    def get_bridges(self):
        rc = C.c_size_t()
        m = lib().coresg_hip_graph_get_bridges(self.h, None, None, None, 0, C.byref(rc))
        u = np.empty(m, np.uint32)
        v = np.empty(m, np.uint32)
        w = np.empty(m, np.float32)
        lib().coresg_hip_graph_get_bridges(self.h, ptr(u), ptr(v), ptr(w), m, C.byref(rc))
        return u, v, w, rc.value

    def device_bytes(self) -> int:
        return lib().coresg_hip_graph_device_bytes(self.h)

    # ---- build ----------------------------------------------------------------
    def build(self, params: BuildParams | None = None) -> dict:
        p = params or build_params()
        st = BuildStats()
        check(lib().coresg_hip_build(self.h, C.byref(p), C.byref(st)))
        return st.asdict()

    def insert(self, rows: np.ndarray, params: BuildParams | None = None) -> dict:
        """CoreSG::insert: append rows (n_new x dim, or x stride padded) and
        update lists, hubness, selection and bridges."""
        p = params or build_params()
        rows = _f32(rows)
        if rows.ndim != 2 or rows.shape[1] not in (self.dim, self.stride):
            raise ValueError(f"rows must be (m, {self.dim}) or (m, {self.stride})")
        padded = int(rows.shape[1] == self.stride and self.stride != self.dim)
        st = BuildStats()
        check(lib().coresg_hip_insert(self.h, ptr(rows), rows.shape[0], HOST, padded,
                                      C.byref(p), C.byref(st)))
        self.n = int(lib().coresg_hip_graph_size(self.h))
        return st.asdict()

    def build_knn(self, params: BuildParams | None = None) -> dict:
        p = params or build_params()
        st = BuildStats()
        check(lib().coresg_hip_build_knn(self.h, C.byref(p), C.byref(st)))
        return st.asdict()

    def build_bridges(self, params: BuildParams | None = None) -> dict:
        p = params or build_params()
        st = BuildStats()
        check(lib().coresg_hip_build_bridges(self.h, C.byref(p), C.byref(st)))
        return st.asdict()

    def selected(self) -> np.ndarray:
        m = lib().coresg_hip_graph_get_selected(self.h, None, 0)
        out = np.empty(m, np.uint32)
        lib().coresg_hip_graph_get_selected(self.h, ptr(out), m)
        return out

    def hubness(self):
        rd = np.empty(self.n, np.int32)
        sc = np.empty(self.n, np.float64)
        lab = np.empty(self.n, np.uint32)
        nc = C.c_uint32()
        check(lib().coresg_hip_graph_get_hubness(self.h, ptr(rd), ptr(sc), ptr(lab),
                                                 C.byref(nc)))
        return rd, sc, lab, nc.value

    # ---- extraction -----------------------------------------------------------
    def prepare_edges(self):
        ne = C.c_uint64()
        ke = C.c_uint64()
        check(lib().coresg_hip_prepare_edges(self.h, C.byref(ne), C.byref(ke)))
        return ne.value, ke.value

    def edges(self, mpts: int):
        ne, _ = self.prepare_edges()
        u = np.empty(ne, np.uint32)
        v = np.empty(ne, np.uint32)
        w = np.empty(ne, np.float32)
        check(lib().coresg_hip_edges(self.h, mpts, ptr(u), ptr(v), ptr(w)))
        return u, v, w

    def extract(self, mpts_list, edges: bool = False, labels: bool = False):
        mp = _u32(list(mpts_list))
        q = mp.size
        info = (MstInfo * q)()
        stride = max(self.n - 1, 0)
        ou = np.empty(q * stride, np.uint32) if edges else None
        ov = np.empty(q * stride, np.uint32) if edges else None
        ow = np.empty(q * stride, np.float32) if edges else None
        ol = np.empty(q * self.n, np.uint32) if labels else None
        check(lib().coresg_hip_extract(self.h, ptr(mp), q, info, ptr(ou), ptr(ov), ptr(ow),
                                       ptr(ol)))
        out = []
        for i in range(q):
            r = dict(mpts=info[i].mpts, num_edges=info[i].num_edges,
                     num_components=info[i].num_components, rounds=info[i].rounds,
                     weight=info[i].weight)
            if edges:
                t = info[i].num_edges
                r["u"] = ou[i * stride:i * stride + t]
                r["v"] = ov[i * stride:i * stride + t]
                r["w"] = ow[i * stride:i * stride + t]
            if labels:
                r["labels"] = ol[i * self.n:(i + 1) * self.n]
            out.append(r)
        return out

    def exact_mst(self, mpts: int, labels: bool = False) -> dict:
        """CoreSG::exact_mst (coresg.cpp:419-430): Prim over the complete MRD
        graph of the resident points; edges in acceptance order."""
        info = MstInfo()
        n1 = max(self.n - 1, 1)
        ou = np.empty(n1, np.uint32)
        ov = np.empty(n1, np.uint32)
        ow = np.empty(n1, np.float32)
        ol = np.empty(max(self.n, 1), np.uint32) if labels else None
        check(lib().coresg_hip_exact_mst(self.h, mpts, C.byref(info), ptr(ou), ptr(ov), ptr(ow),
                                         ptr(ol)))
        t = info.num_edges
        r = dict(mpts=info.mpts, num_edges=t, num_components=info.num_components,
                 rounds=info.rounds, weight=info.weight, u=ou[:t], v=ov[:t], w=ow[:t])
        if labels:
            r["labels"] = ol[:self.n]
        return r

    def sampled_recall(self, queries: int, seed: int):
        r = C.c_double()
        ratio = C.c_double()
        check(lib().coresg_hip_sampled_recall(self.h, queries, seed, C.byref(r), C.byref(ratio)))
        return r.value, ratio.value

    # ---- profiling ------------------------------------------------------------
    @classmethod
    def load(cls, path: str, device: int = 0) -> "Graph":
        """CoreSG::load (io.cpp:262-330) of a .coresg file onto the device."""
        h = lib().coresg_hip_graph_load(device, os.fsencode(path))
        if not h:
            raise CoresgHipError(last_error())
        g = cls.__new__(cls)
        g.h = C.c_void_p(h)
        g.n = int(lib().coresg_hip_graph_size(g.h))
        with open(path, "rb") as f:
            head = f.read(20)
        g.k, g.dim, g.stride = (int(x) for x in np.frombuffer(head[8:20], np.uint32))
        return g

    def save(self, path: str) -> None:
        """CoreSG::save (io.cpp:232-260) in the reference's .coresg format."""
        check(lib().coresg_hip_graph_save(self.h, os.fsencode(path)))

    def autotune_sweep(self, mpts) -> dict:
        """Time the sweep's launch shapes with these mpts values (<= 32 used)
        and keep the fastest for this graph and its size bucket."""
        mp = _u32(list(mpts))
        r = TuneResult()
        check(lib().coresg_hip_autotune_sweep(self.h, ptr(mp), mp.size, C.byref(r)))
        return r.asdict()

    def set_exact_launch(self, mode: int):
        """exact_mst's launches: 0 one cooperative launch for every Prim step,
        1 one launch per step (captured batches); same edges."""
        check(lib().coresg_hip_graph_set_exact_launch(self.h, int(mode)))

    def set_plain(self, on: bool):
        """The plain one-launch-per-step form (host sync after every launch,
        one mpts per sweep pass, no captured graphs); same results."""
        check(lib().coresg_hip_graph_set_plain(self.h, int(on)))

    def set_profiling(self, on: bool):
        check(lib().coresg_hip_graph_set_profiling(self.h, int(on)))

    def kernel_time(self, name: str):
        ms = C.c_double()
        n = C.c_uint64()
        check(lib().coresg_hip_graph_kernel_times(self.h, name.encode(), C.byref(ms),
                                                  C.byref(n)))
        return ms.value, n.value

    def kernel_work(self, name: str):
        w = (C.c_uint64 * 4)()
        check(lib().coresg_hip_graph_kernel_work(self.h, name.encode(), w))
        return [int(v) for v in w]

    def set_instances(self, offsets):
        """Independent instances: instance i holds points [offsets[i], offsets[i+1])."""
        off = _u32(offsets)
        self.instance_offsets = off
        check(lib().coresg_hip_graph_set_instances(self.h, ptr(off), max(off.size - 1, 0)))

    def extract_instances(self, mpts) -> list:
        """Per-instance MST summaries: out[i][q] for instance i, mpts[q]."""
        mp = _u32(list(mpts))
        ninst = self.instance_offsets.size - 1
        info = (MstInfo * (ninst * mp.size))()
        check(lib().coresg_hip_extract_instances(self.h, ptr(mp), mp.size, info))
        return [[dict(mpts=info[i * mp.size + q].mpts, num_edges=info[i * mp.size + q].num_edges,
                      num_components=info[i * mp.size + q].num_components,
                      rounds=info[i * mp.size + q].rounds, weight=info[i * mp.size + q].weight)
                 for q in range(mp.size)] for i in range(ninst)]

    def set_audit(self, on: bool):
        check(lib().coresg_hip_graph_set_audit(self.h, int(bool(on))))

    def reset_timers(self):
        check(lib().coresg_hip_graph_reset_timers(self.h))


# ---- reference-shaped host-pointer calls -------------------------------------

def mrd_block(pts_padded: np.ndarray, dim: int, core: np.ndarray, r0: int = 0,
              r1: int | None = None, device: int = 0, metric: int = 0) -> np.ndarray:
    pts = _f32(pts_padded)
    m, stride = pts.shape
    r1 = m if r1 is None else r1
    out = np.empty((r1 - r0) * m, np.float32)
    core = _f32(core)
    check(lib().coresg_hip_mrd_block(device, ptr(pts), m, dim, stride, metric, ptr(core), r0, r1,
                                     ptr(out)))
    return out.reshape(r1 - r0, m)


def probe_gather(rows: int, stride: int, gathers: int, device: int = 0) -> float:
    """Device ms of `gathers` random whole-row reads of a rows x stride array."""
    ms = C.c_double()
    check(lib().coresg_hip_probe_gather(device, rows, stride, gathers, C.byref(ms)))
    return ms.value


def probe_mfma(a: np.ndarray, b: np.ndarray, bf16: bool, device: int = 0) -> np.ndarray:
    """a @ b.T for 16 x 32 operands on one wavefront of the filter join's
    matrix-core instruction and lane layout."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    if a.shape != (16, 32) or b.shape != (16, 32):
        raise ValueError("probe_mfma: operands are 16 x 32")
    c = np.empty((16, 16), np.float32)
    check(lib().coresg_hip_probe_mfma(device, int(bool(bf16)), ptr(a), ptr(b), ptr(c)))
    return c


def tune_table() -> str:
    """The process-wide table of tuned sweep shapes ("log2n k slots lanes
    max_blocks" per line)."""
    m = lib().coresg_hip_tune_table(None, 0)
    buf = C.create_string_buffer(m + 1)
    lib().coresg_hip_tune_table(buf, m + 1)
    return buf.value.decode()


def tune_table_load(text: str) -> None:
    check(lib().coresg_hip_tune_table_load(text.encode()))


def write_edges_tsv(path: str, u, v, w) -> None:
    """write_edges_tsv (io.cpp:131-138), host-only."""
    u, v, w = _u32(u), _u32(v), _f32(w)
    check(lib().coresg_hip_write_edges_tsv(os.fsencode(path), ptr(u), ptr(v), ptr(w), u.size))


def write_labels_tsv(path: str, labels) -> None:
    """write_labels_tsv (io.cpp:140-145), host-only."""
    lab = _u32(labels)
    check(lib().coresg_hip_write_labels_tsv(os.fsencode(path), ptr(lab), lab.size))


def mrd_block_noalloc(pts_padded: np.ndarray, dim: int, core: np.ndarray, r0: int = 0,
                      r1: int | None = None, device: int = 0, metric: int = 0) -> np.ndarray:
    """mrd_block through the allocation-free entry point (fixed device scratch)."""
    pts = _f32(pts_padded)
    m, stride = pts.shape
    r1 = m if r1 is None else r1
    out = np.empty((r1 - r0) * m, np.float32)
    core = _f32(core)
    check(lib().coresg_hip_mrd_block_noalloc(device, ptr(pts), m, dim, stride, metric, ptr(core),
                                             r0, r1, ptr(out)))
    return out.reshape(r1 - r0, m)


METRIC_L2, METRIC_COSINE = 0, 1


def nndescent(vectors_padded: np.ndarray, dim: int, k: int, sample: int = 16, iters: int = 10,
              seed: int = 42, delta: float = 0.001, graph=None, device: int = 0,
              metric: int = METRIC_L2):
    vec = _f32(vectors_padded)
    n, stride = vec.shape
    cfg = NndConfig(n=n, dim=dim, stride=stride, k=k, sample=sample, iters=iters, seed=seed,
                    metric=metric, delta=delta, random_fill=1)
    if graph is None:
        ids = np.full(n * k, 0xFFFFFFFF, np.uint32)
        dist = np.full(n * k, np.inf, np.float32)
        flags = np.zeros(n * k, np.uint8)
        count = np.zeros(n, np.uint32)
    else:
        ids, dist, flags, count = [np.ascontiguousarray(a).copy() for a in graph]
    upd = C.c_uint64()
    check(lib().coresg_hip_nndescent(device, C.byref(cfg), ptr(vec), ptr(ids), ptr(dist),
                                     ptr(flags), ptr(count), C.byref(upd)))
    return (ids, dist, flags, count), upd.value


class Session:
    """A serving session over one device-resident CORE-SG: inserts applied in
    submission order by one device worker, extracts served concurrently on the
    latest published version by another (waiting extracts share one fused
    sweep); every result reports the version it read."""

    def __init__(self, rows: np.ndarray, k_max: int, params: BuildParams | None = None,
                 hbm_budget: int = 0, device: int = 0, graph: "Graph | None" = None):
        p = params or build_params()
        self.params = p
        if graph is not None:
            h = lib().coresg_hip_session_adopt(graph.h, C.byref(p), hbm_budget)
            if not h:
                raise CoresgHipError(last_error())
            graph.h = None  # owned by the session now
            self.dim, self.k = graph.dim, graph.k
            self.build_stats = None
        else:
            rows = _f32(rows)
            self.dim, self.k = rows.shape[1], int(k_max)
            st = BuildStats()
            h = lib().coresg_hip_session_create(device, ptr(rows), rows.shape[0], rows.shape[1],
                                                self.k, C.byref(p), hbm_budget, C.byref(st))
            if not h:
                raise CoresgHipError(last_error())
            self.build_stats = st.asdict()
        self.h = C.c_void_p(h)

    def submit_insert(self, rows: np.ndarray) -> int:
        rows = _f32(rows)
        if rows.ndim != 2 or rows.shape[1] != self.dim:
            raise ValueError(f"rows must be (m, {self.dim})")
        rid = lib().coresg_hip_session_submit_insert(self.h, ptr(rows), rows.shape[0])
        if not rid:
            raise CoresgHipError(last_error())
        return int(rid)

    def submit_extract(self, mpts, edges: bool = False, labels: bool = False) -> int:
        mp = _u32(list(mpts))
        rid = lib().coresg_hip_session_submit_extract(self.h, ptr(mp), mp.size, int(edges),
                                                      int(labels))
        if not rid:
            raise CoresgHipError(last_error())
        return int(rid)

    def wait(self, rid: int, timeout: float = -1.0) -> dict | None:
        st = RequestStatus()
        r = lib().coresg_hip_session_wait(self.h, rid, timeout, C.byref(st))
        if r < 0:
            raise CoresgHipError(last_error())
        return st.asdict() if r == 1 else None

    def result(self, rid: int, edges: bool = False, labels: bool = False) -> list:
        st = self.wait(rid)
        if st["state"] != REQ_DONE:
            raise CoresgHipError(st["error"])
        q, n = st["num_msts"], st["graph_n"]
        info = (MstInfo * q)()
        es = max(n - 1, 0)
        ou = np.empty(q * es, np.uint32) if edges else None
        ov = np.empty(q * es, np.uint32) if edges else None
        ow = np.empty(q * es, np.float32) if edges else None
        ol = np.empty(q * n, np.uint32) if labels else None
        check(lib().coresg_hip_session_extract_result(self.h, rid, info, ptr(ou), ptr(ov), ptr(ow),
                                                      ptr(ol)))
        out = []
        for i in range(q):
            r = dict(mpts=info[i].mpts, num_edges=info[i].num_edges,
                     num_components=info[i].num_components, rounds=info[i].rounds,
                     weight=info[i].weight)
            if edges:
                t = info[i].num_edges
                r["u"], r["v"], r["w"] = (ou[i * es:i * es + t], ov[i * es:i * es + t],
                                          ow[i * es:i * es + t])
            if labels:
                r["labels"] = ol[i * n:(i + 1) * n]
            out.append(r)
        return out

    def release(self, rid: int) -> None:
        check(lib().coresg_hip_session_release(self.h, rid))

    def graph_info(self) -> dict:
        """Points and device bytes of the served graph (call once the queue
        has drained)."""
        gh = lib().coresg_hip_session_graph(self.h)
        if not gh:
            raise CoresgHipError(last_error())
        gh = C.c_void_p(gh)
        return dict(n=int(lib().coresg_hip_graph_size(gh)),
                    device_bytes=int(lib().coresg_hip_session_device_bytes(self.h)))

    def close(self):
        if getattr(self, "h", None):
            lib().coresg_hip_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
