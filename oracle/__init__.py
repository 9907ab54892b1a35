"""ctypes view of the C restatement of the reference CPU path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker. The product library never
imports it.

Parity of each function is pinned by tests/test_oracle_golden.py against
vectors the reference itself wrote (tests/golden/), plus closed forms and
brute force where those leave gaps, including the connectivity repair's
representative cut above max_reps (libstdc++'s partial_sort restated step for
step, against reference builds with tied scores, tests/golden/maxreps.npz).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile the C restatement (and nothing else)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


class OrcParams(C.Structure):
    _fields_ = [
        ("k_max", C.c_uint32), ("sample", C.c_uint32), ("iters", C.c_uint32),
        ("delta", C.c_double), ("nnd_seed", C.c_uint64), ("brute_force_limit", C.c_uint32),
        ("antihub_multiplier", C.c_double), ("max_pair_ops", C.c_double),
        ("min_selected", C.c_size_t), ("max_selected", C.c_size_t),
        ("min_per_component", C.c_size_t), ("ensure_connected", C.c_int),
        ("reps_per_component", C.c_size_t), ("max_reps", C.c_size_t),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_l2sqr.restype = C.c_float
        L.orc_l2sqr.argtypes = [_f32p, _f32p, C.c_uint32]
        L.orc_pair_distances.argtypes = [_f32p, C.c_uint32, _u32p, C.c_size_t, _f32p, _f32p]
        L.orc_noderng_first.restype = C.c_uint64
        L.orc_noderng_first.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                        C.c_uint32, _u32p]
        L.orc_mt64_first.argtypes = [C.c_uint64, C.c_size_t, _u64p]
        L.orc_try_insert.restype = C.c_int
        L.orc_try_insert.argtypes = [_u32p, _f32p, _u8p, _u32p, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.c_float, C.c_int, C.POINTER(C.c_uint32)]
        L.orc_init_random.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64,
                                      _u32p, _f32p, _u8p, _u32p]
        L.orc_nndescent.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_double, _u32p, _f32p, _u8p, _u32p, _u64p]
        L.orc_brute_force.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, _f32p,
                                      _u8p, _u32p]
        L.orc_validate.restype = C.c_int
        L.orc_validate.argtypes = [_u32p, _f32p, _u32p, C.c_uint32, C.c_uint32]
        L.orc_core_distances.argtypes = [_f32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        L.orc_undirected_edges.restype = C.c_size_t
        L.orc_undirected_edges.argtypes = [_u32p, _f32p, _u32p, C.c_uint32, C.c_uint32,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_mst_kruskal.restype = C.c_size_t
        L.orc_mst_kruskal.argtypes = [C.c_uint32, _u32p, _u32p, _f32p, C.c_size_t, _u32p, _u32p,
                                      _f32p, _u32p, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_double)]
        L.orc_reverse_degrees.argtypes = [_u32p, _u32p, C.c_uint32, C.c_uint32, _i32p]
        L.orc_score.restype = C.c_double
        L.orc_score.argtypes = [_u32p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, _i32p,
                                C.c_uint32]
        L.orc_antihub_budget.restype = C.c_size_t
        L.orc_antihub_budget.argtypes = [C.c_double, C.c_uint32, C.c_size_t, C.c_uint32,
                                         C.c_double, C.c_size_t, C.c_size_t,
                                         C.POINTER(C.c_size_t), C.POINTER(C.c_int)]
        L.orc_select.restype = C.c_size_t
        L.orc_select.argtypes = [C.c_uint32, _f64p, _i32p, C.c_size_t, C.c_void_p, C.c_uint32,
                                 C.c_size_t, _u32p]
        L.orc_components.restype = C.c_uint32
        L.orc_components.argtypes = [_u32p, _u32p, C.c_uint32, C.c_uint32, _u32p]
        L.orc_mrd_block.argtypes = [_f32p, C.c_size_t, C.c_uint32, _f32p, C.c_size_t,
                                    C.c_size_t, _f32p]
        L.orc_prim_dense.restype = C.c_size_t
        L.orc_prim_dense.argtypes = [_f32p, C.c_size_t, _u32p, _u32p, _f32p]
        L.orc_default_params.argtypes = [C.POINTER(OrcParams)]
        L.orc_sg_build.restype = C.c_void_p
        L.orc_sg_build.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.POINTER(OrcParams)]
        L.orc_sg_load.restype = C.c_void_p
        L.orc_sg_load.argtypes = [C.c_char_p, C.POINTER(OrcParams)]
        L.orc_sg_save.restype = C.c_int
        L.orc_sg_save.argtypes = [C.c_void_p, C.c_char_p]
        L.orc_sg_free.argtypes = [C.c_void_p]
        L.orc_sg_rebuild_bridges.argtypes = [C.c_void_p]
        for nm in ("orc_sg_n", "orc_sg_k", "orc_sg_dim", "orc_sg_stride"):
            getattr(L, nm).restype = C.c_uint32
            getattr(L, nm).argtypes = [C.c_void_p]
        L.orc_sg_points.restype = C.POINTER(C.c_float)
        L.orc_sg_points.argtypes = [C.c_void_p]
        L.orc_sg_graph.argtypes = [C.c_void_p] + [C.POINTER(C.c_void_p)] * 4
        L.orc_sg_bridges.restype = C.c_size_t
        L.orc_sg_bridges.argtypes = [C.c_void_p] + [C.POINTER(C.c_void_p)] * 3
        L.orc_sg_repair_count.restype = C.c_size_t
        L.orc_sg_repair_count.argtypes = [C.c_void_p]
        L.orc_sg_selected.restype = C.c_size_t
        L.orc_sg_selected.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
        L.orc_sg_scores.restype = C.POINTER(C.c_double)
        L.orc_sg_scores.argtypes = [C.c_void_p]
        L.orc_sg_rdeg.restype = C.POINTER(C.c_int32)
        L.orc_sg_rdeg.argtypes = [C.c_void_p]
        L.orc_sg_conn_labels.restype = C.POINTER(C.c_uint32)
        L.orc_sg_conn_labels.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
        L.orc_sg_knn_edges.restype = C.c_size_t
        L.orc_sg_knn_edges.argtypes = [C.c_void_p]
        L.orc_sg_stats.argtypes = [C.c_void_p, _f64p]
        L.orc_sg_edges.restype = C.c_size_t
        L.orc_sg_edges.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sg_mst.restype = C.c_size_t
        L.orc_sg_mst.argtypes = [C.c_void_p, C.c_uint32, _u32p, _u32p, _f32p, _u32p,
                                 C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_sg_exact_mst.restype = C.c_size_t
        L.orc_sg_exact_mst.argtypes = [C.c_void_p, C.c_uint32, _u32p, _u32p, _f32p, _u32p,
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_double)]
        L.orc_sg_sampled_recall.restype = C.c_double
        L.orc_sg_sampled_recall.argtypes = [C.c_void_p, C.c_size_t, C.c_uint64,
                                            C.POINTER(C.c_double)]
        _lib = L
    return _lib


def params(**kw) -> OrcParams:
    p = OrcParams()
    lib().orc_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def stride_of(d: int) -> int:
    return (d + 7) & ~7


def pad(x: np.ndarray) -> np.ndarray:
    n, d = x.shape
    out = np.zeros((n, stride_of(d)), dtype=np.float32)
    out[:, :d] = x
    return out


# ---- primitives --------------------------------------------------------------
def l2sqr_pairs(padded: np.ndarray, pairs: np.ndarray):
    padded = np.ascontiguousarray(padded, np.float32)
    pairs = np.ascontiguousarray(pairs, np.uint32)
    m = pairs.size // 2
    r = np.empty(m, np.float32)
    d = np.empty(m, np.float32)
    lib().orc_pair_distances(padded.ravel(), padded.shape[1], pairs.ravel(), m, r, d)
    return r, d


def empty_graph(n: int, k: int):
    return (np.full(n * k, 0xFFFFFFFF, np.uint32), np.full(n * k, np.inf, np.float32),
            np.zeros(n * k, np.uint8), np.zeros(n, np.uint32))


def init_random(padded: np.ndarray, k: int, seed: int):
    n, stride = padded.shape
    g = empty_graph(n, k)
    lib().orc_init_random(np.ascontiguousarray(padded).ravel(), n, stride, k, seed, *g)
    return g


def nndescent(padded: np.ndarray, k: int, sample: int, iters: int, seed: int,
              delta: float = 0.001):
    n, stride = padded.shape
    g = init_random(padded, k, seed)
    st = np.zeros(3, np.uint64)
    lib().orc_nndescent(np.ascontiguousarray(padded).ravel(), n, stride, k, sample, iters, delta,
                        *g, st)
    return g, st


def brute_force(padded: np.ndarray, k: int):
    n, stride = padded.shape
    g = empty_graph(n, k)
    lib().orc_brute_force(np.ascontiguousarray(padded).ravel(), n, stride, k, *g)
    return g


def undirected_edges(ids, dist, count, n, k):
    m = lib().orc_undirected_edges(ids, dist, count, n, k, None, None, None)
    u = np.empty(m, np.uint32)
    v = np.empty(m, np.uint32)
    w = np.empty(m, np.float32)
    lib().orc_undirected_edges(ids, dist, count, n, k, u.ctypes.data, v.ctypes.data,
                               w.ctypes.data)
    return u, v, w


def kruskal(n, u, v, w):
    m = len(u)
    ou = np.empty(max(n, 1), np.uint32)
    ov = np.empty(max(n, 1), np.uint32)
    ow = np.empty(max(n, 1), np.float32)
    lab = np.empty(max(n, 1), np.uint32)
    nc = C.c_uint32()
    wt = C.c_double()
    t = lib().orc_mst_kruskal(n, np.ascontiguousarray(u, np.uint32),
                              np.ascontiguousarray(v, np.uint32),
                              np.ascontiguousarray(w, np.float32), m, ou, ov, ow, lab,
                              C.byref(nc), C.byref(wt))
    return ou[:t], ov[:t], ow[:t], lab[:n], nc.value, wt.value


def reverse_degrees(ids, count, n, k):
    r = np.empty(n, np.int32)
    lib().orc_reverse_degrees(ids, count, n, k, r)
    return r


def scores(ids, count, n, k, rdeg, k_for_score=None):
    kf = k if k_for_score is None else k_for_score
    return np.array([lib().orc_score(ids, count, n, k, kf, rdeg, v) for v in range(n)],
                    np.float64)


def budget(multiplier, dim, num_edges, n, max_pair_ops=2e9, max_selected=0, min_selected=2):
    unc = C.c_size_t()
    cap = C.c_int()
    b = lib().orc_antihub_budget(multiplier, dim, num_edges, n, max_pair_ops, max_selected,
                                 min_selected, C.byref(unc), C.byref(cap))
    return b, unc.value, bool(cap.value)


def select(score, rdeg, bud, comp=None, ncomp=0, per=2):
    n = len(score)
    out = np.empty(bud + ncomp * per + 1, np.uint32)
    cptr = None if comp is None else np.ascontiguousarray(comp, np.uint32).ctypes.data
    m = lib().orc_select(n, np.ascontiguousarray(score, np.float64),
                         np.ascontiguousarray(rdeg, np.int32), bud, cptr, ncomp,
                         per if comp is not None else 0, out)
    return out[:m]


def components(ids, count, n, k):
    lab = np.empty(n, np.uint32)
    nc = lib().orc_components(ids, count, n, k, lab)
    return lab, nc


def mrd_block(pts_padded, core):
    m, stride = pts_padded.shape
    out = np.empty(m * m, np.float32)
    lib().orc_mrd_block(np.ascontiguousarray(pts_padded).ravel(), m, stride,
                        np.ascontiguousarray(core, np.float32), 0, m, out)
    return out.reshape(m, m)


def prim_dense(mat):
    m = mat.shape[0]
    u = np.empty(max(m, 1), np.uint32)
    v = np.empty(max(m, 1), np.uint32)
    w = np.empty(max(m, 1), np.float32)
    t = lib().orc_prim_dense(np.ascontiguousarray(mat, np.float32).ravel(), m, u, v, w)
    return u[:t], v[:t], w[:t]


def mt64_first(seed, count):
    out = np.empty(count, np.uint64)
    lib().orc_mt64_first(seed, count, out)
    return out


def noderng_first(seed, node, rnd, count, bound):
    out = np.empty(count, np.uint32)
    lib().orc_noderng_first(seed, node, rnd, count, bound, out)
    return out


# ---- whole CORE-SG -----------------------------------------------------------
class CoreSG:
    """Oracle CORE-SG (single-threaded reference schedule)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle)
        L = lib()
        self.n = L.orc_sg_n(self.h)
        self.k = L.orc_sg_k(self.h)
        self.dim = L.orc_sg_dim(self.h)
        self.stride = L.orc_sg_stride(self.h)

    @classmethod
    def build(cls, rows: np.ndarray, **kw):
        rows = np.ascontiguousarray(rows, np.float32)
        p = params(**kw)
        h = lib().orc_sg_build(rows.ravel(), rows.shape[0], rows.shape[1], C.byref(p))
        return cls(h)

    @classmethod
    def load(cls, path: str, **kw):
        p = params(**kw)
        h = lib().orc_sg_load(path.encode(), C.byref(p))
        if not h:
            raise IOError(f"cannot load {path}")
        return cls(h)

    def save(self, path: str) -> None:
        if not lib().orc_sg_save(self.h, path.encode()):
            raise IOError(f"cannot save {path}")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_sg_free(self.h)
            self.h = None

    def rebuild_bridges(self) -> None:
        lib().orc_sg_rebuild_bridges(self.h)

    def points(self) -> np.ndarray:
        p = lib().orc_sg_points(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n * self.stride,)).reshape(
            self.n, self.stride).copy()

    def graph(self):
        ptrs = [C.c_void_p() for _ in range(4)]
        lib().orc_sg_graph(self.h, *[C.byref(p) for p in ptrs])
        nk = self.n * self.k

        def arr(p, ct, cnt):
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(cnt,)).copy()

        return (arr(ptrs[0], C.c_uint32, nk), arr(ptrs[1], C.c_float, nk),
                arr(ptrs[2], C.c_uint8, nk), arr(ptrs[3], C.c_uint32, self.n))

    def bridges(self):
        u, v, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        m = lib().orc_sg_bridges(self.h, C.byref(u), C.byref(v), C.byref(w))
        if m == 0:
            return (np.empty(0, np.uint32), np.empty(0, np.uint32), np.empty(0, np.float32))
        raw = np.ctypeslib.as_array(C.cast(u, C.POINTER(C.c_uint32)), shape=(3 * m,)).copy()
        rec = raw.reshape(m, 3)
        return rec[:, 0].copy(), rec[:, 1].copy(), rec[:, 2].copy().view(np.float32)

    def repair_count(self) -> int:
        return lib().orc_sg_repair_count(self.h)

    def selected(self) -> np.ndarray:
        p = C.c_void_p()
        m = lib().orc_sg_selected(self.h, C.byref(p))
        if m == 0:
            return np.empty(0, np.uint32)
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint32)), shape=(m,)).copy()

    def scores(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_sg_scores(self.h), shape=(self.n,)).copy()

    def rdeg(self) -> np.ndarray:
        return np.ctypeslib.as_array(lib().orc_sg_rdeg(self.h), shape=(self.n,)).copy()

    def conn(self):
        nc = C.c_uint32()
        p = lib().orc_sg_conn_labels(self.h, C.byref(nc))
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy(), nc.value

    def knn_edges(self) -> int:
        return lib().orc_sg_knn_edges(self.h)

    def stats(self) -> dict:
        s = np.zeros(7, np.float64)
        lib().orc_sg_stats(self.h, s)
        return dict(knn_s=s[0], hub_s=s[1], bridge_s=s[2], total_s=s[3], rounds=int(s[4]),
                    updates=int(s[5]), evals=int(s[6]))

    def edges(self, mpts: int):
        m = lib().orc_sg_edges(self.h, mpts, None, None, None)
        if m == C.c_size_t(-1).value:
            raise ValueError("mpts exceeds k_max")
        u = np.empty(m, np.uint32)
        v = np.empty(m, np.uint32)
        w = np.empty(m, np.float32)
        lib().orc_sg_edges(self.h, mpts, u.ctypes.data, v.ctypes.data, w.ctypes.data)
        return u, v, w

    def mst(self, mpts: int):
        n = max(self.n, 1)
        u = np.empty(n, np.uint32)
        v = np.empty(n, np.uint32)
        w = np.empty(n, np.float32)
        lab = np.empty(n, np.uint32)
        nc = C.c_uint32()
        wt = C.c_double()
        t = lib().orc_sg_mst(self.h, mpts, u, v, w, lab, C.byref(nc), C.byref(wt))
        if t == C.c_size_t(-1).value:
            raise ValueError("mpts exceeds k_max")
        return dict(u=u[:t], v=v[:t], w=w[:t], labels=lab[:self.n], ncomp=nc.value,
                    weight=wt.value)

    def exact_mst(self, mpts: int):
        """CoreSG::exact_mst: Prim over the complete MRD graph (coresg.cpp:419-430)."""
        n = max(self.n, 1)
        u = np.empty(n, np.uint32)
        v = np.empty(n, np.uint32)
        w = np.empty(n, np.float32)
        lab = np.empty(n, np.uint32)
        nc = C.c_uint32()
        wt = C.c_double()
        t = lib().orc_sg_exact_mst(self.h, mpts, u, v, w, lab, C.byref(nc), C.byref(wt))
        return dict(u=u[:t], v=v[:t], w=w[:t], labels=lab[:self.n], ncomp=nc.value,
                    weight=wt.value)

    def sampled_recall(self, queries: int, seed: int):
        ratio = C.c_double()
        r = lib().orc_sg_sampled_recall(self.h, queries, seed, C.byref(ratio))
        return r, ratio.value
