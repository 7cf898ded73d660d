"""ctypes binding of oracle/oracle.c -- the CPU checker (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It builds oracle/_build/liboracle.so on first use when a C
compiler is present and the library is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_hash_values.restype = C.c_uint64
        L.orc_hash_values.argtypes = [C.c_uint64, u32p, C.c_uint32,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        L.orc_read_header.argtypes = [C.c_char_p, C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_read_graph.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_uint64,
                                     u64p, u32p, C.c_void_p, C.c_void_p]
        L.orc_write_graph.argtypes = [C.c_char_p, C.c_int, C.c_uint64, C.c_uint64,
                                      u64p, u32p, C.c_void_p, C.c_void_p]
        L.orc_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, u64p]
        L.orc_rmat_tuples.argtypes = [C.c_int, C.c_uint64, u64p, C.c_uint64,
                                      C.c_uint64, u32p, u32p]
        L.orc_weight.restype = C.c_uint32
        L.orc_weight.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_rmat_csr.restype = C.c_uint64
        L.orc_rmat_csr.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_double,
                                   C.c_double, C.c_double, C.c_int, u64p, u32p]
        L.orc_transpose.argtypes = [C.c_uint64, u64p, u32p, u64p, u32p]
        for name in ("orc_bfs",):
            getattr(L, name).argtypes = [C.c_uint64, u64p, u32p, C.c_uint32, u32p]
        L.orc_bfs_mt.restype = C.c_uint64
        L.orc_bfs_mt.argtypes = [C.c_uint64, u64p, u32p, C.c_uint32, u32p]
        L.orc_cc.argtypes = [C.c_uint64, u64p, u32p, u32p]
        L.orc_cc_minreach.argtypes = [C.c_uint64, u64p, u32p, u32p]
        L.orc_cc_mt.restype = C.c_uint64
        L.orc_cc_mt.argtypes = [C.c_uint64, u64p, u32p, u32p]
        L.orc_sssp.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_uint32, u32p]
        L.orc_sssp_bf.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_uint32, u32p]
        L.orc_sssp_mt.restype = C.c_uint64
        L.orc_sssp_mt.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_uint32, u32p]
        L.orc_sssp_delta_mt.restype = C.c_uint64
        L.orc_sssp_delta_mt.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_uint32,
                                        C.c_uint32, u32p]
        L.orc_bfs_multi.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_uint32, C.c_int,
                                    u64p, u64p, u32p, C.c_void_p]
        L.orc_pr_exact.restype = C.c_int
        L.orc_pr_exact.argtypes = [C.c_uint64, u64p, u32p, C.c_double, C.c_double,
                                   C.c_int, f64p]
        L.orc_pr_exact_mt.restype = C.c_int
        L.orc_pr_exact_mt.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_double, C.c_double,
                                      C.c_int, f64p]
        L.orc_pr_pull_ref.restype = C.c_int
        L.orc_pr_pull_ref.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_double,
                                      C.c_double, C.c_int, f64p]
        L.orc_pr_push_ref.restype = C.c_int
        L.orc_pr_push_ref.argtypes = [C.c_uint64, u64p, u32p, C.c_double, C.c_double,
                                      C.c_int, f64p]
        L.orc_pr_pull_mt.restype = C.c_int
        L.orc_pr_pull_mt.argtypes = [C.c_uint64, u64p, u32p, u32p, C.c_float,
                                     C.c_float, C.c_int, f32p]
        L.orc_pr_push_mt.restype = C.c_int
        L.orc_pr_push_mt.argtypes = [C.c_uint64, u64p, u32p, C.c_float, C.c_float,
                                     C.c_int, f32p]
        L.orc_threads.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class CSR:
    """A graph in host memory: off (n+1 u64), adj (m u32), w (m u32 or None),
    outdeg (n u32, CSC files only)."""

    def __init__(self, off, adj, w=None, outdeg=None):
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.adj = np.ascontiguousarray(adj, dtype=np.uint32)
        self.w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
        self.outdeg = None if outdeg is None else np.ascontiguousarray(outdeg, dtype=np.uint32)

    @property
    def n(self):
        return len(self.off) - 1

    @property
    def m(self):
        return len(self.adj)

    def degrees(self):
        return np.diff(self.off).astype(np.uint32)


FMT = {"csr": 0, "csc": 1, "wcsr": 2}


def read_graph(path, fmt):
    L = lib()
    n, m, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = L.orc_read_header(path.encode(), C.byref(n), C.byref(m), C.byref(b))
    if rc:
        raise IOError(f"cannot read {path}")
    n, m = n.value, m.value
    off = np.zeros(n + 1, np.uint64)
    adj = np.zeros(m, np.uint32)
    w = np.zeros(m, np.uint32)
    od = np.zeros(n, np.uint32) if FMT[fmt] == 1 else None
    rc = L.orc_read_graph(path.encode(), FMT[fmt], n, m, off, adj, _ptr(w), _ptr(od))
    if rc < 0:
        raise IOError(f"bad file {path}: {rc}")
    return CSR(off, adj, w if rc == 1 else None, od)


def write_graph(path, fmt, g):
    rc = lib().orc_write_graph(path.encode(), FMT[fmt], g.n, g.m, g.off, g.adj,
                               _ptr(g.w), _ptr(g.outdeg))
    if rc:
        raise IOError(path)


def mix64(x):
    return lib().orc_mix64(x)


def hash_values(val, sentinel):
    r, d = C.c_uint64(), C.c_uint32()
    val = np.ascontiguousarray(val, dtype=np.uint32)
    h = lib().orc_hash_values(len(val), val, sentinel, C.byref(r), C.byref(d))
    return h, r.value, d.value


RMAT_ABC = (0.57, 0.19, 0.19)


def rmat(scale, edge_factor, seed=1, symmetric=True, weights=False, abc=RMAT_ABC):
    """R-MAT graph with 2^scale vertices and edge_factor * 2^scale directed
    edges before clean-up (symmetric: edge_factor * 2^scale / 2 tuples,
    each stored both ways); self-loops and duplicates removed."""
    n = 1 << scale
    ntup = (edge_factor * n) // 2 if symmetric else edge_factor * n
    off = np.zeros(n + 1, np.uint64)
    adj = np.zeros(2 * ntup if symmetric else ntup, np.uint32)
    m = lib().orc_rmat_csr(scale, ntup, seed, abc[0], abc[1], abc[2], 1 if symmetric else 0,
                           off, adj)
    adj = adj[:m].copy()
    w = None
    if weights:
        w = edge_weights(off, adj, seed)
    return CSR(off, adj, w)


def edge_weights(off, adj, seed):
    src = np.repeat(np.arange(len(off) - 1, dtype=np.uint64), np.diff(off).astype(np.int64))
    dst = adj.astype(np.uint64)
    lo = np.minimum(src, dst)
    hi = np.maximum(src, dst)
    with np.errstate(over="ignore"):
        key = np.uint64((seed * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF) + ((lo << np.uint64(32)) | hi)
    return (np.uint32(1) + (mix64_np(key) & np.uint64(63)).astype(np.uint32)).astype(np.uint32)


def mix64_np(z):
    z = np.asarray(z, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        z += np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def transpose(g):
    toff = np.zeros(g.n + 1, np.uint64)
    tadj = np.zeros(g.m, np.uint32)
    lib().orc_transpose(g.n, g.off, g.adj, toff, tadj)
    return CSR(toff, tadj, None, g.degrees())


def bfs(g, src):
    out = np.zeros(g.n, np.uint32)
    lib().orc_bfs(g.n, g.off, g.adj, src, out)
    return out


def bfs_mt(g, src):
    out = np.zeros(g.n, np.uint32)
    trav = lib().orc_bfs_mt(g.n, g.off, g.adj, src, out)
    return out, trav


def cc(g):
    out = np.zeros(g.n, np.uint32)
    lib().orc_cc(g.n, g.off, g.adj, out)
    return out


def cc_minreach(g):
    out = np.zeros(g.n, np.uint32)
    lib().orc_cc_minreach(g.n, g.off, g.adj, out)
    return out


def cc_mt(g):
    out = np.zeros(g.n, np.uint32)
    lib().orc_cc_mt(g.n, g.off, g.adj, out)
    return out


def _weights(g):
    if g.w is None:
        raise ValueError("graph has no weights")
    return g.w


def sssp(g, src):
    out = np.zeros(g.n, np.uint32)
    lib().orc_sssp(g.n, g.off, g.adj, _weights(g), src, out)
    return out


def sssp_bf(g, src):
    out = np.zeros(g.n, np.uint32)
    lib().orc_sssp_bf(g.n, g.off, g.adj, _weights(g), src, out)
    return out


def sssp_mt(g, src):
    out = np.zeros(g.n, np.uint32)
    lib().orc_sssp_mt(g.n, g.off, g.adj, _weights(g), src, out)
    return out


def sssp_delta_mt(g, src, delta=16):
    """Delta-stepping SSSP on the host cores (same values as sssp)."""
    out = np.zeros(g.n, np.uint32)
    lib().orc_sssp_delta_mt(g.n, g.off, g.adj, _weights(g), src, delta, out)
    return out


def bfs_multi(g, srcs, symmetric=False, full=False):
    """k <= 64 BFS queries at once (one bit per query): per-query digests
    (hash_values with sentinel n+1), reached counts, depths, and optionally
    the k x n value arrays."""
    srcs = np.ascontiguousarray(srcs, np.uint32)
    k = len(srcs)
    if k > 64:
        raise ValueError("at most 64 queries")
    h = np.zeros(max(k, 1), np.uint64)
    r = np.zeros(max(k, 1), np.uint64)
    d = np.zeros(max(k, 1), np.uint32)
    f = np.zeros((k, g.n), np.uint32) if full else None
    lib().orc_bfs_multi(g.n, g.off, g.adj, srcs, k, int(bool(symmetric)), h, r, d, _ptr(f))
    return {"hash": h[:k], "reached": r[:k], "depth": d[:k], "full": f}


ALPHA_F32 = float(np.float32(0.85))


def pr_exact(g, alpha=ALPHA_F32, tol=1e-13, max_iter=10000):
    out = np.zeros(g.n, np.float64)
    it = lib().orc_pr_exact(g.n, g.off, g.adj, alpha, tol, max_iter, out)
    return out, it


def pr_exact_mt(csc, alpha=ALPHA_F32, tol=1e-12, max_iter=10000):
    """FP64 fixpoint from the CSC (multithreaded Jacobi)."""
    out = np.zeros(csc.n, np.float64)
    it = lib().orc_pr_exact_mt(csc.n, csc.off, csc.adj, csc.outdeg, alpha, tol, max_iter, out)
    return out, it


def pr_pull_ref(csc, alpha=ALPHA_F32, tol=1e-3, max_iter=1000):
    out = np.zeros(csc.n, np.float64)
    it = lib().orc_pr_pull_ref(csc.n, csc.off, csc.adj, csc.outdeg, alpha, tol, max_iter, out)
    return out, it


def pr_push_ref(g, alpha=ALPHA_F32, tol=1e-3, max_iter=1000):
    out = np.zeros(g.n, np.float64)
    it = lib().orc_pr_push_ref(g.n, g.off, g.adj, alpha, tol, max_iter, out)
    return out, it


def pr_pull_mt(csc, alpha=0.85, tol=1e-7, max_iter=100):
    out = np.zeros(csc.n, np.float32)
    it = lib().orc_pr_pull_mt(csc.n, csc.off, csc.adj, csc.outdeg, alpha, tol, max_iter, out)
    return out, it


def pr_push_mt(g, alpha=0.85, tol=1e-7, max_iter=100):
    out = np.zeros(g.n, np.float32)
    it = lib().orc_pr_push_mt(g.n, g.off, g.adj, alpha, tol, max_iter, out)
    return out, it


def normalized_l1(p, q):
    p = np.asarray(p, np.float64)
    q = np.asarray(q, np.float64)
    return float(np.abs(p / p.sum() - q / q.sum()).sum())


def threads():
    return lib().orc_threads()


def set_threads(t):
    lib().orc_set_threads(t)
