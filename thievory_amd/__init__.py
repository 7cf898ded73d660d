"""thievory_amd -- MI355X-native out-of-memory graph engine (host mirror).

A ctypes binding of libthv_hip.so (include/thv.h).  The engine itself is
native: HIP kernels for gfx950 behind a C ABI.  This module only moves
arguments and numpy buffers across that boundary; it has no compute path of
its own and raises if the native library is missing.

    g = Graph.rmat(20, 16)                 # or Graph.open("g.csr", "csr")
    g.store_config(budget_bytes=..., partition_edges=1 << 16)
    levels, stats = g.bfs(1)

Value conventions are the reference's (qqueke/Thievory): BFS level = 1 +
hops with n+1 for unreached (src/graph.cuh:240-246), CC = least id reaching
the vertex (src/cc/cc_kernels.cu:50-66), SSSP = min(n+1, 1 + distance),
PageRank unnormalised (src/pr/pr_kernels.cu:449-476, 846-860).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libthv_hip.so")
CLI_PATH = os.path.join(HERE, "_lib", "thv_main")

FMT = {"csr": 0, "csc": 1, "wcsr": 2}
ALGO = {"bfs": 0, "cc": 1, "sssp": 2, "pr": 3}
PR_MODE = {"push": 0, "pull": 1}
OPTIONS = {"symmetric": 1, "direction": 2, "do_alpha": 3, "do_beta": 4, "do_vertex_ns": 5,
           "do_link_gbps": 6, "do_pull_degree": 7, "parent_hints": 8, "hint_bytes": 9,
           "sssp_delta": 10, "cc_union": 11,
           "cc_uf_steps": 12, "kernel_timing": 13, "store_layout": 14,
           "do_summary": 15}
ERRORS = {0: "OK", -1: "EINVAL", -2: "EIO", -3: "ENOMEM", -4: "EHIP", -5: "EFORMAT",
          -6: "EBUSY", -7: "ENOTSUP", -8: "ESTATE"}

RMAT_ABC = (0.57, 0.19, 0.19)


class ThvError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


TUNE_WORKLOAD = {"bfs": 0, "sssp": 1, "pr": 2, "batch-bfs": 3, "batch-sssp": 4}


class TuneCfg(C.Structure):
    _fields_ = [("workload", C.c_int), ("nsources", C.c_uint32),
                ("sources", C.c_void_p), ("partitions", C.c_int), ("repeats", C.c_int),
                ("pr_iterations", C.c_int), ("cache_path", C.c_char_p)]


class TuneResult(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int), ("edges_per_thread", C.c_int),
                ("probe_blocks_per_cu", C.c_int), ("batch_blocks_per_cu", C.c_int),
                ("partition_edges", C.c_uint32), ("sssp_delta", C.c_uint32),
                ("pull_degree", C.c_uint32), ("candidates", C.c_uint32),
                ("default_ms", C.c_double), ("tuned_ms", C.c_double), ("from_cache", C.c_int),
                ("store_layout", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("edges_traversed", C.c_uint64), ("hbm_bytes", C.c_uint64),
                ("host_bytes", C.c_uint64), ("moved_bytes", C.c_uint64),
                ("reached", C.c_uint64), ("iterations", C.c_uint32),
                ("host_partitions_touched", C.c_uint32),
                ("host_partitions_skipped", C.c_uint32), ("expand_launches", C.c_uint32),
                ("ms", C.c_double), ("expand_ms", C.c_double),
                ("edges_examined", C.c_uint64), ("bottom_up_levels", C.c_uint32),
                ("pad", C.c_uint32), ("topdown_host_bytes", C.c_uint64),
                ("topdown_hbm_bytes", C.c_uint64), ("bottom_up_ms", C.c_double),
                ("bottom_up_host_bytes", C.c_uint64), ("hint_hits", C.c_uint64),
                ("probe_ms", C.c_double), ("probe_launches", C.c_uint32),
                ("pad2", C.c_uint32), ("probe_hbm_bytes", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("pad")}


class ServeCfg(C.Structure):
    _fields_ = [("max_lanes", C.c_uint32), ("scratch_bytes", C.c_uint64),
                ("keep_full", C.c_uint32), ("sssp_engines", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("algo", C.c_int32), ("status", C.c_int32),
                ("src", C.c_uint32), ("depth", C.c_uint32), ("reached", C.c_uint64),
                ("hash", C.c_uint64), ("submit_ms", C.c_double), ("start_ms", C.c_double),
                ("done_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
_vp = C.c_void_p


def lib():
    """Loads the native library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run `python -m thievory_amd.build` "
                          "(the engine has no non-native fallback)")
    L = C.CDLL(LIB_PATH)
    G = C.POINTER(_vp)
    sig = {
        "thv_graph_open": (C.c_int, [C.c_char_p, C.c_int, C.c_int, G]),
        "thv_graph_from_arrays": (C.c_int, [C.c_uint64, C.c_uint64, _vp, _vp, _vp, _vp,
                                            C.c_int, G]),
        "thv_graph_rmat": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, C.c_int, C.c_int, C.c_int, G]),
        "thv_graph_info": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "thv_graph_export": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
        "thv_graph_close": (C.c_int, [_vp]),
        "thv_graph_set_option": (C.c_int, [_vp, C.c_int, C.c_double]),
        "thv_graph_get_option": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_double)]),
        "thv_store_config": (C.c_int, [_vp, C.c_uint64, C.c_uint32]),
        "thv_store_info": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint64)]),
        "thv_bfs": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(Stats)]),
        "thv_bfs_noalloc": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(Stats)]),
        "thv_cc": (C.c_int, [_vp, _vp, C.POINTER(Stats)]),
        "thv_sssp": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(Stats)]),
        "thv_pr": (C.c_int, [_vp, C.c_int, C.c_float, C.c_float, C.c_int, _vp,
                             C.POINTER(Stats)]),
        "thv_batch": (C.c_int, [_vp, C.c_int, _vp, C.c_uint32, _vp, _vp, _vp, _vp,
                                C.POINTER(Stats)]),
        "thv_autotune": (C.c_int, [_vp, C.c_char_p, _vp]),
        "thv_tune": (C.c_int, [_vp, C.POINTER(TuneCfg), C.POINTER(TuneResult)]),
        "thv_link_probe": (C.c_int, [C.c_uint64, _vp]),
        "thv_graph64_from_lists": (C.c_int, [C.c_uint64, C.c_uint64, _vp, _vp, _vp, _vp,
                                             C.POINTER(_vp)]),
        "thv_graph64_store_config": (C.c_int, [_vp, C.c_uint64, C.c_uint32]),
        "thv_graph64_info": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
        "thv_bfs64": (C.c_int, [_vp, C.c_uint64, _vp, C.POINTER(Stats), _vp]),
        "thv_cc64": (C.c_int, [_vp, _vp, C.POINTER(Stats), _vp]),
        "thv_sssp64": (C.c_int, [_vp, C.c_uint64, _vp, C.POINTER(Stats), _vp]),
        "thv_graph64_values": (C.c_int, [_vp, _vp, C.c_uint64, _vp]),
        "thv_graph64_close": (None, [_vp]),
        "thv_serve_start": (C.c_int, [_vp, C.POINTER(ServeCfg), G]),
        "thv_serve_submit": (C.c_int, [_vp, C.c_int, C.c_uint32, C.POINTER(C.c_uint64)]),
        "thv_serve_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(Result), _vp]),
        "thv_serve_drain": (C.c_int, [_vp]),
        "thv_serve_stop": (None, [_vp]),
        "thv_last_error": (C.c_char_p, []),
        "thv_version": (C.c_char_p, []),
        "thv_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                      C.POINTER(C.c_uint64)]),
        "thv_device_memory": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().thv_last_error()
        raise ThvError(rc, msg.decode() if msg else "")
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _out(out, n, dtype):
    """out=None allocates a host array; out=False keeps the result in HBM
    only (no readback: the call then measures just the traversal)."""
    if out is False:
        return None
    return np.zeros(n, dtype) if out is None else out


def link_probe(nbytes=1 << 30):
    """thv_link_probe: GB/s of a zero-copy kernel read of mapped pinned host
    memory, pinned host -> device DMA, a kernel read of HBM and a device
    copy (read + written), over buffers of nbytes."""
    o = np.zeros(4, np.float64)
    check(lib().thv_link_probe(int(nbytes), _p(o)))
    return {"zero_copy_read_GBps": float(o[0]), "dma_h2d_GBps": float(o[1]),
            "hbm_read_GBps": float(o[2]), "hbm_copy_GBps": float(o[3]), "bytes": int(nbytes)}


def device_info():
    name = C.create_string_buffer(64)
    cus, hbm = C.c_int(), C.c_uint64()
    check(lib().thv_device_info(name, 64, C.byref(cus), C.byref(hbm)))
    return {"arch": name.value.decode(), "cus": cus.value, "hbm_bytes": hbm.value}


def device_memory():
    fr, tot = C.c_uint64(), C.c_uint64()
    check(lib().thv_device_memory(C.byref(fr), C.byref(tot)))
    return {"free": fr.value, "total": tot.value}


class Graph:
    """A graph resident in the engine: pinned host edge array, per-vertex
    tables and the HBM share of the split edge store."""

    def __init__(self, handle):
        self._h = handle
        n, m, eb, csc = C.c_uint64(), C.c_uint64(), C.c_int(), C.c_int()
        check(lib().thv_graph_info(self._h, C.byref(n), C.byref(m), C.byref(eb), C.byref(csc)))
        self.n, self.m, self.edge_bytes, self.is_csc = n.value, m.value, eb.value, bool(csc.value)

    # ---- construction -------------------------------------------------------
    @classmethod
    def open(cls, path, fmt="csr", edge_bytes=4):
        h = _vp()
        check(lib().thv_graph_open(os.fsencode(path), FMT[fmt], edge_bytes, C.byref(h)))
        return cls(h)

    @classmethod
    def from_arrays(cls, off, dst, w=None, outdeg=None, edge_bytes=None):
        off = np.ascontiguousarray(off, np.uint64)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = None if w is None else np.ascontiguousarray(w, np.uint32)
        outdeg = None if outdeg is None else np.ascontiguousarray(outdeg, np.uint32)
        if edge_bytes is None:
            edge_bytes = 8 if w is not None else 4
        h = _vp()
        check(lib().thv_graph_from_arrays(len(off) - 1, len(dst), _p(off), _p(dst), _p(w),
                                          _p(outdeg), edge_bytes, C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor, seed=1, symmetric=True, weights=False, transpose=False,
             abc=RMAT_ABC):
        h = _vp()
        check(lib().thv_graph_rmat(scale, edge_factor, abc[0], abc[1], abc[2], seed,
                                   int(symmetric), int(weights), int(transpose), C.byref(h)))
        return cls(h)

    def close(self):
        if self._h:
            check(lib().thv_graph_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- store ----------------------------------------------------------------
    def store_config(self, budget_bytes=None, partition_edges=1 << 16):
        b = (1 << 64) - 1 if budget_bytes is None else int(budget_bytes)
        check(lib().thv_store_config(self._h, b, partition_edges))

    def store_info(self):
        p, pd, de = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().thv_store_info(self._h, C.byref(p), C.byref(pd), C.byref(de)))
        return {"partitions": p.value, "device_partitions": pd.value,
                "device_edges": de.value}

    def set_option(self, name, value):
        check(lib().thv_graph_set_option(self._h, OPTIONS[name], float(value)))

    def get_option(self, name):
        v = C.c_double()
        check(lib().thv_graph_get_option(self._h, OPTIONS[name], C.byref(v)))
        return v.value

    def offsets(self):
        off = np.zeros(self.n + 1, np.uint64)
        check(lib().thv_graph_export(self._h, _p(off), None, None, None))
        return off

    def export(self):
        off = np.zeros(self.n + 1, np.uint64)
        dst = np.zeros(self.m, np.uint32)
        w = np.zeros(self.m, np.uint32) if self.edge_bytes == 8 else None
        od = np.zeros(self.n, np.uint32) if self.is_csc else None
        check(lib().thv_graph_export(self._h, _p(off), _p(dst), _p(w), _p(od)))
        return off, dst, w, od

    # ---- algorithms -------------------------------------------------------------
    def bfs(self, src, out=None):
        out = _out(out, self.n, np.uint32)
        st = Stats()
        check(lib().thv_bfs(self._h, src, _p(out), C.byref(st)))
        return out, st

    def bfs_noalloc(self, src, out=None):
        """BFS levels with every per-call buffer in the library's static
        device array (thv_bfs_noalloc): the call allocates nothing."""
        out = _out(out, self.n, np.uint32)
        st = Stats()
        check(lib().thv_bfs_noalloc(self._h, src, _p(out), C.byref(st)))
        return out, st

    def cc(self, out=None):
        out = _out(out, self.n, np.uint32)
        st = Stats()
        check(lib().thv_cc(self._h, _p(out), C.byref(st)))
        return out, st

    def sssp(self, src, out=None):
        out = _out(out, self.n, np.uint32)
        st = Stats()
        check(lib().thv_sssp(self._h, src, _p(out), C.byref(st)))
        return out, st

    def pr(self, mode="push", damping=0.85, tol=1e-7, max_iter=100, out=None):
        out = _out(out, self.n, np.float32)
        st = Stats()
        check(lib().thv_pr(self._h, PR_MODE[mode], damping, tol, max_iter, _p(out),
                           C.byref(st)))
        return out, st

    def batch(self, algo, srcs, full=False):
        srcs = np.ascontiguousarray(srcs, np.uint32)
        k = len(srcs)
        h = np.zeros(k, np.uint64)
        r = np.zeros(k, np.uint64)
        d = np.zeros(k, np.uint32)
        f = np.zeros((k, self.n), np.uint32) if full else None
        st = Stats()
        check(lib().thv_batch(self._h, ALGO[algo], _p(srcs), k, _p(h), _p(r), _p(d), _p(f),
                              C.byref(st)))
        return {"hash": h, "reached": r, "depth": d, "full": f, "stats": st}

    def autotune(self, cache_path=None):
        ch = np.zeros(2, np.int32)
        check(lib().thv_autotune(self._h, None if cache_path is None else os.fsencode(cache_path),
                                 _p(ch)))
        return {"blocks_per_cu": int(ch[0]), "edges_per_thread": int(ch[1])}

    def tune(self, workload="bfs", sources=None, partitions=False, repeats=3, pr_iterations=10,
             cache_path=None):
        """thv_tune: times `workload` under candidate settings (launch shapes,
        probe and batch grids, full-scan degree, SSSP bucket width and, with
        partitions, partition sizes), keeps the fastest on the handle and
        returns the choice with the starting and tuned median times."""
        src = None if sources is None else np.ascontiguousarray(sources, np.uint32)
        cfg = TuneCfg(TUNE_WORKLOAD[workload], 0 if src is None else len(src),
                      None if src is None else src.ctypes.data, int(partitions), repeats,
                      pr_iterations, None if cache_path is None else os.fsencode(cache_path))
        r = TuneResult()
        check(lib().thv_tune(self._h, C.byref(cfg), C.byref(r)))
        return {f: getattr(r, f) for f, _ in TuneResult._fields_}


class Graph64:
    """A graph with 64-bit vertex ids (thv_graph64_*): the sorted ids of the
    vertices with edges, their list offsets and u64 destinations (optional
    u32 weights); values are n x u64."""

    def __init__(self, n, ids, loff, dst, w=None):
        self._h = None
        self.n = int(n)
        ids = np.ascontiguousarray(ids, np.uint64)
        loff = np.ascontiguousarray(loff, np.uint64)
        dst = np.ascontiguousarray(dst, np.uint64)
        w = None if w is None else np.ascontiguousarray(w, np.uint32)
        h = _vp()
        check(lib().thv_graph64_from_lists(self.n, len(ids), _p(ids), _p(loff), _p(dst),
                                           _p(w), C.byref(h)))
        self._h = h
        self.m = int(loff[-1]) if len(loff) else 0

    def store_config(self, budget_bytes=None, partition_edges=1 << 16):
        check(lib().thv_graph64_store_config(
            self._h, 2 ** 64 - 1 if budget_bytes is None else int(budget_bytes), partition_edges))

    def store_info(self):
        p, pd, de = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().thv_graph64_info(self._h, None, None, C.byref(p), C.byref(pd), C.byref(de)))
        return {"partitions": p.value, "device_partitions": pd.value,
                "device_edges": de.value}

    def _run(self, fn, *args, out=False):
        o = np.zeros(self.n, np.uint64) if out else None
        st = Stats()
        dg = np.zeros(3, np.uint64)
        check(fn(self._h, *args, _p(o), C.byref(st), _p(dg)))
        return {"values": o, "hash": int(dg[0]), "reached": int(dg[1]),
                "depth": int(dg[2]), "stats": st}

    def bfs(self, src, out=False):
        return self._run(lib().thv_bfs64, int(src), out=out)

    def cc(self, out=False):
        return self._run(lib().thv_cc64, out=out)

    def sssp(self, src, out=False):
        return self._run(lib().thv_sssp64, int(src), out=out)

    def values(self, vertices):
        at = np.ascontiguousarray(vertices, np.uint64)
        o = np.zeros(len(at), np.uint64)
        check(lib().thv_graph64_values(self._h, _p(at), len(at), _p(o)))
        return o

    def close(self):
        if self._h:
            lib().thv_graph64_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching query server over a resident graph."""

    def __init__(self, graph, max_lanes=64, scratch_bytes=0, keep_full=False, sssp_engines=0):
        self.graph = graph
        cfg = ServeCfg(max_lanes, scratch_bytes, 1 if keep_full else 0, sssp_engines)
        h = _vp()
        check(lib().thv_serve_start(graph._h, C.byref(cfg), C.byref(h)))
        self._h = h

    def submit(self, algo, src):
        t = C.c_uint64()
        check(lib().thv_serve_submit(self._h, ALGO[algo], src, C.byref(t)))
        return t.value

    def poll(self, ticket, full=None):
        r = Result()
        check(lib().thv_serve_poll(self._h, ticket, C.byref(r), _p(full)))
        return r.as_dict()

    def drain(self):
        check(lib().thv_serve_drain(self._h))

    def stop(self):
        if self._h:
            lib().thv_serve_stop(self._h)
            self._h = None

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
