"""ctypes loader for the CPU oracle (oracle/build/libktoracle.so).

TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg.  The product never imports this module.
The C sources restate the reference's Host path; see ktoracle.h for the map
from each function to the reference lines it follows.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libktoracle.so")

UNDIRECTED, DIRECTED_BY_INDEX, DIRECTED_BY_DEGREE = 0, 1, 2
COOCSR, TILED_COOCSR = 0, 1
RECOUNT_ALL, RECOUNT_AFFECTED = 0, 1
SCHED_PLAIN, SCHED_JUMP, SCHED_REUSE = 0, 1, 2

DIRECTEDNESS = {"undirected": 0, "directed-by-index": 1, "directed-by-degree": 2}
FORMATS = {"coocsr": 0, "tiled-coocsr": 1}
RECOUNTS = {"all": 0, "affected": 1}


class Variant(ctypes.Structure):
    _fields_ = [("directedness", ctypes.c_int), ("format", ctypes.c_int),
                ("recount", ctypes.c_int), ("num_edge_remove_iter", ctypes.c_uint32),
                ("tiles_per_dim", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("count_passes", ctypes.c_uint64), ("recount_edges", ctypes.c_uint64),
                ("merge_elems", ctypes.c_uint64), ("live_in_passes", ctypes.c_uint64),
                ("compactions", ctypes.c_uint64), ("compact_elems", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("phases", ctypes.c_uint64),
                ("triangles", ctypes.c_uint64), ("k_max", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class TraceRec(ctypes.Structure):
    _fields_ = [("k", ctypes.c_uint32), ("iter", ctypes.c_uint32),
                ("deleted", ctypes.c_uint64), ("recounted", ctypes.c_uint64)]


class Trace(ctypes.Structure):
    _fields_ = [("recs", ctypes.POINTER(TraceRec)), ("cap", ctypes.c_uint64),
                ("len", ctypes.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.kto_ktruss.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, ctypes.c_uint32,
                                 ctypes.POINTER(Variant), u32p, u32p, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(Stats),
                                 ctypes.POINTER(Trace)]
        L.kto_decompose.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p,
                                    ctypes.POINTER(Variant), ctypes.c_uint32, u32p,
                                    ctypes.POINTER(Stats), ctypes.POINTER(Trace)]
        L.kto_bruteforce_tau.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p]
        L.kto_balg_bytes.argtypes = [ctypes.POINTER(Stats)]
        L.kto_balg_bytes.restype = ctypes.c_double
        L.kto_read_tsv.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(u32p),
                                   ctypes.POINTER(u32p)]
        L.kto_write_tsv.argtypes = [ctypes.c_char_p, ctypes.c_uint64, u32p, u32p]
        L.kto_verify.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p,
                                 ctypes.c_uint64, u32p, u32p]
        L.kto_free.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def make_variant(directedness="directed-by-degree", fmt="coocsr", recount="all",
                 e_iters=0, tiles=8):
    d = DIRECTEDNESS[directedness] if isinstance(directedness, str) else directedness
    f = FORMATS[fmt] if isinstance(fmt, str) else fmt
    r = RECOUNTS[recount] if isinstance(recount, str) else recount
    return Variant(d, f, r, e_iters, tiles)


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with code {code}")
        self.code = code


def _trace(cap):
    recs = (TraceRec * max(cap, 1))()
    tr = Trace(ctypes.cast(recs, ctypes.POINTER(TraceRec)), cap, 0)
    return tr, recs


def _trace_list(tr, recs):
    n = min(tr.len, tr.cap)
    return [(recs[i].k, recs[i].iter, recs[i].deleted, recs[i].recounted) for i in range(n)]


def ktruss(n, src, dst, k, variant=None, trace_cap=1 << 16):
    """Reference -k semantics.  Returns (out_src, out_dst, stats, trace)."""
    variant = variant or make_variant()
    s, sp = _u32(src)
    d, dp = _u32(dst)
    m = len(s)
    os_ = np.zeros(max(m, 1), np.uint32)
    od_ = np.zeros(max(m, 1), np.uint32)
    out_m = ctypes.c_uint64(0)
    st = Stats()
    tr, recs = _trace(trace_cap)
    rc = lib().kto_ktruss(n, m, sp, dp, k, ctypes.byref(variant),
                          os_.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                          od_.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                          m, ctypes.byref(out_m), ctypes.byref(st), ctypes.byref(tr))
    if rc != 0:
        raise OracleError(rc, "kto_ktruss")
    mm = out_m.value
    return os_[:mm].copy(), od_[:mm].copy(), st.as_dict(), _trace_list(tr, recs)


def decompose(n, src, dst, variant=None, schedule=SCHED_JUMP | SCHED_REUSE, trace_cap=1 << 20):
    """Truss number per input entry.  Returns (tau, stats, trace)."""
    variant = variant or make_variant()
    s, sp = _u32(src)
    d, dp = _u32(dst)
    m = len(s)
    tau = np.zeros(max(m, 1), np.uint32)
    st = Stats()
    tr, recs = _trace(trace_cap)
    rc = lib().kto_decompose(n, m, sp, dp, ctypes.byref(variant), schedule,
                             tau.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                             ctypes.byref(st), ctypes.byref(tr))
    if rc != 0:
        raise OracleError(rc, "kto_decompose")
    return tau[:m].copy(), st.as_dict(), _trace_list(tr, recs)


def balg_bytes(stats):
    return (4.0 * stats["merge_elems"] + 24.0 * stats["recount_edges"] +
            12.0 * stats["live_in_passes"] + 16.0 * stats["compact_elems"])


def bruteforce_tau(n, src, dst):
    s, sp = _u32(src)
    d, dp = _u32(dst)
    m = len(s)
    tau = np.zeros(max(m, 1), np.uint32)
    rc = lib().kto_bruteforce_tau(n, m, sp, dp, tau.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    if rc != 0:
        raise OracleError(rc, "kto_bruteforce_tau")
    return tau[:m].copy()


def read_tsv(path):
    n = ctypes.c_uint32(0)
    m = ctypes.c_uint64(0)
    sp = ctypes.POINTER(ctypes.c_uint32)()
    dp = ctypes.POINTER(ctypes.c_uint32)()
    rc = lib().kto_read_tsv(path.encode(), ctypes.byref(n), ctypes.byref(m),
                            ctypes.byref(sp), ctypes.byref(dp))
    if rc != 0:
        raise OracleError(rc, "kto_read_tsv")
    mm = m.value
    src = np.ctypeslib.as_array(sp, shape=(mm,)).copy() if mm else np.zeros(0, np.uint32)
    dst = np.ctypeslib.as_array(dp, shape=(mm,)).copy() if mm else np.zeros(0, np.uint32)
    lib().kto_free(ctypes.cast(sp, ctypes.c_void_p))
    lib().kto_free(ctypes.cast(dp, ctypes.c_void_p))
    return n.value, src, dst


def write_tsv(path, src, dst):
    s, sp = _u32(src)
    d, dp = _u32(dst)
    rc = lib().kto_write_tsv(path.encode(), len(s), sp, dp)
    if rc != 0:
        raise OracleError(rc, "kto_write_tsv")


def verify(n, src_a, dst_a, src_b, dst_b):
    a, ap = _u32(src_a)
    b, bp = _u32(dst_a)
    c, cp = _u32(src_b)
    d, dp = _u32(dst_b)
    return lib().kto_verify(n, len(a), ap, bp, len(c), cp, dp) == 0
