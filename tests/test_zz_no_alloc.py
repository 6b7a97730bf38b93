"""kt_ktruss_noalloc (one k of the directed COO+CSR peel on a caller-held
device layout, peel state in the library's fixed device array) against
kt_plan_ktruss, the entry point it is a variant of: the same k-truss edge
set, bit-exact, on two calls in a row, the second on a smaller graph (the
array keeps the first call's contents).  The caller's layout is put in
device memory with the HIP runtime the library itself links (torch ships a
HIP runtime of its own, and two runtimes in one process do not share the
device)."""
import ctypes

import numpy as np
import pytest

import ktruss_amd as K

pytestmark = pytest.mark.gpu


def oriented_layout(n, s, d):
    """The degree-ordered oriented COO+CSR of undirected pairs (s, d): vertices
    relabelled by ascending (degree, id), each edge stored once from its lower
    to its higher label, rows ascending.  Returns ptr, src, dst and the map
    new label -> old vertex."""
    deg = np.bincount(np.concatenate([s, d]).astype(np.int64), minlength=n)
    order = np.lexsort((np.arange(n), deg)).astype(np.uint32)
    new_of = np.empty(n, np.uint32)
    new_of[order] = np.arange(n, dtype=np.uint32)
    a, b = new_of[s], new_of[d]
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    idx = np.argsort((lo.astype(np.uint64) << np.uint64(32)) | hi.astype(np.uint64), kind="stable")
    lo, hi = lo[idx], hi[idx]
    ptr = np.zeros(n + 1, np.uint32)
    np.cumsum(np.bincount(lo.astype(np.int64), minlength=n), out=ptr[1:])
    return ptr, lo, hi, order


def undirected_keys(src, dst):
    a = np.minimum(src, dst).astype(np.uint64)
    b = np.maximum(src, dst).astype(np.uint64)
    return np.unique((a << np.uint64(32)) | b)


_hip = None


def hip():
    """The HIP runtime libktruss_hip.so is linked against (already loaded by it)."""
    global _hip
    if _hip is None:
        K.lib()
        L = ctypes.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip = L
    return _hip


class DeviceArray:
    """A uint32 array in device memory, owned by the test (the caller)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a, np.uint32)
        self.n = len(a)
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), max(4 * self.n, 4)) == 0
        if self.n:
            assert hip().hipMemcpy(self.p, a.ctypes.data, 4 * self.n, 1) == 0  # host to device

    def get(self):
        out = np.zeros(self.n, np.uint32)
        if self.n:
            assert hip().hipMemcpy(out.ctypes.data, self.p, 4 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def run_noalloc(n, s, d, k):
    ptr, lo, hi, order = oriented_layout(n, s, d)
    arrs = [DeviceArray(ptr), DeviceArray(lo), DeviceArray(hi)]
    try:
        tp, ts, td = arrs
        live = K.ktruss_noalloc(n, len(lo), tp.p.value, ts.p.value, td.p.value, k)
        out = td.get()
    finally:
        for a in arrs:
            a.free()
    # ids unchanged, bit 31 marks the entries outside the k-truss
    assert np.array_equal(out & np.uint32(0x7FFFFFFF), hi)
    alive = (out & np.uint32(0x80000000)) == 0
    assert int(alive.sum()) == live
    return undirected_keys(order[lo[alive]], order[hi[alive]])


def plan_ktruss(n, s, d, k):
    g = K.Graph(n, s, d, pairs=True)
    p = K.Plan(g, K.default_variant())
    es, ed, _ = p.ktruss(k)
    return undirected_keys(es, ed)


def test_noalloc_matches_plan_ktruss_twice_second_smaller():
    K.init(0)
    cases = [(13, 6), (10, 4)]  # (R-MAT scale, k): the second call is smaller
    for scale, k in cases:
        n, s, d = K.generate_rmat(scale, 16, seed=3)
        print(f"noalloc: R-MAT {scale}, {len(s)} edges, k = {k}", flush=True)
        got = run_noalloc(n, s, d, k)
        want = plan_ktruss(n, s, d, k)
        assert len(want) > 0
        assert np.array_equal(got, want), f"R-MAT {scale}, k = {k}: {len(got)} vs {len(want)} edges"


def test_noalloc_k3_and_k2():
    K.init(0)
    n, s, d = K.generate_rmat(11, 8, seed=5)
    for k in (3, 2):
        assert np.array_equal(run_noalloc(n, s, d, k), plan_ktruss(n, s, d, k))


def test_noalloc_rejects_oversize_and_bad_layout():
    K.init(0)
    with pytest.raises(K.KtError) as e:
        K.ktruss_noalloc(10, K.NOALLOC_MAX_ENTRIES + 1, 0, 0, 0, 4)
    assert e.value.status == K.ERR_CAPACITY
    n, s, d = K.generate_rmat(9, 8, seed=2)
    ptr, lo, hi, _ = oriented_layout(n, s, d)
    r = int(np.argmax(np.diff(ptr.astype(np.int64))))  # the longest row
    b = int(ptr[r])
    bad = hi.copy()
    bad[b], bad[b + 1] = hi[b + 1], hi[b]  # row no longer ascending
    arrs = [DeviceArray(ptr), DeviceArray(lo), DeviceArray(bad)]
    try:
        tp, ts, td = arrs
        with pytest.raises(K.KtError) as e:
            K.ktruss_noalloc(n, len(lo), tp.p.value, ts.p.value, td.p.value, 4)
        assert e.value.status == K.ERR_ARG
        assert f"entry {b + 1} " in str(e.value)
        assert np.array_equal(td.get(), bad)  # nothing changed
    finally:
        for a in arrs:
            a.free()
