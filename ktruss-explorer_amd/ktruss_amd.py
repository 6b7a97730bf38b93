"""Python binding of the MI355X k-truss engine (build/libktruss_hip.so).

A thin ctypes layer over the C ABI in include/ktruss.h; every call goes to
the HIP library, which fails loudly when no gfx950 device is visible.  The
names follow the reference's interface (KTrussExplorer):

  Graph(n, src, dst)             the COOGraph handed to ktruss()  (kernel.cu:9)
  Variant(...)                   Config's variant fields          (config.h:163-181)
  Plan(graph, variant).ktruss(k) one k, ./ktruss -k k             (kernel_ddc.cu:116-214)
  Plan.decompose()               all k: truss number per edge     (k++ loop over the same peel)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libktruss_hip.so")
CLI_PATH = os.path.join(_HERE, "build", "ktruss")

OK = 0
ERR_ARG, ERR_SELF_LOOP, ERR_ASYMMETRIC, ERR_DUPLICATE, ERR_RANGE = -1, -2, -3, -4, -5
ERR_NOMEM, ERR_CAPACITY, ERR_IO, ERR_HIP, ERR_NO_DEVICE, ERR_STATE = -6, -7, -8, -9, -10, -11

UNDIRECTED, DIRECTED_BY_INDEX, DIRECTED_BY_DEGREE = 0, 1, 2
COOCSR, TILED_COOCSR = 0, 1
RECOUNT_ALL, RECOUNT_AFFECTED = 0, 1
SCHED_JUMP, SCHED_REUSE = 1, 2
INPUT_SYMMETRIC, INPUT_PAIRS = 0, 1

DIRECTEDNESS = {"undirected": 0, "directed-by-index": 1, "directed-by-degree": 2}
FORMATS = {"coocsr": 0, "tiled-coocsr": 1}
RECOUNTS = {"all": 0, "affected": 1}

# C symbols declared in include/ktruss.h
EXPORTS = [
    "kt_abi_version", "kt_last_error", "kt_status_string", "kt_init", "kt_device_synchronize",
    "kt_free",
    "kt_variant_default", "kt_graph_create", "kt_graph_destroy", "kt_graph_info",
    "kt_read_tsv", "kt_write_tsv", "kt_read_truss_tsv", "kt_write_truss_tsv", "kt_write_edges_bin",
    "kt_read_edges_bin", "kt_is_edges_bin", "kt_verify_edges", "kt_generate_rmat",
    "kt_generate_chung_lu", "kt_generate_erdos_renyi", "kt_plan_create", "kt_plan_destroy",
    "kt_plan_info", "kt_plan_ktruss", "kt_plan_decompose", "kt_plan_truss_numbers",
    "kt_plan_trace", "kt_plan_set_candidates", "kt_plan_set_pick", "kt_plan_phase_log",
    "kt_plan_set_timing", "kt_ktruss_noalloc", "kt_plan_iteration_log", "kt_plan_stage_times",
    "kt_plan_destroy_timed", "kt_plan_set_sweep_sampling", "kt_plan_autotune", "kt_plan_tune_log",
    "kt_mlp_create", "kt_mlp_destroy", "kt_mlp_num_params", "kt_mlp_get_params",
    "kt_mlp_set_params", "kt_mlp_forward", "kt_mlp_loss_grad", "kt_mlp_train",
    "kt_mlp_set_input_range", "kt_mlp_get_input_range",
    "kt_batch_create", "kt_batch_destroy", "kt_batch_info", "kt_batch_set_tuning",
    "kt_batch_get_tuning", "kt_batch_autotune", "kt_batch_tune_log", "kt_batch_decompose",
    "kt_batch_truss_numbers", "kt_batch_graph_info", "kt_batch_set_timing",
]

MLP_IN, MLP_HIDDEN, MLP_MAX_OUT, MAX_CANDIDATES, PHASE_REC = 16, 64, 16, 16, 36
PICK_FIXED, PICK_SWEEP, PICK_MODEL = 0, 1, 2


class KtError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class Variant(ctypes.Structure):
    _fields_ = [("directedness", ctypes.c_int32), ("format", ctypes.c_int32),
                ("recount", ctypes.c_int32), ("num_edge_remove_iter", ctypes.c_uint32),
                ("block_size", ctypes.c_uint32), ("tiles_per_dim", ctypes.c_uint32),
                ("num_parallel_tiles", ctypes.c_uint32), ("tile_size", ctypes.c_uint32),
                ("chunk", ctypes.c_uint32), ("schedule", ctypes.c_uint32),
                ("count_kernel", ctypes.c_uint32), ("cap_lds", ctypes.c_uint32)]

    def describe(self):
        names = {v: k for k, v in DIRECTEDNESS.items()}
        fmt = {v: k for k, v in FORMATS.items()}
        rec = {v: k for k, v in RECOUNTS.items()}
        s = f"{names[self.directedness]}/{fmt[self.format]}/{rec[self.recount]}"
        if self.num_edge_remove_iter:
            s += f"/e{self.num_edge_remove_iter}"
        if self.format == TILED_COOCSR:
            s += f"/t{self.tiles_per_dim}" if not self.tile_size else f"/S{self.tile_size}"
            s += f"/y{self.num_parallel_tiles}"
        return s


class CountConfig(ctypes.Structure):
    _fields_ = [("count_kernel", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
                ("cap_lds", ctypes.c_uint32), ("chunk", ctypes.c_uint32)]

    def describe(self):
        names = {0: "auto", 1: "row", 2: "search", 3: "merge", 4: "delta"}
        chunk = self.chunk or (64 if self.count_kernel == COUNT_KERNELS["delta"] else 128)
        return f"{names[self.count_kernel]}/b{self.block_size}/c{self.cap_lds or 'auto'}/k{chunk}"


class TuneTrial(ctypes.Structure):
    _fields_ = [("tile_size", ctypes.c_uint32), ("num_parallel_tiles", ctypes.c_uint32),
                ("count_kernel", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
                ("cap_lds", ctypes.c_uint32), ("chunk", ctypes.c_uint32), ("ms", ctypes.c_float),
                ("pad", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


class Stats(ctypes.Structure):
    _fields_ = [("count_passes", ctypes.c_uint64), ("recount_edges", ctypes.c_uint64),
                ("merge_elems", ctypes.c_uint64), ("live_in_passes", ctypes.c_uint64),
                ("compactions", ctypes.c_uint64), ("compact_elems", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64), ("phases", ctypes.c_uint64),
                ("k_max", ctypes.c_uint32), ("k_last", ctypes.c_uint32),
                ("ms_total", ctypes.c_double), ("ms_count", ctypes.c_double),
                ("count_launches", ctypes.c_uint64), ("balg_count_bytes", ctypes.c_double),
                ("delta_passes", ctypes.c_uint64), ("delta_edges", ctypes.c_uint64),
                ("delta_elems", ctypes.c_uint64), ("delta_triangles", ctypes.c_uint64),
                ("ms_delta", ctypes.c_double), ("balg_delta_bytes", ctypes.c_double),
                ("delta_live", ctypes.c_uint64), ("ms_setup", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchTuning(ctypes.Structure):
    _fields_ = [("block_size", ctypes.c_uint32), ("chunk", ctypes.c_uint32),
                ("cap_lds", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("mark_chunk", ctypes.c_uint32), ("steps_per_replay", ctypes.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("graphs", ctypes.c_uint64), ("entries", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("replays", ctypes.c_uint64),
                ("count_passes", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("phases", ctypes.c_uint64), ("recount_edges", ctypes.c_uint64),
                ("merge_elems", ctypes.c_uint64), ("live_in_passes", ctypes.c_uint64),
                ("k_max", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("ms_total", ctypes.c_double), ("balg_count_bytes", ctypes.c_double),
                ("ms_count", ctypes.c_double), ("count_launches", ctypes.c_uint64),
                ("delta_elems", ctypes.c_uint64), ("delta_triangles", ctypes.c_uint64),
                ("ms_delta", ctypes.c_double), ("balg_delta_bytes", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


_lib = None
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)


def lib():
    """Load the HIP library; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.kt_last_error.restype = ctypes.c_char_p
        L.kt_status_string.restype = ctypes.c_char_p
        L.kt_status_string.argtypes = [ctypes.c_int]
        L.kt_init.argtypes = [ctypes.c_int]
        L.kt_free.argtypes = [ctypes.c_void_p]
        L.kt_variant_default.argtypes = [ctypes.POINTER(Variant)]
        L.kt_variant_default.restype = None
        L.kt_graph_create.argtypes = [ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p,
                                      ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)]
        L.kt_graph_destroy.argtypes = [ctypes.c_void_p]
        L.kt_graph_destroy.restype = None
        L.kt_graph_info.argtypes = [ctypes.c_void_p, _u32p, _u64p, _u64p, _u32p]
        L.kt_read_tsv.argtypes = [ctypes.c_char_p, _u32p, _u64p, ctypes.POINTER(_u32p),
                                  ctypes.POINTER(_u32p)]
        L.kt_write_tsv.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u32p, _u32p]
        L.kt_read_truss_tsv.argtypes = [ctypes.c_char_p, _u32p, _u64p, ctypes.POINTER(_u32p),
                                        ctypes.POINTER(_u32p), ctypes.POINTER(_u32p)]
        L.kt_write_truss_tsv.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u32p, _u32p, _u32p]
        L.kt_write_edges_bin.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, _u32p, _u32p,
                                         ctypes.c_uint32]
        L.kt_read_edges_bin.argtypes = [ctypes.c_char_p, _u32p, _u64p, ctypes.POINTER(_u32p),
                                        ctypes.POINTER(_u32p), _u32p]
        L.kt_is_edges_bin.argtypes = [ctypes.c_char_p]
        L.kt_verify_edges.argtypes = [ctypes.c_uint64, _u32p, _u32p, ctypes.c_uint64, _u32p, _u32p]
        L.kt_generate_rmat.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_uint64, _u32p,
                                       _u64p, ctypes.POINTER(_u32p), ctypes.POINTER(_u32p)]
        L.kt_generate_chung_lu.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_uint32, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint64, _u64p,
                                           ctypes.POINTER(_u32p), ctypes.POINTER(_u32p)]
        L.kt_generate_erdos_renyi.argtypes = [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64,
                                              _u64p, ctypes.POINTER(_u32p), ctypes.POINTER(_u32p)]
        L.kt_plan_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(Variant),
                                     ctypes.POINTER(ctypes.c_void_p)]
        L.kt_plan_destroy.argtypes = [ctypes.c_void_p]
        L.kt_plan_destroy.restype = None
        L.kt_plan_info.argtypes = [ctypes.c_void_p, _u64p, _u32p]
        L.kt_plan_ktruss.argtypes = [ctypes.c_void_p, ctypes.c_uint32, _u32p, _u32p,
                                     ctypes.c_uint64, _u64p, ctypes.POINTER(Stats)]
        L.kt_plan_decompose.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.kt_plan_truss_numbers.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_uint64]
        L.kt_plan_trace.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_uint64, _u64p]
        L.kt_plan_set_candidates.argtypes = [ctypes.c_void_p, ctypes.POINTER(CountConfig), ctypes.c_uint32]
        L.kt_plan_set_pick.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        L.kt_plan_set_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.kt_plan_phase_log.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint64, _u64p]
        L.kt_mlp_create.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]
        L.kt_mlp_destroy.argtypes = [ctypes.c_void_p]
        L.kt_mlp_destroy.restype = None
        L.kt_mlp_num_params.argtypes = [ctypes.c_void_p, _u32p]
        L.kt_mlp_get_params.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32]
        L.kt_mlp_set_params.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32]
        L.kt_mlp_forward.argtypes = [ctypes.c_void_p, _f32p, ctypes.c_uint32, _f32p]
        L.kt_mlp_loss_grad.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                       _f32p, _f32p]
        L.kt_mlp_train.argtypes = [ctypes.c_void_p, _f32p, _f32p, _f32p, _f32p, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64,
                                   _f32p]
        L.kt_mlp_set_input_range.argtypes = [ctypes.c_void_p, _f32p, _f32p]
        L.kt_mlp_get_input_range.argtypes = [ctypes.c_void_p, _f32p, _f32p]
        _vp = ctypes.c_void_p
        L.kt_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.POINTER(Variant),
                                      ctypes.POINTER(_vp)]
        L.kt_batch_destroy.argtypes = [_vp]
        L.kt_batch_destroy.restype = None
        L.kt_batch_info.argtypes = [_vp, _u32p, _u64p, _u32p]
        L.kt_batch_set_tuning.argtypes = [_vp, ctypes.POINTER(BatchTuning)]
        L.kt_batch_get_tuning.argtypes = [_vp, ctypes.POINTER(BatchTuning)]
        L.kt_batch_autotune.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(BatchTuning)]
        L.kt_batch_tune_log.argtypes = [_vp, ctypes.POINTER(BatchTuning), _f32p, ctypes.c_uint64, _u64p]
        L.kt_batch_decompose.argtypes = [_vp, ctypes.POINTER(BatchStats)]
        L.kt_batch_set_timing.argtypes = [_vp, ctypes.c_int]
        L.kt_batch_truss_numbers.argtypes = [_vp, ctypes.c_uint32, _u32p, ctypes.c_uint64]
        L.kt_batch_graph_info.argtypes = [_vp, ctypes.c_uint32, _u32p]
        L.kt_plan_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(Variant)]
        L.kt_plan_tune_log.argtypes = [ctypes.c_void_p, ctypes.POINTER(TuneTrial), ctypes.c_uint64, _u64p]
        L.kt_plan_set_sweep_sampling.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.kt_ktruss_noalloc.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp, ctypes.c_uint32,
                                        _vp, _u64p]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != OK:
        msg = lib().kt_last_error().decode(errors="replace")
        raise KtError(rc, f"{what}: {msg}")
    return rc


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


def _take_array(p, count):
    """Copy a malloc'ed uint32 array returned by the library, then free it."""
    out = np.ctypeslib.as_array(p, shape=(count,)).copy() if count else np.zeros(0, np.uint32)
    lib().kt_free(ctypes.cast(p, ctypes.c_void_p))
    return out


def init(device=0):
    _check(lib().kt_init(device), "kt_init")


NOALLOC_MAX_ENTRIES = 1 << 23


def ktruss_noalloc(n, num_entries, ptr, src, dst, k, stream=0):
    """kt_ktruss_noalloc: one k of the directed COO+CSR peel on a device
    layout the caller holds (device addresses ptr, src, dst as integers, dst
    rewritten in place: entries outside the k-truss get bit 31), with no
    device allocation.  Returns the number of entries left."""
    live = ctypes.c_uint64(0)
    _check(lib().kt_ktruss_noalloc(n, num_entries, ptr, src, dst, k, stream, ctypes.byref(live)),
           "kt_ktruss_noalloc")
    return live.value


def device_synchronize():
    _check(lib().kt_device_synchronize(), "kt_device_synchronize")


COUNT_AUTO, COUNT_ROW, COUNT_SEARCH, COUNT_MERGE, COUNT_DELTA = 0, 1, 2, 3, 4
COUNT_KERNELS = {"auto": 0, "row": 1, "search": 2, "merge": 3, "delta": 4}


def make_variant(directedness="directed-by-degree", fmt="coocsr", recount="all", e_iters=0,
                 block_size=512, tiles=8, parallel_tiles=8, tile_size=0, chunk=0,
                 schedule=SCHED_JUMP | SCHED_REUSE, count_kernel=COUNT_AUTO, cap_lds=0):
    v = Variant()
    lib().kt_variant_default(ctypes.byref(v))
    v.directedness = DIRECTEDNESS[directedness] if isinstance(directedness, str) else directedness
    v.format = FORMATS[fmt] if isinstance(fmt, str) else fmt
    v.recount = RECOUNTS[recount] if isinstance(recount, str) else recount
    v.num_edge_remove_iter = e_iters
    v.block_size = block_size
    v.tiles_per_dim = tiles
    v.num_parallel_tiles = parallel_tiles
    v.tile_size = tile_size
    v.chunk = chunk
    v.schedule = schedule
    v.count_kernel = COUNT_KERNELS[count_kernel] if isinstance(count_kernel, str) else count_kernel
    v.cap_lds = cap_lds
    return v


def default_variant():
    return make_variant()


def all_variants(tiles=8, parallel_tiles=8):
    """The 12 combinations of the reference's design space (3 orientations x 2
    layouts x 2 recount policies)."""
    out = []
    for d in ("undirected", "directed-by-index", "directed-by-degree"):
        for f in ("coocsr", "tiled-coocsr"):
            for r in ("all", "affected"):
                out.append(make_variant(d, f, r, tiles=tiles, parallel_tiles=parallel_tiles))
    return out


class Graph:
    """Validated undirected graph.  `pairs=False`: symmetric COO (both
    directions, as read from a file); `pairs=True`: each edge once."""

    def __init__(self, n, src, dst, pairs=False):
        s, sp = _u32(src)
        d, dp = _u32(dst)
        if len(s) != len(d):
            raise ValueError("src and dst differ in length")
        h = ctypes.c_void_p()
        _check(lib().kt_graph_create(n, len(s), sp, dp, INPUT_PAIRS if pairs else INPUT_SYMMETRIC,
                                     ctypes.byref(h)), "kt_graph_create")
        self._h = h
        self.n = n
        self.src, self.dst = s, d
        nn = ctypes.c_uint32()
        mu = ctypes.c_uint64()
        mi = ctypes.c_uint64()
        md = ctypes.c_uint32()
        _check(lib().kt_graph_info(h, ctypes.byref(nn), ctypes.byref(mu), ctypes.byref(mi),
                                   ctypes.byref(md)), "kt_graph_info")
        self.m = mu.value
        self.m_input = mi.value
        self.max_degree = md.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.kt_graph_destroy(h)
            self._h = None


class Plan:
    """Device-resident layout of one variant of one graph."""

    def __init__(self, graph, variant=None):
        variant = variant or default_variant()
        h = ctypes.c_void_p()
        _check(lib().kt_plan_create(graph._h, ctypes.byref(variant), ctypes.byref(h)),
               "kt_plan_create")
        self._h = h
        self.graph = graph
        self.variant = variant
        e = ctypes.c_uint64()
        mr = ctypes.c_uint32()
        _check(lib().kt_plan_info(h, ctypes.byref(e), ctypes.byref(mr)), "kt_plan_info")
        self.entries = e.value
        self.max_row = mr.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.kt_plan_destroy(h)
            self._h = None

    def ktruss(self, k):
        """The k-truss as a symmetric COO in the graph's labels, plus stats."""
        cap = 2 * self.graph.m
        s = np.zeros(max(cap, 1), np.uint32)
        d = np.zeros(max(cap, 1), np.uint32)
        out_m = ctypes.c_uint64()
        st = Stats()
        _check(lib().kt_plan_ktruss(self._h, k, s.ctypes.data_as(_u32p), d.ctypes.data_as(_u32p),
                                    cap, ctypes.byref(out_m), ctypes.byref(st)), "kt_plan_ktruss")
        m = out_m.value
        return s[:m].copy(), d[:m].copy(), st.as_dict()

    def decompose(self):
        st = Stats()
        _check(lib().kt_plan_decompose(self._h, ctypes.byref(st)), "kt_plan_decompose")
        return st.as_dict()

    def truss_numbers(self):
        m = self.graph.m_input
        t = np.zeros(max(m, 1), np.uint32)
        _check(lib().kt_plan_truss_numbers(self._h, t.ctypes.data_as(_u32p), m),
               "kt_plan_truss_numbers")
        return t[:m].copy()

    def trace(self, with_time=False):
        """(k, iteration, edges deleted, entries recounted[, count ns]) per iteration."""
        n = ctypes.c_uint64()
        _check(lib().kt_plan_trace(self._h, None, 0, ctypes.byref(n)), "kt_plan_trace")
        buf = np.zeros(max(5 * n.value, 1), np.uint64)
        _check(lib().kt_plan_trace(self._h, buf.ctypes.data_as(_u64p), n.value, ctypes.byref(n)),
               "kt_plan_trace")
        w = 5 if with_time else 4
        return [tuple(int(x) for x in buf[5 * i:5 * i + w]) for i in range(n.value)]


    def set_candidates(self, configs):
        """configs: CountConfig list (at most 16) tried per phase by SWEEP / MODEL."""
        arr = (CountConfig * max(len(configs), 1))(*configs)
        _check(lib().kt_plan_set_candidates(self._h, arr, len(configs)), "kt_plan_set_candidates")
        self.candidates = list(configs)

    def set_timing(self, on=True):
        """hipEvents around every support pass (ms statistics and trace times); off
        saves the device the gaps the events cost, results unchanged."""
        _check(lib().kt_plan_set_timing(self._h, 1 if on else 0), "kt_plan_set_timing")

    def set_pick(self, mode, model=None):
        _check(lib().kt_plan_set_pick(self._h, mode, model._h if model is not None else None),
               "kt_plan_set_pick")

    def autotune(self, k=3, repeats=1):
        """kt_plan_autotune: applies and returns the fastest variant found."""
        best = Variant()
        _check(lib().kt_plan_autotune(self._h, k, repeats, ctypes.byref(best)), "kt_plan_autotune")
        return best

    def tune_log(self):
        n = ctypes.c_uint64()
        _check(lib().kt_plan_tune_log(self._h, None, 0, ctypes.byref(n)), "kt_plan_tune_log")
        recs = (TuneTrial * max(n.value, 1))()
        _check(lib().kt_plan_tune_log(self._h, recs, n.value, ctypes.byref(n)), "kt_plan_tune_log")
        return [recs[i].as_dict() for i in range(n.value)]

    def set_sweep_sampling(self, every, fill=0, offset=0):
        """SWEEP times every candidate only in the phases of index `offset`
        modulo `every`; the others run candidate `fill` alone."""
        _check(lib().kt_plan_set_sweep_sampling(self._h, every, offset, fill), "kt_plan_set_sweep_sampling")

    def phase_log(self):
        """(phases, 36) float32: k, iterations, chosen, phase count ms, 16 features,
        16 per-candidate count ms (SWEEP)."""
        n = ctypes.c_uint64()
        _check(lib().kt_plan_phase_log(self._h, None, 0, ctypes.byref(n)), "kt_plan_phase_log")
        buf = np.zeros((max(n.value, 1), PHASE_REC), np.float32)
        _check(lib().kt_plan_phase_log(self._h, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       n.value, ctypes.byref(n)), "kt_plan_phase_log")
        return buf[: n.value].copy()


class Batch:
    """Many graphs decomposed together on the device (directed COO+CSR,
    recount all): per-graph peel state on the device, peel steps replayed
    from a captured HIP graph."""

    def __init__(self, graphs, variant=None):
        variant = variant or default_variant()
        self.graphs = list(graphs)
        arr = (ctypes.c_void_p * max(len(self.graphs), 1))(*[g._h.value for g in self.graphs])
        h = ctypes.c_void_p()
        _check(lib().kt_batch_create(arr, len(self.graphs), ctypes.byref(variant), ctypes.byref(h)),
               "kt_batch_create")
        self._h = h
        self.variant = variant
        g = ctypes.c_uint32()
        e = ctypes.c_uint64()
        mr = ctypes.c_uint32()
        _check(lib().kt_batch_info(h, ctypes.byref(g), ctypes.byref(e), ctypes.byref(mr)),
               "kt_batch_info")
        self.entries = e.value
        self.max_row = mr.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.kt_batch_destroy(h)
            self._h = None

    def tuning(self):
        t = BatchTuning()
        _check(lib().kt_batch_get_tuning(self._h, ctypes.byref(t)), "kt_batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = BatchTuning(**kw)
        _check(lib().kt_batch_set_tuning(self._h, ctypes.byref(t)), "kt_batch_set_tuning")
        return self.tuning()

    def autotune(self, trial_steps=8, repeats=2):
        t = BatchTuning()
        _check(lib().kt_batch_autotune(self._h, trial_steps, repeats, ctypes.byref(t)),
               "kt_batch_autotune")
        return t.as_dict()

    def tune_log(self):
        n = ctypes.c_uint64()
        _check(lib().kt_batch_tune_log(self._h, None, None, 0, ctypes.byref(n)), "kt_batch_tune_log")
        cfg = (BatchTuning * max(n.value, 1))()
        ms = np.zeros(max(n.value, 1), np.float32)
        _check(lib().kt_batch_tune_log(self._h, cfg, ms.ctypes.data_as(_f32p), n.value,
                                       ctypes.byref(n)), "kt_batch_tune_log")
        return [(cfg[i].as_dict(), float(ms[i])) for i in range(n.value)]

    def set_timing(self, on=True):
        """Eager steps with every count launch timed (stats["ms_count"])."""
        _check(lib().kt_batch_set_timing(self._h, 1 if on else 0), "kt_batch_set_timing")

    def decompose(self):
        st = BatchStats()
        _check(lib().kt_batch_decompose(self._h, ctypes.byref(st)), "kt_batch_decompose")
        return st.as_dict()

    def truss_numbers(self, i):
        m = self.graphs[i].m_input
        t = np.zeros(max(m, 1), np.uint32)
        _check(lib().kt_batch_truss_numbers(self._h, i, t.ctypes.data_as(_u32p), m),
               "kt_batch_truss_numbers")
        return t[:m].copy()

    def graph_info(self, i):
        """(largest truss number, iterations, phases, count passes) of graph i."""
        info = np.zeros(4, np.uint32)
        _check(lib().kt_batch_graph_info(self._h, i, info.ctypes.data_as(_u32p)),
               "kt_batch_graph_info")
        return tuple(int(x) for x in info)


def count_config(kernel="row", block_size=512, cap_lds=0, chunk=0):
    return CountConfig(COUNT_KERNELS[kernel] if isinstance(kernel, str) else kernel, block_size,
                       cap_lds, chunk)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Mlp:
    """The count-kernel cost model: 16 -> 64 -> 64 -> n_out, FP32, on the GPU."""

    def __init__(self, n_out, seed=1):
        h = ctypes.c_void_p()
        _check(lib().kt_mlp_create(n_out, seed, ctypes.byref(h)), "kt_mlp_create")
        self._h = h
        self.n_out = n_out
        n = ctypes.c_uint32()
        _check(lib().kt_mlp_num_params(h, ctypes.byref(n)), "kt_mlp_num_params")
        self.num_params = n.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.kt_mlp_destroy(h)
            self._h = None

    def get_params(self):
        p = np.zeros(self.num_params, np.float32)
        _check(lib().kt_mlp_get_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       self.num_params), "kt_mlp_get_params")
        return p

    def set_params(self, p):
        p, pp = _f32(p)
        _check(lib().kt_mlp_set_params(self._h, pp, len(p)), "kt_mlp_set_params")

    def set_input_range(self, lo, hi):
        """Inference inputs are clamped to [lo, hi] per feature."""
        lo, lp = _f32(lo)
        hi, hp = _f32(hi)
        _check(lib().kt_mlp_set_input_range(self._h, lp, hp), "kt_mlp_set_input_range")

    def get_input_range(self):
        lo = np.zeros(MLP_IN, np.float32)
        hi = np.zeros(MLP_IN, np.float32)
        _check(lib().kt_mlp_get_input_range(self._h, lo.ctypes.data_as(_f32p), hi.ctypes.data_as(_f32p)),
               "kt_mlp_get_input_range")
        return lo, hi

    def forward(self, x):
        x, xp = _f32(x)
        b = x.shape[0]
        y = np.zeros((b, self.n_out), np.float32)
        _check(lib().kt_mlp_forward(self._h, xp, b, y.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "kt_mlp_forward")
        return y

    def loss_grad(self, x, t, mask=None, w=None):
        x, xp = _f32(x)
        t, tp = _f32(t)
        mp = _f32(mask)[1] if mask is not None else None
        m_keep = _f32(mask)[0] if mask is not None else None
        w_keep = _f32(w)[0] if w is not None else None
        wp = w_keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if w is not None else None
        if m_keep is not None:
            mp = m_keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        loss = ctypes.c_float()
        g = np.zeros(self.num_params, np.float32)
        _check(lib().kt_mlp_loss_grad(self._h, xp, tp, mp, wp, x.shape[0], ctypes.byref(loss),
                                      g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "kt_mlp_loss_grad")
        return float(loss.value), g

    def train(self, x, t, mask=None, w=None, epochs=50, batch=256, lr=1e-3, seed=1):
        x, xp = _f32(x)
        t, tp = _f32(t)
        m_keep = _f32(mask)[0] if mask is not None else None
        w_keep = _f32(w)[0] if w is not None else None
        mp = m_keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if m_keep is not None else None
        wp = w_keep.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if w_keep is not None else None
        loss = ctypes.c_float()
        _check(lib().kt_mlp_train(self._h, xp, tp, mp, wp, x.shape[0], epochs, batch, lr, seed,
                                  ctypes.byref(loss)), "kt_mlp_train")
        return float(loss.value)


def _gen_out():
    return ctypes.c_uint64(), _u32p(), _u32p()


def generate_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    """R-MAT edge pairs (u < v, sorted, cleaned).  Returns (n, src, dst)."""
    n = ctypes.c_uint32()
    m, sp, dp = _gen_out()
    _check(lib().kt_generate_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(n),
                                  ctypes.byref(m), ctypes.byref(sp), ctypes.byref(dp)),
           "kt_generate_rmat")
    return n.value, _take_array(sp, m.value), _take_array(dp, m.value)


def generate_chung_lu(n, gamma=2.1, avg_degree=16.0, max_weight=0.0, num_cliques=0,
                      clique_min=20, clique_max=200, seed=1):
    m, sp, dp = _gen_out()
    _check(lib().kt_generate_chung_lu(n, gamma, avg_degree, max_weight, num_cliques, clique_min,
                                      clique_max, seed, ctypes.byref(m), ctypes.byref(sp),
                                      ctypes.byref(dp)), "kt_generate_chung_lu")
    return n, _take_array(sp, m.value), _take_array(dp, m.value)


def generate_erdos_renyi(n, avg_degree=8.0, seed=1):
    m, sp, dp = _gen_out()
    _check(lib().kt_generate_erdos_renyi(n, avg_degree, seed, ctypes.byref(m), ctypes.byref(sp),
                                         ctypes.byref(dp)), "kt_generate_erdos_renyi")
    return n, _take_array(sp, m.value), _take_array(dp, m.value)


def symmetrize(src, dst):
    """Pairs -> symmetric COO (both directions), the reference's input shape."""
    s = np.concatenate([src, dst]).astype(np.uint32)
    d = np.concatenate([dst, src]).astype(np.uint32)
    return s, d


def read_tsv(path):
    n = ctypes.c_uint32()
    m, sp, dp = _gen_out()
    _check(lib().kt_read_tsv(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(sp),
                             ctypes.byref(dp)), "kt_read_tsv")
    return n.value, _take_array(sp, m.value), _take_array(dp, m.value)


def write_tsv(path, src, dst):
    s, sp = _u32(src)
    d, dp = _u32(dst)
    _check(lib().kt_write_tsv(path.encode(), len(s), sp, dp), "kt_write_tsv")


def read_truss_tsv(path):
    """(n, src, dst, tau) of a file written by write_truss_tsv (or any TSV:
    tau is then its weight column)."""
    n = ctypes.c_uint32()
    m, sp, dp = _gen_out()
    tp = _u32p()
    _check(lib().kt_read_truss_tsv(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(sp),
                                   ctypes.byref(dp), ctypes.byref(tp)), "kt_read_truss_tsv")
    return (n.value, _take_array(sp, m.value), _take_array(dp, m.value), _take_array(tp, m.value))


def write_truss_tsv(path, src, dst, tau):
    s, sp = _u32(src)
    d, dp = _u32(dst)
    t, tp = _u32(tau)
    if not (len(s) == len(d) == len(t)):
        raise ValueError("src, dst and tau differ in length")
    _check(lib().kt_write_truss_tsv(path.encode(), len(s), sp, dp, tp), "kt_write_truss_tsv")


def write_edges_bin(path, n, src, dst, pairs=False):
    s, sp = _u32(src)
    d, dp = _u32(dst)
    if len(s) != len(d):
        raise ValueError("src and dst differ in length")
    _check(lib().kt_write_edges_bin(path.encode(), n, len(s), sp, dp,
                                    INPUT_PAIRS if pairs else INPUT_SYMMETRIC), "kt_write_edges_bin")


def read_edges_bin(path):
    """(n, src, dst, pairs) of a binary edge cache."""
    n = ctypes.c_uint32()
    m, sp, dp = _gen_out()
    fl = ctypes.c_uint32()
    _check(lib().kt_read_edges_bin(path.encode(), ctypes.byref(n), ctypes.byref(m), ctypes.byref(sp),
                                   ctypes.byref(dp), ctypes.byref(fl)), "kt_read_edges_bin")
    return n.value, _take_array(sp, m.value), _take_array(dp, m.value), fl.value == INPUT_PAIRS


def is_edges_bin(path):
    return bool(lib().kt_is_edges_bin(path.encode()))


def verify_edges(src_a, dst_a, src_b, dst_b):
    a, ap = _u32(src_a)
    b, bp = _u32(dst_a)
    c, cp = _u32(src_b)
    d, dp = _u32(dst_b)
    rc = lib().kt_verify_edges(len(a), ap, bp, len(c), cp, dp)
    if rc < 0:
        _check(rc, "kt_verify_edges")
    return rc == 0
