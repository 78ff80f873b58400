"""gcx -- Python bindings (ctypes) for libgcx.so, the MI355X colouring library.

The library is C host code over hand-written HIP kernels for gfx950. This
module only marshals arguments: every ordering, GNN pass, colouring and check
runs in the library on the GPU. Importing works without a GPU (host-side
generators and file I/O are usable); GPU entry points raise GcxError with
status GCX_ERR_NODEV when no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GCX_LIB_PATH selects another build of the same library (the host
# sanitizer build used by the CPU tests)
LIB_PATH = os.environ.get("GCX_LIB_PATH") or os.path.join(_HERE, "libgcx.so")
CLI_PATH = os.path.join(_HERE, "coloring")
MODEL_DIR = os.path.normpath(os.path.join(_HERE, "..", "..", "models"))

OK, ERR_ARG, ERR_NOMEM, ERR_HIP, ERR_CAPACITY, ERR_IO, ERR_FORMAT, ERR_NODEV, ERR_INVALID, ERR_BUDGET = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)

FF, LF, LLF, SL, SLL, GNN = range(6)
FP32, BF16, BF16X3 = 0, 1, 2
CULB_REVERSE, CULB_LARGEST_FIRST, CULB_RANDOM, CULB_MIXED = 0, 1, 2, 3
CULB_NO_GRAPH = 0x100
SCHED_ROUNDS, SCHED_ORDERED, SCHED_AUTO = 0, 1, 2
HEURISTICS = {"FF": FF, "LF": LF, "LLF": LLF, "SL": SL, "SLL": SLL}
STAGES = ["keys", "sl", "gnn_in", "gnn_hid", "gnn_out", "jp_init", "jp_rounds"]

u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


class GcxError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {_lib.gcx_status_string(status).decode()} ({status})")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C gnn-coloring_amd` "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "gcx_status_string": (C.c_char_p, [C.c_int]),
        "gcx_version": (C.c_char_p, []),
        "gcx_init": (C.c_int, []),
        "gcx_shutdown": (C.c_int, []),
        "gcx_stream_create": (C.c_int, [C.POINTER(vp)]),
        "gcx_stream_destroy": (C.c_int, [vp]),
        "gcx_stream_sync": (C.c_int, [vp]),
        "gcx_device_sync": (C.c_int, []),
        "gcx_event_create": (C.c_int, [C.POINTER(vp)]),
        "gcx_event_destroy": (C.c_int, [vp]),
        "gcx_event_record": (C.c_int, [vp, vp]),
        "gcx_event_elapsed_ms": (C.c_int, [vp, vp, C.POINTER(C.c_float)]),
        "gcx_mem_info": (C.c_int, [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "gcx_graph_from_csr": (C.c_int, [C.c_uint32, u32p, u32p, vp, C.POINTER(vp)]),
        "gcx_graph_from_pairs": (C.c_int, [C.c_uint32, C.c_uint64, u32p, u32p, vp, C.POINTER(vp)]),
        "gcx_graph_rmat": (C.c_int, [C.c_int, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                     C.c_uint64, vp, C.POINTER(vp)]),
        "gcx_graph_er": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, vp, C.POINTER(vp)]),
        "gcx_graph_rgg": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, vp, C.POINTER(vp)]),
        "gcx_graph_free": (C.c_int, [vp]),
        "gcx_graph_num_vertices": (C.c_uint32, [vp]),
        "gcx_graph_num_entries": (C.c_uint64, [vp]),
        "gcx_graph_max_degree": (C.c_uint32, [vp]),
        "gcx_graph_download": (C.c_int, [vp, u32p, u32p]),
        "gcx_graph_check": (C.c_int, [vp, vp, i32p]),
        "gcx_graph_footprint": (C.c_uint64, [C.c_uint32, C.c_uint64]),
        "gcx_graph_pack": (C.c_int, [C.POINTER(vp), C.c_uint32, vp, C.POINTER(vp)]),
        "gcx_graph_packed_count": (C.c_uint32, [vp]),
        "gcx_graph_prepare": (C.c_int, [vp, vp]),
        "gcx_packed_colors": (C.c_int, [vp, vp, i32p, C.POINTER(C.c_uint64), vp]),
        "gcx_model_load": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "gcx_model_from_params": (C.c_int, [C.c_int, f32p, C.c_int, C.POINTER(vp)]),
        "gcx_model_free": (C.c_int, [vp]),
        "gcx_model_layers": (C.c_int, [vp]),
        "gcx_workspace_create": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(vp)]),
        "gcx_workspace_free": (C.c_int, [vp]),
        "gcx_workspace_footprint": (C.c_uint64, [C.c_uint32, C.c_uint64]),
        "gcx_order": (C.c_int, [vp, vp, C.c_int, vp, C.c_int, vp]),
        "gcx_set_priorities": (C.c_int, [vp, vp, f64p, vp]),
        "gcx_color": (C.c_int, [vp, vp, vp]),
        "gcx_recolor": (C.c_int, [vp, vp, C.c_int, vp]),
        "gcx_recolor_ex": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_uint64, vp]),
        "gcx_set_schedule": (C.c_int, [vp, C.c_int, C.c_uint32]),
        "gcx_set_ordered_head": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "gcx_last_head": (C.c_int, [vp, u32p]),
        "gcx_validate": (C.c_int, [vp, vp, vp, i32p]),
        "gcx_mex_check": (C.c_int, [vp, vp, vp, C.POINTER(C.c_uint64)]),
        "gcx_color_digest": (C.c_int, [vp, vp, vp, C.POINTER(C.c_uint64)]),
        "gcx_server_create": (C.c_int, [C.c_int, C.c_uint64, C.c_char_p, C.POINTER(vp)]),
        "gcx_server_run": (C.c_int, [vp, vp, C.c_size_t, vp]),
        "gcx_server_destroy": (C.c_int, [vp]),
        "gcx_request_footprint": (C.c_uint64, [vp]),
        "gcx_server_peak_reserved": (C.c_uint64, [vp]),
        "gcx_server_add_graph": (C.c_int, [vp, vp, u32p]),
        "gcx_server_set_batching": (C.c_int, [vp, C.c_int]),
        "gcx_server_set_express": (C.c_int, [vp, C.c_int, C.c_uint64]),
        "gcx_set_ordered_launch": (C.c_int, [C.c_int]),
        "gcx_degree_signature": (C.c_uint32, [vp]),
        "gcx_tune": (C.c_int, [vp, vp, C.c_int, vp, vp]),
        "gcx_tuning_lookup": (C.c_int, [vp, vp]),
        "gcx_tuner_reset": (None, []),
        "gcx_tuner_save": (C.c_int, [C.c_char_p]),
        "gcx_tuner_load": (C.c_int, [C.c_char_p]),
        "gcx_graph_set_hub_degree": (C.c_int, [vp, C.c_uint32, vp]),
        "gcx_graph_hub_degree": (C.c_uint32, [vp]),
        "gcx_get_ordered_launch": (C.c_int, []),
        "gcx_download_colors": (C.c_int, [vp, C.c_uint32, i32p, vp]),
        "gcx_upload_colors": (C.c_int, [vp, C.c_uint32, i32p, vp]),
        "gcx_download_priorities": (C.c_int, [vp, C.c_uint32, f64p, vp]),
        "gcx_last_rounds": (C.c_int, [vp, u32p]),
        "gcx_last_launches": (C.c_int, [vp, u32p]),
        "gcx_last_fallback": (C.c_int, [vp, u32p]),
        "gcx_profile_enable": (C.c_int, [C.c_int]),
        "gcx_calibration_copy": (C.c_int, [C.c_uint64, C.POINTER(C.c_float)]),
        "gcx_profile_reset": (C.c_int, []),
        "gcx_profile_read": (C.c_int, [f64p, C.POINTER(C.c_uint64)]),
        "gcx_rmat_thresholds": (None, [C.c_double, C.c_double, C.c_double, u32p]),
        "gcx_gen_rmat_pairs": (C.c_int, [C.c_int, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                         C.c_uint64, u32p, u32p]),
        "gcx_gen_er_pairs": (C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, u32p, u32p]),
        "gcx_gen_permutation": (C.c_int, [C.c_uint32, C.c_uint64, u32p]),
        "gcx_stencil27_pair_count": (C.c_uint64, [C.c_uint32]),
        "gcx_gen_stencil27_pairs": (C.c_int, [C.c_uint32, C.c_uint64, u32p, u32p]),
        "gcx_gen_rgg2d_pairs": (C.c_int, [C.c_uint32, C.c_double, C.c_uint64, C.c_uint64, u32p,
                                          u32p, C.POINTER(C.c_uint64)]),
        # reference API
        "jones_plassmann_coloring": (i32p, [C.c_uint32, u32p, u32p, f64p]),
        "greedy_coloring": (i32p, [C.c_uint32, u32p, u32p, f64p]),
        "jones_plassmann_setup": (vp, [C.c_uint32, C.c_uint32]),
        "jones_plassmann_cleanup": (None, [vp]),
        "jones_plassmann_internal": (None, [C.c_uint32, u32p, u32p, vp, f64p, i32p]),
        "jones_plassmann_internal_noalloc": (C.c_int, [C.c_uint32, u32p, u32p, f64p, i32p]),
        "validate_coloring_par": (C.c_int, [C.c_uint32, u32p, u32p, i32p]),
        "validate_coloring": (C.c_int, [C.c_uint32, u32p, u32p, i32p]),
        "cgnn_graph_sage_16_par": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f32p, f32p,
                                          f64p]),
        "cgnn_graph_sage_16": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f32p, f32p,
                                      f64p]),
        "gnn_ordering": (None, [C.c_uint32, u32p, u32p, f64p, C.c_int]),
        "first_fit_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "first_fit_ordering_par": (None, [C.c_uint32, u32p, u32p, f64p]),
        "largest_degree_first_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "largest_log_degree_first_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "smallest_degree_last_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "smallest_log_degree_last_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "incidence_degree_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "saturation_degree_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
        "saturation_degree_ordering_alt": (None, [C.c_uint32, u32p, u32p, f64p]),
        "first_fit_coloring_par": (i32p, [C.c_uint32, u32p, u32p]),
        "largest_degree_first_coloring_par": (i32p, [C.c_uint32, u32p, u32p]),
        "largest_log_degree_first_coloring_par": (i32p, [C.c_uint32, u32p, u32p]),
        "smallest_degree_last_coloring_par": (i32p, [C.c_uint32, u32p, u32p]),
        "smallest_log_degree_last_coloring_par": (i32p, [C.c_uint32, u32p, u32p]),
        "graph_sort_edges_par": (None, [C.c_uint32, u32p, u32p]),
        "graph_make_simple": (None, [C.c_uint32, u32p, C.POINTER(u32p)]),
        "graph_validate_par": (C.c_uint32, [C.c_uint32, u32p, u32p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = _load()
_libc = C.CDLL(None)
_libc.free.argtypes = [vp]
_libc.malloc.restype = vp
_libc.malloc.argtypes = [C.c_size_t]


def lib():
    return _lib


def _chk(rc: int, what: str):
    if rc != OK:
        raise GcxError(rc, what)


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(u32p)


def _take_malloced_ints(ptr, n):
    out = np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].copy()
    _libc.free(C.cast(ptr, vp))
    return out


def init():
    _chk(_lib.gcx_init(), "gcx_init")


def shutdown():
    """Wait for the device, then release the library's cached blocks, pinned
    buffers, streams and events (gcx_shutdown). Free graphs, workspaces,
    models and servers first."""
    _chk(_lib.gcx_shutdown(), "gcx_shutdown")


def gpu_available() -> bool:
    return _lib.gcx_init() == OK


def version() -> str:
    return _lib.gcx_version().decode()


# ------------------------------------------------------------------ streams

class Stream:
    def __init__(self):
        h = vp()
        _chk(_lib.gcx_stream_create(C.byref(h)), "stream")
        self.handle = h

    def sync(self):
        _chk(_lib.gcx_stream_sync(self.handle), "stream sync")

    def close(self):
        if self.handle:
            _lib.gcx_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = vp()
        _chk(_lib.gcx_event_create(C.byref(h)), "event")
        self.handle = h

    def record(self, stream=None):
        _chk(_lib.gcx_event_record(self.handle, stream.handle if stream else None), "event record")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        _chk(_lib.gcx_event_elapsed_ms(self.handle, later.handle, C.byref(ms)), "event elapsed")
        return float(ms.value)

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            _lib.gcx_event_destroy(self.handle)
        except Exception:
            pass


def _sh(stream):
    return stream.handle if stream is not None else None


class Tuning(C.Structure):
    """gcx_tuning (gcx.h)."""
    _fields_ = [("signature", C.c_uint32), ("hub_degree", C.c_uint32), ("schedule", C.c_int),
                ("order_ms", C.c_float), ("color_ms", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def tune(g, model=None, precision=FP32, stream=None) -> Tuning:
    """Autotune hub degree and JP schedule for g's degree signature."""
    t = Tuning()
    _chk(_lib.gcx_tune(g.handle, model.handle if model else None, precision, _sh(stream),
                       C.byref(t)), "tune")
    return t


def tuning_lookup(g):
    t = Tuning()
    return t if _lib.gcx_tuning_lookup(g.handle, C.byref(t)) else None


def tuner_reset():
    _lib.gcx_tuner_reset()


def tuner_save(path):
    _chk(_lib.gcx_tuner_save(str(path).encode()), "tuner save")


def tuner_load(path):
    _chk(_lib.gcx_tuner_load(str(path).encode()), "tuner load")


def set_ordered_launch(slots: int):
    """0: cooperative ORDERED launches, one at a time; k >= 1: plain
    launches sharing the device, at most k at once (gcx_set_ordered_launch)."""
    _chk(_lib.gcx_set_ordered_launch(slots), "set ordered launch")


def get_ordered_launch() -> int:
    return int(_lib.gcx_get_ordered_launch())


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    _chk(_lib.gcx_mem_info(C.byref(f), C.byref(t)), "mem info")
    return f.value, t.value


def device_sync():
    _chk(_lib.gcx_device_sync(), "device sync")


# ------------------------------------------------------------------- graphs

class Graph:
    """Device-resident canonical CSR (symmetric, sorted rows, simple)."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def from_csr(cls, V, E, stream=None):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        h = vp()
        _chk(_lib.gcx_graph_from_csr(len(V) - 1, Vp, Ep, _sh(stream), C.byref(h)), "graph from CSR")
        return cls(h)

    @classmethod
    def from_pairs(cls, N, I, J, stream=None):
        I, Ip = _u32(I)
        J, Jp = _u32(J)
        h = vp()
        _chk(_lib.gcx_graph_from_pairs(N, len(I), Ip, Jp, _sh(stream), C.byref(h)), "graph from pairs")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, stream=None):
        h = vp()
        _chk(_lib.gcx_graph_rmat(scale, edge_factor, a, b, c, seed, _sh(stream), C.byref(h)), "R-MAT graph")
        return cls(h)

    @classmethod
    def er(cls, n, m, seed=1, stream=None):
        h = vp()
        _chk(_lib.gcx_graph_er(n, m, seed, _sh(stream), C.byref(h)), "ER graph")
        return cls(h)

    @classmethod
    def rgg(cls, n, radius, seed=1, stream=None):
        h = vp()
        _chk(_lib.gcx_graph_rgg(n, radius, seed, _sh(stream), C.byref(h)), "RGG graph")
        return cls(h)

    @classmethod
    def pack(cls, graphs, stream=None):
        """One block-diagonal graph holding `graphs` in order (gcx_graph_pack)."""
        arr = (vp * len(graphs))(*[g.handle for g in graphs])
        h = vp()
        _chk(_lib.gcx_graph_pack(arr, len(graphs), _sh(stream), C.byref(h)), "graph pack")
        return cls(h)

    def prepare(self, stream=None):
        """Redo the one-time per-graph preparation (degree order, hub tables,
        remapped columns); for timing it."""
        _chk(_lib.gcx_graph_prepare(self.handle, _sh(stream)), "graph prepare")

    @property
    def hub_degree(self) -> int:
        return int(_lib.gcx_graph_hub_degree(self.handle))

    def set_hub_degree(self, d, stream=None):
        """Block-per-vertex from degree d up (0 = tuned value or default)."""
        _chk(_lib.gcx_graph_set_hub_degree(self.handle, d, _sh(stream)), "set hub degree")

    @property
    def signature(self) -> int:
        return int(_lib.gcx_degree_signature(self.handle))

    @property
    def packed_count(self) -> int:
        return int(_lib.gcx_graph_packed_count(self.handle))

    @property
    def N(self) -> int:
        return int(_lib.gcx_graph_num_vertices(self.handle))

    @property
    def M(self) -> int:
        return int(_lib.gcx_graph_num_entries(self.handle))

    @property
    def max_degree(self) -> int:
        return int(_lib.gcx_graph_max_degree(self.handle))

    def download(self):
        V = np.empty(self.N + 1, np.uint32)
        E = np.empty(max(self.M, 1), np.uint32)
        _chk(_lib.gcx_graph_download(self.handle, V.ctypes.data_as(u32p), E.ctypes.data_as(u32p)),
             "graph download")
        return V, E[: self.M]

    def check(self, stream=None) -> bool:
        v = C.c_int()
        _chk(_lib.gcx_graph_check(self.handle, _sh(stream), C.byref(v)), "graph check")
        return bool(v.value)

    def free(self):
        if self.handle:
            _lib.gcx_graph_free(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            self.free()
        except Exception:
            pass


def graph_footprint(N, M) -> int:
    return int(_lib.gcx_graph_footprint(N, M))


def workspace_footprint(N, M) -> int:
    return int(_lib.gcx_workspace_footprint(N, M))


# ------------------------------------------------------------------- models

class Model:
    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def load(cls, path):
        h = vp()
        _chk(_lib.gcx_model_load(str(path).encode(), C.byref(h)), f"model {path}")
        return cls(h)

    @classmethod
    def layers_file(cls, layers, model_dir=None):
        return cls.load(os.path.join(model_dir or MODEL_DIR, f"m{layers}.gnn"))

    @classmethod
    def from_params(cls, param, md=100):
        param = np.ascontiguousarray(param, np.float32)
        h = vp()
        _chk(_lib.gcx_model_from_params(len(param), param.ctypes.data_as(f32p), md, C.byref(h)),
             "model from params")
        return cls(h)

    @property
    def layers(self) -> int:
        return int(_lib.gcx_model_layers(self.handle))

    def free(self):
        if self.handle:
            _lib.gcx_model_free(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- workspace

class Workspace:
    """Scratch for ordering + colouring of graphs up to (N, M) vertices/entries."""

    def __init__(self, N, M=0):
        h = vp()
        _chk(_lib.gcx_workspace_create(N, M, C.byref(h)), "workspace")
        self.handle = h
        self.capN = N

    def order(self, g: Graph, heuristic, model: Model | None = None, precision=FP32, stream=None):
        if isinstance(heuristic, str):
            heuristic = HEURISTICS[heuristic]
        _chk(_lib.gcx_order(self.handle, g.handle, heuristic, model.handle if model else None,
                            precision, _sh(stream)), "order")

    def set_priorities(self, g: Graph, p, stream=None):
        p = np.ascontiguousarray(p, np.float64)
        _chk(_lib.gcx_set_priorities(self.handle, g.handle, p.ctypes.data_as(f64p), _sh(stream)),
             "set priorities")

    def color(self, g: Graph, stream=None):
        _chk(_lib.gcx_color(self.handle, g.handle, _sh(stream)), "color")

    def set_schedule(self, schedule, grid=0):
        """JP schedule: SCHED_ROUNDS (launch per DAG level) or SCHED_ORDERED
        (greedy-order items, one persistent launch of `grid` workgroups, 0 =
        all that are resident at once)."""
        _chk(_lib.gcx_set_schedule(self.handle, schedule, grid), "schedule")

    def recolor_ex(self, g: Graph, iterations, strategy, seed=0, stream=None):
        """Iterated greedy with a class order (CULB_REVERSE / LARGEST_FIRST /
        RANDOM / MIXED), each pass a captured HIP graph replay."""
        _chk(_lib.gcx_recolor_ex(self.handle, g.handle, iterations, strategy, seed, _sh(stream)),
             "recolor")

    def recolor(self, g: Graph, iterations, stream=None):
        _chk(_lib.gcx_recolor(self.handle, g.handle, iterations, _sh(stream)), "recolor")

    def validate(self, g: Graph, stream=None) -> int:
        c = C.c_int()
        _chk(_lib.gcx_validate(self.handle, g.handle, _sh(stream), C.byref(c)), "validate")
        return int(c.value)

    def mex_check(self, g: Graph, stream=None) -> int:
        c = C.c_uint64()
        _chk(_lib.gcx_mex_check(self.handle, g.handle, _sh(stream), C.byref(c)), "mex check")
        return int(c.value)

    def digest(self, g: Graph, stream=None) -> int:
        """Order-independent 64-bit digest of the colours (gcx_color_digest)."""
        c = C.c_uint64()
        _chk(_lib.gcx_color_digest(self.handle, g.handle, _sh(stream), C.byref(c)), "digest")
        return int(c.value)

    def packed_colors(self, g: Graph, stream=None):
        """Per packed graph: (colour count, digest of the graph alone)."""
        n = g.packed_count
        c = np.zeros(n, np.int32)
        d = np.zeros(n, np.uint64)
        _chk(_lib.gcx_packed_colors(self.handle, g.handle, c.ctypes.data_as(i32p),
                                    d.ctypes.data_as(C.POINTER(C.c_uint64)), _sh(stream)),
             "packed colours")
        return c, d

    def colors(self, N, stream=None):
        out = np.empty(N, np.int32)
        _chk(_lib.gcx_download_colors(self.handle, N, out.ctypes.data_as(i32p), _sh(stream)), "colors")
        return out

    def upload_colors(self, color, stream=None):
        color = np.ascontiguousarray(color, np.int32)
        _chk(_lib.gcx_upload_colors(self.handle, len(color), color.ctypes.data_as(i32p), _sh(stream)),
             "upload colors")

    def priorities(self, N, stream=None):
        out = np.empty(N, np.float64)
        _chk(_lib.gcx_download_priorities(self.handle, N, out.ctypes.data_as(f64p), _sh(stream)),
             "priorities")
        return out

    def rounds(self) -> int:
        r = C.c_uint32()
        _chk(_lib.gcx_last_rounds(self.handle, C.byref(r)), "rounds")
        return int(r.value)

    def launches(self) -> int:
        r = C.c_uint32()
        _chk(_lib.gcx_last_launches(self.handle, C.byref(r)), "launches")
        return int(r.value)

    def set_ordered_head(self, max_positions=32768, min_density=64):
        """Head of the ORDERED schedule (0 positions = off, the default);
        colours do not change."""
        _chk(_lib.gcx_set_ordered_head(self.handle, max_positions, min_density), "ordered head")

    def head(self) -> int:
        """Greedy-order positions the head coloured in the last color()."""
        r = C.c_uint32()
        _chk(_lib.gcx_last_head(self.handle, C.byref(r)), "last head")
        return int(r.value)

    def fell_back(self) -> bool:
        """True when the last colour() redid an ordered launch with rounds."""
        r = C.c_uint32()
        _chk(_lib.gcx_last_fallback(self.handle, C.byref(r)), "fallback")
        return bool(r.value)

    def free(self):
        if self.handle:
            _lib.gcx_workspace_free(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- profiling

def calibration_copy(nbytes):
    """Run the streaming copy kernel once over `nbytes`; returns milliseconds."""
    ms = C.c_float()
    _chk(_lib.gcx_calibration_copy(nbytes, C.byref(ms)), "calibration copy")
    return float(ms.value)


def profile_enable(on=True):
    _chk(_lib.gcx_profile_enable(1 if on else 0), "profile")


def profile_read():
    ms = (C.c_double * len(STAGES))()
    n = (C.c_uint64 * len(STAGES))()
    _chk(_lib.gcx_profile_read(ms, n), "profile read")
    return {s: (ms[i], int(n[i])) for i, s in enumerate(STAGES)}


# --------------------------------------------------------- host generators

RMAT_ABC = (0.57, 0.19, 0.19)


def rmat_pairs(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    L = edge_factor << scale
    I = np.empty(L, np.uint32)
    J = np.empty(L, np.uint32)
    _chk(_lib.gcx_gen_rmat_pairs(scale, edge_factor, a, b, c, seed, I.ctypes.data_as(u32p),
                                 J.ctypes.data_as(u32p)), "R-MAT pairs")
    return 1 << scale, I, J


def er_pairs(n, m, seed=1):
    I = np.empty(m, np.uint32)
    J = np.empty(m, np.uint32)
    _chk(_lib.gcx_gen_er_pairs(n, m, seed, I.ctypes.data_as(u32p), J.ctypes.data_as(u32p)), "ER pairs")
    return n, I, J


def permutation(n, seed=1):
    p = np.empty(n, np.uint32)
    _chk(_lib.gcx_gen_permutation(n, seed, p.ctypes.data_as(u32p)), "permutation")
    return p


def stencil27_pairs(g, seed=1):
    L = int(_lib.gcx_stencil27_pair_count(g))
    I = np.empty(L, np.uint32)
    J = np.empty(L, np.uint32)
    _chk(_lib.gcx_gen_stencil27_pairs(g, seed, I.ctypes.data_as(u32p), J.ctypes.data_as(u32p)),
         "stencil pairs")
    return g * g * g, I, J


def rgg2d_pairs(n, radius, seed=1, cap=None):
    cap = cap or max(16, int(n * n * 3.3 * radius * radius) + 1024)
    I = np.empty(cap, np.uint32)
    J = np.empty(cap, np.uint32)
    L = C.c_uint64()
    _chk(_lib.gcx_gen_rgg2d_pairs(n, radius, seed, cap, I.ctypes.data_as(u32p), J.ctypes.data_as(u32p),
                                  C.byref(L)), "RGG pairs")
    return n, I[: L.value].copy(), J[: L.value].copy()


# ------------------------------------------------- reference API (host arrays)

class ref:
    """The reference's C API (coloring.h, fast_gnn.h, gnn.h, ordering
    headers, fast_io.h) called through the drop-in library."""

    @staticmethod
    def jones_plassmann_coloring(V, E, p):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        p = np.ascontiguousarray(p, np.float64)
        N = len(V) - 1
        return _take_malloced_ints(_lib.jones_plassmann_coloring(N, Vp, Ep, p.ctypes.data_as(f64p)), N)

    @staticmethod
    def jones_plassmann_internal(V, E, p):
        """jones_plassmann_setup + jones_plassmann_internal + cleanup."""
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        p = np.ascontiguousarray(p, np.float64)
        N = len(V) - 1
        color = np.zeros(max(N, 1), np.int32)
        h = _lib.jones_plassmann_setup(N, int(V[N]) if N else 0)
        _lib.jones_plassmann_internal(N, Vp, Ep, h, p.ctypes.data_as(f64p), color.ctypes.data_as(i32p))
        _lib.jones_plassmann_cleanup(h)
        return color[:N]

    @staticmethod
    def jones_plassmann_internal_noalloc(V, E, p):
        """The no-allocation variant: returns (status, colours)."""
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        p = np.ascontiguousarray(p, np.float64)
        N = len(V) - 1
        color = np.zeros(max(N, 1), np.int32)
        rc = _lib.jones_plassmann_internal_noalloc(N, Vp, Ep, p.ctypes.data_as(f64p),
                                                   color.ctypes.data_as(i32p))
        return int(rc), color[:N]

    @staticmethod
    def greedy_coloring(V, E, p):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        p = np.ascontiguousarray(p, np.float64)
        N = len(V) - 1
        return _take_malloced_ints(_lib.greedy_coloring(N, Vp, Ep, p.ctypes.data_as(f64p)), N)

    @staticmethod
    def validate_coloring_par(V, E, color):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        color = np.ascontiguousarray(color, np.int32)
        return int(_lib.validate_coloring_par(len(V) - 1, Vp, Ep, color.ctypes.data_as(i32p)))

    @staticmethod
    def cgnn_graph_sage_16_par(param, V, E, md=100):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        param = np.ascontiguousarray(param, np.float32)
        N = len(V) - 1
        p = np.zeros(N + 8, np.float64)
        _lib.cgnn_graph_sage_16_par(len(param), param.ctypes.data_as(f32p), N, Vp, Ep, md, None, None,
                                    p.ctypes.data_as(f64p))
        return p[:N]

    @staticmethod
    def ordering(name, V, E, layers=None):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        N = len(V) - 1
        p = np.zeros(N + 8, np.float64)
        pp = p.ctypes.data_as(f64p)
        fn = {"FF": "first_fit_ordering_par", "LF": "largest_degree_first_ordering",
              "LLF": "largest_log_degree_first_ordering", "SL": "smallest_degree_last_ordering",
              "SLL": "smallest_log_degree_last_ordering",
              # sequential heap walks on the host CPU (incidence_degree.h)
              "ID": "incidence_degree_ordering", "SD": "saturation_degree_ordering",
              "SDA": "saturation_degree_ordering_alt"}
        if name.startswith("GNN"):
            _lib.gnn_ordering(N, Vp, Ep, pp, int(name[3:]) if layers is None else layers)
        else:
            getattr(_lib, fn[name])(N, Vp, Ep, pp)
        return p[:N]

    @staticmethod
    def coloring_par(name, V, E):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        N = len(V) - 1
        fn = {"FF": "first_fit_coloring_par", "LF": "largest_degree_first_coloring_par",
              "LLF": "largest_log_degree_first_coloring_par", "SL": "smallest_degree_last_coloring_par",
              "SLL": "smallest_log_degree_last_coloring_par"}[name]
        return _take_malloced_ints(getattr(_lib, fn)(N, Vp, Ep), N)

    @staticmethod
    def graph_sort_edges_par(V, E):
        V, Vp = _u32(V)
        E = np.array(E, dtype=np.uint32, copy=True)
        _lib.graph_sort_edges_par(len(V) - 1, Vp, E.ctypes.data_as(u32p))
        return E

    @staticmethod
    def graph_make_simple(V, E):
        V = np.array(V, dtype=np.uint32, copy=True)
        N = len(V) - 1
        # The library may realloc E: hand it a malloc'ed copy.
        raw = _libc.malloc(max(len(E), 1) * 4)
        if len(E):
            C.memmove(raw, np.ascontiguousarray(E, np.uint32).ctypes.data, len(E) * 4)
        box = C.pointer(C.cast(raw, u32p))
        _lib.graph_make_simple(N, V.ctypes.data_as(u32p), box)
        M = int(V[N])
        out = np.ctypeslib.as_array(box.contents, (max(M, 1),))[:M].copy() if M else np.empty(0, np.uint32)
        _libc.free(C.cast(box.contents, vp))
        return V, out

    @staticmethod
    def graph_validate_par(V, E):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        return int(_lib.graph_validate_par(len(V) - 1, Vp, Ep))


# ---------------------------------------------------------------- serving

GEN_RMAT, GEN_ER, GEN_RGG, GEN_RESIDENT = 0, 1, 2, 3


class Request(C.Structure):
    """gcx_request (gcx_serve.h)."""
    _fields_ = [("generator", C.c_int), ("n", C.c_uint32), ("param", C.c_double),
                ("seed", C.c_uint64), ("heuristic", C.c_int), ("layers", C.c_int),
                ("precision", C.c_int), ("schedule", C.c_int), ("arrival", C.c_double)]


class Reply(C.Structure):
    """gcx_reply (gcx_serve.h)."""
    _fields_ = [("status", C.c_int), ("batch", C.c_int), ("N", C.c_uint32), ("M", C.c_uint64),
                ("colors", C.c_int),
                ("digest", C.c_uint64), ("footprint", C.c_uint64), ("worker", C.c_int),
                ("t_arrival", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def make_request(generator, n, param=0.0, seed=1, heuristic=FF, layers=3, precision=FP32,
                 schedule=SCHED_ORDERED, arrival=0.0) -> Request:
    if isinstance(heuristic, str):
        if heuristic.startswith("GNN"):
            layers, heuristic = int(heuristic[3:]), GNN
        else:
            heuristic = HEURISTICS[heuristic]
    return Request(generator, n, param, seed, heuristic, layers, precision, schedule, arrival)


def request_footprint(req: Request) -> int:
    return int(_lib.gcx_request_footprint(C.byref(req)))


def serving_trace(k, seed=1, nmin=1000, nmax=4_000_000, rate=None, precision=BF16,
                  ordered_from=1 << 15):
    """Seeded request mix: 40% R-MAT (edge factor 16), 30% Erdos-Renyi
    (average degree uniform in [8, 64]), 30% 2D random geometric (about 12
    neighbours); n log-uniform in [nmin, nmax]; heuristic uniform over FF, LF,
    SL, GNN2, GNN3; GNN in `precision`. Arrivals are Poisson at `rate` per
    second (None = all at t = 0). Graphs with n >= ordered_from use the
    ordered JP schedule, smaller ones the round schedule."""
    import numpy as np
    rng = np.random.default_rng(seed)
    heur = ["FF", "LF", "SL", "GNN2", "GNN3"]
    reqs, t = [], 0.0
    for i in range(k):
        u = rng.random()
        gen = GEN_RMAT if u < 0.4 else (GEN_ER if u < 0.7 else GEN_RGG)
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        param = {GEN_RMAT: 16.0, GEN_ER: float(rng.uniform(8, 64)), GEN_RGG: 0.0}[gen]
        h = heur[int(rng.integers(0, len(heur)))]
        if rate:
            t += float(rng.exponential(1.0 / rate))
        sched = SCHED_ORDERED if n >= ordered_from else SCHED_ROUNDS
        reqs.append(make_request(gen, n, param, seed=int(rng.integers(1, 1 << 62)), heuristic=h,
                                 precision=precision, schedule=sched, arrival=t))
    return reqs


def with_arrivals(reqs, rate, seed=2, precision=None):
    """Copy of reqs with Poisson arrivals at `rate` per second (None = all
    at t = 0), optionally with another GNN precision."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for q in reqs:
        if rate:
            t += float(rng.exponential(1.0 / rate))
        out.append(Request(q.generator, q.n, q.param, q.seed, q.heuristic, q.layers,
                           q.precision if precision is None else precision, q.schedule, t))
    return out


class Server:
    """Request loop: `workers` threads/streams, HBM admission budget in bytes
    (0 = 3/4 of free HBM), models m2/m3/m4 loaded once from model_dir."""

    def __init__(self, workers=4, budget=0, model_dir=None):
        h = vp()
        md = (model_dir or MODEL_DIR).encode()
        _chk(_lib.gcx_server_create(workers, budget, md, C.byref(h)), "server create")
        self.handle = h
        self._resident = []

    def add_graph(self, g: Graph) -> int:
        """Register a resident graph; requests name it as
        make_request(GEN_RESIDENT, id, ...). Keep g alive while the server is."""
        gid = C.c_uint32()
        _chk(_lib.gcx_server_add_graph(self.handle, g.handle, C.byref(gid)), "server add graph")
        self._resident.append(g)
        return int(gid.value)

    def set_batching(self, max_batch: int):
        """Pack up to max_batch compatible queued resident-graph jobs into one
        pass (0 or 1 = off)."""
        _chk(_lib.gcx_server_set_batching(self.handle, max_batch), "server batching")

    def set_express(self, workers: int, max_footprint: int):
        """Reserve workers 0 .. workers-1 for jobs of at most max_footprint
        bytes, taken in arrival order ahead of larger queued jobs (0 = off)."""
        _chk(_lib.gcx_server_set_express(self.handle, workers, max_footprint), "server express")

    def run(self, requests):
        """Run requests (sorted by arrival); returns the list of Reply."""
        n = len(requests)
        arr = (Request * max(n, 1))(*requests)
        rep = (Reply * max(n, 1))()
        _chk(_lib.gcx_server_run(self.handle, arr, n, rep), "server run")
        return [rep[i] for i in range(n)]

    def peak_reserved(self) -> int:
        return int(_lib.gcx_server_peak_reserved(self.handle))

    def close(self):
        if self.handle:
            _lib.gcx_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if sys.is_finalizing():  # HIP may already be torn down
            return
        try:
            self.close()
        except Exception:
            pass
