"""CPU oracle for the ordering + colouring path (ctypes over liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared with.
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package; the product library never does.

Each function restates the reference file:line named in oracle/oracle.h.
`ref` (below) binds the reference's own sources compiled into
oracle/_ref/libref.so when that library was built (needs /root/reference at
build time); it is used to produce and pin golden vectors.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref.so")

u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


def _load():
    if not os.path.exists(ORACLE_SO):
        raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle oracle`")
    lib = C.CDLL(ORACLE_SO)
    sig = {
        "orc_validate": (C.c_int, [C.c_uint32, u32p, u32p, i32p]),
        "orc_greedy": (None, [C.c_uint32, u32p, u32p, f64p, i32p, u32p]),
        "orc_mex_violations": (C.c_uint64, [C.c_uint32, u32p, u32p, f64p, i32p]),
        "orc_key_ff": (None, [C.c_uint32, f64p]),
        "orc_key_lf": (None, [C.c_uint32, u32p, f64p]),
        "orc_key_llf": (None, [C.c_uint32, u32p, f64p]),
        "orc_sl": (None, [C.c_uint32, u32p, u32p, f64p, u32p]),
        "orc_sll": (None, [C.c_uint32, u32p, u32p, f64p, u32p]),
        "orc_load_model": (C.c_int, [C.c_char_p, i32p, i32p, C.POINTER(f32p)]),
        "orc_gnn_fp32": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f64p]),
        "orc_gnn_fp64": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f64p]),
        "orc_csr_from_pairs": (None, [C.c_uint32, C.c_uint64, u32p, u32p, C.POINTER(u32p),
                                      C.POINTER(u32p)]),
        "orc_graph_validate": (C.c_uint32, [C.c_uint32, u32p, u32p]),
        "orc_culberson": (None, [C.c_uint32, u32p, u32p, C.c_int, f64p, i32p, i32p]),
        "orc_free": (None, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_o = _load()


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(u32p)


def csr_from_pairs(N, I, J):
    """fast_io.c:171-202 + 450-477: canonical symmetric CSR of a pair list."""
    I, Ip = _u32(I)
    J, Jp = _u32(J)
    Vp, Ep = u32p(), u32p()
    _o.orc_csr_from_pairs(N, len(I), Ip, Jp, C.byref(Vp), C.byref(Ep))
    V = np.ctypeslib.as_array(Vp, (N + 1,)).copy()
    M = int(V[N])
    E = np.ctypeslib.as_array(Ep, (max(M, 1),))[:M].copy()
    _o.orc_free(C.cast(Vp, C.c_void_p))
    _o.orc_free(C.cast(Ep, C.c_void_p))
    return V, E


def graph_validate(V, E):
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    return int(_o.orc_graph_validate(len(V) - 1, Vp, Ep))


def greedy(V, E, p):
    """coloring.c:60-125: returns (colours, rounds)."""
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    p = np.ascontiguousarray(p, np.float64)
    N = len(V) - 1
    color = np.zeros(max(N, 1), np.int32)
    r = C.c_uint32()
    _o.orc_greedy(N, Vp, Ep, p.ctypes.data_as(f64p), color.ctypes.data_as(i32p), C.byref(r))
    return color[:N], int(r.value)


def validate(V, E, color):
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    color = np.ascontiguousarray(color, np.int32)
    return int(_o.orc_validate(len(V) - 1, Vp, Ep, color.ctypes.data_as(i32p)))


def mex_violations(V, E, p, color):
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    p = np.ascontiguousarray(p, np.float64)
    color = np.ascontiguousarray(color, np.int32)
    return int(_o.orc_mex_violations(len(V) - 1, Vp, Ep, p.ctypes.data_as(f64p),
                                     color.ctypes.data_as(i32p)))


def key(name, V, E):
    """Priorities p (double) of FF / LF / LLF / SL / SLL; returns (p, rounds)."""
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    N = len(V) - 1
    p = np.zeros(max(N, 1), np.float64)
    pp = p.ctypes.data_as(f64p)
    r = C.c_uint32()
    if name == "FF":
        _o.orc_key_ff(N, pp)
    elif name == "LF":
        _o.orc_key_lf(N, Vp, pp)
    elif name == "LLF":
        _o.orc_key_llf(N, Vp, pp)
    elif name == "SL":
        _o.orc_sl(N, Vp, Ep, pp, C.byref(r))
    elif name == "SLL":
        _o.orc_sll(N, Vp, Ep, pp, C.byref(r))
    else:
        raise ValueError(name)
    return p[:N], int(r.value)


def load_model(path):
    """gnn.c:20-44: returns (params float32 array, md)."""
    params, md, w = C.c_int(), C.c_int(), f32p()
    rc = _o.orc_load_model(str(path).encode(), C.byref(params), C.byref(md), C.byref(w))
    if rc != 0:
        raise IOError(f"model {path}: status {rc}")
    arr = np.ctypeslib.as_array(w, (params.value,)).copy()
    _o.orc_free(C.cast(w, C.c_void_p))
    return arr, md.value


def gnn_fp32(param, V, E, md=100):
    """fast_gnn.c:252-347 FP32 priorities (bit-exact op order)."""
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    param = np.ascontiguousarray(param, np.float32)
    N = len(V) - 1
    p = np.zeros(max(N, 1), np.float64)
    _o.orc_gnn_fp32(len(param), param.ctypes.data_as(f32p), N, Vp, Ep, md, p.ctypes.data_as(f64p))
    return p[:N]


def gnn_fp64(param, V, E, md=100):
    """Same network evaluated in double precision."""
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    param = np.ascontiguousarray(param, np.float32)
    N = len(V) - 1
    p = np.zeros(max(N, 1), np.float64)
    _o.orc_gnn_fp64(len(param), param.ctypes.data_as(f32p), N, Vp, Ep, md, p.ctypes.data_as(f64p))
    return p[:N]


def culberson(V, E, p, K):
    """main.c:32-45: returns (final colours, colour count after each pass)."""
    V, Vp = _u32(V)
    E, Ep = _u32(E)
    p = np.array(p, np.float64, copy=True)
    N = len(V) - 1
    color = np.zeros(max(N, 1), np.int32)
    counts = np.zeros(K, np.int32)
    _o.orc_culberson(N, Vp, Ep, K, p.ctypes.data_as(f64p), color.ctypes.data_as(i32p),
                     counts.ctypes.data_as(i32p))
    return color[:N], counts


def mix64(z):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def class_keys(color, strategy, seed, it):
    """Culberson class orders (no reference counterpart beyond strategy 0,
    main.c:38-40): key per vertex, larger = coloured earlier.
    0 reverse: key = colour; 1 largest class first (ties: higher colour
    first); 2 random: (hi32(mix64((seed + it) ^ c)) << b) | c, b = 12 for
    C < 4096; 3 mixed: strategy it mod 3."""
    color = np.asarray(color, np.int64)
    C = int(color.max(initial=0))
    st = it % 3 if strategy == 3 else strategy
    c = np.arange(C + 1, dtype=np.int64)
    if st == 0:
        ck = c
    elif st == 1:
        size = np.bincount(color, minlength=C + 1)
        order = sorted(range(1, C + 1), key=lambda d: (-size[d], -d))
        ck = np.zeros(C + 1, np.int64)
        for r, d in enumerate(order):
            ck[d] = C - r
    else:
        b = 12 if C < 4096 else int(C).bit_length()
        h = mix64((np.uint64(seed + it) ^ c.astype(np.uint64)).astype(np.uint64))
        hi = (h >> np.uint64(32)).astype(np.uint64)
        ck = (((hi << np.uint64(b)) & np.uint64(0xFFFFFFFF)) | c.astype(np.uint64)).astype(np.int64)
    return ck[color].astype(np.float64)


def culberson_classes(V, E, color, K, strategy, seed=0):
    """K iterated-greedy passes from `color` with a class-order strategy;
    returns (final colours, colour count after each pass)."""
    counts = []
    for it in range(K):
        color, _ = greedy(V, E, class_keys(color, strategy, seed, it))
        counts.append(int(color.max(initial=0)))
    return color, counts


def bf16_bits(x):
    """float32 -> bfloat16 bit patterns (uint16), the library's rule
    (gnn_bf16.hip / bf16.c to_bf16): finite values round to nearest, ties to
    even (values past the largest bf16 round to Inf); Inf stays Inf; a NaN
    keeps its sign and top payload bits and gets the quiet bit 0x40 when any
    dropped payload bit was set, so it never turns into Inf."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    special = (b & 0x7F800000) == 0x7F800000
    rounded = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    kept = (b >> 16) | np.where((b & 0xFFFF) != 0, 0x40, 0).astype(np.uint64)
    return np.where(special, kept, rounded).astype(np.uint16)


def bf16_round(x):
    """Round float32 values to bfloat16 (bf16_bits), returned as float32."""
    return (bf16_bits(x).astype(np.uint32) << np.uint32(16)).view(np.float32)


def gnn_bf16_emulated(param, V, E, md=100, rounding=True, split=False):
    """The BF16 GraphSAGE contract evaluated on the CPU: operands of each
    dense step ([mean | self] and W) rounded to bf16, products accumulated
    exactly (float64 here), FP32 bias, ReLU, next-layer rows rounded to bf16,
    output = sum of the 16 features. With split=True the BF16X3 contract:
    FP32 rows, every operand a and weight w split into bf16 parts
    hi = bf16(a), lo = bf16(a - hi), products hi*w_hi + hi*w_lo + lo*w_hi.
    The reference has no BF16 path, so this is pinned by `rounding=False`,
    which must equal the FP64 network."""
    import scipy.sparse as sp

    def f32(a):
        return np.asarray(a, np.float32).astype(np.float64)

    def rnd(a):  # float32 value rounded to bf16, as float64
        return bf16_round(np.asarray(a, np.float32)).astype(np.float64)

    def parts(a):  # (hi, lo) of the FP32 value of a
        a32 = np.asarray(a, np.float32)
        hi = bf16_round(a32)
        lo = bf16_round((a32 - hi).astype(np.float32))
        return hi.astype(np.float64), lo.astype(np.float64)

    if not rounding:
        def f32(a):  # noqa: F811
            return np.asarray(a, np.float64)
        rnd = f32
    V = np.asarray(V, np.int64)
    E = np.asarray(E, np.int64)
    N = len(V) - 1
    deg = np.diff(V)
    A = sp.csr_matrix((np.ones(len(E)), E, V), shape=(N, N))
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    w = np.asarray(param, np.float32).reshape(-1, 528)
    fdeg = f32(deg.astype(np.float64) / md)
    fidv = f32(np.arange(N, dtype=np.float64) / N)
    x = None
    for l in range(w.shape[0]):
        b = w[l, 512:].astype(np.float64)
        if l == 0:
            sd = f32((A @ fdeg) * inv)
            si = f32((A @ fidv) * inv)
            Aop = np.stack([sd] * 8 + [si] * 8 + [fdeg] * 8 + [fidv] * 8, 1)
        else:
            Aop = np.concatenate([f32((A @ x) * inv[:, None]), x], 1)
        if split and rounding:
            ah, al = parts(Aop)
            wh, wl = parts(w[l, :512])
            wh, wl = wh.reshape(32, 16), wl.reshape(32, 16)
            H = np.maximum(ah @ wh + ah @ wl + al @ wh + b, 0.0)
        else:
            H = np.maximum(rnd(Aop) @ rnd(w[l, :512]).reshape(32, 16) + b, 0.0)
        if l < w.shape[0] - 1:
            x = f32(H) if split else rnd(H)
        else:
            return H.sum(1)
    return None


# ------------------------------------------------------------------ reference

class Ref:
    """The reference's own C sources compiled into oracle/_ref/libref.so."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise ImportError(f"{path} not built (needs the reference tree at build time)")
        lib = C.CDLL(path)
        sig = {
            "jones_plassmann_coloring": (i32p, [C.c_uint32, u32p, u32p, f64p]),
            "greedy_coloring": (i32p, [C.c_uint32, u32p, u32p, f64p]),
            "validate_coloring_par": (C.c_int, [C.c_uint32, u32p, u32p, i32p]),
            "cgnn_graph_sage_16_par": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f32p,
                                              f32p, f64p]),
            "cgnn_graph_sage_16": (None, [C.c_int, f32p, C.c_uint32, u32p, u32p, C.c_int, f32p,
                                          f32p, f64p]),
            "first_fit_ordering_par": (None, [C.c_uint32, u32p, u32p, f64p]),
            "largest_degree_first_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "largest_log_degree_first_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "smallest_degree_last_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "smallest_log_degree_last_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "incidence_degree_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "saturation_degree_ordering": (None, [C.c_uint32, u32p, u32p, f64p]),
            "saturation_degree_ordering_alt": (None, [C.c_uint32, u32p, u32p, f64p]),
            "gnn_setup": (C.c_void_p, [C.c_uint32, C.c_uint32]),
            "gnn_cleanup": (None, [C.c_void_p]),
            "graph_sort_edges_par": (None, [C.c_uint32, u32p, u32p]),
            "graph_make_simple": (None, [C.c_uint32, u32p, C.POINTER(u32p)]),
            "graph_validate_par": (C.c_uint32, [C.c_uint32, u32p, u32p]),
            "mtx_parse_par": (None, [C.c_void_p, u32p, C.POINTER(u32p), C.POINTER(u32p)]),
            "mtx_store": (None, [C.c_void_p, C.c_uint32, u32p, u32p]),
            "jones_plassmann_setup": (C.c_void_p, [C.c_uint32, C.c_uint32]),
            "jones_plassmann_cleanup": (None, [C.c_void_p]),
            "jones_plassmann_internal": (None, [C.c_uint32, u32p, u32p, C.c_void_p, f64p, i32p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self.lib = lib
        self.libc = C.CDLL(None)
        self.libc.free.argtypes = [C.c_void_p]
        self.libc.fopen.restype = C.c_void_p
        self.libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        self.libc.fclose.argtypes = [C.c_void_p]

    def _ints(self, ptr, n):
        out = np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].copy()
        self.libc.free(C.cast(ptr, C.c_void_p))
        return out

    def jp(self, V, E, p):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        p = np.ascontiguousarray(p, np.float64)
        N = len(V) - 1
        return self._ints(self.lib.jones_plassmann_coloring(N, Vp, Ep, p.ctypes.data_as(f64p)), N)

    def validate(self, V, E, color):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        color = np.ascontiguousarray(color, np.int32)
        return int(self.lib.validate_coloring_par(len(V) - 1, Vp, Ep, color.ctypes.data_as(i32p)))

    def ordering(self, name, V, E):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        N = len(V) - 1
        p = np.zeros(N + 8, np.float64)
        fn = {"FF": "first_fit_ordering_par", "LF": "largest_degree_first_ordering",
              "LLF": "largest_log_degree_first_ordering", "SL": "smallest_degree_last_ordering",
              "SLL": "smallest_log_degree_last_ordering", "ID": "incidence_degree_ordering",
              "SD": "saturation_degree_ordering", "SDA": "saturation_degree_ordering_alt"}[name]
        getattr(self.lib, fn)(N, Vp, Ep, p.ctypes.data_as(f64p))
        return p[:N]

    def gnn(self, param, V, E, md=100):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        # The reference loads weights with aligned AVX loads (gnn.c:34 uses
        # aligned_alloc(32)): hand it a 32-byte aligned copy.
        raw = np.zeros(len(param) + 16, np.float32)
        po = (-raw.ctypes.data % 32) // 4
        param_a = raw[po:po + len(param)]
        param_a[:] = np.asarray(param, np.float32)
        param = param_a
        N = len(V) - 1
        x = np.zeros((N + 8) * 16 + 64, np.float32)
        y = np.zeros((N + 8) * 16 + 64, np.float32)
        # 32-byte aligned views, as gnn.c:36-37 allocates them
        xo = (-x.ctypes.data % 32) // 4
        yo = (-y.ctypes.data % 32) // 4
        xa, ya = x[xo:], y[yo:]
        p = np.zeros(N + 8, np.float64)
        self.lib.cgnn_graph_sage_16_par(len(param), param.ctypes.data_as(f32p), N, Vp, Ep, md,
                                        xa.ctypes.data_as(f32p), ya.ctypes.data_as(f32p),
                                        p.ctypes.data_as(f64p))
        return p[:N]

    @staticmethod
    def threads():
        gomp = C.CDLL("libgomp.so.1")
        return int(gomp.omp_get_max_threads())

    def time_gnn_jp(self, param, V, E, md=100, reps=3):
        """Best-of-`reps` seconds of cgnn_graph_sage_16_par + jones_plassmann_internal
        with allocation and setup outside the timed region, as runner.c:99-134
        times f_par. Returns (seconds, colours)."""
        import time

        V, Vp = _u32(V)
        E, Ep = _u32(E)
        N = len(V) - 1
        raw = np.zeros(len(param) + 16, np.float32)
        po = (-raw.ctypes.data % 32) // 4
        w = raw[po:po + len(param)]
        w[:] = np.asarray(param, np.float32)
        x = np.zeros((N + 8) * 16 + 64, np.float32)
        y = np.zeros((N + 8) * 16 + 64, np.float32)
        xa = x[(-x.ctypes.data % 32) // 4:]
        ya = y[(-y.ctypes.data % 32) // 4:]
        p = np.zeros(N + 8, np.float64)
        color = np.zeros(max(N, 1), np.int32)
        jp = self.lib.jones_plassmann_setup(N, int(V[N]))
        best = float("inf")
        try:
            for _ in range(reps):
                t0 = time.perf_counter()
                self.lib.cgnn_graph_sage_16_par(len(w), w.ctypes.data_as(f32p), N, Vp, Ep, md,
                                                xa.ctypes.data_as(f32p), ya.ctypes.data_as(f32p),
                                                p.ctypes.data_as(f64p))
                self.lib.jones_plassmann_internal(N, Vp, Ep, jp, p.ctypes.data_as(f64p),
                                                  color.ctypes.data_as(i32p))
                best = min(best, time.perf_counter() - t0)
        finally:
            self.lib.jones_plassmann_cleanup(jp)
        return best, color[:N]

    def model_params(self, model_dir, layers):
        """gnn_setup's own parse of <model_dir>/models/m<l>.gnn (gnn.c:20-44)."""

        class GnnData(C.Structure):
            _fields_ = [("params", C.c_int), ("md", C.c_int), ("param", f32p),
                        ("x", f32p), ("y", f32p)]

        layers_sym = C.c_int.in_dll(self.lib, "layers")
        layers_sym.value = layers
        cwd = os.getcwd()
        os.chdir(model_dir)
        try:
            h = self.lib.gnn_setup(1, 0)
        finally:
            os.chdir(cwd)
        d = C.cast(h, C.POINTER(GnnData)).contents
        out = np.ctypeslib.as_array(d.param, (d.params,)).copy(), d.md
        self.lib.gnn_cleanup(h)
        return out

    def mtx_csr(self, path):
        """mtx_parse_par + graph_sort_edges_par + graph_make_simple (main.c:114-129)."""
        f = self.libc.fopen(str(path).encode(), b"r")
        if not f:
            raise IOError(path)
        N = C.c_uint32()
        Vp, Ep = u32p(), u32p()
        self.lib.mtx_parse_par(f, C.byref(N), C.byref(Vp), C.byref(Ep))
        self.libc.fclose(f)
        n = N.value
        self.lib.graph_sort_edges_par(n, Vp, Ep)
        self.lib.graph_make_simple(n, Vp, C.byref(Ep))
        V = np.ctypeslib.as_array(Vp, (n + 1,)).copy()
        M = int(V[n])
        E = np.ctypeslib.as_array(Ep, (max(M, 1),))[:M].copy()
        self.libc.free(C.cast(Vp, C.c_void_p))
        self.libc.free(C.cast(Ep, C.c_void_p))
        return V, E

    def mtx_store(self, path, V, E):
        V, Vp = _u32(V)
        E, Ep = _u32(E)
        f = self.libc.fopen(str(path).encode(), b"w")
        self.lib.mtx_store(f, len(V) - 1, Vp, Ep)
        self.libc.fclose(f)
