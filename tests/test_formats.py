"""The reference's adjacency, .gr and METIS formats (fast_io.h:41-122):
known answers for hand-written files (comments, weights, empty rows) and
store -> parse round trips. adj/METIS parsing is host text I/O; .gr goes
through the device CSR builder (needs gfx950)."""
import ctypes as C

import numpy as np
import pytest

import gcx
import oracle

libc = C.CDLL(None)
libc.fopen.restype = C.c_void_p
libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
libc.fclose.argtypes = [C.c_void_p]
libc.free.argtypes = [C.c_void_p]
u32p = C.POINTER(C.c_uint32)
lib = gcx.lib()
for name in ("adj_parse", "gr_parse"):
    getattr(lib, name).argtypes = [C.c_void_p, u32p, C.POINTER(u32p), C.POINTER(u32p)]
    getattr(lib, name).restype = None
lib.metis_parse.argtypes = [C.c_void_p, u32p, C.POINTER(u32p), C.POINTER(u32p), C.POINTER(u32p)]
lib.metis_parse.restype = None
for name in ("adj_store", "gr_store"):
    getattr(lib, name).argtypes = [C.c_void_p, C.c_uint32, u32p, u32p]
    getattr(lib, name).restype = None
lib.metis_store.argtypes = [C.c_void_p, C.c_uint32, u32p, u32p, u32p]
lib.metis_store.restype = None


def arr(a):
    a = np.ascontiguousarray(a, np.uint32)
    return a, a.ctypes.data_as(u32p)


def take(p, n):
    out = np.ctypeslib.as_array(p, (max(n, 1),))[:n].copy() if n else np.empty(0, np.uint32)
    libc.free(C.cast(p, C.c_void_p))
    return out


def parse(name, path):
    f = libc.fopen(str(path).encode(), b"r")
    N, V, E, W = C.c_uint32(), u32p(), u32p(), u32p()
    if name == "metis":
        lib.metis_parse(f, C.byref(N), C.byref(V), C.byref(E), C.byref(W))
    else:
        getattr(lib, name + "_parse")(f, C.byref(N), C.byref(V), C.byref(E))
    libc.fclose(f)
    n = N.value
    Va = take(V, n + 1)
    Ea = take(E, int(Va[n]))
    Wa = take(W, n) if (name == "metis" and bool(W)) else None
    return n, Va, Ea, Wa


def store(name, path, V, E, W=None):
    f = libc.fopen(str(path).encode(), b"w")
    V, Vp = arr(V)
    E, Ep = arr(E)
    if name == "metis":
        Wp = arr(W)[1] if W is not None else None
        lib.metis_store(f, len(V) - 1, Vp, Ep, Wp)
    else:
        getattr(lib, name + "_store")(f, len(V) - 1, Vp, Ep)
    libc.fclose(f)


def test_adj_known_answer(tmp_path):
    p = tmp_path / "g.adj"
    p.write_text("AdjacencyGraph\n4\n5\n0\n2\n2\n4\n1\n3\n0\n1\n2\n")
    n, V, E, _ = parse("adj", p)
    assert n == 4
    assert V.tolist() == [0, 2, 2, 4, 5] and E.tolist() == [1, 3, 0, 1, 2]


def test_adj_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    deg = rng.integers(0, 6, 50)
    V = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint32)
    E = rng.integers(0, 50, int(V[-1])).astype(np.uint32)  # directed, repeats allowed
    store("adj", tmp_path / "g.adj", V, E)
    n, V2, E2, _ = parse("adj", tmp_path / "g.adj")
    assert n == 50 and np.array_equal(V, V2) and np.array_equal(E, E2)


def test_metis_known_answers(tmp_path):
    p = tmp_path / "a.graph"
    # fmt 0, comments, an isolated vertex, trailing spaces
    p.write_text("% comment\n4 3 0\n2 3 \n1\n% inner comment\n1\n\n")
    n, V, E, W = parse("metis", p)
    assert n == 4 and W is None
    assert V.tolist() == [0, 2, 3, 4, 4] and E.tolist() == [1, 2, 0, 0]
    p = tmp_path / "b.graph"
    # fmt 11: vertex weight first, an edge weight after every neighbour
    p.write_text("3 2 11\n7 2 5 3 6\n8 1 5\n9 1 6\n")
    n, V, E, W = parse("metis", p)
    assert V.tolist() == [0, 2, 3, 4] and E.tolist() == [1, 2, 0, 0] and W.tolist() == [7, 8, 9]
    p = tmp_path / "c.graph"
    # fmt 1: edge weights only
    p.write_text("2 1 1\n2 4\n1 4\n")
    n, V, E, W = parse("metis", p)
    assert V.tolist() == [0, 1, 2] and E.tolist() == [1, 0] and W is None


@pytest.mark.parametrize("weights", [False, True])
def test_metis_round_trip(tmp_path, weights):
    V, E = oracle.csr_from_pairs(30, *zip(*[(i, (3 * i + 1) % 30) for i in range(30)
                                             if i != (3 * i + 1) % 30]))
    W = np.arange(30, dtype=np.uint32) + 1 if weights else None
    store("metis", tmp_path / "g.graph", V, E, W)
    text = (tmp_path / "g.graph").read_text().splitlines()
    assert text[0] == f"30 {len(E) // 2} {10 if weights else 0}"
    n, V2, E2, W2 = parse("metis", tmp_path / "g.graph")
    assert np.array_equal(V, V2) and np.array_equal(E, E2)
    assert (W2 is None) if not weights else np.array_equal(W, W2)


def test_gr_store_text(tmp_path):
    V, E = oracle.csr_from_pairs(4, [0, 1, 2], [1, 2, 3])
    store("gr", tmp_path / "g.gr", V, E)
    assert (tmp_path / "g.gr").read_text() == "p td 4 3\n1 2\n2 3\n3 4\n"


@pytest.mark.gpu
def test_gr_parse_canonical(tmp_path):
    gcx.init()
    p = tmp_path / "g.gr"
    # comments, a repeated edge in both orientations, a self loop
    p.write_text("c hello\nc again\np td 5 6\n1 2\nc mid\n2 3\n3 1\n2 1\n4 4\n5 1\n")
    n, V, E, _ = parse("gr", p)
    assert n == 5
    V0, E0 = oracle.csr_from_pairs(5, [0, 1, 2, 1, 3, 4], [1, 2, 0, 0, 3, 0])
    assert np.array_equal(V, V0) and np.array_equal(E, E0)
    # round trip through gr_store
    store("gr", tmp_path / "h.gr", V, E)
    n2, V2, E2, _ = parse("gr", tmp_path / "h.gr")
    assert np.array_equal(V, V2) and np.array_equal(E, E2)


@pytest.mark.parametrize("name,text", [("adj", "AdjacencyGraph\n3\n4\n0\n2\n"),
                                       ("metis", "3 2 0\n2 9\n1\n\n"),
                                       ("metis", "2 1 0\n2\n")])
def test_malformed_files_exit(tmp_path, name, text):
    """Truncated or out-of-range input ends the process with the library's
    error line (the reference reads past the end of the mapping instead)."""
    import subprocess
    import sys
    p = tmp_path / f"bad.{name}"
    p.write_text(text)
    import os
    tests = os.path.dirname(os.path.abspath(__file__))
    paths = [tests, os.path.dirname(tests), os.path.join(os.path.dirname(tests), "gnn-coloring_amd")]
    code = (f"import sys; sys.path[:0] = {paths!r}; "
            f"import test_formats as t; t.parse({name!r}, {str(p)!r})")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=tests)
    if text == "2 1 0\n2\n":  # a valid file: vertex 2 has an empty last line
        assert r.returncode == 0, r.stderr
    else:
        assert r.returncode == 1 and "parse failed" in r.stderr, (r.returncode, r.stderr)
