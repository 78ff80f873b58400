"""The no-allocation entry point jones_plassmann_internal_noalloc against
the existing jones_plassmann_internal (setup handle) and the oracle's greedy,
through the C ABI (needs gfx950).

Two calls in a row, the second on a smaller graph, so the second call runs on
an array that still holds the first call's contents; plus the size limits and
the argument errors, which must come back as status codes.
"""
import numpy as np
import pytest

import gcx
import oracle

pytestmark = pytest.mark.gpu

MAX_N = 1 << 18
MAX_M = 1 << 23


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    yield


def rmat_csr(scale, seed):
    N, I, J = gcx.rmat_pairs(scale, 16, seed=seed)
    return oracle.csr_from_pairs(N, I, J)


def check_same(V, E, p):
    rc, c1 = gcx.ref.jones_plassmann_internal_noalloc(V, E, p)
    assert rc == 0, gcx.lib().gcx_status_string(rc)
    c0 = gcx.ref.jones_plassmann_internal(V, E, p)
    np.testing.assert_array_equal(c1, c0)
    c_or, _ = oracle.greedy(V, E, p)
    np.testing.assert_array_equal(c1, c_or)
    return c1


def test_two_calls_second_smaller():
    print("noalloc: R-MAT 16 (float priorities) then R-MAT 11 (non-float, tied)", flush=True)
    V, E = rmat_csr(16, 3)
    deg = np.diff(V).astype(np.float64)
    check_same(V, E, deg)  # LF priorities: many ties, 32-bit keys
    V2, E2 = rmat_csr(11, 4)
    rng = np.random.default_rng(5)
    p2 = rng.integers(0, 50, len(V2) - 1) + 0.1  # not floats exactly: 64-bit keys, ties
    check_same(V2, E2, p2)


def test_at_the_limits():
    print("noalloc: R-MAT 18 at N = 2^18", flush=True)
    V, E = rmat_csr(18, 1)
    assert len(V) - 1 == MAX_N and V[-1] <= MAX_M
    rng = np.random.default_rng(7)
    check_same(V, E, rng.standard_normal(len(V) - 1))
    # an empty graph and an isolated-vertex graph afterwards
    check_same(np.zeros(9, np.uint32), np.zeros(0, np.uint32), np.arange(8.0))


def test_limits_and_errors():
    print("noalloc: capacity and argument errors", flush=True)
    N = MAX_N + 1
    rc, _ = gcx.ref.jones_plassmann_internal_noalloc(np.zeros(N + 1, np.uint32),
                                                     np.zeros(0, np.uint32), np.zeros(N))
    assert rc == gcx.ERR_CAPACITY
    V = np.array([0, MAX_M + 1], np.uint32)  # refused before E is read
    rc, _ = gcx.ref.jones_plassmann_internal_noalloc(V, np.zeros(MAX_M + 1, np.uint32), np.zeros(1))
    assert rc == gcx.ERR_CAPACITY
    V, E = rmat_csr(8, 2)
    p = np.ones(len(V) - 1)
    p[3] = np.nan
    rc, _ = gcx.ref.jones_plassmann_internal_noalloc(V, E, p)
    assert rc == gcx.ERR_ARG
    Ebad = E.copy()
    Ebad[0] = len(V) + 5  # column out of range: refused before any kernel follows it
    rc, _ = gcx.ref.jones_plassmann_internal_noalloc(V, Ebad, np.ones(len(V) - 1))
    assert rc == gcx.ERR_ARG
    # and the entry point still works after the errors
    check_same(V, E, np.arange(len(V) - 1, dtype=np.float64))
