"""The ordered schedule's head kernel (one workgroup colouring the dense top
of the greedy order in LDS) gives the reference's colours (needs gfx950).

With min_density = 0 the head takes the first min(N, max) positions of any
graph, so every golden graph and heuristic runs through it; with density 64
it engages on skewed graphs only. Colours must equal the golden
vectors / digests and the same colouring with the head off, bit for bit.
The head is off by default; these tests turn it on."""
import hashlib

import numpy as np
import pytest

import gcx
import oracle
from conftest import HEUR, TINY, load_case

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    yield


@pytest.fixture(scope="module")
def gmodels():
    return {l: gcx.Model.layers_file(l) for l in (2, 3, 4)}


def order(ws, g, h, gmodels):
    if h.startswith("GNN"):
        ws.order(g, gcx.GNN, gmodels[int(h[3:])], gcx.FP32)
    else:
        ws.order(g, h)


@pytest.mark.parametrize("name", TINY + ["rmat10"])
def test_head_everywhere_matches_goldens(name, gmodels):
    print(f"head: {name}, every heuristic, head over the whole graph", flush=True)
    c = load_case(name)
    N = int(c["N"])
    g = gcx.Graph.from_csr(c["V"], c["E"])
    ws = gcx.Workspace(N, len(c["E"]))
    ws.set_schedule(gcx.SCHED_ORDERED)
    ws.set_ordered_head(32768, 0)
    for h in HEUR:
        order(ws, g, h, gmodels)
        ws.color(g)
        np.testing.assert_array_equal(ws.colors(N), c[f"color_{h}"], err_msg=h)
        assert ws.head() == min(N, 32768) and not ws.fell_back()
        assert ws.mex_check(g) == 0


@pytest.mark.parametrize("head", [0, 1024, 4096, 32768])
def test_head_lengths_rmat14(head, gmodels):
    print(f"head: R-MAT 14, head of {head}", flush=True)
    g = gcx.Graph.rmat(14, 16, seed=3)
    V, E = g.download()
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(gcx.SCHED_ORDERED)
    ws.set_ordered_head(head, 0)
    for h in ("LF", "GNN3", "FF"):
        order(ws, g, h, gmodels)
        ws.color(g)
        want, _ = oracle.greedy(V, E, ws.priorities(g.N))
        np.testing.assert_array_equal(ws.colors(g.N), want, err_msg=h)
        assert ws.head() == min(head, g.N)


def test_head_density_rule_engages_on_rmat_and_matches_digests(digests, gmodels):
    print("head: R-MAT 16 and 20 with the density rule", flush=True)
    for name, scale in (("rmat16", 16), ("rmat20", 20)):
        d = digests[name]
        g = gcx.Graph.rmat(scale, 16, seed=1)
        ws = gcx.Workspace(g.N, g.M)
        ws.set_schedule(gcx.SCHED_ORDERED)
        ws.set_ordered_head(32768, 64)
        for h in HEUR:
            order(ws, g, h, gmodels)
            ws.color(g)
            assert sha(ws.colors(g.N)) == d[h]["sha_color"], (name, h)
            if h.startswith("GNN"):
                assert ws.head() > 0, (name, h)  # the dense core is there


@pytest.mark.parametrize("n", [1100, 5000])
def test_head_on_cliques(n):
    """Colours past the first 4096-colour mex window (K_5000) and a hub of
    full degree; the head covers the whole clique."""
    print(f"head: clique K_{n}", flush=True)
    a, b = np.triu_indices(n, 1)
    g = gcx.Graph.from_pairs(n, a.astype(np.uint32), b.astype(np.uint32))
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(gcx.SCHED_ORDERED)
    ws.set_ordered_head(32768, 0)
    ws.order(g, "LF")
    ws.color(g)
    col = ws.colors(g.N)
    assert sorted(col.tolist()) == list(range(1, n + 1))  # LF ties: id order
    np.testing.assert_array_equal(col, np.arange(1, n + 1))
    assert ws.head() == n and ws.validate(g) == n


def test_head_with_64bit_priorities():
    """Caller doubles that are not floats (64-bit keys, ties): the head still
    runs on the rank keys of their order."""
    print("head: 64-bit caller priorities", flush=True)
    g = gcx.Graph.rmat(12, 16, seed=8)
    V, E = g.download()
    rng = np.random.default_rng(4)
    p = rng.integers(0, 40, g.N) + 0.1
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(gcx.SCHED_ORDERED)
    ws.set_ordered_head(32768, 0)
    ws.set_priorities(g, p)
    ws.color(g)
    want, _ = oracle.greedy(V, E, p)
    np.testing.assert_array_equal(ws.colors(g.N), want)
    assert ws.head() == g.N
