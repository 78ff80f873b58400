"""The ordered persistent JP schedule must give exactly the colours of the
round schedule and of the reference (needs gfx950)."""
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
        ws.order(g, gcx.GNN, gmodels[int(h[3:])])
    else:
        ws.order(g, h)


# grid 0 = every resident workgroup; small grids give each wave long item
# strides (many waits per wave), grid 1 = a single workgroup of 4 waves
SCHEDULES = [(gcx.SCHED_ORDERED, 0), (gcx.SCHED_ORDERED, 1), (gcx.SCHED_ORDERED, 7)]


@pytest.mark.parametrize("sched", SCHEDULES)
@pytest.mark.parametrize("name", TINY + ["rmat10"])
def test_ordered_matches_reference(name, sched, gmodels):
    c = load_case(name)
    N = int(c["N"])
    g = gcx.Graph.from_csr(c["V"], c["E"])
    ws = gcx.Workspace(N, len(c["E"]))
    ws.set_schedule(*sched)
    for h in HEUR:
        order(ws, g, h, gmodels)
        ws.color(g)
        assert not ws.fell_back(), h
        np.testing.assert_array_equal(ws.colors(N), c[f"color_{h}"], err_msg=h)
        assert ws.validate(g) == int(c[f"ncolors_{h}"])
        assert ws.mex_check(g) == 0


@pytest.mark.parametrize("name", ["rmat16", "rmat20"])
def test_ordered_digests(name, digests, gmodels):
    rec = digests[name]
    gd = rec["graph"]
    g = gcx.Graph.rmat(gd["scale"], gd["edge_factor"], *gd["abc"], seed=gd["seed"])
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(gcx.SCHED_ORDERED)
    for h in HEUR:
        order(ws, g, h, gmodels)
        ws.color(g)
        assert not ws.fell_back(), h
        assert sha(ws.colors(g.N)) == rec[h]["sha_color"], h
    order(ws, g, "GNN3", gmodels)
    ws.color(g)
    ws.recolor(g, 7)
    assert not ws.fell_back()
    assert sha(ws.colors(g.N)) == rec["culb_GNN3_8"]["sha_color"]


@pytest.mark.parametrize("n", [130, 1100])
def test_ordered_cliques(n, gmodels):
    """Cliques push colours far past the state word into the overflow words."""
    V, E = oracle.csr_from_pairs(n, *zip(*[(a, b) for a in range(n) for b in range(a + 1, n)]))
    g = gcx.Graph.from_csr(V, E)
    ws = gcx.Workspace(n, len(E))
    for sched in [(gcx.SCHED_ROUNDS, 0), (gcx.SCHED_ORDERED, 0), (gcx.SCHED_ORDERED, 3)]:
        ws.set_schedule(*sched)
        for h in ["FF", "LF", "GNN2"]:
            order(ws, g, h, gmodels)
            p = ws.priorities(n)
            ws.color(g)
            want, _ = oracle.greedy(V, E, p)
            np.testing.assert_array_equal(ws.colors(n), want, err_msg=f"{h} {sched}")


def test_ordered_star_and_hubs(gmodels):
    rng = np.random.default_rng(5)
    n = 30000
    hubs = rng.choice(n, 8, replace=False)
    I = np.concatenate([np.repeat(hubs, 5000), rng.integers(0, n, 80000)]).astype(np.uint32)
    J = np.concatenate([rng.integers(0, n, 40000), rng.integers(0, n, 80000)]).astype(np.uint32)
    V, E = oracle.csr_from_pairs(n, I, J)
    g = gcx.Graph.from_csr(V, E)
    ws = gcx.Workspace(n, len(E))
    for sched in [(gcx.SCHED_ORDERED, 0), (gcx.SCHED_ORDERED, 2)]:
        ws.set_schedule(*sched)
        for h in ["FF", "LF", "SLL", "GNN4"]:
            order(ws, g, h, gmodels)
            p = ws.priorities(n)
            ws.color(g)
            want, _ = oracle.greedy(V, E, p)
            np.testing.assert_array_equal(ws.colors(n), want, err_msg=f"{h} {sched}")


def test_ordered_caller_priorities_key64():
    """Priorities that are not floats take the 64-bit key path of the sort."""
    c = load_case("rmat10")
    N = int(c["N"])
    g = gcx.Graph.from_csr(c["V"], c["E"])
    ws = gcx.Workspace(N, len(c["E"]))
    rng = np.random.default_rng(3)
    p = rng.integers(0, 50, N).astype(np.float64) + 0.1  # not floats, many ties
    want, _ = oracle.greedy(c["V"], c["E"], p)
    for sched in [gcx.SCHED_ROUNDS, gcx.SCHED_ORDERED]:
        ws.set_schedule(sched)
        ws.set_priorities(g, p)
        ws.color(g)
        np.testing.assert_array_equal(ws.colors(N), want)


def test_rgg_device_equals_host():
    for n, r, seed in [(3000, 0.04, 1), (20000, 0.011, 2), (1, 0.5, 3), (500, 0.9, 4)]:
        n_, I, J = gcx.rgg2d_pairs(n, r, seed=seed)
        Vh, Eh = gcx.Graph.from_pairs(n, I, J).download()
        Vd, Ed = gcx.Graph.rgg(n, r, seed=seed).download()
        np.testing.assert_array_equal(Vh, Vd)
        np.testing.assert_array_equal(Eh, Ed)


@pytest.mark.parametrize("slots", [1, 2, 4])
def test_shared_ordered_launch(slots, digests, gmodels):
    """Plain (non-cooperative) ordered launches, up to `slots` at once, give
    the reference digests; concurrent streams give the same colourings."""
    rec = digests["rmat16"]
    gd = rec["graph"]
    g = gcx.Graph.rmat(gd["scale"], gd["edge_factor"], *gd["abc"], seed=gd["seed"])
    gcx.set_ordered_launch(slots)
    try:
        ws = gcx.Workspace(g.N, g.M)
        ws.set_schedule(gcx.SCHED_ORDERED)
        for h in HEUR:
            order(ws, g, h, gmodels)
            ws.color(g)
            assert not ws.fell_back(), h
            assert sha(ws.colors(g.N)) == rec[h]["sha_color"], h
        srv = gcx.Server(workers=4, budget=8 << 30)
        gid = srv.add_graph(g)
        reqs = [gcx.make_request(gcx.GEN_RESIDENT, gid, heuristic=h, schedule=gcx.SCHED_ORDERED,
                                 arrival=0.0002 * i) for i, h in enumerate(HEUR * 2)]
        out = srv.run(reqs)
        srv.close()
        for i, r in enumerate(out):
            h = (HEUR * 2)[i]
            assert r.status == 0 and r.colors == rec[h]["ncolors"], h
    finally:
        gcx.set_ordered_launch(0)


def _probe_graph():
    """A 300-clique whose members take colours 1..300 in greedy order, plus
    independent probe vertices of lower priority, each joined to clique
    members chosen by colour. Probe predecessor counts sit on both sides of
    every state layout boundary (4-bit fields up to 15, 8-bit fields up to
    255, bit-set words above); the chosen colours fill the fields (mex just
    past them), skip past the fields (counted in the last field) or leave
    gaps."""
    K = 300
    I, J = np.triu_indices(K, 1)
    pairs_i, pairs_j = [I.astype(np.int64)], [J.astype(np.int64)]
    probes = []
    for np_ in (1, 2, 14, 15, 16, 17, 126, 127, 254, 255, 256, 257, 299):
        F = 16 if np_ <= 15 else 8
        top = -(-(np_ + 2) // F) * F - 1  # colours past this go to the last field
        sets = [list(range(1, np_ + 1)),                        # mex = np + 1
                list(range(K - np_ + 1, K + 1)),                # high colours, mex 1
                list(range(1, np_)) + [K]]                      # mex = np
        if np_ >= 3:
            sets.append(list(range(2, np_ + 2)))                # mex 1
            sets.append([c for c in range(1, np_ + 2) if c != np_ // 2][:np_])
        if top < K and np_ < top:
            sets.append(list(range(1, np_)) + [top])            # the last colour field
            sets.append(list(range(1, np_)) + [top + 1])        # one colour just past
        for cols in sets:
            assert len(cols) == np_
            probes.append(cols)
    N = K + len(probes)
    for pi, cols in enumerate(probes):
        v = K + pi
        for c in cols:
            pairs_i.append(np.array([c - 1]))
            pairs_j.append(np.array([v]))
    I = np.concatenate(pairs_i).astype(np.uint32)
    J = np.concatenate(pairs_j).astype(np.uint32)
    p = np.zeros(N, np.float64)
    p[:K] = 1e6 - np.arange(K)  # member i precedes i + 1: colour i + 1
    return N, I, J, p, probes


@pytest.mark.parametrize("hub_degree", [0, 64])
@pytest.mark.parametrize("sched", SCHEDULES + [(gcx.SCHED_ROUNDS, 0)])
def test_state_layout_boundaries(sched, hub_degree):
    N, I, J, p, probes = _probe_graph()
    g = gcx.Graph.from_pairs(N, I, J)
    if hub_degree:
        g.set_hub_degree(hub_degree)  # clique members become hub rows
    V, E = g.download()
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(*sched)
    ws.set_priorities(g, p)
    ws.color(g)
    col = ws.colors(N)
    ref, _ = oracle.greedy(V, E, p)
    np.testing.assert_array_equal(col, ref)
    np.testing.assert_array_equal(col[:300], np.arange(1, 301))
    for pi, cols in enumerate(probes):
        used = set(cols)
        mex = next(c for c in range(1, len(cols) + 2) if c not in used)
        assert col[300 + pi] == mex, (len(cols), pi)
    assert ws.mex_check(g) == 0
