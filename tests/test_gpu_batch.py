"""Packed graphs (gcx_graph_pack): ordering + colouring a block-diagonal pack
gives every graph the colours it gets alone, for every heuristic, both
precisions and both JP schedules (needs gfx950)."""
import numpy as np
import pytest

import gcx
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    yield


def tiny(V, E):
    return gcx.Graph.from_csr(np.asarray(V, np.uint32), np.asarray(E, np.uint32))


def graphs():
    """Mixed sizes and shapes, including a lone vertex, isolated vertices and
    a graph with a hub (degree >= 2048 goes to the hub kernels)."""
    gs = [tiny([0, 0], []),
          tiny([0, 2, 4, 6], [1, 2, 0, 2, 0, 1]),  # triangle
          gcx.Graph.er(3000, 1000, seed=3),          # many isolated vertices
          gcx.Graph.rmat(10, 16, seed=4),
          gcx.Graph.rgg(5000, float(np.sqrt(12.0 / (np.pi * 5000))), seed=5),
          gcx.Graph.er(20000, 320000, seed=6),
          gcx.Graph.rmat(14, 16, seed=7)]
    star = 5000  # vertex 0 joined to all others, plus a path
    I = np.concatenate([np.zeros(star - 1, np.uint32), np.arange(1, star - 1, dtype=np.uint32)])
    J = np.concatenate([np.arange(1, star, dtype=np.uint32), np.arange(2, star, dtype=np.uint32)])
    gs.append(gcx.Graph.from_pairs(star, I, J))
    return gs


CASES = [("FF", gcx.FP32), ("LF", gcx.FP32), ("LLF", gcx.FP32), ("SL", gcx.FP32),
         ("SLL", gcx.FP32), ("GNN2", gcx.FP32), ("GNN3", gcx.FP32), ("GNN3", gcx.BF16)]


def alone(g, heuristic, precision, schedule, models):
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(schedule)
    if heuristic.startswith("GNN"):
        ws.order(g, gcx.GNN, models[int(heuristic[3:])], precision)
    else:
        ws.order(g, heuristic)
    ws.color(g)
    return ws.validate(g), ws.digest(g)


@pytest.mark.parametrize("schedule", [gcx.SCHED_ROUNDS, gcx.SCHED_ORDERED])
@pytest.mark.parametrize("heuristic,precision", CASES)
def test_pack_matches_alone(heuristic, precision, schedule):
    models = {l: gcx.Model.layers_file(l) for l in (2, 3)}
    gs = graphs()
    pack = gcx.Graph.pack(gs)
    assert pack.packed_count == len(gs)
    assert pack.N == sum(g.N for g in gs) and pack.M == sum(g.M for g in gs)
    assert pack.check()
    ws = gcx.Workspace(pack.N, pack.M)
    ws.set_schedule(schedule)
    if heuristic.startswith("GNN"):
        ws.order(pack, gcx.GNN, models[int(heuristic[3:])], precision)
    else:
        ws.order(pack, heuristic)
    ws.color(pack)
    assert ws.mex_check(pack) == 0
    colors, digests = ws.packed_colors(pack)
    assert ws.validate(pack) == colors.max()
    for i, g in enumerate(gs):
        c, d = alone(g, heuristic, precision, schedule, models)
        assert (int(colors[i]), int(digests[i])) == (c, d), (i, heuristic)


def test_pack_layout():
    """The pack's CSR is the block-diagonal union with shifted ids."""
    gs = graphs()[:5]
    pack = gcx.Graph.pack(gs)
    V, E = pack.download()
    off, eoff = 0, 0
    for g in gs:
        v, e = g.download()
        assert np.array_equal(V[off:off + g.N + 1], v + eoff)
        assert np.array_equal(E[eoff:eoff + g.M], e + off)
        off += g.N
        eoff += g.M
    assert V[-1] == pack.M


def test_pack_reuse_after_free():
    """Packing copies: the pack stays valid after its parts are freed."""
    gs = graphs()[2:5]
    ref = [alone(g, "LF", gcx.FP32, gcx.SCHED_ROUNDS, {}) for g in gs]
    pack = gcx.Graph.pack(gs)
    for g in gs:
        g.free()
    ws = gcx.Workspace(pack.N, pack.M)
    ws.order(pack, "LF")
    ws.color(pack)
    colors, digests = ws.packed_colors(pack)
    assert [(int(c), int(d)) for c, d in zip(colors, digests)] == ref


@pytest.mark.parametrize("kind", ["rmat14", "er", "clique130", "star"])
def test_recolor_graph_equals_eager(kind):
    """gcx_recolor replays one captured pass (C + 1 round launches); the
    eager loop sets p = colours and colours again, pass by pass."""
    if kind == "rmat14":
        g = gcx.Graph.rmat(14, 16, seed=9)
    elif kind == "er":
        g = gcx.Graph.er(20000, 400000, seed=10)
    elif kind == "clique130":  # 130 colours: past the 64-bit state word
        n = 130
        I, J = np.triu_indices(n, 1)
        g = gcx.Graph.from_pairs(n, I.astype(np.uint32), J.astype(np.uint32))
    else:
        g = graphs()[-1]
    model = gcx.Model.layers_file(3)
    a = gcx.Workspace(g.N, g.M)
    b = gcx.Workspace(g.N, g.M)
    for ws in (a, b):
        ws.order(g, gcx.GNN, model)
        ws.color(g)
    a.recolor(g, 6)
    counts = [b.validate(g)]
    for _ in range(6):
        b.set_priorities(g, b.colors(g.N).astype(np.float64))
        b.color(g)
        counts.append(b.validate(g))
    np.testing.assert_array_equal(a.colors(g.N), b.colors(g.N))
    assert a.validate(g) == counts[-1]
    assert all(x >= y for x, y in zip(counts, counts[1:]))  # never more classes
    assert a.mex_check(g) == 0


@pytest.mark.parametrize("strategy", [gcx.CULB_REVERSE, gcx.CULB_LARGEST_FIRST, gcx.CULB_RANDOM,
                                      gcx.CULB_MIXED])
@pytest.mark.parametrize("kind", ["er", "rmat12", "clique130"])
def test_recolor_strategies_match_oracle(strategy, kind):
    """Class-order Culberson passes (captured) against the CPU restatement."""
    if kind == "er":
        g = gcx.Graph.er(3000, 30000, seed=4)
    elif kind == "rmat12":
        g = gcx.Graph.rmat(12, 16, seed=6)
    else:
        I, J = np.triu_indices(130, 1)
        g = gcx.Graph.from_pairs(130, I.astype(np.uint32), J.astype(np.uint32))
    V, E = g.download()
    ws = gcx.Workspace(g.N, g.M)
    ws.order(g, "LF")
    ws.color(g)
    c0 = ws.colors(g.N)
    ws.recolor_ex(g, 6, strategy, seed=77)
    col, counts = oracle.culberson_classes(V, E, c0, 6, strategy, seed=77)
    np.testing.assert_array_equal(ws.colors(g.N), col)
    assert ws.validate(g) == counts[-1]
    # a second call restarts the pass counter: same seed, same result; one
    # gcx_color per pass (rounds or ordered) gives the same colours
    for sched, flags in ((gcx.SCHED_ROUNDS, 0), (gcx.SCHED_ROUNDS, gcx.CULB_NO_GRAPH),
                         (gcx.SCHED_ORDERED, 0)):
        ws.set_schedule(sched)
        ws.upload_colors(c0)
        ws.recolor_ex(g, 6, strategy | flags, seed=77)
        np.testing.assert_array_equal(ws.colors(g.N), col)
