"""Pins the CPU oracle (oracle/) before anything is compared with it.

Checks, per oracle function:
  * golden vectors produced by the reference's own compiled sources
    (tests/golden/make_golden.py): CSR, priorities (bit patterns), colours,
    colour counts, Culberson chains, digests of R-MAT scale 16 and 20;
  * brute force: sequential greedy in (p desc, id asc) order, written
    independently below, on the hand-made graphs and on random tiny graphs;
  * hand-derived known answers for SL / SLL (path, star, clique, a
    threshold-jump graph) and ID / SD / SDA (star, clique, path, components,
    the isolated last vertex); a graph on which a textbook priority queue
    gives another ID order than the reference's heap;
  * the GNN in float64 as sparse matrix algebra (numpy/scipy), independent of
    the oracle's loop structure, against the FP64 and FP32 oracles;
  * planted mistakes: a corrupted colouring / CSR must be rejected.
"""
import hashlib
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import HEUR, SEQ, TINY, load_case
from oracle import seqorder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def brute_greedy(V, E, p):
    """Sequential greedy in (p descending, id ascending) order."""
    N = len(V) - 1
    order = sorted(range(N), key=lambda u: (-p[u], u))
    color = np.zeros(N, np.int32)
    for u in order:
        used = {int(color[v]) for v in E[V[u]:V[u + 1]] if color[v] > 0}
        c = 1
        while c in used:
            c += 1
        color[u] = c
    return color


def priorities(case, h, models):
    if h.startswith("GNN"):
        param, md = models[int(h[3:])]
        return oracle.gnn_fp32(param, case["V"], case["E"], md)
    return oracle.key(h, case["V"], case["E"])[0]


@pytest.mark.parametrize("name", TINY + ["rmat10"])
def test_csr_matches_reference(name):
    c = load_case(name)
    V, E = oracle.csr_from_pairs(int(c["N"]), c["I"], c["J"])
    np.testing.assert_array_equal(V, c["V"])
    np.testing.assert_array_equal(E, c["E"])
    assert oracle.graph_validate(V, E) == 1


@pytest.mark.parametrize("name", TINY + ["rmat10"])
@pytest.mark.parametrize("h", HEUR)
def test_priorities_and_colours_match_reference(name, h, models):
    c = load_case(name)
    p = priorities(c, h, models)
    # bit-exact priorities (compare the IEEE bit patterns)
    assert p.view(np.uint64).tolist() == c[f"p_{h}"].view(np.uint64).tolist()
    col, _ = oracle.greedy(c["V"], c["E"], p)
    np.testing.assert_array_equal(col, c[f"color_{h}"])
    assert oracle.validate(c["V"], c["E"], col) == int(c[f"ncolors_{h}"])
    assert oracle.mex_violations(c["V"], c["E"], p, col) == 0


@pytest.mark.parametrize("name", TINY + ["rmat10"])
def test_culberson_matches_reference(name, models):
    c = load_case(name)
    p = priorities(c, "GNN3", models)
    col, counts = oracle.culberson(c["V"], c["E"], p, 5)
    np.testing.assert_array_equal(col, c["culb_color"])
    np.testing.assert_array_equal(counts, c["culb_counts"])
    assert all(counts[i + 1] <= counts[i] for i in range(len(counts) - 1))


@pytest.mark.parametrize("name", TINY + ["rmat10"])
@pytest.mark.parametrize("h", ["FF", "LF", "SL", "GNN3"])
def test_greedy_equals_brute_force(name, h, models):
    c = load_case(name)
    p = priorities(c, h, models)
    col, rounds = oracle.greedy(c["V"], c["E"], p)
    np.testing.assert_array_equal(col, brute_greedy(c["V"], c["E"], p))
    N = len(c["V"]) - 1
    maxdeg = int(np.diff(c["V"]).max()) if N else 0
    assert col.max(initial=0) <= maxdeg + 1
    assert rounds <= N


def test_greedy_random_tiny_graphs():
    rng = np.random.default_rng(11)
    for trial in range(60):
        n = int(rng.integers(1, 40))
        m = int(rng.integers(0, 3 * n + 1))
        I = rng.integers(0, n, m).astype(np.uint32)
        J = rng.integers(0, n, m).astype(np.uint32)
        V, E = oracle.csr_from_pairs(n, I, J)
        # ties on purpose: small integer priorities exercise the id tie-break
        p = rng.integers(0, 4, n).astype(np.float64)
        col, _ = oracle.greedy(V, E, p)
        np.testing.assert_array_equal(col, brute_greedy(V, E, p))
        assert oracle.validate(V, E, col) == (col.max() if n else 1)


@pytest.mark.parametrize("name", ["rmat16", "rmat20"])
def test_digests_match_reference(name, digests, models):
    import gcx  # host generators only

    rec = digests[name]
    N, I, J = gcx.rmat_pairs(rec["graph"]["scale"], rec["graph"]["edge_factor"],
                             *rec["graph"]["abc"], seed=rec["graph"]["seed"])
    V, E = oracle.csr_from_pairs(N, I, J)
    assert sha(V) == rec["sha_V"] and sha(E) == rec["sha_E"]
    for h in HEUR:
        p = priorities({"V": V, "E": E}, h, models)
        assert sha(p) == rec[h]["sha_p"], h
        col, _ = oracle.greedy(V, E, p)
        assert sha(col) == rec[h]["sha_color"], h
        assert oracle.validate(V, E, col) == rec[h]["ncolors"]
    p = priorities({"V": V, "E": E}, "GNN3", models)
    col, counts = oracle.culberson(V, E, p, 8)
    assert sha(col) == rec["culb_GNN3_8"]["sha_color"]
    assert counts.tolist() == rec["culb_GNN3_8"]["counts"]


# ---------------------------------------------------------- known answers

def csr(n, pairs):
    I = np.array([a for a, _ in pairs], np.uint32)
    J = np.array([b for _, b in pairs], np.uint32)
    return oracle.csr_from_pairs(n, I, J)


KNOWN_SL = {
    # path 0-1-2-3-4: ends go at delta=1, the middle at delta=2
    "path5": (5, [(0, 1), (1, 2), (2, 3), (3, 4)], [1, 2, 2, 2, 1], [0, 1, 1, 1, 0]),
    # star: leaves at 1, centre at delta = max(1+1, 0) = 2
    "star6": (6, [(0, i) for i in range(1, 6)], [2, 1, 1, 1, 1, 1], [1, 0, 0, 0, 0, 0]),
    # clique K6: everything in round one at delta = 5
    "clique6": (6, [(a, b) for a in range(6) for b in range(a + 1, 6)], [5] * 6, [3] * 6),
    # path 0-1-2 glued to K5 on 3..7: delta 1 -> 2 -> jumps to 4
    "jump8": (8, [(0, 1), (1, 2), (2, 3)] + [(a, b) for a in range(3, 8) for b in range(a + 1, 8)],
              [1, 2, 2, 4, 4, 4, 4, 4], [0, 1, 1, 2, 2, 2, 2, 2]),
}


@pytest.mark.parametrize("name", list(KNOWN_SL))
def test_sl_sll_known_answers(name):
    n, pairs, sl, sll = KNOWN_SL[name]
    V, E = csr(n, pairs)
    assert oracle.key("SL", V, E)[0].tolist() == sl
    assert oracle.key("SLL", V, E)[0].tolist() == sll


def test_closed_form_keys():
    V, E = csr(6, [(0, 1), (0, 2), (0, 3), (1, 2), (4, 4)])
    deg = np.diff(V)
    assert oracle.key("FF", V, E)[0].tolist() == [6, 5, 4, 3, 2, 1]
    assert oracle.key("LF", V, E)[0].tolist() == deg.tolist()
    assert oracle.key("LLF", V, E)[0].tolist() == [int(d).bit_length() for d in deg]


def test_clique_and_bipartite_colours():
    V, E = csr(9, [(a, b) for a in range(9) for b in range(a + 1, 9)])
    col, rounds = oracle.greedy(V, E, oracle.key("FF", V, E)[0])
    assert sorted(col.tolist()) == list(range(1, 10)) and rounds == 9
    V, E = csr(7, [(a, b) for a in range(3) for b in range(3, 7)])
    col, _ = oracle.greedy(V, E, oracle.key("FF", V, E)[0])
    assert col.tolist() == [1, 1, 1, 2, 2, 2, 2]


# -------------------------------------------------------- GNN as algebra

def gnn_algebra(param, V, E, md):
    """GraphSAGE-16 in float64 as sparse matrix products (row-normalised
    adjacency), independent of the oracle's per-vertex loops."""
    N = len(V) - 1
    deg = np.diff(V).astype(np.float64)
    A = sp.csr_matrix((np.ones(len(E)), E.astype(np.int64), V.astype(np.int64)), shape=(N, N))
    Dinv = sp.diags(np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0))
    P = Dinv @ A  # mean over neighbours (0 for isolated vertices)
    w = np.asarray(param, np.float64).reshape(-1, 528)
    L = w.shape[0]
    fdeg = deg / md
    fid = np.arange(N, dtype=np.float64) / N
    X0 = np.concatenate([np.repeat(fdeg[:, None], 8, 1), np.repeat(fid[:, None], 8, 1)], 1)
    x = X0
    for l in range(L):
        W = w[l, :512].reshape(32, 16)
        b = w[l, 512:]
        H = np.concatenate([P @ x, x], 1) @ W + b
        x = np.maximum(H, 0.0)
    return x.sum(1)


@pytest.mark.parametrize("name", ["er300", "stencil6", "mixed10", "rmat10"])
@pytest.mark.parametrize("layers", [2, 3, 4])
def test_gnn_fp64_matches_algebra(name, layers, models):
    c = load_case(name)
    param, md = models[layers]
    want = gnn_algebra(param, c["V"], c["E"], md)
    got = oracle.gnn_fp64(param, c["V"], c["E"], md)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)
    # FP32 (reference op order) stays within FP32 rounding of the exact value
    p32 = oracle.gnn_fp32(param, c["V"], c["E"], md)
    np.testing.assert_allclose(p32, want, rtol=1e-4, atol=1e-5)


def test_model_parse_matches_reference_loader(models):
    for l in (2, 3, 4):
        param, md = oracle.load_model(os.path.join(ROOT, "models", f"m{l}.gnn"))
        ref_param, ref_md = models[l]
        assert md == ref_md
        assert param.view(np.uint32).tolist() == ref_param.view(np.uint32).tolist()


def test_model_loader_rejects_bad_files(tmp_path):
    bad = tmp_path / "bad.gnn"
    bad.write_text("1056 100\n0.5 0.25\n")
    with pytest.raises(IOError):
        oracle.load_model(bad)
    with pytest.raises(IOError):
        oracle.load_model(tmp_path / "missing.gnn")


# ------------------------------------------------------ planted mistakes

def test_checkers_reject_planted_mistakes(models):
    c = load_case("rmat10")
    V, E = c["V"], c["E"]
    p = priorities(c, "GNN3", models)
    col = c["color_GNN3"].copy()
    # a monochromatic edge
    u = int(np.argmax(np.diff(V)))
    bad = col.copy()
    bad[E[V[u]]] = bad[u]
    assert oracle.validate(V, E, bad) == 0
    # a proper but non-greedy colouring (shift one vertex to a fresh colour)
    bad = col.copy()
    bad[u] = col.max() + 1
    assert oracle.validate(V, E, bad) > 0
    assert oracle.mex_violations(V, E, p, bad) >= 1
    # an uncoloured vertex
    bad = col.copy()
    bad[0] = 0
    assert oracle.validate(V, E, bad) == 0
    # a flipped tie-break (v > u instead of v < u) changes colours here
    q = np.zeros_like(p)
    wrong, _ = oracle.greedy(V, E, q)
    right = brute_greedy(V, E, q)
    np.testing.assert_array_equal(wrong, right)
    flipped = brute_greedy(V, E, -np.arange(len(p), dtype=np.float64) * 0 + np.arange(len(p)))
    assert not np.array_equal(flipped, right)
    # an asymmetric CSR fails the graph check
    E2 = E.copy()
    E2[V[u]] = (E2[V[u]] + 1) % (len(V) - 1)
    assert oracle.graph_validate(V, E2) == 0


def test_bf16_rounding_pinned():
    """bf16_bits against torch.bfloat16 (round to nearest even) on random bit
    patterns of every class, and hand-derived cases a truncating or
    round-half-up rule, or one that lets NaNs become Inf, would fail."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(11)
    bits = np.concatenate([
        rng.integers(0, 1 << 32, 1 << 20, dtype=np.uint64).astype(np.uint32),
        # exact ties and their neighbours: low 16 bits 0x7FFF, 0x8000, 0x8001
        (rng.integers(0, 1 << 16, 1 << 14, dtype=np.uint64).astype(np.uint32) << 16)
        | rng.choice(np.array([0x7FFF, 0x8000, 0x8001], np.uint32), 1 << 14),
    ])
    x = bits.view(np.float32)
    mine = oracle.bf16_bits(x)
    ref = torch.from_numpy(x.copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    nan = np.isnan(x)
    np.testing.assert_array_equal(mine[~nan], ref[~nan])
    back = (mine.astype(np.uint32) << 16).view(np.float32)
    assert np.isnan(back[nan]).all()  # NaNs stay NaN (torch: its canonical NaN)
    cases = {
        0x3F800000: 0x3F80,  # 1.0 exact
        0x3F808000: 0x3F80,  # tie, even neighbour below: down
        0x3F818000: 0x3F82,  # tie, odd below: up (truncation gives 0x3F81)
        0x3F808001: 0x3F81,  # just past the tie: up
        0x3F807FFF: 0x3F80,  # just below the tie: down (half-up would agree; 0x3F818000 does not)
        0xBF818000: 0xBF82,  # the same on a negative value
        0x7F7FFFFF: 0x7F80,  # largest float: rounds to Inf
        0x7F7F7FFF: 0x7F7F,  # largest bf16 below it stays finite
        0x80000000: 0x8000,  # -0
        0x00008000: 0x0000,  # subnormal tie to even (0)
        0x00018000: 0x0002,  # subnormal tie to even (up)
        0x7F800000: 0x7F80,  # +Inf
        0xFF800000: 0xFF80,  # -Inf
        0x7F800001: 0x7FC0,  # NaN with only low payload bits: quiet NaN, not Inf
        0xFFFFFFFF: 0xFFFF,  # all-ones NaN: stays NaN (an unchecked add wraps it to +0)
        0x7FC00000: 0x7FC0,  # quiet NaN
    }
    got = oracle.bf16_bits(np.array(list(cases), np.uint32).view(np.float32))
    assert [hex(v) for v in got] == [hex(v) for v in cases.values()]


@pytest.mark.parametrize("name", ["er300", "rmat10", "mixed10"])
@pytest.mark.parametrize("layers", [2, 3, 4])
def test_bf16_emulation_structure(name, layers, models):
    """With bf16 rounding switched off the BF16 contract is the FP64 network;
    with it on, priorities stay within bf16 error of FP64."""
    c = load_case(name)
    param, md = models[layers]
    exact = oracle.gnn_fp64(param, c["V"], c["E"], md)
    plain = oracle.gnn_bf16_emulated(param, c["V"], c["E"], md, rounding=False)
    np.testing.assert_allclose(plain, exact, rtol=1e-12, atol=1e-12)
    rounded = oracle.gnn_bf16_emulated(param, c["V"], c["E"], md)
    rel = np.abs(rounded - exact) / np.maximum(np.abs(exact), 1e-3)
    assert rel.mean() < 2e-2 and not np.array_equal(rounded, exact)


@pytest.mark.parametrize("name", ["er300", "rmat10", "mixed10"])
@pytest.mark.parametrize("layers", [2, 3, 4])
def test_bf16x3_emulation_structure(name, layers, models):
    """The split (BF16X3) contract: each operand and weight is carried to
    about 2^-18 relative by its two bf16 parts and only the lo*lo product
    (about 2^-18 relative) is dropped, so priorities sit about 2^-17 from
    FP64 -- hundreds of times closer than the one-term BF16 contract."""
    c = load_case(name)
    param, md = models[layers]
    exact = oracle.gnn_fp64(param, c["V"], c["E"], md)
    x3 = oracle.gnn_bf16_emulated(param, c["V"], c["E"], md, split=True)
    one = oracle.gnn_bf16_emulated(param, c["V"], c["E"], md)
    scale = np.maximum(np.abs(exact), 1e-3)
    err3, err1 = np.abs(x3 - exact) / scale, np.abs(one - exact) / scale
    assert err3.max() < 5e-3 and err3.mean() < 3e-5, (err3.max(), err3.mean())
    assert err3.mean() < err1.mean() / 50, (err3.mean(), err1.mean())



def test_class_orders_keep_colour_count():
    """Every class order gives a proper colouring with at most as many
    colours (Culberson); reverse is the reference's p = colour."""
    rng = np.random.default_rng(5)
    pairs = {(int(a), int(b)) for a, b in rng.integers(0, 400, (3000, 2)) if a != b}
    V, E = oracle.csr_from_pairs(400, *zip(*pairs))
    c0, _ = oracle.greedy(V, E, rng.random(400))
    # orc_culberson's first pass colours with p = c0 itself: 5 reverse passes
    ref, ref_counts = oracle.culberson(V, E, c0.astype(np.float64), 5)
    for strategy in range(4):
        col, counts = oracle.culberson_classes(V, E, c0, 5, strategy, seed=9)
        assert oracle.validate(V, E, col) == counts[-1]
        assert all(a >= b for a, b in zip([c0.max()] + counts, counts))
        if strategy == 0:  # the reference chain
            np.testing.assert_array_equal(col, ref)
            assert counts == ref_counts.tolist()
    # largest first puts the biggest class first: it keeps colour 1
    key = oracle.class_keys(c0, 1, 0, 0)
    sizes = np.bincount(c0)
    assert key[c0 == np.argmax(sizes[1:]) + 1].max() == c0.max()


# ------------------------------------------- ID / SD / SDA (incidence_degree.c)

SEQ_FN = {"ID": seqorder.incidence_degree, "SD": seqorder.saturation_degree,
          "SDA": seqorder.saturation_degree_alt}


@pytest.mark.parametrize("name", TINY + ["rmat10"])
@pytest.mark.parametrize("h", SEQ)
def test_seq_orderings_match_reference(name, h):
    c = load_case(name)
    p = SEQ_FN[h](c["V"], c["E"])
    assert p.view(np.uint64).tolist() == c[f"p_{h}"].view(np.uint64).tolist()
    col, _ = oracle.greedy(c["V"], c["E"], p)
    np.testing.assert_array_equal(col, c[f"color_{h}"])


def test_seq_textbook_queue_is_told_apart():
    """The reference's heap removal only sifts down; a correct priority queue
    orders heapquirk10 differently, so the goldens pin the heap rule."""
    c = load_case("heapquirk10")
    assert not np.array_equal(seqorder.textbook_incidence_degree(c["V"], c["E"]), c["p_ID"])
    np.testing.assert_array_equal(seqorder.incidence_degree(c["V"], c["E"]), c["p_ID"])


def _pairs_graph(n, pairs):
    I = np.array([a for a, _ in pairs], np.uint32)
    J = np.array([b for _, b in pairs], np.uint32)
    return oracle.csr_from_pairs(n, I, J)


def test_seq_known_answers():
    # star K1,5 with the centre at 2: centre first (largest degree), then the
    # leaves, all tied on (1 ordered neighbour, degree 1): largest id first
    V, E = _pairs_graph(6, [(2, x) for x in (0, 1, 3, 4, 5)])
    for f in (seqorder.incidence_degree, seqorder.saturation_degree):
        np.testing.assert_array_equal(f(V, E), [1, 2, 6, 3, 4, 5])
    # SDA: the centre alone has the maximum degree; then all leaves share
    # (saturation 1, degree 1) and one priority
    np.testing.assert_array_equal(seqorder.saturation_degree_alt(V, E), [5, 5, 6, 5, 5, 5])
    # clique K5: start at 0 (smallest id of maximum degree), then every
    # vertex has the same count, so the largest id goes next
    V, E = _pairs_graph(5, [(a, b) for a in range(5) for b in range(a + 1, 5)])
    for f in (seqorder.incidence_degree, seqorder.saturation_degree):
        np.testing.assert_array_equal(f(V, E), [5, 1, 2, 3, 4])
    # path 0-1-2-3-4: start at 1 (first vertex of degree 2); 0 and 2 both
    # have one ordered neighbour and 2 has the larger degree, so the walk runs
    # down the path (3, then 4 over 0 by id) and ends at 0
    V, E = _pairs_graph(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    np.testing.assert_array_equal(seqorder.incidence_degree(V, E), [1, 5, 4, 3, 2])
    # components in the order of their smallest vertex; an isolated last
    # vertex is never walked (incidence_degree.c:160-164) and keeps -1
    V, E = _pairs_graph(5, [(0, 1), (1, 2)])
    np.testing.assert_array_equal(seqorder.incidence_degree(V, E), [3, 5, 4, 2, -1])
    V, E = _pairs_graph(4, [(2, 3)])
    np.testing.assert_array_equal(seqorder.incidence_degree(V, E), [4, 3, 2, 1])


def test_seq_oracle_against_compiled_reference():
    """Random small graphs, when oracle/_ref is built (development container)."""
    try:
        ref = oracle.Ref()
    except ImportError:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(21)
    for _ in range(150):
        n = int(rng.integers(1, 50))
        m = int(rng.integers(0, 4 * n + 1))
        V, E = oracle.csr_from_pairs(n, rng.integers(0, n, m).astype(np.uint32),
                                     rng.integers(0, n, m).astype(np.uint32))
        for h, f in SEQ_FN.items():
            np.testing.assert_array_equal(f(V, E), ref.ordering(h, V, E))
