"""BF16 GraphSAGE on MFMA (needs gfx950): priorities against the CPU
emulation of the same BF16 contract and against the FP64 network; colours
proper and greedy-exact for the BF16 priorities."""
import numpy as np
import pytest

import gcx
import oracle
from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    yield


@pytest.mark.parametrize("name", ["er300", "stencil6", "mixed10", "rmat10", "star6", "empty5"])
@pytest.mark.parametrize("layers", [2, 3, 4])
def test_bf16_priorities_match_emulation(name, layers, models):
    c = load_case(name)
    V, E = c["V"], c["E"]
    N = len(V) - 1
    param, md = models[layers]
    g = gcx.Graph.from_csr(V, E)
    ws = gcx.Workspace(N, len(E))
    ws.order(g, gcx.GNN, gcx.Model.from_params(param, md), gcx.BF16)
    p = ws.priorities(N)
    emu = oracle.gnn_bf16_emulated(param, V, E, md)
    rel = np.abs(p - emu) / np.maximum(np.abs(emu), 1e-3)
    # only accumulation order differs (plus rare bf16 rounding flips it causes)
    assert rel.mean() < 1e-3 and rel.max() < 5e-2, (rel.mean(), rel.max())
    ws.color(g)
    assert ws.validate(g) > 0 and ws.mex_check(g) == 0


def test_bf16_hub_vertices(models):
    # a star whose centre is a hub (degree >= 1024) and a 1100-clique
    for V, E in [oracle.csr_from_pairs(3001, np.zeros(3000, np.uint32),
                                       np.arange(1, 3001, dtype=np.uint32)),
                 oracle.csr_from_pairs(1100, *zip(*[(a, b) for a in range(1100)
                                                     for b in range(a + 1, 1100)]))]:
        N = len(V) - 1
        param, md = models[3]
        g = gcx.Graph.from_csr(V, E)
        ws = gcx.Workspace(N, len(E))
        ws.order(g, gcx.GNN, gcx.Model.from_params(param, md), gcx.BF16)
        p = ws.priorities(N)
        emu = oracle.gnn_bf16_emulated(param, V, E, md)
        rel = np.abs(p - emu) / np.maximum(np.abs(emu), 1e-3)
        assert rel.mean() < 1e-3 and rel.max() < 5e-2
        ws.color(g)
        assert ws.validate(g) > 0 and ws.mex_check(g) == 0


def test_bf16_rmat16_close_to_fp32(models):
    g = gcx.Graph.rmat(16, 16, seed=1)
    V, E = g.download()
    N = g.N
    param, md = models[3]
    model = gcx.Model.from_params(param, md)
    ws = gcx.Workspace(N, g.M)
    ws.order(g, gcx.GNN, model, gcx.FP32)
    ws.color(g)
    c32 = ws.validate(g)
    ws.order(g, gcx.GNN, model, gcx.BF16)
    pb = ws.priorities(N)
    ws.color(g)
    cb = ws.validate(g)
    assert cb > 0 and ws.mex_check(g) == 0
    assert abs(cb - c32) <= 3, (cb, c32)
    emu = oracle.gnn_bf16_emulated(param, V, E, md)
    rel = np.abs(pb - emu) / np.maximum(np.abs(emu), 1e-3)
    assert rel.mean() < 1e-3, rel.mean()
    # the BF16 error itself (vs FP64) is what bf16 rounding gives: median ~0.4%
    exact = oracle.gnn_fp64(param, V, E, md)
    rel = np.abs(pb - exact) / np.maximum(np.abs(exact), 1e-3)
    assert np.median(rel) < 1e-2
    # deterministic bits run to run
    ws.order(g, gcx.GNN, model, gcx.BF16)
    assert ws.priorities(N).tobytes() == pb.tobytes()


@pytest.mark.parametrize("name", ["er300", "stencil6", "mixed10", "rmat10", "star6", "empty5"])
@pytest.mark.parametrize("layers", [2, 3, 4])
def test_bf16x3_priorities_match_emulation(name, layers, models):
    """Split operands (three MFMAs per product, FP32 rows) against the CPU
    emulation of the same contract, and far closer to FP64 than BF16."""
    c = load_case(name)
    V, E = c["V"], c["E"]
    N = len(V) - 1
    param, md = models[layers]
    g = gcx.Graph.from_csr(V, E)
    ws = gcx.Workspace(N, len(E))
    ws.order(g, gcx.GNN, gcx.Model.from_params(param, md), gcx.BF16X3)
    p = ws.priorities(N)
    emu = oracle.gnn_bf16_emulated(param, V, E, md, split=True)
    scale = np.maximum(np.abs(emu), 1e-3)
    rel = np.abs(p - emu) / scale
    # The GPU sums neighbours in FP32, the emulation in float64; the ~1e-7
    # difference moves bf16 rounding boundaries, and hi + lo carries each
    # operand only to ~2^-18, so the two agree to the contract's own accuracy
    # (~2^-17), amplified where the priority is a small difference of large
    # terms.
    assert rel.mean() < 3e-5 and rel.max() < 3e-3, (rel.mean(), rel.max())
    exact = oracle.gnn_fp64(param, V, E, md)
    err = np.abs(p - exact) / np.maximum(np.abs(exact), 1e-3)
    assert err.mean() < 2e-4 and err.max() < 5e-3, (err.mean(), err.max())
    ws.color(g)
    assert ws.validate(g) > 0 and ws.mex_check(g) == 0


def test_bf16x3_hubs_and_rmat16(models):
    """Hubs (VALU path) and a larger graph: BF16X3 stays within 2^-15 of the
    FP32 path. The colour count is a chaotic function of the order: any
    perturbation above FP32 rounding moves it by a few colours either way."""
    for g in [gcx.Graph.from_csr(*oracle.csr_from_pairs(3001, np.zeros(3000, np.uint32),
                                                        np.arange(1, 3001, dtype=np.uint32))),
              gcx.Graph.rmat(16, 16, seed=1)]:
        V, E = g.download()
        N = g.N
        param, md = models[3]
        model = gcx.Model.from_params(param, md)
        ws = gcx.Workspace(N, g.M)
        ws.order(g, gcx.GNN, model, gcx.FP32)
        p32 = ws.priorities(N)
        ws.color(g)
        c32 = ws.validate(g)
        ws.order(g, gcx.GNN, model, gcx.BF16X3)
        p3 = ws.priorities(N)
        ws.color(g)
        c3 = ws.validate(g)
        assert c3 > 0 and ws.mex_check(g) == 0
        assert abs(c3 - c32) <= 3, (c3, c32)
        emu = oracle.gnn_bf16_emulated(param, V, E, md, split=True)
        rel = np.abs(p3 - emu) / np.maximum(np.abs(emu), 1e-3)
        assert rel.mean() < 3e-5, rel.mean()
        rel32 = np.abs(p3 - p32) / np.maximum(np.abs(p32), 1e-3)
        assert rel32.mean() < 3e-5, rel32.mean()
        ws.order(g, gcx.GNN, model, gcx.BF16X3)
        assert ws.priorities(N).tobytes() == p3.tobytes()  # deterministic
