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
