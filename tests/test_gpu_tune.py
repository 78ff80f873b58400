"""Autotuner (gcx_tune): hub degree and JP schedule choices change speed
only -- priorities and colours are identical under every choice -- and
recorded choices reach graphs created later with the same degree signature
(needs gfx950)."""
import hashlib

import numpy as np
import pytest

import gcx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    gcx.tuner_reset()
    yield
    gcx.tuner_reset()


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(g, heuristic, precision, schedule, model):
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(schedule)
    if heuristic == "GNN":
        ws.order(g, gcx.GNN, model, precision)
    else:
        ws.order(g, heuristic)
    p = ws.priorities(g.N)
    ws.color(g)
    assert ws.mex_check(g) == 0
    return sha(p), sha(ws.colors(g.N)), ws.validate(g)


@pytest.mark.parametrize("heuristic,precision", [("GNN", gcx.FP32), ("GNN", gcx.BF16),
                                                 ("GNN", gcx.BF16X3), ("LF", gcx.FP32),
                                                 ("SL", gcx.FP32)])
def test_hub_degree_changes_nothing(heuristic, precision):
    model = gcx.Model.layers_file(3)
    g = gcx.Graph.rmat(16, 16, seed=3)
    ref = None
    for d in [256, 1024, 4096, 1 << 30]:
        g.set_hub_degree(d)
        assert g.hub_degree == d
        for sched in (gcx.SCHED_ROUNDS, gcx.SCHED_ORDERED):
            got = run(g, heuristic, precision, sched, model)
            if ref is None:
                ref = got
            assert got == ref, (d, sched)
    assert g.check()


def test_tune_records_and_applies():
    model = gcx.Model.layers_file(3)
    g = gcx.Graph.rmat(15, 16, seed=5)
    before = run(g, "GNN", gcx.FP32, gcx.SCHED_ROUNDS, model)
    t = gcx.tune(g, model, gcx.FP32)
    assert t.signature == g.signature
    assert t.hub_degree in (256, 1024, 4096, 16384)
    assert t.schedule in (gcx.SCHED_ROUNDS, gcx.SCHED_ORDERED)
    assert t.order_ms > 0 and t.color_ms > 0
    assert g.hub_degree == t.hub_degree
    assert gcx.tuning_lookup(g).as_dict() == t.as_dict()
    # results unchanged after tuning; AUTO picks the recorded schedule
    assert run(g, "GNN", gcx.FP32, gcx.SCHED_AUTO, model) == before
    # a new graph with the same degree signature takes the recorded hub degree
    h = gcx.Graph.rmat(15, 16, seed=6)
    if h.signature == g.signature:
        assert h.hub_degree == t.hub_degree


def test_tuner_save_load(tmp_path):
    g = gcx.Graph.er(20000, 200000, seed=8)
    t = gcx.tune(g)  # LF keys when no model is given
    path = tmp_path / "tune.txt"
    gcx.tuner_save(path)
    gcx.tuner_reset()
    assert gcx.tuning_lookup(g) is None
    gcx.tuner_load(path)
    got = gcx.tuning_lookup(g)
    assert (got.signature, got.hub_degree, got.schedule) == (t.signature, t.hub_degree, t.schedule)
    assert got.order_ms == pytest.approx(t.order_ms, abs=1e-5)
    assert got.color_ms == pytest.approx(t.color_ms, abs=1e-5)
