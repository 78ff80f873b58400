"""Serving loop (gcx_serve.h): concurrent jobs give the same colourings as
the same jobs alone and as the direct device API; admission respects the
HBM budget (needs gfx950)."""
import numpy as np
import pytest

import gcx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    gcx.init()
    yield


def trace(k, seed=7, nmax=60000):
    """A small mixed trace: R-MAT / ER / RGG, FF/LF/SL/GNN2/GNN3, BF16 GNN,
    both JP schedules, arrivals spread over ~20 ms."""
    rng = np.random.default_rng(seed)
    gens = [gcx.GEN_RMAT, gcx.GEN_ER, gcx.GEN_RGG]
    heur = ["FF", "LF", "SL", "GNN2", "GNN3"]
    reqs, t = [], 0.0
    for i in range(k):
        g = gens[i % 3]
        n = int(np.exp(rng.uniform(np.log(1000), np.log(nmax))))
        param = {gcx.GEN_RMAT: 16.0, gcx.GEN_ER: float(rng.uniform(8, 64)), gcx.GEN_RGG: 0.0}[g]
        h = heur[int(rng.integers(0, len(heur)))]
        sched = gcx.SCHED_ORDERED if i % 2 else gcx.SCHED_ROUNDS
        reqs.append(gcx.make_request(g, n, param, seed=100 + i, heuristic=h,
                                     precision=gcx.BF16, schedule=sched, arrival=t))
        t += float(rng.exponential(0.001))
    return reqs


def direct(req):
    """The same job through the plain device API."""
    if req.generator == gcx.GEN_RMAT:
        scale = max(1, int(np.ceil(np.log2(req.n))))
        g = gcx.Graph.rmat(scale, int(req.param), seed=req.seed)
    elif req.generator == gcx.GEN_ER:
        g = gcx.Graph.er(req.n, int(np.floor(req.n * req.param / 2 + 0.5)), seed=req.seed)
    else:
        g = gcx.Graph.rgg(req.n, float(np.sqrt(12.0 / (np.pi * req.n))), seed=req.seed)
    ws = gcx.Workspace(g.N, g.M)
    ws.set_schedule(req.schedule)
    if req.heuristic == gcx.GNN:
        ws.order(g, gcx.GNN, gcx.Model.layers_file(req.layers), req.precision)
    else:
        ws.order(g, req.heuristic)
    ws.color(g)
    assert ws.mex_check(g) == 0
    return ws.validate(g), ws.digest(g), g.M


def test_serve_matches_alone_and_direct():
    reqs = trace(24)
    srv = gcx.Server(workers=4, budget=8 << 30)
    mix = srv.run(reqs)
    assert all(r.status == 0 for r in mix), [r.status for r in mix]
    assert len({r.worker for r in mix}) > 1  # jobs really ran on several workers
    for i, q in enumerate(reqs):
        alone = srv.run([gcx.make_request(q.generator, q.n, q.param, q.seed, q.heuristic,
                                          q.layers, q.precision, q.schedule, 0.0)])[0]
        assert alone.status == 0
        assert (alone.colors, alone.digest, alone.M) == (mix[i].colors, mix[i].digest, mix[i].M), i
    for i in (0, 1, 2, 5):
        colors, digest, M = direct(reqs[i])
        assert (colors, digest, M) == (mix[i].colors, mix[i].digest, mix[i].M), i
    for r in mix:
        assert r.t_arrival <= r.t_admit <= r.t_done
    srv.close()


def test_serve_budget_admission():
    reqs = trace(10, seed=3, nmax=20000)
    fps = [gcx.request_footprint(q) for q in reqs]
    budget = max(fps) + min(fps)  # room for the largest job plus at most a small one
    srv = gcx.Server(workers=4, budget=budget)
    tight = srv.run(reqs)
    assert srv.peak_reserved() <= budget
    srv.close()
    srv = gcx.Server(workers=4, budget=64 << 30)
    loose = srv.run(reqs)
    srv.close()
    assert [(r.colors, r.digest) for r in tight] == [(r.colors, r.digest) for r in loose]


def test_serve_rejects_oversized():
    big = gcx.make_request(gcx.GEN_RMAT, 1 << 20, 16.0, heuristic="FF")
    small = gcx.make_request(gcx.GEN_ER, 2000, 8.0, heuristic="LF", arrival=0.0)
    srv = gcx.Server(workers=2, budget=gcx.request_footprint(small) * 2)
    out = srv.run([small, big])
    assert out[0].status == 0 and out[0].colors > 0
    assert out[1].status == gcx.ERR_BUDGET
    srv.close()


def test_serve_resident_graphs():
    """Requests on registered device graphs equal direct runs on them, alone
    and mixed with generated jobs; the graph is only read."""
    models = {l: gcx.Model.layers_file(l) for l in (2, 3)}
    gs = [gcx.Graph.rmat(14, 16, seed=11), gcx.Graph.er(20000, 320000, seed=12)]
    srv = gcx.Server(workers=4, budget=8 << 30)
    ids = [srv.add_graph(g) for g in gs]
    reqs, t = [], 0.0
    for i in range(16):
        h = ["FF", "SL", "GNN2", "GNN3"][i % 4]
        prec = gcx.BF16 if i % 3 == 0 else gcx.FP32
        sched = gcx.SCHED_ORDERED if i % 2 else gcx.SCHED_ROUNDS
        if i % 5 == 4:
            reqs.append(gcx.make_request(gcx.GEN_ER, 5000, 16.0, seed=i, heuristic=h,
                                         precision=prec, schedule=sched, arrival=t))
        else:
            reqs.append(gcx.make_request(gcx.GEN_RESIDENT, ids[i % 2], heuristic=h,
                                         precision=prec, schedule=sched, arrival=t))
        t += 0.0005
    out = srv.run(reqs)
    assert all(r.status == 0 for r in out), [r.status for r in out]
    for q, r in zip(reqs, out):
        if q.generator != gcx.GEN_RESIDENT:
            continue
        g = gs[q.n]
        assert (r.N, r.M) == (g.N, g.M)
        assert r.footprint == gcx.workspace_footprint(g.N, g.M)
        ws = gcx.Workspace(g.N, g.M)
        ws.set_schedule(q.schedule)
        if q.heuristic == gcx.GNN:
            ws.order(g, gcx.GNN, models[q.layers], q.precision)
        else:
            ws.order(g, q.heuristic)
        ws.color(g)
        assert (ws.validate(g), ws.digest(g)) == (r.colors, r.digest)
    bad = srv.run([gcx.make_request(gcx.GEN_RESIDENT, 99, heuristic="FF")])[0]
    assert bad.status == gcx.ERR_ARG
    srv.close()
    for g in gs:
        assert g.check()


def test_serve_batching_matches_alone():
    """Continuous batching packs compatible queued jobs on resident graphs
    into one pass; every reply equals the job run alone, and the packed
    passes stay within the budget."""
    gs = [gcx.Graph.rmat(13, 16, seed=21), gcx.Graph.er(6000, 60000, seed=22),
          gcx.Graph.rgg(8000, float(np.sqrt(12.0 / (np.pi * 8000))), seed=23)]
    reqs = []
    for i in range(24):
        h = ["GNN3", "GNN2", "LF", "SL"][(i // 6) % 4]   # runs of 6 compatible jobs
        prec = gcx.BF16 if i % 12 < 6 else gcx.FP32
        reqs.append(gcx.make_request(gcx.GEN_RESIDENT, i % 3, heuristic=h, precision=prec,
                                     schedule=gcx.SCHED_ORDERED, arrival=0.0))
    results = {}
    for batching in (0, 4):
        srv = gcx.Server(workers=2, budget=16 << 30)
        ids = [srv.add_graph(g) for g in gs]
        assert ids == [0, 1, 2]
        srv.set_batching(batching)
        out = srv.run(reqs)
        assert all(r.status == 0 for r in out), [r.status for r in out]
        assert srv.peak_reserved() <= 16 << 30
        results[batching] = out
        srv.close()
    alone, batched = results[0], results[4]
    assert all(r.batch == 1 for r in alone)
    assert max(r.batch for r in batched) > 1  # some passes really were packed
    for a, b in zip(alone, batched):
        assert (a.N, a.M, a.colors, a.digest) == (b.N, b.M, b.colors, b.digest)
    # a tight budget still admits every job (smaller or no batches)
    srv = gcx.Server(workers=2, budget=2 * gcx.workspace_footprint(gs[0].N, gs[0].M)
                     + 3 * gcx.graph_footprint(gs[0].N, gs[0].M))
    for g in gs:
        srv.add_graph(g)
    srv.set_batching(8)
    out = srv.run(reqs[:8])
    assert [(r.colors, r.digest) for r in out] == [(r.colors, r.digest) for r in alone[:8]]
    srv.close()


def test_serve_express_small_jobs_overtake():
    """Express workers take small jobs ahead of larger queued ones: replies
    equal the first-come-first-served run, large jobs never run on an
    express worker, and the first small job queued behind three large builds
    is admitted before the second large one (first come, first served
    admits it after the third). Completion times are not compared: a small
    job shares the GPU with the large job running beside it."""
    large = [gcx.make_request(gcx.GEN_RMAT, 1 << 20, 16.0, seed=31 + i, heuristic="FF")
             for i in range(3)]
    small = [gcx.make_request(gcx.GEN_ER, 2000 + 500 * i, 8.0, seed=41 + i, heuristic="LF")
             for i in range(4)]
    reqs = large + small  # all queued at t = 0, large first
    limit = max(gcx.request_footprint(q) for q in small)
    assert limit < min(gcx.request_footprint(q) for q in large)
    srv = gcx.Server(workers=2, budget=16 << 30)
    with pytest.raises(gcx.GcxError):
        srv.set_express(2, limit)  # needs at least one general worker
    fifo = srv.run(reqs)
    srv.set_express(1, limit)
    expr = srv.run(reqs)
    srv.set_express(0, 0)
    srv.close()
    assert all(r.status == 0 for r in fifo + expr)
    assert [(r.colors, r.digest) for r in fifo] == [(r.colors, r.digest) for r in expr]
    assert all(r.worker == 1 for r in expr[:3])
    assert expr[3].t_admit < expr[1].t_admit, [(r.t_admit, r.t_done) for r in expr]
    assert fifo[3].t_admit >= fifo[2].t_admit, [(r.t_admit, r.t_done) for r in fifo]
    assert sum(r.t_admit for r in expr[3:]) < sum(r.t_admit for r in fifo[3:])
