#!/usr/bin/env python3
"""Headline benchmark: GNN3 ordering + Jones-Plassmann colouring on one MI355X.

Workload (one "step"): GraphSAGE-16 GNN3 priorities (models/m3.gnn) followed by
Jones-Plassmann colouring of an R-MAT scale-22 graph (4,194,304 vertices,
edge factor 16, a/b/c/d = .57/.19/.19/.05, seed 1, symmetrised, self loops
and repeats removed, int32 CSR), with the CSR already resident in HBM. This
is the FP32 path, bit-exact with the reference. The metric is colored edges/s
= CSR entries (V[N]) / seconds per step. A second leg times the BF16 GNN
(bf16 rows, MFMA layer products, FP32 accumulation) on the same graph and
reports it beside the FP32 value with its colour count (must be within +1).

Prints ONE JSON line. `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gnn-coloring_amd"))

METRIC = "GNN3+JP colored edges/s BF16 vs FP32 at equal #colors; serving p50/p99 ms"
HBM_PEAK_GBS = 8000.0
# kernels behind the timed stages (the JP stage's kernel depends on the schedule)
KERNEL_OF_STAGE = {"jp_init": "pull_orient_quads + pull_orient_hub_chunks",
                   "gnn_hid": "sage_quads<1> + sage_hubs<1>", "gnn_out": "sage_quads<2> + sage_hubs<2>",
                   "gnn_in": "sage_lanes_input + sage_hubs<0>"}
# the JP stage's kernel by schedule
JP_KERNEL = {"ordered": "jp_pull_kernel", "rounds": "jp_round_kernel"}


# Algorithmic bytes per launch (model: every neighbour-row gather counted once,
# 64-byte FP32 rows, 8-byte priorities; N vertices, M CSR entries).
# JP, ordered schedule (pulled colours): the orientation reads the greedy
# order and scatters positions (8N), reads V (4N), E (4M) and each
# neighbour's position (4M), writes the predecessor positions (M/2 of them,
# 2M) and a header and colour slot per position (12N); the colouring kernel
# reads a header and id per position (12N), each predecessor's position and
# colour (M/2 of each, 4M) and writes two colours (8N). The round schedule:
# orientation 8N + 12M, rounds 8N + 8M.
def algorithmic_bytes(stage, N, M, launches=1, R=64, schedule="ordered"):
    per_stage = {
        "gnn_in": 4 * N + 12 * M + R * N,
        "gnn_hid": 4 * N + (4 + R) * M + 2 * R * N,
        "gnn_out": 4 * N + (4 + R) * M + R * N + 8 * N,
        "jp_init": 24 * N + 10 * M if schedule == "ordered" else 8 * N + 12 * M,
        "jp_rounds": 20 * N + 4 * M if schedule == "ordered" else 8 * N + 8 * M,
    }
    return per_stage[stage] / max(launches, 1)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def host_cpu():
    """CPU model, sockets and cores of this host (lscpu), for the baseline line."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except Exception as e:  # lscpu missing: say so
        info["lscpu"] = f"unavailable: {e}"
    info["nproc"] = os.cpu_count()
    return info


def cpu_baseline(g, scale):
    """The reference's own OpenMP CPU path (oracle/_ref/libref.so, its C
    sources compiled here) on the same graph: cgnn_graph_sage_16_par +
    jones_plassmann_internal, setup untimed (runner.c:99-134). Each run is a
    child process (oracle/cpu_baseline.py) started with OMP_PLACES=cores and
    OMP_PROC_BIND=spread (README.md:8-14): at this host's thread share for
    one GPU (OMP_NUM_THREADS, 16 on the GPU boxes; best of 5, as main.c:199)
    and at 1 thread (best of 3, to bound the run to about 25 s)."""
    import subprocess
    import tempfile
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not os.path.exists(ref_so):
        return {"value": None, "unit": "colored edges/s", "cores": 0, "kind": "reference",
                "sample": f"unavailable: {ref_so} was not built"}
    V, E = g.download()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    runs = {}
    with tempfile.TemporaryDirectory() as d:
        np.save(os.path.join(d, "V.npy"), V)
        np.save(os.path.join(d, "E.npy"), E)
        del V
        for threads, reps in ((share, 5), (1, 3)):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
                       OMP_PROC_BIND="spread")
            r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"),
                                "--csr", d, "--model", os.path.join(ROOT, "models", "m3.gnn"),
                                "--reps", str(reps)], env=env, capture_output=True, text=True,
                               timeout=600)
            if r.returncode != 0:
                return {"value": None, "unit": "colored edges/s", "cores": 0, "kind": "reference",
                        "sample": f"failed: {r.stderr[-500:]}"}
            runs[threads] = json.loads(r.stdout.strip().splitlines()[-1])
    M = len(E)
    full, one = runs[share], runs[1]
    return {"value": float(M / full["seconds"]), "unit": "colored edges/s",
            "cores": full["threads"], "kind": "reference",
            "seconds_per_step": full["seconds"], "colors": full["colors"],
            "sample": f"full R-MAT scale-{scale} graph, GNN3+JP; {full['threads']} OpenMP threads "
                      "(this host's share for one GPU) best of 5, OMP_PLACES=cores "
                      "OMP_PROC_BIND=spread",
            "single_thread": {"value": float(M / one["seconds"]), "seconds_per_step": one["seconds"],
                              "colors": one["colors"], "reps": 3},
            "build": "gcc -std=gnu17 -O3 -march=x86-64-v3 -fopenmp (the reference Makefile uses "
                     "-march=native; the library is built in the development container and must "
                     "run on this host; every GNN floating-point operation is an explicit AVX2/FMA "
                     "intrinsic, so results do not change)",
            "host_cpu": host_cpu()}


def serving(args, gcx):
    """Serving leg: the same seeded request mix (R-MAT / ER / RGG, FF LF SL
    GNN2 GNN3) run three ways: BF16 GNN on `workers` streams with all
    requests queued at t = 0 (capacity), BF16 with Poisson arrivals at 70% of
    that capacity (latency), and the plain FP32 path one graph at a time
    (1 worker, all queued, cooperative ordered launches). The BF16 passes
    share the device through plain ordered launches in `slots` slots."""
    import numpy as np
    k = args.serving_requests
    base = gcx.serving_trace(k, seed=11, nmax=args.serving_nmax)

    express_bytes = int(args.serving_express_mb * (1 << 20))

    def summary(rep, wall):
        ok = [r for r in rep if r.status == 0]
        lat = np.array([1000.0 * (r.t_done - r.t_arrival) for r in ok])
        edges = float(sum(r.M for r in ok))
        out = {"requests": len(rep), "ok": len(ok), "seconds": wall,
               "graphs_per_s": len(ok) / wall, "colored_edges_per_s": edges / wall,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "mean_ms": float(lat.mean())}
        # latency by size class: jobs of at most the express footprint, and the rest
        small = np.array([r.footprint <= express_bytes for r in ok])
        for name, m in (("small", small), ("large", ~small)):
            if m.any():
                out[f"{name}_jobs"] = int(m.sum())
                out[f"{name}_p50_ms"] = float(np.percentile(lat[m], 50))
                out[f"{name}_p99_ms"] = float(np.percentile(lat[m], 99))
        out["mean_wait_ms"] = float(np.mean([1000.0 * (r.t_admit - r.t_arrival) for r in ok]))
        return out

    def run(srv, reqs):
        t0 = time.perf_counter()
        rep = srv.run(reqs)
        return rep, time.perf_counter() - t0

    workers = args.serving_mixed_workers
    prev = gcx.get_ordered_launch()
    gcx.set_ordered_launch(args.serving_mixed_slots)
    srv = gcx.Server(workers=workers, budget=int(args.serving_budget_gb * (1 << 30)))
    srv.run(gcx.with_arrivals(base[:min(k, 16)], None))  # warm the pool and the kernels
    if args.serving_rate:
        # one Poisson pass at a given rate (long traces), in segments of 1000
        # requests that each start from an empty queue; a line per segment
        rep2, wall2, seg = [], 0.0, 1000
        for i in range(0, k, seg):
            r, w = run(srv, gcx.with_arrivals(base[i:i + seg], args.serving_rate, seed=12 + i))
            rep2 += r
            wall2 += w
            print(f"[bench] serving segment {i // seg + 1}/{(k + seg - 1) // seg}: "
                  f"{sum(1 for x in r if x.status == 0)}/{len(r)} ok in {w:.1f} s",
                  file=sys.stderr, flush=True)
        srv.close()
        gcx.set_ordered_launch(prev)
        return {"requests": k, "workers": workers, "ordered_launch_slots": args.serving_mixed_slots,
                "budget_gb": args.serving_budget_gb,
                "trace": f"seed 11, n log-uniform [1e3, {args.serving_nmax:.0e}], 40% R-MAT / 30% ER "
                         "(degree 8-64) / 30% RGG (degree ~12), FF/LF/SL/GNN2/GNN3, GNN in BF16",
                "bf16_poisson": dict(summary(rep2, wall2), rate_per_s=args.serving_rate,
                                     segments=f"{(k + 999) // 1000} x 1000 requests, each from an "
                                              "empty queue")}
    rep, wall = run(srv, gcx.with_arrivals(base, None))
    cap = summary(rep, wall)
    rate = 0.7 * cap["graphs_per_s"]
    rep2, wall2 = run(srv, gcx.with_arrivals(base, rate, seed=12))
    lat = summary(rep2, wall2)
    peak = srv.peak_reserved()
    # the same arrivals with express workers: small jobs may overtake large ones
    srv.set_express(args.serving_express_workers, express_bytes)
    rep4, wall4 = run(srv, gcx.with_arrivals(base, rate, seed=12))
    lat_x = summary(rep4, wall4)
    srv.set_express(0, 0)
    srv.close()
    gcx.set_ordered_launch(0)
    plain = gcx.Server(workers=1, budget=int(args.serving_budget_gb * (1 << 30)))
    plain.run(gcx.with_arrivals(base[:min(k, 16)], None, precision=gcx.FP32))
    rep3, wall3 = run(plain, gcx.with_arrivals(base, None, precision=gcx.FP32))
    fp32 = summary(rep3, wall3)
    plain.close()
    gcx.set_ordered_launch(prev)
    same = sum(1 for a, b, c in zip(rep, rep2, rep4)
               if (a.colors, a.digest) == (b.colors, b.digest) == (c.colors, c.digest))
    worse = [b.colors - a.colors for a, b in zip(rep3, rep) if a.status == 0 and b.status == 0]
    return {"requests": k, "workers": workers, "ordered_launch_slots": args.serving_mixed_slots,
            "budget_gb": args.serving_budget_gb, "peak_reserved_gb": peak / (1 << 30),
            "trace": f"seed 11, n log-uniform [1e3, {args.serving_nmax:.0e}], 40% R-MAT / 30% ER "
                     "(degree 8-64) / 30% RGG (degree ~12), FF/LF/SL/GNN2/GNN3, GNN in BF16",
            "bf16_capacity": cap, "bf16_poisson": dict(lat, rate_per_s=rate),
            "bf16_poisson_express": dict(lat_x, rate_per_s=rate,
                                         express_workers=args.serving_express_workers,
                                         express_max_footprint_mb=args.serving_express_mb),
            "plain_fp32_one_at_a_time": fp32,
            "capacity_speedup_vs_plain_fp32": cap["colored_edges_per_s"] / fp32["colored_edges_per_s"],
            "same_result_queued_vs_poisson_vs_express": f"{same}/{k}",
            "bf16_minus_fp32_colors_max": max(worse) if worse else None}


def serving_rmat(args, gcx):
    """Headline serving leg: GNN3 + JP jobs on resident R-MAT graphs of the
    bench scale (seeds 1..4, CSR kept in HBM), all queued at t = 0. BF16 on
    `workers` streams against the plain FP32 path (one job at a time on one
    stream), then BF16 with Poisson arrivals at 70% of its capacity."""
    import numpy as np
    k = args.rmat_jobs
    gs = [gcx.Graph.rmat(args.scale, 16, seed=s) for s in range(1, 5)]

    def jobs(ids, precision, rate=None, seed=5):
        rng = np.random.default_rng(seed)
        out, t = [], 0.0
        for i in range(k):
            if rate:
                t += float(rng.exponential(1.0 / rate))
            out.append(gcx.make_request(gcx.GEN_RESIDENT, ids[i % len(ids)], heuristic="GNN3",
                                        precision=precision, schedule=gcx.SCHED_ORDERED,
                                        arrival=t))
        return out

    def measure(workers, precision, rate=None, batch=0):
        srv = gcx.Server(workers=workers, budget=int(args.serving_budget_gb * (1 << 30)))
        ids = [srv.add_graph(g) for g in gs]
        srv.set_batching(batch)
        srv.run(jobs(ids, precision)[:min(k, 2 * workers)])  # warm the pool and the kernels
        reqs = jobs(ids, precision, rate)
        t0 = time.perf_counter()
        rep = srv.run(reqs)
        wall = time.perf_counter() - t0
        srv.close()
        ok = [r for r in rep if r.status == 0]
        lat = np.array([1000.0 * (r.t_done - r.t_arrival) for r in ok])
        return rep, {"jobs": k, "ok": len(ok), "workers": workers, "batch": batch,
                     "mean_batch": float(np.mean([r.batch for r in ok])) if ok else 0.0,
                     "seconds": wall,
                     "graphs_per_s": len(ok) / wall,
                     "colored_edges_per_s": float(sum(r.M for r in ok)) / wall,
                     "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99))}

    prev = gcx.get_ordered_launch()
    gcx.set_ordered_launch(0)
    plain_rep, plain = measure(1, gcx.FP32)
    sweep = {}
    for slots in args.ordered_slots:
        gcx.set_ordered_launch(slots)
        sweep[slots] = measure(args.serving_workers, gcx.BF16)
    best = max(sweep, key=lambda k: sweep[k][1]["colored_edges_per_s"])
    gcx.set_ordered_launch(best)
    bf_rep, bf = sweep[best]
    lat_rep, lat = measure(args.serving_workers, gcx.BF16, rate=0.7 * bf["graphs_per_s"])
    x3_rep, x3 = measure(args.serving_workers, gcx.BF16X3)
    batched = {}
    for b in args.serving_batch:
        batched[b] = measure(args.serving_workers, gcx.BF16, batch=b)
    gcx.set_ordered_launch(prev)
    fp32_col = {r.digest: r.colors for r in plain_rep}
    per_graph = {}
    for q, a, b, c in zip(range(k), plain_rep, bf_rep, x3_rep):
        per_graph.setdefault(q % len(gs), set()).add((a.colors, b.colors, c.colors, a.digest,
                                                      b.digest, c.digest))
    same = all(len(v) == 1 for v in per_graph.values())  # every job on a graph gives one result
    delta = [next(iter(v))[1] - next(iter(v))[0] for v in per_graph.values()]
    delta3 = [next(iter(v))[2] - next(iter(v))[0] for v in per_graph.values()]
    batched_same = all((a.colors, a.digest) == (b.colors, b.digest)
                       for rb, _ in batched.values() for a, b in zip(bf_rep, rb))
    best_b = max(batched, key=lambda b: batched[b][1]["colored_edges_per_s"]) if batched else None
    for g in gs:
        g.free()
    return {"graphs": f"R-MAT scale {args.scale} (edge factor 16, a/b/c .57/.19/.19), seeds 1-4, "
                      "resident in HBM",
            "job": "GNN3 priorities + ordered JP + validate + digest",
            "plain_fp32_one_at_a_time": plain, "bf16_capacity": bf,
            "ordered_launch_slots": best,
            "bf16_graphs_per_s_by_slots": {str(k): v[1]["graphs_per_s"] for k, v in sweep.items()},
            "bf16_poisson_70pct": lat,
            "speedup_vs_plain_fp32": bf["colored_edges_per_s"] / plain["colored_edges_per_s"],
            "bf16x3_capacity": x3,
            "bf16_batched": {str(b): v[1] for b, v in batched.items()},
            "bf16_batched_same_results": batched_same,
            "bf16_batched_speedup_vs_plain_fp32": (batched[best_b][1]["colored_edges_per_s"]
                                                   / plain["colored_edges_per_s"]) if best_b else None,
            "bf16x3_speedup_vs_plain_fp32": x3["colored_edges_per_s"] / plain["colored_edges_per_s"],
            "bf16x3_minus_fp32_colors": delta3,
            "same_result_every_job_on_a_graph": same,
            "fp32_colors_by_graph": sorted(set(fp32_col.values())),
            "bf16_minus_fp32_colors": delta,
            "same_result_in_poisson_run": sum(1 for a, b in zip(bf_rep, lat_rep)
                                              if (a.colors, a.digest) == (b.colors, b.digest))}


def batching(args, gcx):
    """Batching leg: `batch_graphs` seeded Erdos-Renyi graphs (n log-uniform
    in [1e3, 2e5], average degree uniform in [4, 64]) coloured with GNN3 FP32
    priorities one graph at a time and packed into one block-diagonal graph
    (one launch sequence, one persistent JP launch), schedules and hub
    degrees from the autotuner; plus `culberson_passes` iterated-greedy
    passes (mixed class orders, seed per pass) of an R-MAT scale-20 graph as
    captured graph replays vs one gcx_color per pass."""
    import numpy as np
    rng = np.random.default_rng(21)
    k = args.batch_graphs
    model = gcx.Model.layers_file(3)
    specs = []
    for i in range(k):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        d = float(rng.uniform(4, 64))
        specs.append((n, int(round(n * d / 2)), 1000 + i))
    graphs = [gcx.Graph.er(n, m, seed=sd) for n, m, sd in specs]
    # autotuner: one measurement per new degree signature
    t0 = time.perf_counter()
    sigs = {}
    for g in graphs:
        if g.signature not in sigs:
            sigs[g.signature] = gcx.tune(g, model, gcx.FP32).as_dict()
    tune_s = time.perf_counter() - t0
    for g in graphs:  # graphs made before their signature was tuned
        g.set_hub_degree(0)
    maxN = max(g.N for g in graphs)
    maxM = max(g.M for g in graphs)
    ws = gcx.Workspace(maxN, maxM)
    ws.set_schedule(gcx.SCHED_AUTO)
    stream = gcx.Stream()

    def one_at_a_time():
        for g in graphs:
            ws.order(g, gcx.GNN, model, gcx.FP32, stream)
            ws.color(g, stream)
        stream.sync()

    one_at_a_time()
    t0 = time.perf_counter()
    one_at_a_time()
    single_s = time.perf_counter() - t0
    alone = []
    for g in graphs:
        ws.order(g, gcx.GNN, model, gcx.FP32, stream)
        ws.color(g, stream)
        alone.append((ws.validate(g, stream), ws.digest(g, stream)))
    t0 = time.perf_counter()
    pack = gcx.Graph.pack(graphs, stream)
    pack_s = time.perf_counter() - t0
    tp = gcx.tune(pack, model, gcx.FP32).as_dict()
    wsp = gcx.Workspace(pack.N, pack.M)
    wsp.set_schedule(gcx.SCHED_AUTO)

    def packed():
        wsp.order(pack, gcx.GNN, model, gcx.FP32, stream)
        wsp.color(pack, stream)
        stream.sync()

    packed()
    t0 = time.perf_counter()
    packed()
    packed_s = time.perf_counter() - t0
    colors, digests = wsp.packed_colors(pack, stream)
    same = sum(1 for i, a in enumerate(alone) if a == (int(colors[i]), int(digests[i])))
    M = sum(g.M for g in graphs)
    out = {"graphs": k, "vertices": int(pack.N), "entries": int(M),
           "one_at_a_time_s": single_s, "packed_s": packed_s, "pack_build_s": pack_s,
           "packed_speedup": single_s / packed_s,
           "packed_colored_edges_per_s": M / packed_s,
           "one_at_a_time_colored_edges_per_s": M / single_s,
           "same_colours_as_alone": f"{same}/{k}",
           "tuned_signatures": len(sigs), "tune_s": tune_s, "pack_tuning": tp,
           "schedules_chosen": {"rounds": sum(1 for t in sigs.values() if t["schedule"] == 0),
                                "ordered": sum(1 for t in sigs.values() if t["schedule"] == 1)}}
    pack.free()
    for g in graphs:
        g.free()
    # Culberson chain on R-MAT 20
    g = gcx.Graph.rmat(20, 16, seed=1)
    w = gcx.Workspace(g.N, g.M)
    w.order(g, gcx.GNN, model, gcx.FP32, stream)
    w.color(g, stream)
    c0 = w.colors(g.N)
    k0 = w.validate(g, stream)
    res = {}
    for name, sched, flags in (("rounds_captured", gcx.SCHED_ROUNDS, 0),
                               ("rounds_one_call_per_pass", gcx.SCHED_ROUNDS, gcx.CULB_NO_GRAPH),
                               ("ordered_one_call_per_pass", gcx.SCHED_ORDERED, 0)):
        w.set_schedule(sched)
        w.upload_colors(c0)
        w.recolor_ex(g, 2, gcx.CULB_MIXED | flags, seed=7, stream=stream)  # warm
        w.upload_colors(c0)
        stream.sync()
        t0 = time.perf_counter()
        w.recolor_ex(g, args.culberson_passes, gcx.CULB_MIXED | flags, seed=7, stream=stream)
        stream.sync()
        dt = time.perf_counter() - t0
        res[name] = {"ms_per_pass": 1000.0 * dt / args.culberson_passes,
                     "colors": w.validate(g, stream), "digest": w.digest(g, stream)}
    out["culberson"] = {"graph": "R-MAT scale 20, GNN3 FP32 start", "passes": args.culberson_passes,
                        "strategy": "mixed (reverse / largest first / random, seed 7 + pass)",
                        "start_colors": k0, **res,
                        "same_result": len({r["digest"] for r in res.values()}) == 1,
                        "captured_speedup_rounds": res["rounds_one_call_per_pass"]["ms_per_pass"]
                        / res["rounds_captured"]["ms_per_pass"]}
    return out


# ------------------------------------------------ rocprofv3 database summaries
#
# `bench.py --profile-summary kernels.db [fetch.db write.db]` prints the
# per-kernel table of a --kernel-trace run and, for --pmc runs, each counter
# per dispatch (FETCH_SIZE / WRITE_SIZE in KiB); given the FETCH_SIZE and
# WRITE_SIZE runs of the same command, also the HBM bytes per step of every
# timed stage (JSON for profiles/traffic.json).
# Counter calibration: the streaming copy kernel (copy_u32x4, 1 GiB read and
# 1 GiB written per dispatch) gives the factor from each counter's KiB to
# bytes.

# kernels of each timed stage (FP32 GNN3 + ordered JP); the JP sort's radix
# passes are left out (their rocPRIM kernels also build the CSR)
STAGE_KERNELS = {
    "gnn_in": ("sage_lanes_input", "sage_hubs<0>"),
    "gnn_hid": ("sage_quads<1>", "sage_hubs<1>"),
    "gnn_out": ("sage_quads<2>", "sage_hubs<2>"),
    "jp_init": ("pull_clear", "pull_rank", "pull_orient_quads", "pull_orient_hub_chunks",
                "pull_hub_headers", "jp_order_keys"),
    "jp_rounds": ("jp_pull_kernel",),
}


def _short_kernel(n):
    n = n.replace("(anonymous namespace)::", "")
    depth = 0
    for i in range(len(n) - 1, -1, -1):  # drop the argument list
        if n[i] == ")":
            depth += 1
        elif n[i] == "(":
            depth -= 1
            if depth == 0:
                n = n[:i]
                break
    n = n.replace("void ", "")
    if "rocprim::" in n:
        n = "rocprim " + ("radix sort" if "sort" in n else "scan" if "scan" in n else "kernel")
    return n


def profile_summary(paths):
    import sqlite3
    out = {}
    for p in paths:
        c = sqlite3.connect(p)
        views = {r[0] for r in c.execute("select name from sqlite_master")}
        pmc = "pmc_events" in views and c.execute("select count(*) from pmc_events").fetchone()[0]
        if pmc:
            # a counter may come as several rows per dispatch (one per hardware
            # instance): sum them per dispatch, then average over dispatches
            rows = c.execute("select name, counter_name, count(*), avg(v) from (select name, "
                             "counter_name, dispatch_id, sum(counter_value) as v from pmc_events "
                             "group by name, counter_name, dispatch_id) group by name, "
                             "counter_name").fetchall()
            agg = {}
            for n, cn, k, v in rows:
                key = (_short_kernel(n), cn)
                a = agg.setdefault(key, [0, 0.0])
                a[1] = (a[1] * a[0] + v * k) / (a[0] + k)
                a[0] += k
            print(f"\n### {os.path.basename(p)}: counter averages per dispatch\n")
            print("| kernel | counter | dispatches | value per dispatch |")
            print("|---|---|---|---|")
            for (n, cn), (k, v) in sorted(agg.items(), key=lambda x: -x[1][1] * x[1][0]):
                print(f"| `{n}` | {cn} | {k} | {v:.1f} |")
            for (n, cn), (k, v) in agg.items():
                out.setdefault(cn, {})[n] = v
        else:
            rows = c.execute("select name, count(*), sum(duration), avg(duration), min(duration), "
                             "max(duration) from kernels group by name").fetchall()
            agg = {}
            for n, k, tot, a, mi, ma in rows:
                sn = _short_kernel(n)
                g = agg.setdefault(sn, [0, 0.0, float("inf"), 0.0])
                g[0] += k
                g[1] += tot
                g[2] = min(g[2], mi)
                g[3] = max(g[3], ma)
            total = sum(g[1] for g in agg.values())
            print(f"\n### {os.path.basename(p)}: kernel times\n")
            print("| kernel | calls | total ms | avg us | min us | max us | % of GPU time |")
            print("|---|---|---|---|---|---|---|")
            for n, (k, tot, mi, ma) in sorted(agg.items(), key=lambda x: -x[1][1])[:45]:
                print(f"| `{n}` | {k} | {tot / 1e6:.3f} | {tot / k / 1e3:.1f} | {mi / 1e3:.1f} | "
                      f"{ma / 1e3:.1f} | {100 * tot / total:.1f} |")
    if "FETCH_SIZE" in out and "WRITE_SIZE" in out:
        f, w = out["FETCH_SIZE"], out["WRITE_SIZE"]
        ff = (1 << 30) / (1024.0 * f["copy_u32x4"]) if f.get("copy_u32x4") else 1.0
        fw = (1 << 30) / (1024.0 * w["copy_u32x4"]) if w.get("copy_u32x4") else 1.0
        print(f"\ncalibration (copy_u32x4, 1 GiB each way): FETCH_SIZE x{ff:.3f}, WRITE_SIZE x{fw:.3f}")
        traffic = {}
        for st, ks in STAGE_KERNELS.items():
            b = 0.0
            for kn in set(f) | set(w):
                if any(kn.startswith(k) for k in ks):
                    b += 1024.0 * (f.get(kn, 0.0) * ff + w.get(kn, 0.0) * fw)
            traffic[st] = int(b)
        print("stage HBM bytes per step (profiles/traffic.json, key rmat<scale>_<stage>):")
        print(json.dumps(traffic, indent=1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=1000)
    ap.add_argument("--rmat-jobs", type=int, default=32)
    ap.add_argument("--batch-graphs", type=int, default=512)
    ap.add_argument("--culberson-passes", type=int, default=64)
    ap.add_argument("--ordered-slots", type=int, nargs="+", default=[0, 1, 2, 4])
    ap.add_argument("--serving-batch", type=int, nargs="*", default=[4, 8],
                    help="continuous-batching sizes for the resident R-MAT serving leg")
    ap.add_argument("--serving-nmax", type=float, default=4e6)
    ap.add_argument("--serving-workers", type=int, default=8,
                    help="worker streams of the resident R-MAT serving leg")
    ap.add_argument("--serving-mixed-workers", type=int, default=16,
                    help="worker streams of the mixed-trace serving leg")
    ap.add_argument("--serving-mixed-slots", type=int, default=2,
                    help="ordered-launch slots of the mixed-trace BF16 passes")
    ap.add_argument("--serving-budget-gb", type=float, default=48.0)
    ap.add_argument("--serving-express-workers", type=int, default=4,
                    help="mixed-trace workers reserved for small jobs in the express Poisson pass")
    ap.add_argument("--serving-express-mb", type=float, default=256.0,
                    help="largest job footprint (MiB) the express workers admit; also the "
                         "small/large split of the latency report")
    ap.add_argument("--serving-rate", type=float, default=0.0,
                    help="run only a Poisson pass at this many requests/s (0 = capacity, "
                         "70%% Poisson and plain FP32 passes)")
    ap.add_argument("--schedule", choices=["auto", "rounds", "ordered"], default="auto")
    ap.add_argument("--head", type=int, default=None,
                    help="longest head prefix of the ordered schedule (0 = off; default: the "
                         "library's)")
    ap.add_argument("--head-density", type=int, default=64)
    ap.add_argument("--profile-summary", nargs="+", metavar="DB",
                    help="summarise rocprofv3 databases (kernel trace; FETCH_SIZE and WRITE_SIZE "
                         "runs) and exit")
    args = ap.parse_args()
    if args.profile_summary:
        profile_summary(args.profile_summary)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)

    import gcx

    gcx.init()
    t = time.time()
    g = gcx.Graph.rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1)
    N, M = g.N, g.M
    print(f"[bench] graph R-MAT scale {args.scale}: N={N} M={M} maxdeg={g.max_degree} "
          f"({time.time() - t:.1f}s to generate + build CSR on the GPU)", file=sys.stderr, flush=True)
    model = gcx.Model.layers_file(3)
    ws = gcx.Workspace(N, M)
    if args.head is not None:
        ws.set_ordered_head(args.head, args.head_density)
    stream = gcx.Stream()
    # one-time per-graph preparation (degree order, hub tables, Er), outside
    # the timed step: best of 3
    prep_ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        g.prepare(stream)
        prep_ms.append(1000.0 * (time.perf_counter() - t0))
    graph_prep_ms = min(prep_ms)
    # streaming-copy reference: 1 GiB read + 1 GiB written by a plain kernel
    copy_ms = min(gcx.calibration_copy(1 << 30) for _ in range(3))
    copy_gbs = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9

    def step():
        ws.order(g, gcx.GNN, model, gcx.FP32, stream)
        ws.color(g, stream)

    # schedule choice: time each JP schedule on warm-up steps, keep the faster
    sched_names = {"rounds": (gcx.SCHED_ROUNDS, 0), "ordered": (gcx.SCHED_ORDERED, 0)}
    cands = list(sched_names) if args.schedule == "auto" else [args.schedule]
    sched_ms = {}
    depth = None
    for name in cands:
        ws.set_schedule(*sched_names[name])
        for _ in range(max(args.warmup, 1)):
            t0 = time.perf_counter()
            step()
            stream.sync()
            sched_ms[name] = min(sched_ms.get(name, 1e30), 1000 * (time.perf_counter() - t0))
        if name == "rounds":
            depth = ws.rounds()
    best = min(sched_ms, key=sched_ms.get)
    ws.set_schedule(*sched_names[best])
    stream.sync()
    gcx.profile_enable(True)
    gcx.profile_read()  # clears
    e0, e1 = gcx.Event(), gcx.Event()
    gcx.device_sync()
    e0.record(stream)
    w0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(stream)
    stream.sync()
    wall = time.perf_counter() - w0
    ev_ms = e0.elapsed_ms(e1)
    prof = gcx.profile_read()
    gcx.profile_enable(False)
    launches = ws.launches()
    fell_back = ws.fell_back()

    ms_per_step = 1000.0 * wall / args.steps
    value = M / (wall / args.steps)

    # correctness of what was timed: proper, greedy rule, and (when a golden
    # digest exists for this graph) bit-exact colours vs the reference
    colors = ws.validate(g, stream)
    viol = ws.mex_check(g, stream)
    col = ws.colors(N)
    golden = None
    dpath = os.path.join(ROOT, "tests", "golden", "digests.json")
    if os.path.exists(dpath):
        d = json.load(open(dpath)).get(f"rmat{args.scale}")
        if d:
            golden = (sha(col) == d["GNN3"]["sha_color"]) and colors == d["GNN3"]["ncolors"]
    if colors == 0 or viol != 0 or golden is False:
        print(f"[bench] VALIDATION FAILED colors={colors} mex_violations={viol} golden={golden}",
              file=sys.stderr)
        sys.exit(1)

    # roofline of the dominant stage (largest total time over the timed steps)
    stage_ms = {k: v[0] for k, v in prof.items()}
    counts = {k: v[1] for k, v in prof.items()}
    dominant = max(stage_ms, key=lambda k: stage_ms[k])
    if dominant == "jp_rounds":
        n_launch = launches * args.steps
    else:
        n_launch = counts[dominant]
    avg_ms = stage_ms[dominant] / max(n_launch, 1)
    per_launch = algorithmic_bytes(dominant, N, M, launches if dominant == "jp_rounds" else 1,
                                   schedule=best)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    prof_path = os.path.join(ROOT, "profiles", "traffic.json")
    tmap = json.load(open(prof_path)) if os.path.exists(prof_path) else {}
    traffic = tmap.get(f"rmat{args.scale}_{dominant}")
    # every stage against the same model (ms per launch from the HIP events)
    stage_roof = {}
    for k in stage_ms:
        if stage_ms[k] <= 0 or k not in ("gnn_in", "gnn_hid", "gnn_out", "jp_init", "jp_rounds"):
            continue
        nl = launches * args.steps if k == "jp_rounds" else counts[k]
        per = algorithmic_bytes(k, N, M, launches if k == "jp_rounds" else 1, schedule=best)
        ms_l = stage_ms[k] / max(nl, 1)
        t = tmap.get(f"rmat{args.scale}_{k}")
        stage_roof[k] = {"ms_per_launch": ms_l, "algorithmic_gbs": per / (ms_l * 1e-3) / 1e9,
                         "frac": per / (ms_l * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "traffic_bytes": t, "traffic_gbs": (t / (ms_l * 1e-3) / 1e9) if t else None}

    # BF16 legs: same graph and schedule, GNN layer products on MFMA. BF16:
    # bf16 rows and operands; BF16X3: FP32 rows, operands split into bf16
    # hi + lo (three MFMAs per product). Colours are compared with FP32.
    def mfma_leg(precision):
        def step16():
            ws.order(g, gcx.GNN, model, precision, stream)
            ws.color(g, stream)
        for _ in range(max(args.warmup, 1)):
            step16()
        stream.sync()
        gcx.profile_enable(True)
        gcx.profile_read()
        w0 = time.perf_counter()
        for _ in range(args.steps):
            step16()
        stream.sync()
        wall16 = time.perf_counter() - w0
        prof16 = gcx.profile_read()
        gcx.profile_enable(False)
        c16 = ws.validate(g, stream)
        v16 = ws.mex_check(g, stream)
        if c16 == 0 or v16 != 0:
            print(f"[bench] MFMA LEG VALIDATION FAILED precision={precision} colors={c16} "
                  f"mex_violations={v16}", file=sys.stderr)
            sys.exit(1)
        return {"value": M / (wall16 / args.steps), "ms_per_step": 1000.0 * wall16 / args.steps,
                "colors": c16, "mex_violations": v16, "colors_minus_fp32": c16 - colors,
                "colors_within_1_of_fp32": 0 < c16 <= colors + 1,
                "speedup_vs_fp32": ms_per_step / (1000.0 * wall16 / args.steps),
                "stage_ms_per_step": {k: v[0] / args.steps for k, v in prof16.items()}}

    bf16 = bf16x3 = None
    if not args.no_bf16:
        bf16 = mfma_leg(gcx.BF16)
        bf16x3 = mfma_leg(gcx.BF16X3)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "colored edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded R-MAT generated on the GPU; weights models/m3.gnn",
        "config": {"workload": f"rmat{args.scale}_gnn3_jp_fp32",
                   "graph": f"R-MAT scale {args.scale}, edge factor 16, a/b/c/d=.57/.19/.19/.05, "
                            "seed 1, symmetrised, simple, int32 CSR",
                   "N": N, "M": M, "heuristic": "GNN3", "precision": "fp32 (bit-exact)",
                   "jp_schedule": best, "graph_prep_ms": graph_prep_ms,
                   "graph_prep": "one-time per graph, not in the step: degree order, hub list "
                                 "and chunk table, rank and Er = rank[E]"},
        "schedule_warmup_ms": sched_ms,
        "roofline": {"bound": "hbm",
                     "kernel": JP_KERNEL[best] if dominant == "jp_rounds"
                     else KERNEL_OF_STAGE.get(dominant, dominant),
                     "stage": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "avg_launch_ms": avg_ms, "launches": n_launch,
                     "algorithmic_bytes_per_launch": per_launch,
                     "note": ("the JP chain is latency-bound: DAG depth x per-level hop time; "
                              "its HBM fraction is reported because it is the dominant kernel")
                     if dominant == "jp_rounds" else None},
        "copy_kernel_gbs": copy_gbs,
        "colors": colors,
        "jp_dag_depth": depth,
        "jp_launches_per_step": launches,
        "jp_ordered_fell_back": fell_back,
        "stage_ms_per_step": {k: v / args.steps for k, v in stage_ms.items()},
        "stage_rooflines": stage_roof,
        "bf16": bf16,
        "bf16x3": bf16x3,
        "event_ms_per_step": ev_ms / args.steps,
        "bit_exact_vs_reference": golden,
    }
    if args.batch_graphs > 0:
        print(f"[bench] batching {args.batch_graphs} graphs ...", file=sys.stderr, flush=True)
        out["batching"] = batching(args, gcx)
    if args.rmat_jobs > 0:
        print(f"[bench] serving {args.rmat_jobs} GNN3 jobs on resident R-MAT graphs ...",
              file=sys.stderr, flush=True)
        out["serving_rmat"] = serving_rmat(args, gcx)
    if args.serving_requests > 0:
        print(f"[bench] serving {args.serving_requests} requests ...", file=sys.stderr, flush=True)
        out["serving"] = serving(args, gcx)
    if not args.no_cpu_baseline:
        print("[bench] timing the reference CPU path on the host ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(g, args.scale)
    else:
        out["cpu_baseline"] = None
    # explicit teardown before exit: objects first, then the library's pools
    ws.free()
    g.free()
    model.free()
    stream.close()
    gcx.shutdown()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
