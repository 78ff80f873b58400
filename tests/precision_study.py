"""GNN precision study (test infrastructure: uses the FP64 CPU oracle).

GNN priorities from the device in FP32, BF16 and BF16X3 against the FP64
network evaluated by the oracle on the CPU, on an R-MAT graph, for
models/m4.gnn and seeded Glorot-uniform weight sets of the same widths:
max / mean relative priority error, edge-orientation flips vs FP64 (the
ordering inversions JP sees) and the colours each precision's order uses.

    python tests/precision_study.py --scale 24 --sets 8 --out report.json

tests/test_gpu_precision.py runs the same functions on a small graph.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, ".."))
sys.path.insert(0, os.path.join(HERE, "..", "gnn-coloring_amd"))

import gcx  # noqa: E402
import oracle  # noqa: E402


def glorot_params(layers, seed):
    """Glorot-uniform GraphSAGE-16 weights (fan in 32, fan out 16:
    U(+-sqrt(6/48))), zero biases, in the models/*.gnn layout."""
    rng = np.random.default_rng(seed)
    lim = float(np.sqrt(6.0 / 48.0))
    out = []
    for _ in range(layers):
        out.append(rng.uniform(-lim, lim, 512).astype(np.float32))
        out.append(np.zeros(16, np.float32))
    return np.concatenate(out)


def edge_pairs(V, E):
    """(u, v) with u < v for every undirected edge of the CSR."""
    V = np.asarray(V, np.int64)
    u = np.repeat(np.arange(len(V) - 1, dtype=np.uint32), np.diff(V))
    v = np.asarray(E, np.uint32)
    keep = u < v
    return u[keep], v[keep]


def orientation_flips(V, E, p, q, pairs=None):
    """Undirected edges whose JP orientation (p descending, id ascending)
    differs between priority arrays p and q."""
    u_all, v_all = pairs if pairs is not None else edge_pairs(V, E)
    flips = 0
    step = 1 << 26
    for k in range(0, len(u_all), step):
        u, v = u_all[k:k + step], v_all[k:k + step]
        pu, pv, qu, qv = p[u], p[v], q[u], q[v]
        # u < v always here, so ties order u first under both arrays
        a = pu >= pv
        b = qu >= qv
        flips += int(np.count_nonzero(a != b))
    return flips


def weight_sets(layers, glorot):
    param, md = oracle.load_model(os.path.join(gcx.MODEL_DIR, f"m{layers}.gnn"))
    sets = [(f"m{layers}", param, md)]
    for i in range(glorot):
        sets.append((f"glorot{i}", glorot_params(layers, 100 + i), 100))
    return sets


def study(scale, glorot, layers=4, log=None):
    gcx.init()
    g = gcx.Graph.rmat(scale, 16, seed=1)
    V, E = g.download()
    N = g.N
    pairs = edge_pairs(V, E)
    ws = gcx.Workspace(N, g.M)
    ws.set_schedule(gcx.SCHED_ORDERED)
    rows = []
    for name, param, md in weight_sets(layers, glorot):
        t0 = time.perf_counter()
        exact = oracle.gnn_fp64(param, V, E, md)
        row = {"weights": name, "fp64_cpu_s": time.perf_counter() - t0}
        model = gcx.Model.from_params(param, md)
        scale_ = np.maximum(np.abs(exact), 1e-3)
        ws.set_priorities(g, exact)
        ws.color(g)
        row["fp64_colors"] = ws.validate(g)
        for pname, prec in (("fp32", gcx.FP32), ("bf16", gcx.BF16), ("bf16x3", gcx.BF16X3)):
            ws.order(g, gcx.GNN, model, prec)
            p = ws.priorities(N)
            ws.color(g)
            rel = np.abs(p - exact) / scale_
            row[pname] = {"max_rel_err": float(rel.max()), "mean_rel_err": float(rel.mean()),
                          "orientation_flips_vs_fp64": orientation_flips(V, E, exact, p, pairs),
                          "colors": ws.validate(g)}
        for pname in ("fp32", "bf16", "bf16x3"):
            row[pname]["colors_minus_fp32"] = row[pname]["colors"] - row["fp32"]["colors"]
        rows.append(row)
        if log:
            print(json.dumps(row), file=log, flush=True)
    return {"graph": f"R-MAT scale {scale} (edge factor 16, seed 1), GNN{layers}",
            "vertices": int(N), "entries": int(g.M), "undirected_edges": int(g.M // 2),
            "relative_error": "|p - p64| / max(|p64|, 1e-3)", "sets": rows}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--sets", type=int, default=8, help="Glorot-uniform weight sets")
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    res = study(a.scale, a.sets, a.layers, log=sys.stderr)
    text = json.dumps(res, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    print(text)


if __name__ == "__main__":
    main()
