"""CPU baseline timing (test and measurement infrastructure, not product code).

Times the reference's own OpenMP path -- its C sources compiled into
oracle/_ref/libref.so -- on a CSR saved by bench.py: cgnn_graph_sage_16_par
followed by jones_plassmann_internal, best of `reps`, with model load and
allocation outside the timed region as the reference's runner does
(runner.c:99-134, best of it = 5 in main.c:199-208). Run as a child process
so that the OpenMP placement variables (OMP_PLACES=cores,
OMP_PROC_BIND=spread, README.md:8-14) and the thread count take effect before
the OpenMP runtime starts:

    OMP_NUM_THREADS=16 OMP_PLACES=cores OMP_PROC_BIND=spread \
        python oracle/cpu_baseline.py --csr DIR --model models/m3.gnn --reps 5

Prints one JSON object: seconds, colours, threads.
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--csr", required=True, help="directory holding V.npy and E.npy")
    ap.add_argument("--model", required=True)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    import oracle

    ref = oracle.Ref()
    V = np.load(os.path.join(args.csr, "V.npy"), mmap_mode="r", allow_pickle=False)
    E = np.load(os.path.join(args.csr, "E.npy"), mmap_mode="r", allow_pickle=False)
    param, md = oracle.load_model(args.model)
    secs, col = ref.time_gnn_jp(param, np.asarray(V), np.asarray(E), md, reps=args.reps)
    print(json.dumps({"seconds": secs, "colors": int(col.max(initial=0)),
                      "threads": ref.threads()}))


if __name__ == "__main__":
    main()
