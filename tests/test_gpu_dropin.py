"""The drop-in check: the reference's own driver (src/main.c, unchanged),
compiled against include/ and linked against libgcx.so (oracle/Makefile
`dropin`), prints what the same driver linked against the reference's own
library printed (tests/golden/cli_ref.json, from oracle/_ref/coloring_ref):
the -q colour-count lines exactly (main.c:137-183), and in the full runner
mode (main.c:188-207, runner.c:156-169) the same line names, colour counts
and number of timing columns (needs gfx950)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, load_case

pytestmark = pytest.mark.gpu

DROPIN = os.path.join(ROOT, "oracle", "_ref", "coloring_dropin")

with open(os.path.join(GOLDEN, "cli_ref.json")) as f:
    EXPECTED = json.load(f)


def write_mtx(path, name):
    c = load_case(name)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n")
        f.write(f"{int(c['N'])} {int(c['N'])} {len(c['I'])}\n")
        np.savetxt(f, np.stack([c["I"].astype(np.int64) + 1, c["J"].astype(np.int64) + 1], 1),
                   fmt="%d")


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="oracle/_ref/coloring_dropin is built from the reference's main.c, "
                           "which is not present here")
@pytest.mark.parametrize("key", sorted(EXPECTED))
def test_reference_driver_on_libgcx(tmp_path, key):
    name, *args = key.split()
    path = tmp_path / f"{name}.mtx"
    write_mtx(path, name)
    print(f"drop-in: {name} {' '.join(args)}", flush=True)
    # gnn_setup reads models/m<l>.gnn relative to the working directory (gnn.c:27)
    out = subprocess.run([DROPIN, *args, "-g", str(path)], capture_output=True, text=True,
                         cwd=ROOT, timeout=300)
    assert out.returncode == 0, out.stderr
    want = EXPECTED[key]
    if "-q" in args:
        assert out.stdout == want
        return
    got_lines = out.stdout.strip().splitlines()
    want_lines = want.strip().splitlines()
    assert len(got_lines) == len(want_lines)
    for g, w in zip(got_lines, want_lines):
        gt, wt = g.split(), w.split()
        assert gt[:2] == wt[:2], (g, w)  # "<heuristic>_<pass> <colours>"
        assert len(gt) == len(wt)  # sequential column + one per thread count
        for t in gt[2:]:
            assert float(t) >= 0.0
