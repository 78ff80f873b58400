"""Shared pytest setup: import paths, the `gpu` marker and fixture loaders."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, os.path.join(ROOT, "gnn-coloring_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HEUR = ["FF", "LF", "LLF", "SL", "SLL", "GNN2", "GNN3", "GNN4"]
SEQ = ["ID", "SDA", "SD"]  # sequential heap orderings, incidence_degree.c
TINY = ["path5", "star6", "clique6", "cycle7", "bip34", "empty5", "mixed10", "jump8",
        "er300", "stencil6", "heapquirk10"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_case(name):
    fn = f"{name}.npz" if name.startswith("rmat") else f"tiny_{name}.npz"
    with np.load(os.path.join(GOLDEN, fn), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_models():
    with np.load(os.path.join(GOLDEN, "models.npz"), allow_pickle=False) as z:
        return {l: (z[f"m{l}"], int(z[f"md{l}"])) for l in (2, 3, 4)}


def load_digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def models():
    return load_models()


@pytest.fixture(scope="session")
def digests():
    return load_digests()
