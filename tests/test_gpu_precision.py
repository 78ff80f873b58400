"""The precision study on a small graph (the ladder needs gfx950): FP32 sits at FP32
rounding from the FP64 network, BF16X3 about 2^-17, BF16 about 2^-9, and the
orientation flips JP sees follow the same order."""
import numpy as np
import pytest

import precision_study as ps


@pytest.mark.gpu
def test_precision_ladder():
    res = ps.study(scale=14, glorot=2, layers=4)
    assert len(res["sets"]) == 3
    for row in res["sets"]:
        f32, x3, b16 = row["fp32"], row["bf16x3"], row["bf16"]
        assert f32["mean_rel_err"] < 1e-6, row
        assert x3["mean_rel_err"] < 1e-4, row
        assert b16["mean_rel_err"] < 5e-2, row
        assert f32["mean_rel_err"] < x3["mean_rel_err"] < b16["mean_rel_err"], row
        assert f32["orientation_flips_vs_fp64"] <= x3["orientation_flips_vs_fp64"] \
            <= b16["orientation_flips_vs_fp64"], row
        for k in ("fp32", "bf16", "bf16x3"):
            assert row[k]["colors"] > 0


def test_orientation_flips_counts_edges():
    """Swapping the priorities of the two ends of one edge flips exactly
    the edges whose order between those ends changes."""
    V = np.array([0, 1, 3, 4], np.uint32)  # path 0 - 1 - 2
    E = np.array([1, 0, 2, 1], np.uint32)
    p = np.array([3.0, 2.0, 1.0])
    assert ps.orientation_flips(V, E, p, p) == 0
    assert ps.orientation_flips(V, E, p, np.array([2.0, 3.0, 1.0])) == 1
    assert ps.orientation_flips(V, E, p, np.array([1.0, 2.0, 3.0])) == 2
    # ties fall back to the id order (here the same as p's order)
    assert ps.orientation_flips(V, E, p, np.array([1.0, 1.0, 1.0])) == 0
    assert ps.orientation_flips(V, E, np.array([1.0, 2.0, 3.0]), np.ones(3)) == 2
