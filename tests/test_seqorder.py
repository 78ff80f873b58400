"""The library's ID / SD / SDA orderings (incidence_degree.h), which run on the
host CPU: bit-exact against the reference's golden vectors and digests and
against the oracle's step-for-step restatement (no GPU needed)."""
import hashlib

import numpy as np
import pytest

import gcx
import oracle
from conftest import SEQ, TINY, load_case
from oracle import seqorder

SEQ_ORACLE = {"ID": seqorder.incidence_degree, "SD": seqorder.saturation_degree,
              "SDA": seqorder.saturation_degree_alt}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits(p):
    return np.ascontiguousarray(p, np.float64).view(np.uint64).tolist()


@pytest.mark.parametrize("name", TINY + ["rmat10"])
@pytest.mark.parametrize("h", SEQ)
def test_goldens(name, h):
    c = load_case(name)
    assert bits(gcx.ref.ordering(h, c["V"], c["E"])) == bits(c[f"p_{h}"])


def test_random_graphs_against_oracle():
    rng = np.random.default_rng(33)
    for trial in range(120):
        n = int(rng.integers(1, 70))
        m = int(rng.integers(0, 5 * n + 1))
        V, E = oracle.csr_from_pairs(n, rng.integers(0, n, m).astype(np.uint32),
                                     rng.integers(0, n, m).astype(np.uint32))
        for h in SEQ:
            assert bits(gcx.ref.ordering(h, V, E)) == bits(SEQ_ORACLE[h](V, E)), (trial, h)


def test_many_components_and_cliques():
    """Disjoint cliques of sizes 1..12 on shuffled ids: many components, each
    walked from its own first vertex of maximum degree, and many ties."""
    rng = np.random.default_rng(4)
    ids = rng.permutation(78).astype(np.uint32)
    I, J, off = [], [], 0
    for k in range(1, 13):
        block = ids[off:off + k]
        for a in range(k):
            for b in range(a + 1, k):
                I.append(block[a])
                J.append(block[b])
        off += k
    V, E = oracle.csr_from_pairs(78, np.array(I, np.uint32), np.array(J, np.uint32))
    for h in SEQ:
        assert bits(gcx.ref.ordering(h, V, E)) == bits(SEQ_ORACLE[h](V, E)), h


@pytest.mark.parametrize("name", ["rmat16", "rmat20"])
def test_digests(name, digests):
    """R-MAT scale 16 and 20: the SHA-256 of the priority bits equals the
    compiled reference's."""
    rec = digests[name]
    N, I, J = gcx.rmat_pairs(rec["graph"]["scale"], rec["graph"]["edge_factor"],
                             *rec["graph"]["abc"], seed=rec["graph"]["seed"])
    V, E = oracle.csr_from_pairs(N, I, J)
    assert sha(V) == rec["sha_V"] and sha(E) == rec["sha_E"]
    for h in SEQ:
        p = gcx.ref.ordering(h, V, E)
        assert sha(p) == rec[h]["sha_p"], h


def test_empty_graph():
    V = np.zeros(1, np.uint32)
    E = np.zeros(0, np.uint32)
    for h in SEQ:
        assert gcx.ref.ordering(h, V, E).shape == (0,)
