"""The reference's synthetic suite (tests/synthetic_suite.py:20-119) on the GPU:
social, circuit and disconnected graphs of 16384 vertices built by the same
generator rules, each partitioned for k in {2, 4, 8, 32} at both widths, twice.
Every partition is checked by the GPU verifier and by the oracle verifier: the
capacity holds, the reported cut equals the verifier's, and the labels are the
same across repeats and widths."""
import hashlib
import random

import numpy as np
import pytest

import quality_oracle as qo

pytestmark = pytest.mark.gpu


def generate(kind, n, seed=17):
    """Symmetric CSR (int64 indptr, indices) by the suite's rules
    (tests/synthetic_suite.py:20-61): 8 communities; social graphs get community
    triangles and hubs plus weak random cross edges, circuits get local distance-1,
    2, 4 links and sparse long nets, disconnected graphs are 8 paths."""
    rng = random.Random(seed)
    rows = [set() for _ in range(n)]

    def add(u, v):
        if u != v:
            rows[u].add(v)
            rows[v].add(u)

    size = n // 8
    for v in range(n):
        if kind == "social":
            base = v // size * size
            add(v, base)
            add(v, base + (v - base + 1) % size)
            for _ in range(8):
                add(v, base + rng.randrange(size))
            if rng.random() < 0.6:
                add(v, rng.randrange(n))
        elif kind == "circuit":
            for distance in (1, 2, 4):
                if v + distance < n:
                    add(v, v + distance)
            if v % 16 == 0:
                add(v, rng.randrange(n))
        elif kind == "disconnected":
            if v % size:
                add(v - 1, v)
        else:
            raise ValueError(kind)
    indptr = [0]
    indices = []
    for row in rows:
        indices.extend(sorted(row))
        indptr.append(len(indices))
    return np.array(indptr, np.int64), np.array(indices, np.int64)


@pytest.mark.parametrize("kind", ["social", "circuit", "disconnected"])
def test_synthetic_suite(hx, kind):
    indptr, indices = generate(kind, 16384)
    graphs = {w: hx.Graph(indptr, indices, width=w) for w in ("standard", "big")}
    for k in (2, 4, 8, 32):
        digests = {}
        for width, g in graphs.items():
            for repeat in range(2):
                labels, rep = g.partition(k, 1.03, 0)
                result = g.verify(labels, k)
                assert result == qo.verify(indptr, indices, labels, k)
                assert result["capacity_ok"], (kind, k, width, result)
                assert rep["final_cut"] == result["cut"], (kind, k, width)
                digests[f"{width}-{repeat}"] = hashlib.sha256(labels.tobytes()).hexdigest()
        assert len(set(digests.values())) == 1, (kind, k, digests)
