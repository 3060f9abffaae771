"""The verifier / quality oracle (oracle/quality_oracle.py) pinned by what the
mathematics fixes, and the library's exact decimal capacity against Python's
Decimal.

- A grid cut into its top and bottom halves has closed-form counters.
- On random graphs the per-part counters equal the block sums of the dense
  adjacency matrix (1_p^T A 1_q), an independent formulation.
- The capacity is Decimal's floor(ratio * n / parts), including ratios whose
  binary double lies below the decimal value (1.15 * 3200 / 32 = 115 exactly, while
  the double expression floors to 114).
"""
from decimal import Decimal, ROUND_FLOOR

import numpy as np
import pytest

import quality_oracle as qo


def grid_csr(rows, cols):
    """4-neighbour grid, vertex r * cols + c, rows sorted ascending."""
    indptr, indices = [0], []
    for r in range(rows):
        for c in range(cols):
            nbrs = []
            if r > 0:
                nbrs.append((r - 1) * cols + c)
            if c > 0:
                nbrs.append(r * cols + c - 1)
            if c + 1 < cols:
                nbrs.append(r * cols + c + 1)
            if r + 1 < rows:
                nbrs.append((r + 1) * cols + c)
            indices.extend(nbrs)
            indptr.append(len(indices))
    return np.array(indptr, np.int64), np.array(indices, np.int64)


def dense(indptr, indices, n):
    a = np.zeros((n, n), dtype=np.int64)
    for v in range(n):
        for e in range(indptr[v], indptr[v + 1]):
            a[v, indices[e]] += 1
    return a


def random_symmetric(n, density, seed):
    rng = np.random.default_rng(seed)
    upper = np.triu(rng.random((n, n)) < density, 1)
    a = upper | upper.T
    indptr = np.concatenate([[0], np.cumsum(a.sum(1))]).astype(np.int64)
    indices = np.concatenate([np.flatnonzero(a[v]) for v in range(n)]).astype(np.int64)
    return indptr, indices


def test_grid_halves_closed_form():
    rows, cols = 8, 6
    indptr, indices = grid_csr(rows, cols)
    n = rows * cols
    labels = np.array([0 if v // cols < rows // 2 else 1 for v in range(n)], dtype=np.int32)
    v = qo.verify(indptr, indices, labels, 2, "1.03")
    # the cut is the cols vertical edges between rows 3 and 4
    assert v["cut"] == cols
    assert v["undirected_edges"] == rows * (cols - 1) + (rows - 1) * cols
    assert v["loads"] == [n // 2, n // 2]
    assert v["capacity"] == (103 * n) // 200
    a = qo.analyze(indptr, indices, labels)
    half_rows = rows // 2
    internal = 2 * (half_rows * (cols - 1) + (half_rows - 1) * cols)
    assert list(a["per_part"]["internal"]) == [internal, internal]
    assert list(a["per_part"]["outgoing"]) == [cols, cols]
    assert list(a["per_part"]["incoming"]) == [cols, cols]
    assert a["max_local_cut"] == 2 * cols
    assert a["edge_cut"] == 2 * cols
    assert a["edge_set_avg"] == (len(indices) + 2 * cols) / 2


@pytest.mark.parametrize("seed", range(6))
def test_counters_equal_dense_block_sums(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(8, 60))
    k = int(rng.integers(2, 9))
    indptr, indices = random_symmetric(n, 0.2, seed)
    labels = rng.integers(0, k, n).astype(np.int32)
    a = dense(indptr, indices, n)
    ind = np.zeros((k, n), dtype=np.int64)
    ind[labels, np.arange(n)] = 1
    block = ind @ a @ ind.T  # block[p, q] = entries from part p to part q
    res = qo.analyze(indptr, indices, labels, k)
    pp = res["per_part"]
    assert np.array_equal(pp["internal"], np.diag(block))
    assert np.array_equal(pp["outgoing"], block.sum(1) - np.diag(block))
    assert np.array_equal(pp["incoming"], block.sum(0) - np.diag(block))
    assert np.array_equal(pp["vertices"], ind.sum(1))
    assert res["edge_cut"] == block.sum() - np.trace(block)
    # invariants
    assert pp["internal"].sum() + pp["outgoing"].sum() == len(indices)
    assert pp["outgoing"].sum() == pp["incoming"].sum() == res["edge_cut"]
    v = qo.verify(indptr, indices, labels, k)
    assert v["cut"] * 2 == res["edge_cut"]
    assert v["loads"] == list(ind.sum(1))


def test_weighted_counters_equal_dense_block_sums():
    rng = np.random.default_rng(7)
    n = 30
    indptr, indices = random_symmetric(n, 0.3, 7)
    weights = rng.integers(1, 50, len(indices))
    labels = rng.integers(0, 4, n).astype(np.int32)
    a = np.zeros((n, n), dtype=np.int64)
    for v in range(n):
        for e in range(indptr[v], indptr[v + 1]):
            a[v, indices[e]] += weights[e]
    ind = np.zeros((4, n), dtype=np.int64)
    ind[labels, np.arange(n)] = 1
    block = ind @ a @ ind.T
    res = qo.analyze(indptr, indices, labels, 4, weights=weights)
    assert np.array_equal(res["per_part"]["weighted_internal"], np.diag(block))
    assert np.array_equal(res["per_part"]["weighted_outgoing"], block.sum(1) - np.diag(block))
    assert res["weighted_edge_cut"] == block.sum() - np.trace(block)


def test_verify_error_order_and_messages():
    indptr, indices = grid_csr(3, 3)
    labels = np.zeros(9, np.int32)
    with pytest.raises(qo.VerifyError, match="invalid parts or imbalance ratio"):
        qo.verify(indptr, indices, labels, 1)
    with pytest.raises(qo.VerifyError, match="invalid parts or imbalance ratio"):
        qo.verify(indptr, indices, labels, 2, "0.99")
    bad = labels.copy()
    bad[4], bad[7] = 5, -1
    with pytest.raises(qo.VerifyError, match="invalid label: 5"):
        qo.verify(indptr, indices, bad, 2)
    directed = (np.array([0, 1, 1, 1]), np.array([1]))  # 0 -> 1 only
    with pytest.raises(qo.VerifyError, match="odd undirected"):
        qo.verify(*directed, np.array([0, 1, 0]), 2)
    cap = qo.verify(indptr, indices, np.array([0] * 8 + [1], np.int32), 2)
    assert not cap["capacity_ok"] and cap["max_load"] == 8


def test_analyze_parts_inference_and_errors():
    indptr, indices = grid_csr(4, 4)
    labels = np.array([0, 0, 2, 2] * 4, np.int32)
    assert qo.analyze(indptr, indices, labels)["parts"] == 3
    r = qo.analyze(indptr, indices, labels, 5)
    assert r["parts"] == 5 and r["min_vertices"] == 0
    with pytest.raises(qo.VerifyError):
        qo.analyze(indptr, indices, labels, 2)


CAPACITY_CASES = [("1.03", 1000, 4), ("1.15", 640, 32), ("1.15", 3200, 32), ("1.15", 1500, 3), ("1.1", 300, 3), ("1", 97, 2), ("1.0300", 4096, 8),
                  ("  1.03 ", 12345, 7), ("+1.5", 10, 3), ("2E0", 9, 2), ("103e-2", 800, 8), ("1.7", 1 << 40, 32),
                  ("1.000000000000000000000001", 10 ** 12, 7)]


@pytest.mark.parametrize("ratio,n,parts", CAPACITY_CASES)
def test_decimal_capacity_matches_decimal(hx, ratio, n, parts):
    expect = int((Decimal(ratio) * n / parts).to_integral_value(rounding=ROUND_FLOOR))
    assert qo.capacity(ratio, n, parts) == expect
    assert hx.decimal_capacity(ratio, n, parts) == expect


def test_decimal_capacity_is_not_the_double_formula(hx):
    import math
    assert math.floor(1.15 * 3200 / 32) == 114
    assert hx.decimal_capacity("1.15", 3200, 32) == 115


def test_decimal_capacity_random_against_decimal(hx):
    rng = np.random.default_rng(3)
    for _ in range(2000):
        digits = int(rng.integers(0, 6))
        frac = "".join(str(int(d)) for d in rng.integers(0, 10, digits))
        whole = int(rng.integers(1, 4))
        ratio = f"{whole}.{frac}" if digits else str(whole)
        parts = int(rng.integers(2, 33))
        n = int(rng.integers(0, 10 ** int(rng.integers(1, 12))))
        expect = int((Decimal(ratio) * n / parts).to_integral_value(rounding=ROUND_FLOOR))
        assert hx.decimal_capacity(ratio, n, parts) == expect, (ratio, n, parts)


@pytest.mark.parametrize("ratio", ["0.99", "-1.03", "abc", "", "1.03x", "inf", "NaN", "1e", "."])
def test_decimal_capacity_rejects(hx, ratio):
    with pytest.raises(hx.GpartError, match="invalid parts or imbalance ratio"):
        hx.decimal_capacity(ratio, 100, 4)
