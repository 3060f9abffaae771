"""CPU restatement of the partition verifier and the partition-quality metrics.

TEST INFRASTRUCTURE ONLY: this is the parity checker for the GPU verifier and
analyser. Only tests/ import it; the product package never does.

verify() follows tools/verify_partition.py:22-65: the argument checks (:24-25),
CSR checks (:27-30), label count and range (:33-45, first invalid label in vertex
order), neighbour range, directed cut and non-loop entries (:46-54), the symmetry
parity check (:55-56), the capacity floor(Decimal(ratio) * n / parts) (:57) and the
result fields (:59-65).

analyze() follows graph_untils/analyze_partition.cpp:219-294: parts inferred as
max label + 1 unless given (:176-183), per-part vertices, internal / outgoing /
incoming entries and their weights (:236-254), the edge-set total m + directed cut
(:257-258), local cuts, minima and maxima (:262-275), and the imbalance and
deviation ratios (:281-291, :185-192). Weights are the entries' edge weights
(the tool's --values; unit when absent, :158-160), summed exactly as integers.
"""
from __future__ import annotations

from decimal import Decimal, InvalidOperation, ROUND_FLOOR

import numpy as np


class VerifyError(ValueError):
    pass


def capacity(ratio: str, n: int, parts: int) -> int:
    """floor(Decimal(ratio) * n / parts), tools/verify_partition.py:57."""
    return int((Decimal(ratio) * n / parts).to_integral_value(rounding=ROUND_FLOOR))


def _check_ratio(ratio: str) -> bool:
    try:
        d = Decimal(ratio)
    except (InvalidOperation, ValueError, TypeError):
        return False
    return d.is_finite() and d >= 1


def verify(indptr, indices, labels, parts: int, ratio: str = "1.03") -> dict:
    if not 2 <= parts <= 32 or not _check_ratio(ratio):
        raise VerifyError("invalid parts or imbalance ratio")
    offsets = np.asarray(indptr, dtype=np.int64)
    idx = np.asarray(indices, dtype=np.int64)
    if offsets.size == 0 or offsets[0] != 0 or offsets[-1] != idx.size:
        raise VerifyError("invalid CSR endpoints")
    if np.any(offsets[1:] < offsets[:-1]) or np.any(offsets < 0):
        raise VerifyError("nonmonotone CSR offsets")
    n = offsets.size - 1
    lab = np.asarray(labels)
    if lab.size != n:
        raise VerifyError("partition label count does not match vertices")
    lab = lab.astype(np.int64)
    bad = np.flatnonzero((lab < 0) | (lab >= parts))
    if bad.size:
        raise VerifyError(f"invalid label: {int(lab[bad[0]])}")
    loads = np.bincount(lab, minlength=parts)[:parts]
    if np.any((idx < 0) | (idx >= n)):
        raise VerifyError("invalid neighbor id")
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(offsets))
    nonloop = int(np.count_nonzero(idx != src))
    directed_cut = int(np.count_nonzero(lab[src] != lab[idx]))
    if directed_cut % 2 or nonloop % 2:
        raise VerifyError("odd undirected edge/cut count; input must be symmetric")
    cap = capacity(ratio, n, parts)
    cut, edges = directed_cut // 2, nonloop // 2
    max_load = int(loads.max()) if loads.size else 0
    return {
        "vertices": n, "undirected_edges": edges, "cut": cut,
        "cut_ratio": cut / edges if edges else 0.0, "loads": [int(x) for x in loads],
        "max_load": max_load, "max_load_ratio": max_load * parts / n if n else 0.0,
        "capacity": cap, "capacity_ok": max_load <= cap,
    }


def _deviation(value: float, average: float) -> float:
    if average <= 0.0:
        return 0.0 if value == 0.0 else float("inf")
    return abs(value - average) / average


def _imbalance(max_value: float, average: float) -> float:
    return max_value / average if average > 0.0 else 0.0


def analyze(indptr, indices, labels, parts: int = 0, weights=None) -> dict:
    offsets = np.asarray(indptr, dtype=np.int64)
    idx = np.asarray(indices, dtype=np.int64)
    n = offsets.size - 1
    if np.any((idx < 0) | (idx >= n)):
        raise VerifyError("CSR index out of vertex range")
    lab = np.asarray(labels).astype(np.int64)
    if np.any(lab < 0):
        raise VerifyError("partition id must be non-negative")
    if lab.size != n:
        raise VerifyError("partition label count must equal CSR vertex count")
    inferred = int(lab.max()) + 1 if n else 0
    if parts == 0:
        k = inferred
    elif inferred > parts:
        raise VerifyError("partition contains part id >= parts")
    else:
        k = parts
    if k <= 0:
        raise VerifyError("partition is empty or has no parts")
    w = np.ones(idx.size, dtype=np.int64) if weights is None else np.asarray(weights).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(offsets))
    ps, pd = lab[src], lab[idx]
    same = ps == pd
    vertices = np.bincount(lab, minlength=k).astype(np.int64)
    internal = np.bincount(ps[same], minlength=k).astype(np.int64)
    outgoing = np.bincount(ps[~same], minlength=k).astype(np.int64)
    incoming = np.bincount(pd[~same], minlength=k).astype(np.int64)

    def wsum(keys, vals):
        out = np.zeros(k, dtype=np.int64)
        np.add.at(out, keys, vals)
        return out

    w_internal = wsum(ps[same], w[same])
    w_outgoing = wsum(ps[~same], w[~same])
    w_incoming = wsum(pd[~same], w[~same])
    directed_cut = int(np.count_nonzero(~same))
    weighted_cut = int(w[~same].sum())
    edge_set = internal + outgoing + incoming
    local = outgoing + incoming
    w_local = w_outgoing + w_incoming
    vertex_avg = n / k
    edge_set_avg = float(idx.size + directed_cut) / k
    res = {
        "parts": k, "vertices": n, "directed_edges": int(idx.size),
        "edge_cut": directed_cut, "weighted_edge_cut": weighted_cut,
        "edge_cut_ratio": directed_cut / idx.size if idx.size else 0.0,
        "avg_local_cut": float(local.sum()) / k, "max_local_cut": int(local.max()),
        "weighted_avg_local_cut": float(w_local.sum()) / k, "weighted_max_local_cut": int(w_local.max()),
        "vertex_avg": vertex_avg, "min_vertices": int(vertices.min()), "max_vertices": int(vertices.max()),
        "edge_set_avg": edge_set_avg, "min_edge_set": int(edge_set.min()), "max_edge_set": int(edge_set.max()),
    }
    res["vertex_imbalance"] = _imbalance(res["max_vertices"], vertex_avg)
    res["vertex_max_deviation"] = max(_deviation(res["max_vertices"], vertex_avg),
                                      _deviation(res["min_vertices"], vertex_avg))
    res["edge_set_imbalance"] = _imbalance(res["max_edge_set"], edge_set_avg)
    res["edge_set_max_deviation"] = max(_deviation(res["max_edge_set"], edge_set_avg),
                                        _deviation(res["min_edge_set"], edge_set_avg))
    res["per_part"] = {
        "vertices": vertices, "internal": internal, "outgoing": outgoing, "incoming": incoming,
        "weighted_internal": w_internal, "weighted_outgoing": w_outgoing, "weighted_incoming": w_incoming,
    }
    return res
