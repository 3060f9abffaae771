"""Partition-quality metrics of one or more partitions of one CSR graph, on the GPU.

Command line and report of graph_untils/analyze_partition.cpp (its options,
the "Partition Quality" table, the balance limits and, with --detail, the per-part
tables); the per-part counters come from the device (Graph.analyze,
gpart_graph_analyze). Weighted metrics use integer edge weights: unit weights
without --values, as the tool does; --values must hold non-negative integral
values.

    python -m horkxg_amd.analyze_partition PART.txt [PART2.txt ...] --indptr I --indices J
        [--values V] [--parts K] [--balance-ratio 1.05] [--edge-balance-ratio 1.05] [--detail]
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

import numpy as np


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("partitions", nargs="*", type=Path)
    p.add_argument("--partition", "--parts-file", dest="extra", action="append", type=Path, default=[])
    p.add_argument("--indptr", type=Path)
    p.add_argument("--indices", type=Path)
    p.add_argument("--values", type=Path)
    p.add_argument("--graph-dir", type=Path)
    p.add_argument("--name")
    p.add_argument("--parts", type=int, default=0)
    p.add_argument("--balance-ratio", type=float, default=1.05)
    p.add_argument("--edge-balance-ratio", type=float, default=1.05)
    p.add_argument("--detail", action="store_true")
    p.add_argument("--width", choices=("standard", "big"), default="standard")
    a = p.parse_args(argv)
    a.paths = list(a.partitions) + list(a.extra)
    if a.graph_dir is not None:
        if not a.name:
            p.error("--graph-dir requires --name")
        a.indptr = a.indptr or a.graph_dir / f"{a.name}_indptr.bin"
        a.indices = a.indices or a.graph_dir / f"{a.name}_indices.bin"
        values = a.graph_dir / f"{a.name}_values.bin"
        if a.values is None and values.exists():
            a.values = values
    if not a.paths:
        p.error("Provide at least one partition txt file")
    if a.indptr is None or a.indices is None:
        p.error("Provide original graph CSR with --indptr and --indices, or --graph-dir and --name")
    if a.parts < 0:
        p.error("--parts must be non-negative")
    if a.balance_ratio <= 0 or a.edge_balance_ratio <= 0:
        p.error("balance ratios must be positive")
    return a


def read_partition(path: Path) -> np.ndarray:
    labels = np.array(path.read_text().split(), dtype=np.int64)
    if labels.size and labels.min() < 0:
        raise ValueError(f"Partition id must be non-negative in {path}")
    return labels


def analyze_file(graph, path: Path, parts: int) -> dict:
    labels = read_partition(path)
    if labels.size != graph.n:
        raise ValueError(f"Partition line count must equal CSR vertex count for {path}")
    if labels.size and labels.max() > np.iinfo(np.int32).max:
        raise ValueError(f"Partition file contains part id >= --parts")
    return graph.analyze(labels.astype(np.int32), parts)


def _row(cells, widths):
    return "".join(str(c).rjust(w) for c, w in zip(cells, widths))


def main(argv=None) -> int:
    from ._native import Graph

    a = parse_args(argv)
    try:
        indptr = np.fromfile(a.indptr, dtype=np.int64)
        indices = np.fromfile(a.indices, dtype=np.int64)
        weights = None
        if a.values is not None:
            values = np.fromfile(a.values, dtype=np.float32)
            if values.size != indices.size:
                raise ValueError("values.size() must equal indices.size()")
            if np.any(values < 0) or np.any(values != np.floor(values)):
                raise ValueError("--values must hold non-negative integral values")
            weights = values.astype(np.uint64)
        graph = Graph(indptr, indices, width=a.width, edge_weights=weights)
        results = [(p, analyze_file(graph, p, a.parts)) for p in a.paths]
    except Exception as e:  # the tool's single error path
        print(f"error: {e}", file=sys.stderr)
        return 1
    f = lambda x: f"{x:.6f}"
    print("\n=========== Graph Summary ===========")
    print(f"vertices: {graph.n}\ndirected_edges: {graph.m}")
    print("\n=========== Partition Quality ===========")
    pw = max(len("partition_file"), *(len(str(p)) for p, _ in results)) + 2
    print("partition_file".ljust(pw) + _row(["parts", "edge_cut", "edge_cut_ratio", "vertex_imb", "edge_imb",
                                              "avg_local_cut", "max_local_cut"], [7, 16, 18, 18, 18, 18, 18]))
    for p, r in results:
        print(str(p).ljust(pw) + _row([r["parts"], r["edge_cut"], f(r["edge_cut_ratio"]), f(r["vertex_imbalance"]),
                                       f(r["edge_set_imbalance"]), f(r["avg_local_cut"]), r["max_local_cut"]],
                                      [7, 16, 18, 18, 18, 18, 18]))
    print("\n=========== Balance Limits ===========")
    print(f"vertex_imb_limit: {f(a.balance_ratio)}\nedge_imb_limit: {f(a.edge_balance_ratio)}")
    print("imbalance values closer to 1.0 are better")
    if a.detail:
        for p, r in results:
            pp = r["per_part"]
            print("\n=========== Partition Result ===========")
            print(f"partition_file: {p}")
            print("\n=========== Cut Objectives ===========")
            print(f"global_cut |C(G,P)|: {r['edge_cut']}\nweighted_global_cut: {f(r['weighted_edge_cut'])}")
            print(f"avg_local_cut avg_i |C(G,pi_i)|: {f(r['avg_local_cut'])}")
            print(f"max_local_cut max_i |C(G,pi_i)|: {r['max_local_cut']}")
            print(f"weighted_avg_local_cut: {f(r['weighted_avg_local_cut'])}")
            print(f"weighted_max_local_cut: {f(r['weighted_max_local_cut'])}")
            print("\n=========== Per-Part Metrics: Counts ===========")
            widths = [6, 14, 14, 16, 15, 15, 14]
            print(_row(["part", "|V(pi)|", "|E(pi)|", "internal_edges", "outgoing_cut", "incoming_cut",
                        "local_cut"], widths))
            for i in range(r["parts"]):
                edge_set = pp["internal"][i] + pp["outgoing"][i] + pp["incoming"][i]
                print(_row([i, pp["vertices"][i], edge_set, pp["internal"][i], pp["outgoing"][i], pp["incoming"][i],
                            pp["outgoing"][i] + pp["incoming"][i]], widths))
    return 0


if __name__ == "__main__":
    sys.exit(main())
