"""Independent CSR cut/load check of an int32 partition, on the GPU.

Command-line and result format of tools/verify_partition.py (its arguments, the
JSON line with sorted keys, the ValueError messages and the exit on a capacity
violation); the label, neighbour, cut and load passes run on the device
(Graph.verify, gpart_graph_verify).

    python -m horkxg_amd.verify_partition --indptr I --indices J --partition P --parts K
        [--imbalance-ratio 1.03] [--partition-format binary|text]
"""
from __future__ import annotations

import argparse
import json
from decimal import Decimal, InvalidOperation
from pathlib import Path

import numpy as np


def _read_i64(path: Path) -> np.ndarray:
    size = path.stat().st_size
    if size % 8:
        raise ValueError(f"invalid int64 CSR size: {path}")
    return np.fromfile(path, dtype=np.int64)


def _ratio_ok(imbalance: str) -> bool:
    try:
        d = Decimal(imbalance)
    except (InvalidOperation, ValueError):
        return False
    return d.is_finite() and d >= 1


def verify(indptr: Path, indices_path: Path, partition: Path, parts: int, imbalance: str = "1.03",
           partition_format: str = "binary", width: str = "standard") -> dict:
    """tools/verify_partition.py:22-65: the same checks in the same order, the same
    result dictionary."""
    from ._native import Graph, GpartError

    if not 2 <= parts <= 32 or not _ratio_ok(imbalance):
        raise ValueError("invalid parts or imbalance ratio")
    offsets, indices = _read_i64(Path(indptr)), _read_i64(Path(indices_path))
    if offsets.size == 0 or offsets[0] != 0 or offsets[-1] != indices.size:
        raise ValueError("invalid CSR endpoints")
    if np.any(offsets[1:] < offsets[:-1]) or np.any(offsets < 0):
        raise ValueError("nonmonotone CSR offsets")
    n = offsets.size - 1
    raw = Path(partition).read_bytes()
    if partition_format == "text":
        labels = np.array([int(x) for x in raw.decode().split()], dtype=np.int64)
    else:
        if len(raw) != n * 4:
            raise ValueError("partition byte length does not match vertices")
        labels = np.frombuffer(raw, dtype=np.int32).astype(np.int64)
    if labels.size != n:
        raise ValueError("partition label count does not match vertices")
    try:
        graph = Graph(offsets, indices, width=width)
    except GpartError as e:
        if "invalid neighbor" not in str(e):
            raise
        # The upload refuses out-of-range neighbours; the reference reports an
        # invalid label first when there is one.
        bad = np.flatnonzero((labels < 0) | (labels >= parts))
        if bad.size:
            raise ValueError(f"invalid label: {int(labels[bad[0]])}") from None
        raise ValueError("invalid neighbor id") from None
    if np.any((labels < np.iinfo(np.int32).min) | (labels > np.iinfo(np.int32).max)):
        bad = np.flatnonzero((labels < 0) | (labels >= parts))
        raise ValueError(f"invalid label: {int(labels[bad[0]])}")
    try:
        return graph.verify(labels.astype(np.int32), parts, imbalance)
    except GpartError as e:
        raise ValueError(str(e)) from None


def main() -> None:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--indptr", required=True, type=Path)
    parser.add_argument("--indices", required=True, type=Path)
    parser.add_argument("--partition", required=True, type=Path)
    parser.add_argument("--parts", required=True, type=int)
    parser.add_argument("--imbalance-ratio", default="1.03")
    parser.add_argument("--partition-format", choices=("binary", "text"), default="binary")
    parser.add_argument("--width", choices=("standard", "big"), default="standard",
                        help="graph width on the device (big for more than 2^31 entries)")
    args = parser.parse_args()
    result = verify(args.indptr, args.indices, args.partition, args.parts, args.imbalance_ratio,
                    args.partition_format, args.width)
    print(json.dumps(result, sort_keys=True))
    if not result["capacity_ok"]:
        raise SystemExit("partition violates the requested capacity")


if __name__ == "__main__":
    main()
