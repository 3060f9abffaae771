#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden from the CPU oracle.

  golden.json   per-level SHA-256 digests of coarsening hierarchies (grid 256^2 at
                k = 4 and R-MAT scale 12 at k = 8, seeds 0-4, with the partition
                CLI's options and with the library defaults), and the full
                partition results (cut, loads, label digest) for the same inputs
  spectral.npz  the FP64 diffusion vector and labels of the initial partition of
                the coarsest grid level (seed 0, k = 4)

Inputs come from the package's own generators (host code only). Run from the
repository root after building: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import horkxg_amd as hx  # noqa: E402
import oracle_py as O  # noqa: E402

SEEDS = range(5)
INPUTS = {
    "grid256": (lambda: hx.generate_grid2d(256, 256), 4),
    "rmat12": (lambda: hx.generate_rmat(12, 16), 8),
}


def digest(array, dtype):
    return hashlib.sha256(np.ascontiguousarray(array, dtype=dtype).tobytes()).hexdigest()


def level_record(h, i):
    off, adj, ew, vw = h.level(i).arrays()
    rec = {"n": len(off) - 1, "m": len(adj), "offsets": digest(off, "<i8"), "neighbors": digest(adj, "<i4"),
           "edge_weights": digest(ew, "<u8"), "vertex_weights": digest(vw, "<u8")}
    if i + 1 < len(h):
        rec["map"] = digest(h.map(i), "<i4")
    return rec


def production_factor(k):
    return max(64, -(-512 // k))


def main():
    out = {"hierarchy": {}, "partition": {}}
    spectral = {}
    for name, (make, k) in INPUTS.items():
        ip, ix = make()
        g = O.Graph(ip, ix.astype(np.int32))
        for seed in SEEDS:
            for mode, factor, pack in (("production", production_factor(k), True), ("defaults", 256, False)):
                h = O.coarsen(g, k, seed, 0.90, 24, factor, pack)
                out["hierarchy"][f"{name}_k{k}_s{seed}_{mode}"] = {
                    "parts": k, "seed": seed, "cluster_size_factor": factor, "pack_isolated": pack,
                    "stop_reason": h.stop_reason, "levels": [level_record(h, i) for i in range(len(h))]}
                if name == "grid256" and seed == 0 and mode == "production":
                    coarsest = h.level(len(h) - 1)
                    labels, coords, _ = O.initial_partition(coarsest, k, 1.03, 0, with_coords=True)
                    off, adj, ew, vw = coarsest.arrays()
                    spectral = dict(offsets=off, neighbors=adj, edge_weights=ew, vertex_weights=vw,
                                    coords=coords, labels=labels)
            labels, res = O.partition(g, k, 1.03, seed)
            out["partition"][f"{name}_k{k}_s{seed}"] = {
                "parts": k, "seed": seed, "cut": int(res.final_cut),
                "loads": [int(x) for x in O.part_weights(g, labels, k)],
                "labels": digest(labels, "<i4")}
    (HERE / "golden.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    np.savez_compressed(HERE / "spectral.npz", **spectral)
    print("wrote", HERE / "golden.json", HERE / "spectral.npz")


if __name__ == "__main__":
    main()
