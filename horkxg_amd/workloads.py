"""Synthetic workloads: the single-graph configurations and the batched graph mix.

The mix is 256 independent graphs. Graph i draws, from numpy's default_rng(i):
- n_i = round(exp(U(ln 1e4, ln 2e6))) vertices (log-uniform);
- k_i uniform from {2, 4, 8, 32}.
Its type is i mod 3: a 2D 4-neighbour grid of side round(sqrt(n_i)), a 3D 6-neighbour
mesh of side round(cbrt(n_i)), or an R-MAT graph of scale round(log2(n_i)), edge
factor 8, (a, b, c) = (0.57, 0.19, 0.19), generator seed i. Ratio 1.03, partition
seed i. All weights are unit weights.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor

import numpy as np

MIX_COUNT = 256
MIX_MIN_N = 10_000
MIX_MAX_N = 2_000_000
MIX_PARTS = (2, 4, 8, 32)


def mix_spec(i: int, min_n: int = MIX_MIN_N, max_n: int = MIX_MAX_N) -> dict:
    rng = np.random.default_rng(i)
    n = int(round(math.exp(rng.uniform(math.log(min_n), math.log(max_n)))))
    k = int(MIX_PARTS[int(rng.integers(len(MIX_PARTS)))])
    kind = ("grid2d", "mesh3d", "rmat")[i % 3]
    if kind == "grid2d":
        side = max(2, int(round(math.sqrt(n))))
        size = (side, side)
    elif kind == "mesh3d":
        side = max(2, int(round(n ** (1.0 / 3.0))))
        size = (side, side, side)
    else:
        size = (max(2, int(round(math.log2(n)))), 8)
    return dict(index=i, kind=kind, size=size, k=k, ratio=1.03, seed=i)


def mix_specs(count: int = MIX_COUNT, min_n: int = MIX_MIN_N, max_n: int = MIX_MAX_N) -> list:
    return [mix_spec(i, min_n, max_n) for i in range(count)]


def generate(hx, spec: dict):
    """(indptr, indices) int64 arrays of one workload graph."""
    if spec["kind"] == "mesh3d":
        return hx.generate_mesh3d(*spec["size"])
    if spec["kind"] == "grid2d":
        return hx.generate_grid2d(*spec["size"])
    scale, edge_factor = spec["size"]
    return hx.generate_rmat(scale, edge_factor, 0.57, 0.19, 0.19, spec.get("seed", 0))


def device_graph(hx, spec: dict):
    """The workload graph of `spec`, generated on the GPU (identical to generate())."""
    if spec["kind"] == "mesh3d":
        return hx.Graph.generate_mesh3d(*spec["size"])
    if spec["kind"] == "grid2d":
        return hx.Graph.generate_grid2d(*spec["size"])
    scale, edge_factor = spec["size"]
    return hx.Graph.generate_rmat(scale, edge_factor, 0.57, 0.19, 0.19, spec.get("seed", 0))


def assign_largest_first(specs, sizes, workers: int):
    """Spec indices per worker: graphs by decreasing size, each to the least loaded
    worker (size = vertices + entries)."""
    order = sorted(range(len(specs)), key=lambda i: -sizes[i])
    load = [0] * workers
    out = [[] for _ in range(workers)]
    for i in order:
        w = min(range(workers), key=lambda j: load[j])
        out[w].append(i)
        load[w] += sizes[i]
    return out


def upload_mix(hx, specs, threads: int = 8):
    """Generate and upload every graph of a mix; host arrays are dropped once the
    graph holds its own copies. Returns the hx.Graph list in spec order."""
    def one(spec):
        indptr, indices = generate(hx, spec)
        return indptr, indices

    graphs = [None] * len(specs)
    with ThreadPoolExecutor(max_workers=threads) as pool:
        for i, (indptr, indices) in enumerate(pool.map(one, specs)):
            graphs[i] = hx.Graph(indptr, indices)
            del indptr, indices
    return graphs
