"""The persistent bisection queue (hip/spectral_queue.hip).

One launch serves every bisection task of at most 4096 vertices of one or many
graphs. Its labels, order and coordinates must be bit-identical to the depth by
depth launches (gpart_spectral_depth) in both precisions and, in FP64, to the
oracle's recursive bisection (src/initial_partition_gpu.cu:170-276), for every
team size the tuning allows: teams of 1 to 64 workgroups meet at barriers.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


def coarsest(hx, oracle, indptr, indices, k):
    """The oracle's coarsest level of a graph (weighted vertices and edges)."""
    og = oracle.Graph(indptr, indices.astype(np.int32))
    oh = oracle.coarsen(og, k, cluster_size_factor=max(64, (512 + k - 1) // k), pack_isolated=True)
    return oh.level(len(oh) - 1).arrays()


INPUTS = {
    "grid64": lambda hx: hx.generate_grid2d(64, 64),     # 4096 vertices: the largest queue root
    "mesh12": lambda hx: hx.generate_mesh3d(12, 12, 12),
    "rmat11": lambda hx: hx.generate_rmat(11, 8),
    "grid90": lambda hx: hx.generate_grid2d(90, 90),     # 8100: the root goes plain, its children queue
}


@pytest.fixture
def restore_queue(hx):
    saved = hx.spectral_queue_settings()
    yield
    hx.set_spectral_queue(saved["enabled"], saved["own_blocks"], saved["max_task"])


@pytest.mark.parametrize("name", list(INPUTS))
@pytest.mark.parametrize("mode", ["batched_fp64", "batched_fp32"])
def test_queue_matches_depth_form(hx, oracle, restore_queue, name, mode):
    indptr, indices = INPUTS[name](hx)
    g = hx.Graph(indptr, indices)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    for k in (2, 5, 8, 32):
        say(f"case queue graph={name} mode={mode} k={k}")
        hx.set_spectral_queue(False)
        ref_labels, ref_coords = g.initial_partition(k, 1.03, 3, mode, with_coords=True)
        # team sizes 1 to 64, and a queue that takes only the deeper tasks
        for own_blocks, max_task in ((64, 4096), (8, 4096), (128, 4096), (16, 600)):
            hx.set_spectral_queue(True, own_blocks, max_task)
            labels, coords = g.initial_partition(k, 1.03, 3, mode, with_coords=True)
            assert np.array_equal(labels, ref_labels), (k, own_blocks, max_task)
            assert np.array_equal(coords.view(np.uint64), ref_coords.view(np.uint64)), (k, own_blocks, max_task)
        if mode == "batched_fp64":
            ol, ocoords, _ = oracle.initial_partition(og, k, 1.03, 3, with_coords=True)
            assert np.array_equal(ref_labels, ol), k
            assert np.array_equal(ref_coords.view(np.uint64), ocoords.view(np.uint64)), k


@pytest.mark.parametrize("width", ["standard", "big"])
def test_queue_weighted_coarsest_levels(hx, oracle, restore_queue, width):
    """Weighted coarsest levels of the oracle's hierarchies, FP64 against the oracle."""
    for name, k in (("grid256", 8), ("mesh24", 32), ("rmat12", 4)):
        indptr, indices = {"grid256": lambda: hx.generate_grid2d(256, 256),
                           "mesh24": lambda: hx.generate_mesh3d(24, 24, 24),
                           "rmat12": lambda: hx.generate_rmat(12, 16)}[name]()
        off, adj, ew, vw = coarsest(hx, oracle, indptr, indices, k)
        say(f"case queue weighted graph={name} width={width} k={k} n={len(off) - 1}")
        g = hx.Graph(off, adj, width=width, edge_weights=ew, vertex_weights=vw)
        oc = oracle.Graph(off, adj, ew, vw)
        labels, coords = g.initial_partition(k, 1.03, 0, "batched_fp64", with_coords=True)
        ol, ocoords, _ = oracle.initial_partition(oc, k, 1.03, 0, with_coords=True)
        assert np.array_equal(labels, ol)
        assert np.array_equal(coords.view(np.uint64), ocoords.view(np.uint64))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_queue_many_graphs_one_launch(hx, restore_queue, precision):
    """Several graphs in one launch give what each gives alone; the task count is
    the bisection tree's internal nodes."""
    from horkxg_amd import device

    # (graph, parts, weighted): weights 1..2 where the 3% slack of a part holds them
    specs = [(hx.generate_grid2d(40, 40), 8, False), (hx.generate_mesh3d(10, 10, 10), 8, True),
             (hx.generate_rmat(10, 8), 4, False), (hx.generate_grid2d(64, 64), 5, True),
             (hx.generate_grid2d(9, 7), 2, False), (hx.generate_mesh3d(12, 12, 12), 32, False)]
    graphs, parts, caps, seeds, weights = [], [], [], [], []
    rng = np.random.default_rng(7)
    for i, ((indptr, indices), k, weighted) in enumerate(specs):
        n = len(indptr) - 1
        w = rng.integers(1, 3, n).astype(np.uint64) if weighted else np.ones(n, np.uint64)
        graphs.append(device.DeviceGraph(indptr, indices, vertex_weights=w))
        parts.append(k)
        caps.append(hx.partition_capacity(int(w.sum()), k, 1.03))
        seeds.append(11 * i + 1)
        weights.append(w)
    for own_blocks in (64, 16):
        hx.set_spectral_queue(True, own_blocks)
        say(f"case queue many graphs precision={precision} own_blocks={own_blocks}")
        together, rep = device.spectral_queue(graphs, parts, caps, seeds, precision, weights)
        assert rep["tasks"] == sum(k - 1 for k in parts)
        for i in range(len(graphs)):
            alone, rep1 = device.spectral_queue([graphs[i]], [parts[i]], [caps[i]], [seeds[i]], precision,
                                                [weights[i]])
            assert rep1["tasks"] == parts[i] - 1
            labels, order, coords, status = together[i]
            assert status == 0
            assert np.array_equal(labels, alone[0][0]) and np.array_equal(order, alone[0][1])
            assert np.array_equal(coords.view(np.uint64), alone[0][2].view(np.uint64))
            loads = np.bincount(labels, weights=weights[i].astype(np.float64), minlength=parts[i])
            assert loads.max() <= caps[i] and len(np.unique(labels)) == parts[i]
            assert np.array_equal(np.sort(order), np.arange(len(order)))
