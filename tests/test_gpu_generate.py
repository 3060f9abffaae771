"""GPU: graphs generated on the device equal the host generators' graphs entry for
entry (R-MAT sampling, dedup and symmetrisation; grids; meshes), in both widths,
and partition identically."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")


def same_csr(g, indptr, indices):
    off, nbr, ew, vw = g.arrays()
    assert np.array_equal(off, indptr)
    assert np.array_equal(nbr.astype(np.int64), indices)
    assert (ew == 1).all() and (vw == 1).all()


@pytest.mark.parametrize("width", ["standard", "big"])
@pytest.mark.parametrize("scale,ef,seed", [(1, 1, 0), (4, 3, 7), (10, 8, 1), (14, 16, 0), (16, 8, 5)])
def test_rmat_matches_host(hx, width, scale, ef, seed):
    say(f"case generate rmat scale={scale} ef={ef} seed={seed} width={width}")
    g = hx.Graph.generate_rmat(scale, ef, 0.57, 0.19, 0.19, seed, width=width)
    indptr, indices = hx.generate_rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    assert g.n == len(indptr) - 1 and g.m == len(indices)
    same_csr(g, indptr, indices)


@pytest.mark.parametrize("width", ["standard", "big"])
@pytest.mark.parametrize("shape", [(1, 1), (1, 7), (5, 1), (37, 53), (256, 256)])
def test_grid_matches_host(hx, width, shape):
    say(f"case generate grid {shape} width={width}")
    g = hx.Graph.generate_grid2d(*shape, width=width)
    same_csr(g, *hx.generate_grid2d(*shape))


@pytest.mark.parametrize("shape", [(2, 2, 2), (1, 3, 4), (7, 5, 3), (32, 32, 32)])
def test_mesh_matches_host(hx, shape):
    say(f"case generate mesh {shape}")
    g = hx.Graph.generate_mesh3d(*shape)
    same_csr(g, *hx.generate_mesh3d(*shape))


def test_generated_graph_partitions_like_uploaded(hx):
    say("case generate partition")
    g = hx.Graph.generate_rmat(14, 8, 0.57, 0.19, 0.19, 3)
    h = hx.Graph(*hx.generate_rmat(14, 8, 0.57, 0.19, 0.19, 3))
    a, ra = g.partition(8, 1.03, 0)
    b, rb = h.partition(8, 1.03, 0)
    assert np.array_equal(a, b) and ra["final_cut"] == rb["final_cut"]


def test_generate_rejects_bad_arguments(hx):
    say("case generate errors")
    for call in (lambda: hx.Graph.generate_rmat(0, 4), lambda: hx.Graph.generate_rmat(8, 4, 0.9, 0.2, 0.1),
                 lambda: hx.Graph.generate_grid2d(0, 5), lambda: hx.Graph.generate_mesh3d(2, -1, 2)):
        with pytest.raises(hx.GpartError) as info:
            call()
        assert info.value.code == 1
