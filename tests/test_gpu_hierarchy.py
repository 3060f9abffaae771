"""GPU: refinement from an imported hierarchy.

A hierarchy written by the oracle in the int32 file format (include/hierarchy.hpp:42-126
of the reference) is read back and uncoarsened on the GPU from the oracle's labels of
its coarsest level, through both host-side entries of the reference:
refine_hierarchy_device(const Hierarchy&, host labels, options)
(include/refine.hpp:69-73) and uncoarsen(const Hierarchy&, host labels, options)
(include/uncoarsen.hpp:10-14). The labels must equal the oracle's uncoarsen of the
same hierarchy; the cut is recounted independently.
"""
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


CASES = {
    "grid256": lambda hx: hx.generate_grid2d(256, 256),
    "rmat12": lambda hx: hx.generate_rmat(12, 16),
    "mesh24": lambda hx: hx.generate_mesh3d(24, 24, 24),
}


@pytest.mark.parametrize("width", ["standard", "big"])
@pytest.mark.parametrize("name,k", [("grid256", 4), ("rmat12", 8), ("mesh24", 32)])
def test_uncoarsen_imported_hierarchy(hx, oracle, tmp_path, name, k, width):
    say(f"case imported hierarchy graph={name} k={k} width={width}")
    indptr, indices = CASES[name](hx)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    csf = max(64, (512 + k - 1) // k)
    oh = oracle.coarsen(og, k, seed=0, cluster_size_factor=csf, pack_isolated=True)
    path = tmp_path / "imported.hierarchy"
    oh.write(path)
    coarsest = oh.level(len(oh) - 1)
    start = oracle.initial_partition(coarsest, k, 1.03, 0)
    expect, expect_cut = oracle.uncoarsen(oh, start, k, 1.03, 32, 2.0)

    h = hx.Hierarchy.read(path, width=width)
    assert len(h) == len(oh)
    g = hx.Graph(indptr, indices, width=width)
    cap = hx.partition_capacity(len(indptr) - 1, k, 1.03)
    for entry in ("refine_hierarchy_device", "uncoarsen"):
        labels, cut = h.uncoarsen(start, k, 1.03, 32, 2.0, entry=entry)
        assert np.array_equal(labels, expect), entry
        assert g.cut(labels) == expect_cut == oracle.cut(og, labels)
        if cut is not None:
            assert cut == expect_cut
        assert np.bincount(labels, minlength=k).max() <= cap


def test_uncoarsen_imported_hierarchy_strict(hx, oracle, tmp_path):
    """Strict verification (projection cut and weights checked per level) passes on an
    imported hierarchy, and wrong-sized coarsest labels are refused."""
    say("case imported hierarchy strict")
    indptr, indices = hx.generate_grid2d(128, 128)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    oh = oracle.coarsen(og, 4, seed=1, cluster_size_factor=128, pack_isolated=True)
    path = tmp_path / "strict.hierarchy"
    oh.write(path)
    start = oracle.initial_partition(oh.level(len(oh) - 1), 4, 1.03, 1)
    expect, _ = oracle.uncoarsen(oh, start, 4, 1.03, 32, 2.0)
    h = hx.Hierarchy.read(path)
    labels, _ = h.uncoarsen(start, 4, 1.03, 32, 2.0, verify=True)
    assert np.array_equal(labels, expect)
    with pytest.raises(hx.GpartError):
        h.uncoarsen(start[:-1], 4)
