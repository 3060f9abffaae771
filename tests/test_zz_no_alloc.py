"""GPU: gpart_cut_noalloc, the cut entry point that allocates nothing.

It runs gpart_cut's kernels over a static device array instead of context
allocations. Checked through the C ABI against gpart_cut (and the oracle's cut)
on two calls in a row, the second on a smaller graph so that stale scratch from
the first call would show; the context allocator's counters must not move during
a no-allocation call; a CSR beyond the array's limit is refused on the host.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")


def _labels(n, k, seed):
    return np.random.default_rng(seed).integers(0, k, n).astype(np.int32)


@pytest.mark.parametrize("width", ["standard", "big"])
def test_noalloc_cut_equals_cut_twice_shrinking(hx, oracle, width):
    from horkxg_amd import device as D

    cases = [("rmat14", hx.generate_rmat(14, 16), 8, 1), ("grid64", hx.generate_grid2d(64, 64), 4, 2)]
    for name, (indptr, indices), k, seed in cases:
        say(f"case cut_noalloc graph={name} width={width} k={k}")
        g = D.DeviceGraph(indptr, indices, width=width)
        labels = _labels(g.n, k, seed)
        dl = D.DeviceArray(labels, np.int32)
        expect = D.cut(g, dl)
        before = hx.cache_info()
        got = D.cut(g, dl, noalloc=True)
        after = hx.cache_info()
        assert got == expect == oracle.cut(oracle.Graph(indptr, indices.astype(np.int32)), labels)
        for key in ("segment_allocations", "live_bytes", "live_blocks", "reserved_bytes"):
            assert after[key] == before[key], key


def test_noalloc_cut_benchmark_mesh(hx):
    """The 128^3 mesh (12,484,608 entries) with random labels, then a much smaller
    graph: both equal gpart_cut."""
    from horkxg_amd import device as D

    for name, (indptr, indices) in (("mesh128", hx.generate_mesh3d(128, 128, 128)),
                                    ("grid16", hx.generate_grid2d(16, 16))):
        say(f"case cut_noalloc graph={name}")
        g = D.DeviceGraph(indptr, indices)
        dl = D.DeviceArray(_labels(g.n, 32, 7), np.int32)
        assert D.cut(g, dl, noalloc=True) == D.cut(g, dl)


def test_noalloc_cut_refuses_oversized_csr(hx):
    """A view claiming more entries than the array covers is refused before any
    kernel runs (the pointers are never read)."""
    from horkxg_amd import device as D

    say("case cut_noalloc oversized")
    limit = D.cut_noalloc_max_entries()
    assert limit == (1 << 20) * 4096
    view = D.CsrView(16, limit + 1, 0, None, None, None, None)
    out = C.c_uint64()
    H = hx._native.hip_lib()
    rc = H.gpart_cut_noalloc(hx._native._ctx(), C.byref(view), None, C.byref(out))
    assert rc == 1
    assert "exceed its limit" in H.gpart_last_error().decode()
