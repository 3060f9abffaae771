"""The single-launch refinement engine (hip/refine_engine.hip).

CPU: the engine's stale-round test, (long double)current < 0.999L * best of the
reference's loop (src/refine_incremental.cuh:349), evaluated by the library's
integer restatement, against the x87 unit itself through numpy.longdouble, on
the cases where rounding decides: best a multiple of 1000 with current at
0.999 * best and its neighbours, and random values around the threshold.

GPU: partitions and single-level refinements with the engine at several grid
sizes (one workgroup, a few, the maximum) and with the engine switched off are
label-identical to each other and to the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest


def _x87_keep(current: int, best: int) -> bool:
    ratio = np.longdouble(999) / np.longdouble(1000)  # correctly rounded, as the literal 0.999L
    return bool(np.longdouble(current) < ratio * np.longdouble(best))


def test_long_double_is_x87_extended():
    assert np.finfo(np.longdouble).nmant == 63


def test_keep_ratio_matches_x87_at_exact_multiples(hx):
    rng = np.random.default_rng(5)
    bests = [1000, 2000, 999000, 10 ** 9, 123456789000, 2 ** 40 * 1000 // 1000 * 1000]
    bests += [int(x) * 1000 for x in rng.integers(1, 2 ** 40, 200)]
    checked = 0
    for best in bests:
        edge = best * 999 // 1000
        for current in (edge - 1, edge, edge + 1):
            assert hx.refine_keep_ratio(current, best) == _x87_keep(current, best), (current, best)
            checked += 1
    assert checked == 3 * len(bests)


def test_keep_ratio_matches_x87_random(hx):
    rng = np.random.default_rng(11)
    for _ in range(3000):
        best = int(rng.integers(1, 2 ** 52))
        current = best * 999 // 1000 + int(rng.integers(-3, 4))
        current = max(current, 0)
        assert hx.refine_keep_ratio(current, best) == _x87_keep(current, best), (current, best)
    assert not hx.refine_keep_ratio(0, 0)
    assert not hx.refine_keep_ratio(5, 5)
    assert hx.refine_keep_ratio(0, 5)


def say(*args):
    print(*args, flush=True)


GRAPHS = {
    "grid256": lambda hx: hx.generate_grid2d(256, 256),
    "rmat12": lambda hx: hx.generate_rmat(12, 16),
    "mesh24": lambda hx: hx.generate_mesh3d(24, 24, 24),
}

# (max_vertices, vertices_per_group, max_grid): engine off, one workgroup, a few, many
ENGINE_SETTINGS = [(0, 8192, 64), (1 << 20, 1 << 30, 1), (1 << 20, 8192, 64), (1 << 20, 512, 128)]


@pytest.mark.gpu
@pytest.mark.parametrize("name,k", [("grid256", 4), ("grid256", 32), ("rmat12", 8), ("mesh24", 32), ("mesh24", 2)])
def test_engine_partitions_match_oracle(hx, oracle, name, k):
    if hx.device_count() < 1:
        pytest.skip("no GPU")
    indptr, indices = GRAPHS[name](hx)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    expect, _ = oracle.partition(og, k, 1.03, 0)
    g = hx.Graph(indptr, indices)
    try:
        for setting in ENGINE_SETTINGS:
            say(f"case engine partition graph={name} k={k} setting={setting}")
            hx.set_refine_engine(setting[0], setting[1], setting[2])
            labels, rep = g.partition(k, 1.03, 0)
            assert np.array_equal(labels, expect), setting
            assert rep["final_cut"] == oracle.cut(og, labels)
    finally:
        hx.set_refine_engine()


@pytest.mark.gpu
@pytest.mark.parametrize("width", ["standard", "big"])
def test_engine_refine_partition_matches_oracle(hx, oracle, width):
    """refine_partition from balanced random labels (many rounds and rebalance
    iterations) at each engine setting."""
    if hx.device_count() < 1:
        pytest.skip("no GPU")
    indptr, indices = hx.generate_rmat(11, 8)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    g = hx.Graph(indptr, indices, width=width)
    try:
        for k in (4, 32):
            start = (np.random.default_rng(k).permutation(len(indptr) - 1) % k).astype(np.int32)
            expect, st = oracle.refine(og, start, k, 1.03)
            for setting in ENGINE_SETTINGS:
                say(f"case engine refine width={width} k={k} setting={setting}")
                hx.set_refine_engine(setting[0], setting[1], setting[2])
                labels, rep = g.refine(start, k, 1.03)
                assert np.array_equal(labels, expect), (k, setting)
                assert rep["final_cut"] == st.final_cut
                assert rep["rounds"] == st.rounds and rep["admission_iterations"] == st.admission_iterations
    finally:
        hx.set_refine_engine()
