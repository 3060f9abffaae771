"""Refinement rounds recorded as hipGraphs and replayed (hip/refine.hip, round
graphs): partitions equal those of the kernel-by-kernel launches, on the first
partition of a graph (recording) and on the next ones (replays), with and without
a timed family splitting the rounds into segments; and refinement from given
labels equals the oracle's with graphs on."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


@pytest.fixture
def graphs_restored(hx):
    yield
    hx.set_refine_graphs(True)
    for family in ("rebalance", "cache_update", "refine_filter"):
        hx.profile_enable(family, False)


CASES = [("mesh24", lambda hx: hx.generate_mesh3d(24, 24, 24), 32, "batched_fp64"),
         ("grid128", lambda hx: hx.generate_grid2d(128, 128), 4, "batched_fp64"),
         ("rmat13", lambda hx: hx.generate_rmat(13, 16), 8, "batched_fp32")]


@pytest.mark.parametrize("name,make,k,mode", CASES, ids=[c[0] for c in CASES])
def test_round_graphs_equal_direct_launches(hx, graphs_restored, name, make, k, mode):
    indptr, indices = make(hx)
    g = hx.Graph(indptr, indices)
    say(f"case round graphs direct graph={name} k={k}")
    hx.set_refine_graphs(False)
    direct, drep = g.partition(k, 1.03, 0, spectral_mode=mode)
    assert hx.refine_graph_stats()["captures"] == 0
    hx.set_refine_graphs(True)
    for attempt in range(3):
        say(f"case round graphs replay graph={name} k={k} attempt={attempt}")
        labels, rep = g.partition(k, 1.03, 0, spectral_mode=mode)
        assert np.array_equal(labels, direct), attempt
        assert rep["final_cut"] == drep["final_cut"] == g.cut(labels)
    stats = hx.refine_graph_stats()
    assert stats["captures"] > 0 and stats["replays"] > stats["captures"]
    # timed families run outside the graphs: the rounds split into segments
    for family in ("rebalance", "cache_update", "refine_filter"):
        hx.profile_enable(family, True)
    for attempt in range(2):
        say(f"case round graphs profiled graph={name} k={k} attempt={attempt}")
        labels, _ = g.partition(k, 1.03, 0, spectral_mode=mode)
        assert np.array_equal(labels, direct), attempt


@pytest.mark.parametrize("width", ["standard", "big"])
def test_round_graphs_refine_matches_oracle(hx, oracle, graphs_restored, width):
    """Many rounds and rebalance iterations from balanced random labels."""
    indptr, indices = hx.generate_rmat(11, 8)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    g = hx.Graph(indptr, indices, width=width)
    hx.set_refine_graphs(True)
    for k in (4, 32):
        start = (np.random.default_rng(k).permutation(len(indptr) - 1) % k).astype(np.int32)
        expect, st = oracle.refine(og, start, k, 1.03)
        for attempt in range(2):
            say(f"case round graphs refine width={width} k={k} attempt={attempt}")
            labels, rep = g.refine(start, k, 1.03)
            assert np.array_equal(labels, expect), (k, attempt)
            assert rep["final_cut"] == st.final_cut
            assert rep["rounds"] == st.rounds and rep["admission_iterations"] == st.admission_iterations
