"""GPU: the batched graph mix. Several host threads (one stream and allocator each)
partition independent graphs concurrently; every partition must equal the one the
same graph gets on its own, which in the FP64 forms equals the oracle's."""
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


@pytest.fixture(scope="module")
def small_mix(hx):
    from horkxg_amd import workloads as W

    specs = W.mix_specs(12, min_n=2_000, max_n=40_000)
    graphs = [None] * len(specs)
    arrays = []
    for i, s in enumerate(specs):
        indptr, indices = W.generate(hx, s)
        arrays.append((indptr, indices))
        graphs[i] = hx.Graph(indptr, indices)
    return specs, graphs, arrays


CONCURRENT_OFF = ("concurrent batch workers are switched off: the full 256-graph mix with four "
                  "workers ended in a host segmentation fault whose cause is not established")


@pytest.mark.parametrize("workers", [1, pytest.param(3, marks=pytest.mark.skip(reason=CONCURRENT_OFF))])
def test_batch_fp64_matches_oracle(hx, oracle, small_mix, workers):
    specs, graphs, arrays = small_mix
    say(f"case batch fp64 workers={workers} graphs={len(graphs)}")
    labels, reports, wall = hx.partition_batch(graphs, [s["k"] for s in specs], [s["ratio"] for s in specs],
                                               [s["seed"] for s in specs], spectral_mode="batched_fp64",
                                               workers=workers)
    assert wall > 0
    for s, g, lab, rep, (indptr, indices) in zip(specs, graphs, labels, reports, arrays):
        og = oracle.Graph(indptr, indices.astype(np.int32))
        ol, orep = oracle.partition(og, s["k"], s["ratio"], s["seed"])
        assert rep["final_cut"] == orep.final_cut, s
        assert np.array_equal(lab, ol), s
        assert g.cut(lab) == rep["final_cut"]
        assert np.bincount(lab, minlength=s["k"]).max() <= rep["capacity"]


def test_batch_fp32_equals_single_graph_runs(hx, small_mix):
    """The batched FP32 form: valid partitions, identical to each graph's own run."""
    specs, graphs, _ = small_mix
    say(f"case batch fp32 graphs={len(graphs)}")
    labels, reports, _ = hx.partition_batch(graphs, [s["k"] for s in specs], [s["ratio"] for s in specs],
                                            [s["seed"] for s in specs], spectral_mode="batched_fp32", workers=1)
    for s, g, lab, rep in zip(specs, graphs, labels, reports):
        single, srep = g.partition(s["k"], s["ratio"], s["seed"], spectral_mode="batched_fp32")
        assert np.array_equal(lab, single), s
        assert rep["final_cut"] == srep["final_cut"] == g.cut(lab)
        assert np.bincount(lab, minlength=s["k"]).max() <= rep["capacity"]


def test_batch_error_is_reported(hx, small_mix):
    """A bad part count in one graph fails the whole call with its error code; more
    than one worker is refused while concurrent workers are switched off."""
    specs, graphs, _ = small_mix
    say("case batch error")
    parts = [s["k"] for s in specs]
    parts[5] = 64
    with pytest.raises(hx.GpartError) as info:
        hx.partition_batch(graphs, parts, workers=1, want_labels=False)
    assert info.value.code == 1
    with pytest.raises(hx.GpartError) as info:
        hx.partition_batch(graphs[:2], [s["k"] for s in specs[:2]], workers=2, want_labels=False)
    assert info.value.code == 1
    # the library stays usable
    _, reports, _ = hx.partition_batch(graphs[:2], [s["k"] for s in specs[:2]], workers=1, want_labels=False)
    assert len(reports) == 2


def test_batch_groups_equal_one_at_a_time(hx, small_mix):
    """Groups of graphs whose initial partitions share one persistent launch give
    every graph the labels it gets alone, in both batched precisions."""
    specs, graphs, _ = small_mix
    args = ([s["k"] for s in specs], [s["ratio"] for s in specs], [s["seed"] for s in specs])
    saved = hx.batch_group()
    try:
        for mode in ("batched_fp64", "batched_fp32"):
            hx.set_batch_group(1)
            say(f"case batch groups mode={mode} group=1")
            alone, arep, _ = hx.partition_batch(graphs, *args, spectral_mode=mode)
            for group in (5, 16):
                hx.set_batch_group(group)
                say(f"case batch groups mode={mode} group={group}")
                together, trep, _ = hx.partition_batch(graphs, *args, spectral_mode=mode)
                for s, a, t, ra, rt in zip(specs, alone, together, arep, trep):
                    assert np.array_equal(a, t), (s, group)
                    assert ra["final_cut"] == rt["final_cut"] and ra["initial_cut"] == rt["initial_cut"]
    finally:
        hx.set_batch_group(saved)
