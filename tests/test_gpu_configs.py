"""GPU: the benchmark configurations at their full sizes.

Every partition is checked the way the reference's own harness checks a run
(tests/synthetic_suite.py:94-103 of the reference): an independent recount of the cut
by the verifier equals the reported cut, every part is within
floor(1.03 * W / k), and every label lies in [0, k).

- mesh 128^3, k = 32: labels equal the CPU oracle's (seed 0);
- R-MAT scale 22, k = 8: FP32 bisection (the configuration's form) and FP64, the
  FP32 vector error per recursion depth, plain and batched FP64 labels identical;
- the 256-graph mix: batched FP64 labels equal to plain FP64 graph by graph, the
  batched FP32 form valid on every graph;
- R-MAT scale 26 at big width (64-bit offsets and weights), k = 32.
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


def check_partition(g, labels, rep, k, n_expected=None):
    """Verifier recount == reported cut, loads within capacity, labels in range."""
    n = g.n if n_expected is None else n_expected
    assert labels.shape == (n,)
    assert labels.min() >= 0 and labels.max() < k
    ver = g.verify(labels, k, "1.03")
    assert ver["capacity_ok"], ver
    assert ver["cut"] == rep["final_cut"], (ver["cut"], rep["final_cut"])
    assert ver["max_load"] <= rep["capacity"] == ver["capacity"]
    assert sum(ver["loads"]) == n
    return ver


def test_mesh128_k32_matches_oracle(hx, oracle):
    say("case config mesh128 k=32 vs oracle")
    indptr, indices = hx.generate_mesh3d(128, 128, 128)
    assert len(indptr) - 1 == 2_097_152 and len(indices) == 12_484_608
    g = hx.Graph(indptr, indices)
    labels, rep = g.partition(32, 1.03, 0)
    check_partition(g, labels, rep, 32)
    assert rep["capacity"] == 67_502
    og = oracle.Graph(indptr, indices.astype(np.int32))
    ol, orep = oracle.partition(og, 32, 1.03, 0)
    assert rep["final_cut"] == orep.final_cut
    assert np.array_equal(labels, ol)


@pytest.fixture(scope="module")
def rmat22(hx):
    g = hx.Graph.generate_rmat(22, 16, 0.57, 0.19, 0.19, 0)
    yield g
    del g
    hx.trim()


def test_rmat22_k8_fp32(hx, rmat22):
    say("case config rmat22 k=8 batched_fp32")
    g = rmat22
    assert g.n == 4_194_304 and g.m > 100_000_000
    labels, rep = g.partition(8, 1.03, 0, spectral_mode="batched_fp32")
    check_partition(g, labels, rep, 8)
    assert rep["capacity"] == 540_016


def test_rmat22_k8_fp64_plain_equals_batched(hx, rmat22):
    say("case config rmat22 k=8 plain_fp64 vs batched_fp64")
    g = rmat22
    plain, prep = g.partition(8, 1.03, 0, spectral_mode="plain_fp64")
    batched, brep = g.partition(8, 1.03, 0, spectral_mode="batched_fp64")
    check_partition(g, plain, prep, 8)
    assert np.array_equal(plain, batched)
    assert prep["final_cut"] == brep["final_cut"]


def test_rmat22_k8_fp32_error_per_depth(hx, rmat22):
    """FP32 diffusion against FP64 on the coarsest level of the production
    hierarchy: every depth's vector error below 1e-3, both complete recursions
    within capacity."""
    say("case config rmat22 k=8 fp32 error per depth")
    k = 8
    h = rmat22.coarsen(k, seed=0, cluster_size_factor=64, pack_isolated=True)
    off, adj, ew, vw = h.level(len(h) - 1)
    del h
    coarse = hx.Graph(off, adj, edge_weights=ew, vertex_weights=vw)
    rep = coarse.spectral_compare(k, 1.03, 0)
    assert len(rep["depths"]) == 3
    for d in rep["depths"]:
        say(f"  depth {d['depth']}: max {d['max_abs_error']:.3g} rms {d['rms_error']:.3g} "
            f"splits differ {d['splits_differ']} vertices {d['labels_differ']}")
        assert 0.0 <= d["rms_error"] <= d["max_abs_error"] < 1e-3, d
    cap = hx.partition_capacity(int(vw.astype(object).sum()), k, 1.03)
    assert max(rep["max_load_fp64"], rep["max_load_fp32"]) <= cap


def test_mix256(hx):
    """All 256 graphs of the mix: the batched FP64 form equals plain FP64 graph by
    graph; the batched FP32 form is valid on every graph."""
    from horkxg_amd import workloads as W

    specs = W.mix_specs()
    assert len(specs) == 256
    say("case config mix256: generating")
    graphs = [W.device_graph(hx, s) for s in specs]
    parts = [s["k"] for s in specs]
    ratios = [s["ratio"] for s in specs]
    seeds = [s["seed"] for s in specs]
    say("case config mix256: batched fp64")
    lab64, rep64, _ = hx.partition_batch(graphs, parts, ratios, seeds, spectral_mode="batched_fp64")
    say("case config mix256: batched fp32")
    lab32, rep32, wall32 = hx.partition_batch(graphs, parts, ratios, seeds, spectral_mode="batched_fp32")
    assert wall32 > 0
    say("case config mix256: plain fp64 per graph")
    for s, g, l64, r64, l32, r32 in zip(specs, graphs, lab64, rep64, lab32, rep32):
        plain, prep = g.partition(s["k"], s["ratio"], s["seed"], spectral_mode="plain_fp64")
        assert np.array_equal(plain, l64), s
        assert prep["final_cut"] == r64["final_cut"], s
        check_partition(g, l64, r64, s["k"])
        check_partition(g, l32, r32, s["k"])
    del graphs
    hx.trim()


def test_rmat26_k32_big(hx):
    """R-MAT scale 26 (67,108,864 vertices, about 2.1e9 directed entries) at big
    width, generated on the GPU."""
    hx.trim()
    say("case config rmat26 k=32 big: generating")
    g = hx.Graph.generate_rmat(26, 16, 0.57, 0.19, 0.19, 0, width="big")
    try:
        assert g.n == 67_108_864 and g.m > 2_000_000_000
        say("case config rmat26 k=32 big: partition")
        labels, rep = g.partition(32, 1.03, 0, spectral_mode="batched_fp32")
        say(f"  cut {rep['final_cut']} seconds {rep['algorithm_seconds']:.2f}")
        check_partition(g, labels, rep, 32)
        assert rep["capacity"] == 2_160_066
    finally:
        del g
        hx.trim()
