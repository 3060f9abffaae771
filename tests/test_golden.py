"""Golden fixtures (tests/golden, written by tests/golden/make_golden.py from the oracle).

CPU: the oracle still reproduces every fixture (a change to the restatement shows
up here). GPU: the HIP path reproduces the same digests and partitions without
running the oracle at all.
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = json.loads((GOLDEN / "golden.json").read_text())
SPECTRAL = dict(np.load(GOLDEN / "spectral.npz"))


def digest(array, dtype):
    return hashlib.sha256(np.ascontiguousarray(array, dtype=dtype).tobytes()).hexdigest()


def make_input(hx, name):
    if name == "grid256":
        return hx.generate_grid2d(256, 256)
    if name == "rmat12":
        return hx.generate_rmat(12, 16)
    raise KeyError(name)


def level_digests(arrays, mp):
    off, adj, ew, vw = arrays
    rec = {"n": len(off) - 1, "m": len(adj), "offsets": digest(off, "<i8"), "neighbors": digest(adj, "<i4"),
           "edge_weights": digest(ew, "<u8"), "vertex_weights": digest(vw, "<u8")}
    if mp is not None:
        rec["map"] = digest(mp, "<i4")
    return rec


# ---------------------------------------------------------------- CPU: oracle

@pytest.mark.parametrize("key", sorted(FIXTURES["hierarchy"]))
def test_oracle_hierarchy_golden(hx, oracle, key):
    spec = FIXTURES["hierarchy"][key]
    ip, ix = make_input(hx, key.split("_")[0])
    g = oracle.Graph(ip, ix.astype(np.int32))
    h = oracle.coarsen(g, spec["parts"], spec["seed"], 0.90, 24, spec["cluster_size_factor"],
                       spec["pack_isolated"])
    assert h.stop_reason == spec["stop_reason"]
    assert len(h) == len(spec["levels"])
    for i, want in enumerate(spec["levels"]):
        got = level_digests(h.level(i).arrays(), h.map(i) if i + 1 < len(h) else None)
        assert got == want, (key, i)


@pytest.mark.parametrize("key", sorted(FIXTURES["partition"]))
def test_oracle_partition_golden(hx, oracle, key):
    spec = FIXTURES["partition"][key]
    ip, ix = make_input(hx, key.split("_")[0])
    g = oracle.Graph(ip, ix.astype(np.int32))
    labels, res = oracle.partition(g, spec["parts"], 1.03, spec["seed"])
    assert int(res.final_cut) == spec["cut"]
    assert digest(labels, "<i4") == spec["labels"]
    cap = oracle.partition_capacity(len(ip) - 1, spec["parts"], 1.03)
    assert max(spec["loads"]) <= cap and sum(spec["loads"]) == len(ip) - 1


def test_oracle_spectral_golden(oracle):
    s = SPECTRAL
    g = oracle.Graph(s["offsets"], s["neighbors"], s["edge_weights"], s["vertex_weights"])
    labels, coords, _ = oracle.initial_partition(g, 4, 1.03, 0, with_coords=True)
    assert np.array_equal(coords.view(np.uint64), s["coords"].view(np.uint64))
    assert np.array_equal(labels, s["labels"])


# ---------------------------------------------------------------- GPU: HIP path

@pytest.fixture()
def gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")
    return hx


@pytest.mark.gpu
@pytest.mark.parametrize("key", sorted(FIXTURES["hierarchy"]))
def test_gpu_hierarchy_golden(gpu, key):
    print(f"case golden hierarchy {key}", flush=True)
    spec = FIXTURES["hierarchy"][key]
    ip, ix = make_input(gpu, key.split("_")[0])
    h = gpu.Graph(ip, ix).coarsen(spec["parts"], spec["seed"], 0.90, 24, spec["cluster_size_factor"],
                                  spec["pack_isolated"])
    assert h.stop_reason == spec["stop_reason"]
    assert len(h) == len(spec["levels"])
    for i, want in enumerate(spec["levels"]):
        got = level_digests(h.level(i), h.map(i) if i + 1 < len(h) else None)
        assert got == want, (key, i)


@pytest.mark.gpu
@pytest.mark.parametrize("key", sorted(FIXTURES["partition"]))
def test_gpu_partition_golden(gpu, key):
    print(f"case golden partition {key}", flush=True)
    spec = FIXTURES["partition"][key]
    ip, ix = make_input(gpu, key.split("_")[0])
    g = gpu.Graph(ip, ix)
    labels, rep = g.partition(spec["parts"], 1.03, spec["seed"])
    assert rep["final_cut"] == spec["cut"]
    assert digest(labels, "<i4") == spec["labels"]
    assert list(np.bincount(labels, minlength=spec["parts"])) == spec["loads"]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["plain_fp64", "batched_fp64", "batched_fp32"])
def test_gpu_spectral_golden(gpu, mode):
    """FP64 forms are bit-exact; the FP32 form keeps integer labels valid and its
    vector within 1e-3 of the FP64 one (reported, not bit-compared)."""
    print(f"case golden spectral {mode}", flush=True)
    s = SPECTRAL
    g = gpu.Graph(s["offsets"], s["neighbors"], edge_weights=s["edge_weights"], vertex_weights=s["vertex_weights"])
    labels, coords = g.initial_partition(4, 1.03, 0, mode, with_coords=True)
    if mode == "batched_fp32":
        cap = gpu.partition_capacity(int(s["vertex_weights"].sum()), 4, 1.03)
        loads = np.bincount(labels, weights=s["vertex_weights"].astype(np.float64), minlength=4)
        assert labels.min() >= 0 and labels.max() < 4 and loads.max() <= cap
        err = float(np.max(np.abs(coords - s["coords"])))
        print(f"fp32 diffusion max abs error vs fp64 oracle: {err:.3e}", flush=True)
        assert err < 1e-3
    else:
        assert np.array_equal(coords.view(np.uint64), s["coords"].view(np.uint64))
        assert np.array_equal(labels, s["labels"])
