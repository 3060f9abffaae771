"""GPU: the CLIs end to end, checked the way the reference's own scripts check them.

- gpart_partition_{standard,big}: int64 CSR files in, int32 labels out, verified
  independently (cut and loads in closed form, capacity floor(ratio * n / k) as
  tools/verify_partition.py:42-65 computes it) and compared with the oracle.
- gpart_coarsen_standard with SCLP_DIAGNOSTICS=1 and --verify on the graph of
  tests/test_sclp_small.py: the log keys that script parses and its checks of
  every level of the exported hierarchy.
"""
from __future__ import annotations

import os
import re
import subprocess
from decimal import Decimal
from pathlib import Path

import numpy as np
import pytest

import kat_graphs as K

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "horkxg_amd" / "bin"


@pytest.fixture(scope="module", autouse=True)
def need_gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")


def write_csr(tmp_path, indptr, indices):
    ip, ix = tmp_path / "indptr.bin", tmp_path / "indices.bin"
    np.asarray(indptr, np.int64).tofile(ip)
    np.asarray(indices, np.int64).tofile(ix)
    return ip, ix


def verify(indptr, indices, labels, k, ratio):
    """tools/verify_partition.py:42-65 restated: labels in range, loads by vertex
    count under floor(Decimal(ratio) * n / k), cut = cross directed entries / 2."""
    n = len(indptr) - 1
    assert len(labels) == n and labels.min() >= 0 and labels.max() < k
    cap = int(Decimal(str(ratio)) * n / k)
    assert np.bincount(labels, minlength=k).max() <= cap
    src = np.repeat(np.arange(n), np.diff(indptr))
    directed = int((labels[src] != labels[indices]).sum())
    assert directed % 2 == 0
    return directed // 2


@pytest.mark.parametrize("cli", ["gpart_partition_standard", "gpart_partition_big"])
@pytest.mark.parametrize("k", [4, 32])
def test_partition_cli(hx, oracle, tmp_path, cli, k):
    print(f"case cli {cli} k={k}", flush=True)
    indptr, indices = hx.generate_grid2d(128, 128)
    ip, ix = write_csr(tmp_path, indptr, indices)
    out = tmp_path / "part.bin"
    run = subprocess.run([str(BIN / cli), str(ip), str(ix), str(k), str(out), "1.03", "0"],
                         capture_output=True, text=True, timeout=300)
    assert run.returncode == 0, run.stderr
    labels = np.fromfile(out, np.int32)
    cut = verify(indptr, indices, labels, k, 1.03)
    logged = int(re.search(r"final_edge_cut=(\d+)", run.stdout).group(1))
    assert logged == cut
    assert float(re.search(r"algorithm_total_seconds=([0-9.e+-]+)", run.stdout).group(1)) > 0
    ol, res = oracle.partition(oracle.Graph(indptr, indices.astype(np.int32)), k, 1.03, 0)
    assert np.array_equal(labels, ol)
    assert int(res.final_cut) == cut


def read_hierarchy(path):
    raw = np.fromfile(path, dtype="<i4")
    pos, levels = 1, []
    for li in range(int(raw[0])):
        n, m = int(raw[pos]), int(raw[pos + 1])
        pos += 2
        rows = raw[pos:pos + n + 1]; pos += n + 1
        cols = raw[pos:pos + m]; pos += m
        ew = raw[pos:pos + m]; pos += m
        vw = raw[pos:pos + n]; pos += n
        mp = None
        if li:
            fn = len(levels[-1][0]) - 1
            mp = raw[pos:pos + fn]; pos += fn
        levels.append((rows, cols, ew, vw, mp))
    assert pos == len(raw)
    return levels


def test_coarsen_cli_sclp_small(tmp_path):
    """tests/test_sclp_small.py:114-151: parts 4, seed 0, stop ratio 1.0, 24 levels,
    --verify, SCLP_DIAGNOSTICS=1."""
    print("case cli coarsen sclp_small", flush=True)
    g = K.sclp_small()
    ip, ix = write_csr(tmp_path, g["indptr"], g["indices"])
    out = tmp_path / "sclp.hierarchy"
    env = dict(os.environ, SCLP_DIAGNOSTICS="1")
    run = subprocess.run([str(BIN / "gpart_coarsen_standard"), str(ip), str(ix), "4", str(out), "0", "1.0", "24",
                          "--verify"], capture_output=True, text=True, timeout=300, env=env)
    assert run.returncode == 0, run.stderr
    text = run.stdout
    capacities = [int(x) for x in re.findall(r"ml_gpu_sclp .*?cluster_cap=(\d+)", text)]
    assert "projection_cut=ok" in text
    assert "predicted_gain=" in text and "mover_fraction=" in text
    gains = [int(x) for x in re.findall(r"actual_gain=(\d+)", text)]
    assert gains and all(x >= 0 for x in gains)
    ratios = [float(x) for x in re.findall(r"ml_gpu_sclp level=0 .*?contraction_ratio=([0-9.e+-]+)", text)]
    assert ratios and 0.45 <= ratios[0] <= 0.60
    levels = read_hierarchy(out)
    assert len(levels) >= 2
    rng = np.random.default_rng(17)
    for li, (rows, cols, ew, vw, mp) in enumerate(levels):
        n = len(rows) - 1
        assert rows[0] == 0 and rows[-1] == len(cols) and np.all(np.diff(rows) >= 0)
        assert len(cols) == 0 or (cols.min() >= 0 and cols.max() < n)
        assert np.all(ew > 0) and np.all(vw > 0)
        src = np.repeat(np.arange(n), np.diff(rows))
        pairs = dict(zip(zip(src.tolist(), cols.tolist()), ew.tolist()))
        for (a, b), w in pairs.items():
            if li:
                assert a != b
            assert pairs.get((b, a)) == w
        if mp is not None:
            frows, fcols, few, fvw, _ = levels[li - 1]
            assert mp.min() >= 0 and mp.max() < n and len(np.unique(mp)) == n
            assert np.array_equal(np.bincount(mp, weights=fvw, minlength=n).astype(np.int64), vw.astype(np.int64))
            labels = rng.integers(0, 7, n)
            fine = dict(indptr=frows, indices=fcols.astype(np.int64), edge_weights=few.astype(np.uint64))
            coarse = dict(indptr=rows, indices=cols.astype(np.int64), edge_weights=ew.astype(np.uint64))
            assert K.independent_cut(fine, labels[mp]) == K.independent_cut(coarse, labels)
            assert int(vw.max()) <= capacities[li - 1]
