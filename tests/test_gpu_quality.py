"""The GPU verifier and quality metrics against the oracle (oracle/quality_oracle.py):
exact equality of every counter and ratio, the reference's error messages, the
LDS and global-counter paths, both widths, and the command-line tools."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import quality_oracle as qo

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def random_symmetric(n, avg_degree, seed, weights=False):
    rng = np.random.default_rng(seed)
    m = n * avg_degree // 2
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    keep = u != v
    a, b = np.minimum(u[keep], v[keep]), np.maximum(u[keep], v[keep])
    key = np.unique(a.astype(np.int64) * n + b)
    a, b = key // n, key % n
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    w = None
    if weights:
        wu = rng.integers(1, 1000, key.size)
        w_all = np.concatenate([wu, wu])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, src + 1, 1)
    indptr = np.cumsum(indptr)
    if weights:
        w = w_all[order].astype(np.uint64)
    return indptr, dst.astype(np.int64), w


def same(a, b):
    assert set(a) == set(b), set(a) ^ set(b)
    for key in a:
        if key == "per_part":
            for name in a[key]:
                assert np.array_equal(np.asarray(a[key][name], np.int64), np.asarray(b[key][name], np.int64)), name
        else:
            assert a[key] == b[key], (key, a[key], b[key])


GRAPHS = [("rmat", 12), ("mesh", 16), ("random", 5000)]


def make(hx, kind, size, seed=0):
    if kind == "rmat":
        indptr, indices = hx.generate_rmat(size, 8, 0.57, 0.19, 0.19, seed)
    elif kind == "mesh":
        indptr, indices = hx.generate_mesh3d(size, size, size)
    else:
        indptr, indices, _ = random_symmetric(size, 12, seed)
    return np.asarray(indptr, np.int64), np.asarray(indices, np.int64)


@pytest.mark.parametrize("width", ["standard", "big"])
@pytest.mark.parametrize("kind,size", GRAPHS)
def test_verify_and_analyze_match_oracle(hx, kind, size, width):
    indptr, indices = make(hx, kind, size)
    g = hx.Graph(indptr, indices, width=width)
    n = len(indptr) - 1
    rng = np.random.default_rng(size)
    for k in (2, 7, 32):
        for labels in (rng.integers(0, k, n).astype(np.int32), np.sort(rng.integers(0, k, n)).astype(np.int32)):
            for ratio in ("1.03", "1.5", "3"):
                assert g.verify(labels, k, ratio) == qo.verify(indptr, indices, labels, k, ratio)
            same(g.analyze(labels), {**qo.analyze(indptr, indices, labels)})
            same(g.analyze(labels, k + 3), qo.analyze(indptr, indices, labels, k + 3))
    labels, rep = g.partition(8, 1.03, 0)
    v = g.verify(labels, 8)
    assert v == qo.verify(indptr, indices, labels, 8)
    assert v["cut"] == rep["final_cut"] and v["capacity_ok"]


def test_weighted_metrics_match_oracle(hx):
    indptr, indices, w = random_symmetric(3000, 10, 4, weights=True)
    g = hx.Graph(indptr, indices, edge_weights=w)
    labels = np.random.default_rng(1).integers(0, 5, 3000).astype(np.int32)
    same(g.analyze(labels), qo.analyze(indptr, indices, labels, weights=w.astype(np.int64)))


@pytest.mark.parametrize("parts", [1024, 1025, 3000])
def test_many_parts_lds_and_global_paths(hx, parts):
    indptr, indices, _ = random_symmetric(20000, 8, parts)
    g = hx.Graph(indptr, indices)
    labels = np.random.default_rng(parts).integers(0, parts, 20000).astype(np.int32)
    same(g.analyze(labels), qo.analyze(indptr, indices, labels))


def test_verify_errors_carry_reference_messages(hx):
    indptr, indices = make(hx, "mesh", 6)
    g = hx.Graph(indptr, indices)
    n = len(indptr) - 1
    labels = np.zeros(n, np.int32)
    for parts, ratio in ((1, "1.03"), (33, "1.03"), (4, "0.9"), (4, "x")):
        with pytest.raises(hx.GpartError, match="invalid parts or imbalance ratio"):
            g.verify(labels, parts, ratio)
    bad = labels.copy()
    bad[17], bad[40], bad[3] = 9, -2, 4
    with pytest.raises(hx.GpartError, match=r"^invalid label: 4$"):
        g.verify(bad, 4)
    # one-directional entry: odd non-loop count
    d = hx.Graph(np.array([0, 1, 1, 1], np.int64), np.array([1], np.int64))
    with pytest.raises(hx.GpartError, match="odd undirected"):
        d.verify(np.array([0, 1, 0], np.int32), 2)
    over = np.zeros(n, np.int32)
    over[: n // 4] = 1
    r = g.verify(over, 2)
    assert not r["capacity_ok"] and r == qo.verify(indptr, indices, over, 2)


def test_analyze_errors(hx):
    indptr, indices = make(hx, "mesh", 5)
    g = hx.Graph(indptr, indices)
    labels = np.arange(len(indptr) - 1, dtype=np.int32) % 6
    with pytest.raises(hx.GpartError, match="part id >= --parts"):
        g.analyze(labels, 3)
    neg = labels.copy()
    neg[10] = -1
    with pytest.raises(hx.GpartError, match="non-negative"):
        g.analyze(neg)


def test_command_line_tools(hx, tmp_path):
    indptr, indices = make(hx, "rmat", 11)
    n = len(indptr) - 1
    indptr.tofile(tmp_path / "g_indptr.bin")
    indices.tofile(tmp_path / "g_indices.bin")
    labels, _ = hx.Graph(indptr, indices).partition(4, 1.03, 0)
    labels.tofile(tmp_path / "p.bin")
    (tmp_path / "p.txt").write_text("\n".join(str(x) for x in labels) + "\n")
    base = [sys.executable, "-m", "horkxg_amd.verify_partition", "--indptr", str(tmp_path / "g_indptr.bin"),
            "--indices", str(tmp_path / "g_indices.bin"), "--parts", "4"]
    expect = qo.verify(indptr, indices, labels, 4)
    for args in (["--partition", str(tmp_path / "p.bin")],
                 ["--partition", str(tmp_path / "p.txt"), "--partition-format", "text"]):
        out = subprocess.run(base + args, capture_output=True, text=True, timeout=120, cwd=ROOT)
        assert out.returncode == 0, out.stderr
        assert json.loads(out.stdout.strip().splitlines()[-1]) == expect
    out = subprocess.run([sys.executable, "-m", "horkxg_amd.analyze_partition", str(tmp_path / "p.txt"),
                          "--graph-dir", str(tmp_path), "--name", "g", "--parts", "4"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 0, out.stderr
    a = qo.analyze(indptr, indices, labels, 4)
    row = [line for line in out.stdout.splitlines() if line.startswith(str(tmp_path / "p.txt"))][0].split()
    assert int(row[1]) == 4 and int(row[2]) == a["edge_cut"] and int(row[7]) == a["max_local_cut"]
    assert float(row[5]) == pytest.approx(a["edge_set_imbalance"], abs=1e-6)
    assert n == len(labels)
