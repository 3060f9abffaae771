"""CPU: the C-ABI libraries, the CLIs' argument handling and the input generators.

Nothing here launches a kernel: library loading and symbol exports, CLI usage and
load errors (which happen before any device call), and the host-side graph
generators against their closed-form sizes and CSR invariants.
"""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "horkxg_amd" / "lib"
BIN = ROOT / "horkxg_amd" / "bin"

HEADERS = {
    "gpart_hip.h": "libgpart_hip.so",
    "gpart.h": "libgpart.so",
}


def declared_functions(header: Path):
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = set()
    for match in re.finditer(r"\b(gpart_\w+)\s*\(", text):
        start = text.rfind("\n", 0, match.start()) + 1
        line = text[start:match.start()]
        if "typedef" in line or "#define" in line:
            continue
        names.add(match.group(1))
    return sorted(names)


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_library_exports_every_declared_symbol(header):
    """Every function include/<header> declares is exported by its library."""
    lib = ctypes.CDLL(str(LIB / HEADERS[header]))
    names = declared_functions(ROOT / "include" / header)
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("cli", ["gpart_partition_standard", "gpart_partition_big",
                                 "gpart_coarsen_standard", "gpart_coarsen_big"])
def test_cli_usage_exit_code(cli):
    """Wrong argument counts print the usage line and exit 2 (app/partition_main.cu:27-33,
    app/coarsen_main.cu:18-23)."""
    run = subprocess.run([str(BIN / cli)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2
    assert "Usage:" in run.stderr


USAGE = {
    "gpart_partition": " <indptr.bin> <indices.bin> <parts> <partition.out> [imbalance_ratio=1.03] [seed] [--verify]\n",
    "gpart_coarsen": " <indptr.bin> <indices.bin> <parts> <hierarchy.out> [seed=0] [stop_contraction_ratio=0.90]"
                     " [max_levels=24] [--verify]\n",
}


@pytest.mark.parametrize("cli,argv", [
    ("gpart_partition_standard", ["a", "b", "4"]),
    ("gpart_partition_big", ["a", "b", "4", "o", "1.03", "0", "x"]),
    ("gpart_coarsen_standard", ["a", "b", "4", "o", "0", "0.9", "24", "x"]),
    # one trailing --verify is a flag for the coarsen tool; a second one counts as a positional
    ("gpart_coarsen_big", ["a", "b", "4", "o", "0", "0.9", "24", "--verify", "--verify"]),
])
def test_cli_usage_text(cli, argv):
    """The usage line of app/partition_main.cu:30-32 and app/coarsen_main.cu:19-21,
    word for word, for too few and too many positionals."""
    run = subprocess.run([str(BIN / cli), *argv], capture_output=True, text=True, timeout=60)
    assert run.returncode == 2
    tool = cli.rsplit("_", 1)[0]
    assert run.stderr == f"Usage: {BIN / cli}{USAGE[tool]}"


@pytest.mark.parametrize("argv,message", [
    (["4", "--verify", "--verify"], "cannot open"),  # the partition tool strips every trailing --verify
    (["x"], "stoi"),
])
def test_partition_cli_arguments(tmp_path, argv, message):
    run = subprocess.run([str(BIN / "gpart_partition_standard"), str(tmp_path / "a"), str(tmp_path / "b"),
                          argv[0], str(tmp_path / "o"), *argv[1:]], capture_output=True, text=True, timeout=60)
    assert run.returncode == 1
    assert run.stderr.startswith("gpart_partition: ") and message in run.stderr


@pytest.mark.parametrize("extra,message", [
    (["1"], "parts must be between 2 and 32"),
    (["33"], "parts must be between 2 and 32"),
    (["4", "0", "0"], "invalid coarsening stop ratio"),
    (["4", "0", "1.5"], "invalid coarsening stop ratio"),
    (["4", "0", "0.9", "0"], "invalid coarsening max levels"),
    (["4", "0", "0.9", "201"], "invalid coarsening max levels"),
])
def test_coarsen_cli_option_checks(tmp_path, extra, message):
    """app/coarsen_main.cu:30-38: option ranges are checked before the input is
    read (the input files here do not exist)."""
    run = subprocess.run([str(BIN / "gpart_coarsen_standard"), str(tmp_path / "a"), str(tmp_path / "b"),
                          extra[0], str(tmp_path / "o"), *extra[1:]], capture_output=True, text=True, timeout=60)
    assert run.returncode == 1
    assert run.stderr == f"gpart_coarsen: {message}\n"


def test_cli_missing_input_exit_code(tmp_path):
    """A load failure is reported on stderr with exit 1, before any device work."""
    run = subprocess.run([str(BIN / "gpart_partition_standard"), str(tmp_path / "none.bin"),
                          str(tmp_path / "none2.bin"), "4", str(tmp_path / "out.bin")],
                         capture_output=True, text=True, timeout=60)
    assert run.returncode == 1
    assert "gpart_partition:" in run.stderr


def check_csr(indptr, indices):
    n = len(indptr) - 1
    assert indptr[0] == 0 and indptr[-1] == len(indices)
    assert np.all(np.diff(indptr) >= 0)
    assert indices.min() >= 0 and indices.max() < n
    src = np.repeat(np.arange(n), np.diff(indptr))
    assert not np.any(src == indices), "self loop"
    key = src.astype(np.int64) * n + indices
    assert np.all(np.diff(key) > 0), "rows not strictly sorted (duplicate or unsorted entry)"
    rev = np.sort(indices.astype(np.int64) * n + src)
    assert np.array_equal(rev, key), "not symmetric"


def test_grid_generator(hx):
    """256 x 256 4-neighbour grid: 2 * 2 * 256 * 255 = 261,120 directed entries."""
    ip, ix = hx.generate_grid2d(256, 256)
    assert len(ip) - 1 == 65536 and len(ix) == 261120
    check_csr(ip, ix)
    v = 300  # r = 1, c = 44
    assert list(ix[ip[v]:ip[v + 1]]) == [44, 299, 301, 556]


def test_mesh_generator(hx):
    """s^3 mesh with 6 face neighbours: 2 * 3 * s^2 * (s - 1) directed entries
    (12,484,608 for s = 128)."""
    for s in (5, 16):
        ip, ix = hx.generate_mesh3d(s, s, s)
        assert len(ip) - 1 == s ** 3 and len(ix) == 6 * s * s * (s - 1)
        check_csr(ip, ix)
    ip, ix = hx.generate_mesh3d(3, 4, 5)
    assert len(ip) - 1 == 60 and len(ix) == 2 * (2 * 4 * 5 + 3 * 3 * 5 + 3 * 4 * 4)


def test_rmat_generator(hx):
    """R-MAT: symmetric, deduplicated, no self loops, deterministic in the seed, and
    skewed (the quadrant probabilities favour low ids)."""
    ip, ix = hx.generate_rmat(12, 16)
    check_csr(ip, ix)
    assert len(ip) - 1 == 4096
    ip2, ix2 = hx.generate_rmat(12, 16)
    assert np.array_equal(ip, ip2) and np.array_equal(ix, ix2)
    ip3, ix3 = hx.generate_rmat(12, 16, seed=1)
    assert not (len(ix3) == len(ix) and np.array_equal(ix3, ix))
    deg = np.diff(ip)
    assert deg[:64].mean() > 4 * deg[-2048:].mean()
    assert 0.5 * 2 * 16 * 4096 < len(ix) <= 2 * 16 * 4096


def test_mix_specs_follow_the_stated_distribution():
    """The batched mix: 256 graphs, types in rotation, log-uniform sizes in
    [1e4, 2e6], parts from {2, 4, 8, 32}, seeds 0..255, deterministic."""
    from horkxg_amd import workloads as W

    specs = W.mix_specs()
    assert len(specs) == 256 and specs == W.mix_specs()
    assert [s["kind"] for s in specs[:3]] == ["grid2d", "mesh3d", "rmat"]
    assert {s["k"] for s in specs} == {2, 4, 8, 32}
    assert [s["seed"] for s in specs] == list(range(256))
    for s in specs:
        if s["kind"] == "grid2d":
            n = s["size"][0] * s["size"][1]
        elif s["kind"] == "mesh3d":
            n = s["size"][0] ** 3
        else:
            n = 2 ** s["size"][0]
            assert s["size"][1] == 8
        assert 5_000 <= n <= 3_000_000, s
