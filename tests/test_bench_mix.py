"""The multi-process mix's parent (bench.py _collect_mix_results): a worker that
dies without posting its result fails the bench with its exit code instead of
leaving the parent blocked on the queue; healthy workers' results all arrive."""
from __future__ import annotations

import importlib.util
import multiprocessing as mp
import os
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def post(queue, rank):
    queue.put((rank, (0.0, 1.0), [], None))


def die(queue, rank):
    os._exit(3)


def test_dead_worker_fails_the_mix():
    bench = load_bench()
    ctx = mp.get_context("fork")
    queue = ctx.Queue()
    workers = [ctx.Process(target=post, args=(queue, 0)), ctx.Process(target=die, args=(queue, 1))]
    for w in workers:
        w.start()
    with pytest.raises(RuntimeError, match="exit code 3"):
        bench._collect_mix_results(queue, workers, poll_seconds=0.1)
    assert all(w.exitcode is not None for w in workers)


def test_all_results_collected():
    bench = load_bench()
    ctx = mp.get_context("fork")
    queue = ctx.Queue()
    workers = [ctx.Process(target=post, args=(queue, r)) for r in range(3)]
    for w in workers:
        w.start()
    results = bench._collect_mix_results(queue, workers, poll_seconds=0.1)
    for w in workers:
        w.join()
    assert sorted(r[0] for r in results) == [0, 1, 2]
