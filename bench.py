#!/usr/bin/env python3
"""Benchmark of the multilevel k-way partitioner on one MI355X.

A step is one complete partition (coarsening, initial partition, uncoarsening with
refinement, hierarchy release) of a graph already resident in HBM, timed with the
reference's scope (app/partition_main.cu:67-93: coarsen start -> hierarchy freed).

Default workload: 3D 7-point mesh 128^3 (2,097,152 vertices, 12,484,608 directed
entries), unit weights, standard width, k = 32, imbalance 1.03, seed 0.

Prints one JSON line: metric algorithm_total_seconds (lower is better) with the edge
cut, the CPU oracle's cut on the same input, the live roofline of the dominant
kernel family and the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"

# The kernel each live-timed family brackets (what a rocprofv3 --pmc run selects).
FAMILY_KERNELS = {
    "rebalance": "balance_proposals_kernel",
    "refine_proposals": "lp_proposals_kernel",
    "refine_filter": "priority_filter_kernel",
    "refine_cache": "build_cache_kernel",
    "spectral": "spectral_depth_kernel",
    "sclp_affinity": "sclp_(low|medium|high|part|split_count|split_scatter|hub_insert|hub_scan|hub_decode)_kernel",
    "contract": "(cross_count|encode|row_offsets)_kernel",
    "cache_update": "update_cache(_tiles)?_kernel",
    "sclp_commit": "admit_apply_kernel",
}

# spectral: the bisection form each configuration is quoted with (FP32 diffusion
# for the R-MAT graphs, FP64 for the mesh and the grid).
WORKLOADS = {
    "mesh128_k32": dict(kind="mesh3d", size=(128, 128, 128), k=32, spectral="batched_fp64"),
    "grid256_k4": dict(kind="grid2d", size=(256, 256), k=4, spectral="batched_fp64"),
    "rmat22_k8": dict(kind="rmat", size=(22, 16), k=8, spectral="batched_fp32"),
    "rmat18_k8": dict(kind="rmat", size=(18, 16), k=8, spectral="batched_fp32"),
    # big width (64-bit offsets and weights), generated on the GPU
    "rmat26_k32": dict(kind="rmat", size=(26, 16), k=32, spectral="batched_fp32", width="big", device=True,
                       steps=2, warmup=1),
}


def make_graph(hx, spec):
    if spec["kind"] == "mesh3d":
        return hx.generate_mesh3d(*spec["size"])
    if spec["kind"] == "grid2d":
        return hx.generate_grid2d(*spec["size"])
    scale, ef = spec["size"]
    return hx.generate_rmat(scale, ef, 0.57, 0.19, 0.19, 0)


def cpu_baseline(indptr, indices, k, ratio, seed, threads):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py as O

    O.set_threads(threads)
    g = O.Graph(indptr, indices.astype(np.int32))
    t = time.perf_counter()
    labels, res = O.partition(g, k, ratio, seed)
    wall = time.perf_counter() - t
    return res, wall, labels


def host_cpu() -> dict:
    """The host's CPU as lscpu reports it (model, sockets, cores, threads per core)."""
    import subprocess

    info = {"logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
            "Thread(s) per core": "threads_per_core"}
    for line in out.splitlines():
        name, _, value = line.partition(":")
        if name.strip() in keys:
            value = value.strip()
            info[keys[name.strip()]] = int(value) if value.isdigit() else value
    return info


def cpu_baseline_seeds(indptr, indices, k, ratio, seeds, threads, gpu_labels_seed0):
    """The CPU oracle with the reference's timed scope: seeds 0-4 on `threads`
    threads (median reported), and seed 0 serially."""
    runs = []
    for seed in seeds:
        res, wall, labels = cpu_baseline(indptr, indices, k, ratio, seed, threads)
        runs.append((res, wall, labels))
        print(f"[bench] cpu oracle seed {seed} threads {threads}: {res.total_seconds:.3f} s cut {res.final_cut}",
              file=sys.stderr, flush=True)
    serial, serial_wall, _ = cpu_baseline(indptr, indices, k, ratio, seeds[0], 1)
    print(f"[bench] cpu oracle seed {seeds[0]} serial: {serial.total_seconds:.3f} s", file=sys.stderr, flush=True)
    seconds = [float(r.total_seconds) for r, _, _ in runs]
    cuts = [int(r.final_cut) for r, _, _ in runs]
    return {
        "value": round(float(np.median(seconds)), 4), "unit": "s", "cores": threads, "kind": "port",
        "sample": f"full partitions of the workload by the CPU oracle, seeds {seeds[0]}-{seeds[-1]} on {threads} "
                  f"threads (median) and seed {seeds[0]} on one thread, same timed scope",
        "seeds": list(seeds), "seconds_per_seed": [round(x, 4) for x in seconds], "cut_per_seed": cuts,
        "cut_median": float(np.median(cuts)),
        "serial": {"cores": 1, "seed": seeds[0], "value": round(float(serial.total_seconds), 4),
                   "wall_s": round(serial_wall, 3)},
        "host": host_cpu(),
        "labels_identical_to_gpu": bool(np.array_equal(runs[0][2], gpu_labels_seed0)),
    }


def stream_copy_gbps(hx, mib: int = 1024, repeats: int = 10) -> float:
    """Device-to-device copy bandwidth (read + write bytes per second) of the
    library's copy kernel over `mib` MiB: the STREAM-copy ceiling beside the
    published HBM peak."""
    import ctypes as C

    from horkxg_amd import _native as N

    nbytes = mib << 20
    src = hx.device.DeviceArray(dtype=np.uint8, count=nbytes)
    dst = hx.device.DeviceArray(dtype=np.uint8, count=nbytes)
    H = N.hip_lib()
    H.gpart_memcpy_dtod.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    N._check(H.gpart_memcpy_dtod(N._ctx(), C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes))
    N._check(H.gpart_ctx_synchronize(N._ctx()))
    t = time.perf_counter()
    for _ in range(repeats):
        N._check(H.gpart_memcpy_dtod(N._ctx(), C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes))
    N._check(H.gpart_ctx_synchronize(N._ctx()))
    elapsed = time.perf_counter() - t
    del src, dst
    return 2.0 * nbytes * repeats / elapsed / 1e9


def copy_calibration(mib: int) -> int:
    """Five device-to-device copies of `mib` MiB on the library stream (the library's
    copy_vec_kernel): a plain copy kernel of known size, whose FETCH_SIZE / WRITE_SIZE per dispatch calibrate the
    counters (run under rocprofv3 --pmc)."""
    import ctypes as C

    import horkxg_amd as hx
    from horkxg_amd import _native as N

    nbytes = mib << 20
    src = hx.device.DeviceArray(dtype=np.uint8, count=nbytes)
    dst = hx.device.DeviceArray(dtype=np.uint8, count=nbytes)
    H = N.hip_lib()
    H.gpart_memcpy_dtod.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    for _ in range(5):
        N._check(H.gpart_memcpy_dtod(N._ctx(), C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes))
    N._check(H.gpart_ctx_synchronize(N._ctx()))
    print(f"[bench] copied {mib} MiB five times", file=sys.stderr, flush=True)
    return 0


def pmc_per_dispatch(csv_path: Path, kernel_regex: str, counter: str):
    """Mean of `counter` per dispatch of kernels matching kernel_regex in a
    rocprofv3 counter_collection.csv (summed over the counter's instances)."""
    import csv
    import re

    per_dispatch = {}
    pattern = re.compile(kernel_regex)
    with open(csv_path, newline="") as f:
        for row in csv.DictReader(f):
            if row.get("Counter_Name") != counter or not pattern.search(row.get("Kernel_Name", "")):
                continue
            key = row.get("Dispatch_Id") or row.get("Correlation_Id")
            per_dispatch[key] = per_dispatch.get(key, 0.0) + float(row["Counter_Value"])
    if not per_dispatch:
        raise SystemExit(f"no {counter} rows for /{kernel_regex}/ in {csv_path}")
    values = list(per_dispatch.values())
    return sum(values) / len(values), len(values)


def pmc_total(csv_path: Path, kernel_regex: str, counter: str) -> float:
    """Sum of `counter` over every dispatch of kernels matching kernel_regex."""
    mean, count = pmc_per_dispatch(csv_path, kernel_regex, counter)
    return mean * count


def traffic_summary(args) -> int:
    """Writes profiles/traffic_<workload>.json from two PMC runs of the bench
    (FETCH_SIZE, WRITE_SIZE) and two of the copy calibration."""
    copy_bytes = args.copy_mib << 20
    cf, _ = pmc_per_dispatch(Path(args.copy_fetch), "copy_vec_kernel", "FETCH_SIZE")
    cw, _ = pmc_per_dispatch(Path(args.copy_write), "copy_vec_kernel", "WRITE_SIZE")
    f_scale = copy_bytes / (cf * 1024.0)
    w_scale = copy_bytes / (cw * 1024.0)
    kf, nf = pmc_per_dispatch(Path(args.fetch), args.kernel, "FETCH_SIZE")
    kw, nw = pmc_per_dispatch(Path(args.write), args.kernel, "WRITE_SIZE")
    if args.family_launches:
        # a family launch runs several kernels: its traffic is the total over the
        # run's dispatches divided by the family's launches in the run
        kf = kf * nf / args.family_launches
        kw = kw * nw / args.family_launches
    entry = {
        "kernel": args.kernel,
        "dispatches": {"fetch": nf, "write": nw},
        "family_launches": args.family_launches or None,
        # per launch: one dispatch, or one family launch when family_launches is given
        "fetch_kib_per_launch": round(kf, 1),
        "write_kib_per_launch": round(kw, 1),
        "copy_calibration": {"bytes": copy_bytes, "fetch_kib": round(cf, 1), "write_kib": round(cw, 1),
                             "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4)},
        "traffic_bytes_per_launch": round(kf * 1024.0 * f_scale + kw * 1024.0 * w_scale),
    }
    # one file per workload, one entry per kernel
    PROFILES.mkdir(exist_ok=True)
    path = PROFILES / f"traffic_{args.workload}.json"
    out = {"workload": args.workload, "kernels": {}}
    if path.exists():
        old = json.loads(path.read_text())
        out["kernels"] = old.get("kernels", {old["kernel"]: old} if "kernel" in old else {})
    out["kernels"][args.kernel] = entry
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(entry))
    return 0


def recorded_traffic(workload: str, family: str):
    """HBM bytes per launch of the family's kernel from the committed PMC summary."""
    path = PROFILES / f"traffic_{workload}.json"
    if not path.exists() or family not in FAMILY_KERNELS:
        return None
    rec = json.loads(path.read_text())
    entries = rec.get("kernels", {rec["kernel"]: rec} if "kernel" in rec else {})
    for kernel, entry in entries.items():
        if kernel == FAMILY_KERNELS[family] or re.search(kernel, FAMILY_KERNELS[family]):
            return entry["traffic_bytes_per_launch"]
    return None


def _mix_worker(rank, specs, order, mode, barrier, counter, queue, hw_queues, group=1):
    """One host process of the concurrent mix. It generates every graph of the mix
    on the GPU (outside the timed region), partitions the three smallest once
    untimed (code objects, allocator), waits at a barrier shared with the other
    processes, then claims groups of `group` graphs from a shared counter over
    `order` (largest first) until none is left, and exits: an idle process that
    keeps its GPU context alive slows the others down. A group is one
    partition_batch call, whose initial partitions share one persistent launch.
    Labels are checked after the clock stops. hw_queues caps the process's
    hardware queues (read at HIP initialisation)."""
    if hw_queues:
        os.environ["GPU_MAX_HW_QUEUES"] = str(hw_queues)
    try:
        import horkxg_amd as hx
        from horkxg_amd import workloads as W

        graphs = [W.device_graph(hx, s) for s in specs]
        for i in order[-3:]:
            graphs[i].partition(specs[i]["k"], specs[i]["ratio"], specs[i]["seed"], spectral_mode=mode,
                                want_labels=False)
        hx.set_batch_group(max(1, group))
        barrier.wait(timeout=MIX_BARRIER_SECONDS)
        t0 = time.perf_counter()
        done = []
        while True:
            with counter.get_lock():
                slot = counter.value
                counter.value += max(1, group)
            if slot >= len(order):
                break
            ids = order[slot:slot + max(1, group)]
            labels, reps, _ = hx.partition_batch([graphs[i] for i in ids], [specs[i]["k"] for i in ids],
                                                 [specs[i]["ratio"] for i in ids], [specs[i]["seed"] for i in ids],
                                                 spectral_mode=mode)
            done.extend(zip(ids, labels, reps))
        t1 = time.perf_counter()
        for i, labels, rep in done:
            assert graphs[i].cut(labels) == rep["final_cut"], specs[i]
            assert np.bincount(labels, minlength=specs[i]["k"]).max() <= rep["capacity"], specs[i]
        queue.put((rank, (t0, t1), [(i, rep["final_cut"]) for i, _, rep in done], None))
    except BaseException as e:  # reported to the parent, which fails the bench
        barrier.abort()
        queue.put((rank, None, None, repr(e)))


MIX_BARRIER_SECONDS = 600  # every worker has its graphs and code objects by then


def _collect_mix_results(queue, workers, poll_seconds=1.0):
    """One result per worker from `queue`. A worker that ends without posting one
    (a hard crash, an out-of-memory kill) is found by its exit code: the remaining
    workers are terminated and the bench fails with that code instead of waiting."""
    import queue as queue_mod

    results = []
    while len(results) < len(workers):
        try:
            results.append(queue.get(timeout=poll_seconds))
            continue
        except queue_mod.Empty:
            pass
        dead = [w for w in workers if w.exitcode not in (None, 0)]
        if dead:
            for w in workers:
                if w.exitcode is None:
                    w.terminate()
            for w in workers:
                w.join()
            raise RuntimeError(f"mix worker died with exit code {dead[0].exitcode} "
                               f"({len(results)} of {len(workers)} results in)")
    return results


def _mix_processes(specs, sizes, procs, mode, steps, hw_queues=0, group=1):
    """Partitions/sec of `procs` processes sharing the GPU, one fresh set of
    processes per step: per step, the wall from the earliest start to the latest
    finish. Returns (walls, {index: cut})."""
    import multiprocessing as mp

    ctx = mp.get_context("spawn")
    order = sorted(range(len(specs)), key=lambda i: -sizes[i])
    walls, cuts = [], {}
    for step in range(steps):
        barrier = ctx.Barrier(procs)
        counter = ctx.Value("i", 0)
        queue = ctx.Queue()
        workers = [ctx.Process(target=_mix_worker,
                               args=(r, specs, order, mode, barrier, counter, queue, hw_queues, group))
                   for r in range(procs)]
        for w in workers:
            w.start()
        results = _collect_mix_results(queue, workers)
        for w in workers:
            w.join()
        errors = [e for _, _, _, e in results if e]
        if errors or any(w.exitcode for w in workers):
            raise RuntimeError(f"mix worker failed: {errors} exit codes {[w.exitcode for w in workers]}")
        start = min(t[0] for _, t, _, _ in results)
        walls.append(max(t[1] for _, t, _, _ in results) - start)
        spans = sorted((r, round(t[0] - start, 3), round(t[1] - t[0], 3), len(p)) for r, t, p, _ in results)
        print(f"[bench] step {step} per process (rank, start, seconds, graphs): {spans}", file=sys.stderr,
              flush=True)
        step_cuts = {i: c for _, _, pairs, _ in results for i, c in pairs}
        assert len(step_cuts) == len(specs)
        if cuts:
            assert step_cuts == cuts, "cuts differ between steps"
        cuts = step_cuts
    return walls, cuts


def run_mix_processes(args) -> int:
    """The batched graph mix with several host processes sharing the GPU: the
    processes claim graphs from one shared queue, largest first, and partition each
    in the batched FP32 form; the plain form (plain FP64 bisection, one graph at a
    time in one process) runs on the same graphs after it. The main process touches
    the GPU only after the workers have ended (CPU baseline)."""
    from horkxg_amd import workloads as W

    specs = W.mix_specs(args.mix_count)
    sizes = []
    for s in specs:  # vertices + entries without generating: closed forms for lattices
        if s["kind"] == "grid2d":
            r, c = s["size"]
            sizes.append(r * c + 2 * (r * (c - 1) + c * (r - 1)))
        elif s["kind"] == "mesh3d":
            x, y, z = s["size"]
            sizes.append(x * y * z + 2 * ((x - 1) * y * z + x * (y - 1) * z + x * y * (z - 1)))
        else:
            scale, ef = s["size"]
            sizes.append((1 << scale) * (1 + 2 * ef))
    t = time.perf_counter()
    walls, cuts = _mix_processes(specs, sizes, args.mix_procs, "batched_fp32", args.steps, args.mix_hw_queues,
                                 args.mix_group)
    for step, wall in enumerate(walls):
        print(f"[bench] mix step {step}: {len(specs)} partitions in {wall:.3f} s on {args.mix_procs} processes",
              file=sys.stderr, flush=True)
    plain_walls, plain_cuts_map = _mix_processes(specs, sizes, 1, "plain_fp64", 1)
    print(f"[bench] mix plain (1 process, plain FP64): {plain_walls[0]:.3f} s "
          f"(all phases {time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    order = sorted(cuts)
    cut_arr = np.array([cuts[i] for i in order], dtype=np.float64)
    plain_arr = np.array([plain_cuts_map[i] for i in order], dtype=np.float64)
    rel = np.where(plain_arr > 0, cut_arr / np.maximum(plain_arr, 1) - 1.0, 0.0)
    wall = float(np.mean(walls))
    value = len(specs) / wall
    plain_value = 1.0 / plain_walls[0] * len(specs)
    cpu = None
    if not args.no_cpu_baseline:
        import horkxg_amd as hx

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        small = sorted(range(len(specs)), key=lambda i: sizes[i])[:16]
        total, same = 0.0, 0
        for i in small:
            indptr, indices = W.generate(hx, specs[i])
            res, _, _ = cpu_baseline(indptr, indices, specs[i]["k"], specs[i]["ratio"], specs[i]["seed"], threads)
            total += res.total_seconds
            _, rep = hx.Graph(indptr, indices).partition(specs[i]["k"], specs[i]["ratio"], specs[i]["seed"],
                                                         spectral_mode="batched_fp64", want_labels=False)
            same += int(rep["final_cut"] == res.final_cut)
        cpu = {"value": round(len(small) / total, 3), "unit": "partitions/s", "cores": threads, "kind": "port",
               "sample": f"the {len(small)} smallest graphs of the mix, one CPU-oracle partition each "
                         f"(batched FP64 bisection), same timed scope",
               "gpu_fp64_cut_equal_to_oracle": f"{same}/{len(small)}"}
    line = {
        "metric": "batched_partitions_per_sec",
        "value": round(value, 3),
        "unit": "partitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"mix{len(specs)}", "graphs": len(specs), "entries_and_vertices_total": int(sum(sizes)),
                   "parts": "k_i from {2,4,8,32}", "imbalance": 1.03, "width": "standard",
                   "processes": args.mix_procs, "hw_queues_per_process": args.mix_hw_queues or "default",
                   "group": args.mix_group, "spectral": "batched_fp32"},
        "plain": {"partitions_per_sec": round(plain_value, 3), "processes": 1, "spectral": "plain_fp64",
                  "wall_s": round(plain_walls[0], 4)},
        "batched_over_plain": round(value / plain_value, 3),
        "cut": {"total_batched_fp32": int(cut_arr.sum()), "total_plain_fp64": int(plain_arr.sum()),
                "graphs_with_equal_cut": int((cut_arr == plain_arr).sum()),
                "max_relative_difference": round(float(np.abs(rel).max()), 5)},
        "all_within_capacity": True,
        "roofline": None,
        "roofline_note": "per-family event timing is per process; the kernels' roofline is reported by the "
                         "single-graph workloads",
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_mix(args) -> int:
    """The batched graph mix: partitions/sec of the batched form (worker threads,
    FP32 diffusion) against the plain form (one thread, plain FP64 bisection)."""
    import horkxg_amd as hx
    from horkxg_amd import workloads as W

    specs = W.mix_specs(args.mix_count)
    t = time.perf_counter()
    graphs = W.upload_mix(hx, specs, threads=min(16, os.cpu_count() or 1))
    sizes = [(g.n, g.m) for g in graphs]
    print(f"[bench] mix of {len(graphs)} graphs: {sum(n for n, _ in sizes)} vertices, "
          f"{sum(m for _, m in sizes)} entries, generated in {time.perf_counter() - t:.1f} s",
          file=sys.stderr, flush=True)
    parts = [s["k"] for s in specs]
    ratios = [s["ratio"] for s in specs]
    seeds = [s["seed"] for s in specs]

    hx.set_batch_group(max(1, args.mix_group))

    def batch(mode, workers, want_labels):
        return hx.partition_batch(graphs, parts, ratios, seeds, spectral_mode=mode, workers=workers,
                                  want_labels=want_labels)

    for _ in range(args.warmup):
        batch("batched_fp32", args.mix_workers, False)
        print("[bench] mix warmup batch done", file=sys.stderr, flush=True)
    walls = []
    labels = reports = None
    for step in range(args.steps):
        labels, reports, wall = batch("batched_fp32", args.mix_workers, step == args.steps - 1)
        walls.append(wall)
        print(f"[bench] mix step {step}: {len(graphs)} partitions in {wall:.3f} s", file=sys.stderr, flush=True)
    _, plain_reports, plain_wall = batch("plain_fp64", 1, False)
    print(f"[bench] mix plain (1 worker, plain FP64): {plain_wall:.3f} s", file=sys.stderr, flush=True)

    for g, lab, r, s in zip(graphs, labels, reports, specs):
        assert g.cut(lab) == r["final_cut"], s
        assert np.bincount(lab, minlength=s["k"]).max() <= r["capacity"], s
    cuts = np.array([r["final_cut"] for r in reports], dtype=np.float64)
    plain_cuts = np.array([r["final_cut"] for r in plain_reports], dtype=np.float64)
    rel = np.where(plain_cuts > 0, cuts / np.maximum(plain_cuts, 1) - 1.0, 0.0)
    wall = float(np.mean(walls))
    value = len(graphs) / wall
    plain_value = len(graphs) / plain_wall

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        order = sorted(range(len(graphs)), key=lambda i: sizes[i][0] + sizes[i][1])[:16]
        total, same = 0.0, 0
        for i in order:
            indptr, indices = W.generate(hx, specs[i])
            res, _, cpu_labels = cpu_baseline(indptr, indices, specs[i]["k"], specs[i]["ratio"], specs[i]["seed"],
                                              threads)
            total += res.total_seconds
            _, rep = graphs[i].partition(specs[i]["k"], specs[i]["ratio"], specs[i]["seed"],
                                         spectral_mode="batched_fp64", want_labels=False)
            same += int(rep["final_cut"] == res.final_cut)
        cpu = {"value": round(len(order) / total, 3), "unit": "partitions/s", "cores": threads, "kind": "port",
               "sample": f"the {len(order)} smallest graphs of the mix, one CPU-oracle partition each "
                         f"(batched FP64 bisection), same timed scope",
               "gpu_fp64_cut_equal_to_oracle": f"{same}/{len(order)}"}

    line = {
        "metric": "batched_partitions_per_sec",
        "value": round(value, 3),
        "unit": "partitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": f"mix{len(graphs)}", "graphs": len(graphs),
                   "vertices_total": int(sum(n for n, _ in sizes)), "entries_total": int(sum(m for _, m in sizes)),
                   "parts": "k_i from {2,4,8,32}", "imbalance": 1.03, "width": "standard",
                   "workers": args.mix_workers, "group": args.mix_group, "spectral": "batched_fp32"},
        "plain": {"partitions_per_sec": round(plain_value, 3), "workers": 1, "spectral": "plain_fp64",
                  "wall_s": round(plain_wall, 4)},
        "batched_over_plain": round(value / plain_value, 3),
        "cut": {"total_batched_fp32": int(cuts.sum()), "total_plain_fp64": int(plain_cuts.sum()),
                "graphs_with_equal_cut": int((cuts == plain_cuts).sum()),
                "max_relative_difference": round(float(np.abs(rel).max()), 5)},
        "all_within_capacity": True,
        "roofline": None,
        "roofline_note": "per-family event timing is per host thread; the kernels' roofline is reported "
                         "by the single-graph workloads",
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=None, help="default 5 (2 for rmat26_k32)")
    p.add_argument("--warmup", type=int, default=None, help="default 1")
    p.add_argument("--workload", default="mesh128_k32", choices=sorted(WORKLOADS) + ["mix"])
    p.add_argument("--mix-count", type=int, default=256, help="graphs of the mix workload")
    p.add_argument("--mix-procs", type=int, default=2,
                   help="host processes sharing the GPU in the batched mix (1: the in-process batch)")
    p.add_argument("--mix-hw-queues", type=int, default=1, choices=range(0, 5),
                   help="GPU_MAX_HW_QUEUES of each mix process (0: the runtime's default)")
    p.add_argument("--mix-group", type=int, default=8,
                   help="graphs per partition_batch group: their initial partitions share one persistent launch")
    p.add_argument("--mix-workers", type=int, default=1,
                   help="host threads (streams) of the batched mix; concurrent workers are switched off (1)")
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cut-seeds", type=int, default=5,
                   help="seeds (from --seed) of the GPU cut median and of the CPU oracle's")
    p.add_argument("--ratio", type=float, default=1.03)
    p.add_argument("--spectral-mode", default=None, choices=["plain_fp64", "batched_fp64", "batched_fp32"],
                   help="default: the workload's own form")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-threads", type=int, default=0, help="0 = all host cores (at most 16)")
    p.add_argument("--no-profile", action="store_true", help="skip the live per-family event timing")
    p.add_argument("--no-round-graphs", action="store_true",
                   help="launch the refinement rounds kernel by kernel instead of replaying their hipGraphs "
                        "(for kernel-trace profiles: the kernels are the same)")
    p.add_argument("--no-precision", action="store_true",
                   help="skip the FP32-against-FP64 comparison (a second coarsening and partition)")
    p.add_argument("--copy-calibration", type=int, metavar="MIB", help="only run the copy calibration")
    p.add_argument("--traffic-summary", action="store_true",
                   help="only summarise rocprofv3 --pmc CSVs into profiles/traffic_<workload>.json")
    p.add_argument("--kernel", default="lp_proposals_kernel", help="kernel regex for --traffic-summary")
    p.add_argument("--family-launches", type=int, default=0,
                   help="--traffic-summary: divide the run's total by this many family launches")
    p.add_argument("--fetch", help="counter_collection.csv of the FETCH_SIZE run")
    p.add_argument("--write", help="counter_collection.csv of the WRITE_SIZE run")
    p.add_argument("--copy-fetch", help="counter_collection.csv of the calibration FETCH_SIZE run")
    p.add_argument("--copy-write", help="counter_collection.csv of the calibration WRITE_SIZE run")
    p.add_argument("--copy-mib", type=int, default=1024)
    args = p.parse_args()
    if args.traffic_summary:
        return traffic_summary(args)
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the partitioner is a single-GPU path)",
              file=sys.stderr)
        return 2

    if args.workload == "mix":
        args.steps = 2 if args.steps is None else args.steps
        args.warmup = 1 if args.warmup is None else args.warmup
        if args.mix_procs > 1:
            # the worker processes are started before this one touches the GPU
            return run_mix_processes(args)
        return run_mix(args)

    import horkxg_amd as hx

    if hx.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.no_round_graphs:
        hx.set_refine_graphs(False)
    spec = WORKLOADS[args.workload]
    k = spec["k"]
    width = spec.get("width", "standard")
    if args.spectral_mode is None:
        args.spectral_mode = spec["spectral"]
    if args.steps is None:
        args.steps = spec.get("steps", 5)
    if args.warmup is None:
        args.warmup = spec.get("warmup", 1)
    if spec.get("device"):
        scale, ef = spec["size"]
        t = time.perf_counter()
        graph = hx.Graph.generate_rmat(scale, ef, 0.57, 0.19, 0.19, 0, width=width)
        indptr = indices = None
        n, m = graph.n, graph.m
        print(f"[bench] generated on the GPU in {time.perf_counter() - t:.2f} s", file=sys.stderr, flush=True)
    else:
        indptr, indices = make_graph(hx, spec)
        n, m = len(indptr) - 1, len(indices)
        graph = hx.Graph(indptr, indices, width=width)
    print(f"[bench] workload={args.workload} vertices={n} entries={m} k={k} width={width}", file=sys.stderr,
          flush=True)

    families = list(hx.PROFILE_FAMILIES)

    def profiled_step():
        """One untimed partition with every family's launches bracketed by events:
        the per-family breakdown (event pairs around every launch slow the step)."""
        hx.profile_reset()
        for f in families:
            hx.profile_enable(f, True)
        _, r = graph.partition(k, args.ratio, args.seed, spectral_mode=args.spectral_mode, want_labels=False)
        out = {f: hx.profile_read(f) for f in families}
        for f in families:
            hx.profile_enable(f, False)
        return r, out

    cold_seconds = None
    breakdown = {}
    for i in range(args.warmup):
        if i == args.warmup - 1 and not args.no_profile:
            rep, breakdown = profiled_step()
        else:
            _, rep = graph.partition(k, args.ratio, args.seed, spectral_mode=args.spectral_mode, want_labels=False)
        if cold_seconds is None:
            cold_seconds = rep["algorithm_seconds"]
        print(f"[bench] warmup step done algorithm_seconds={rep['algorithm_seconds']:.4f}", file=sys.stderr,
              flush=True)
    if not args.no_profile and not breakdown:
        _, breakdown = profiled_step()
    dominant = None
    if breakdown:
        # the dominant family by time; its launches are timed with events over as
        # many steps as the timed ones, right after them
        dominant = max(breakdown, key=lambda f: breakdown[f]["total_ms"] if breakdown[f]["bytes"] > 0 else -1)
    reports = []
    # The context keeps its device segments across runs (freshly recycled device
    # memory is scrubbed by the GPU kernel driver on allocation); the first, cold step is
    # reported beside the steady-state value.
    alloc_before = hx.cache_info()
    t0 = time.perf_counter()
    for step in range(args.steps):
        labels, rep = graph.partition(k, args.ratio, args.seed, spectral_mode=args.spectral_mode,
                                      want_labels=(step == args.steps - 1))
        reports.append(rep)
        print(f"[bench] step {step} algorithm_seconds={rep['algorithm_seconds']:.4f} cut={rep['final_cut']}",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    alloc_after = hx.cache_info()
    graph_stats = hx.refine_graph_stats()
    fam = {}
    roofline_steps = []
    if dominant is not None:
        # Event pairs around every launch of the family cost about 3 ms per mesh
        # partition and run the family outside the refinement's round graphs, so
        # they stay out of the timed steps: the same number of steps of the same
        # workload follows with the family's events on the library stream.
        hx.profile_reset()
        hx.profile_enable(dominant, True)
        for step in range(args.steps):
            _, rep = graph.partition(k, args.ratio, args.seed, spectral_mode=args.spectral_mode, want_labels=False)
            roofline_steps.append(rep["algorithm_seconds"])
        fam[dominant] = hx.profile_read(dominant)
        hx.profile_enable(dominant, False)
        print(f"[bench] {args.steps} steps with {dominant} events: algorithm_seconds mean "
              f"{np.mean(roofline_steps):.4f}", file=sys.stderr, flush=True)

    alg = np.array([r["algorithm_seconds"] for r in reports])
    cut = reports[-1]["final_cut"]
    cap = reports[-1]["capacity"]
    loads = np.bincount(labels, minlength=k)
    independent_cut = graph.cut(labels)
    assert independent_cut == cut, (independent_cut, cut)
    assert loads.max() <= cap, (loads.max(), cap)
    # the GPU verifier and quality metrics on the final labels (host labels uploaded)
    t = time.perf_counter()
    verified = graph.verify(labels, k, str(args.ratio))
    verify_seconds = time.perf_counter() - t
    t = time.perf_counter()
    quality = graph.analyze(labels, k)
    analyze_seconds = time.perf_counter() - t
    # unit weights: the verifier's entry count is the partition's cut
    assert verified["cut"] == cut and verified["cut"] * 2 == quality["edge_cut"], (verified["cut"], cut)
    assert verified["capacity_ok"] and verified["max_load"] == loads.max()

    roofline = None
    if fam:
        d = fam[dominant]
        per_launch_s = d["total_ms"] / 1e3 / max(d["launches"], 1)
        bytes_per_launch = d["bytes"] / max(d["launches"], 1)
        achieved = bytes_per_launch / per_launch_s / 1e9 if per_launch_s > 0 else 0.0
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": recorded_traffic(args.workload, dominant), "family": dominant,
                    "kernel": FAMILY_KERNELS.get(dominant, dominant),
                    "launches": d["launches"], "avg_launch_us": round(per_launch_s * 1e6, 2),
                    "algorithmic_bytes_per_launch": round(bytes_per_launch),
                    "timed_over": f"{args.steps} steps of the workload right after the timed ones, HIP events "
                                  f"around every launch of the family on the library stream (the events are kept "
                                  f"out of the timed steps; those steps took "
                                  f"{np.mean(roofline_steps):.4f} s on average)"}
        # every family over one untimed profiled step
        roofline["families"] = {
            f: {"ms": round(v["total_ms"], 3), "launches": v["launches"],
                "GBps": round(v["bytes"] / (v["total_ms"] / 1e3) / 1e9, 1) if v["total_ms"] > 0 and v["bytes"] > 0
                else None}
            for f, v in breakdown.items()}
        # atomic throughput of the atomic-bound families: the movers' cache updates
        # (R13), the admission commit (K18), the rebalance bucket histogram (R17)
        for f, v in breakdown.items():
            if v.get("atomics", 0) > 0 and v["total_ms"] > 0:
                roofline["families"][f]["atomics_per_step"] = round(v["atomics"])
                roofline["families"][f]["Gatomics_per_s"] = round(v["atomics"] / (v["total_ms"] / 1e3) / 1e9, 3)

    precision = None
    # (skipped for graphs generated on the GPU: their host hierarchy copy is the
    # size of the input; the R-MAT 22 line carries the comparison)
    if args.spectral_mode == "batched_fp32" and indptr is not None and not args.no_precision:
        # FP32 against FP64 diffusion on the coarsest level of the production
        # hierarchy, and the final cut of one FP64 partition (both untimed)
        csf = max(64, (512 + k - 1) // k)
        h = graph.coarsen(k, seed=args.seed, cluster_size_factor=csf, pack_isolated=True)
        off, adj, ew, vw = h.level(len(h) - 1)
        cmp = hx.Graph(off, adj, width=width, edge_weights=ew, vertex_weights=vw).spectral_compare(
            k, args.ratio, args.seed)
        _, rep64 = graph.partition(k, args.ratio, args.seed, spectral_mode="batched_fp64", want_labels=False)
        precision = {
            "coarsest_vertices": len(off) - 1,
            "max_abs_error_per_depth": [d["max_abs_error"] for d in cmp["depths"]],
            "rms_error_per_depth": [d["rms_error"] for d in cmp["depths"]],
            "splits_differ_per_depth": [d["splits_differ"] for d in cmp["depths"]],
            "labels_differ_per_depth": [d["labels_differ"] for d in cmp["depths"]],
            "initial_cut_fp64": cmp["cut_fp64"], "initial_cut_fp32": cmp["cut_fp32"],
            "final_cut_fp64": rep64["final_cut"], "final_cut_fp32": int(cut),
        }

    # the GPU's cut over seeds seed..seed+4 (the reference's protocol: median of 5 seeds)
    seeds = list(range(args.seed, args.seed + max(1, args.cut_seeds)))
    gpu_cuts = [int(cut)]
    for s in seeds[1:]:
        _, rep_s = graph.partition(k, args.ratio, s, spectral_mode=args.spectral_mode, want_labels=False)
        gpu_cuts.append(int(rep_s["final_cut"]))
    gpu_cut_median = float(np.median(gpu_cuts))
    if roofline is not None:
        stream = stream_copy_gbps(hx)
        roofline["stream_copy_gbps"] = round(stream, 1)
        roofline["frac_of_stream"] = round(roofline["achieved"] / stream, 4)

    cpu = None
    reference_cut = None
    if not args.no_cpu_baseline and indptr is not None:
        # the box's CPU share is 16 threads (all host cores when fewer)
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline_seeds(indptr, indices, k, args.ratio, seeds, threads, labels)
        reference_cut = cpu["cut_median"]

    line = {
        "metric": "algorithm_total_seconds",
        "value": round(float(alg.mean()), 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / max(args.steps, 1) * 1e3, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {"workload": args.workload, "vertices": n, "directed_entries": m, "k": k,
                   "imbalance": args.ratio, "seed": args.seed, "width": width,
                   "spectral": args.spectral_mode},
        "edge_cut": int(cut),
        "edge_cut_seeds": gpu_cuts,
        "edge_cut_median": gpu_cut_median,
        "reference_cut": reference_cut,
        "cut_vs_reference": (round(gpu_cut_median / reference_cut, 5) if reference_cut else None),
        "max_part_weight": int(loads.max()),
        "capacity": int(cap),
        "partitions_per_sec": round(1.0 / float(alg.mean()), 3),
        "stage_seconds": {s: round(float(np.mean([r[s] for r in reports])), 5)
                          for s in ("coarsen_seconds", "initial_seconds", "uncoarsen_seconds", "free_seconds")},
        "algorithm_seconds_min": round(float(alg.min()), 6),
        "verify": {"cut": verified["cut"], "capacity": verified["capacity"],
                   "capacity_ok": verified["capacity_ok"], "seconds": round(verify_seconds, 4)},
        "quality": {"edge_cut_ratio": round(quality["edge_cut_ratio"], 6),
                    "vertex_imbalance": round(quality["vertex_imbalance"], 6),
                    "edge_set_imbalance": round(quality["edge_set_imbalance"], 6),
                    "max_local_cut": quality["max_local_cut"], "avg_local_cut": quality["avg_local_cut"],
                    "seconds": round(analyze_seconds, 4)},
        "cold_step_seconds": (round(float(cold_seconds), 6) if cold_seconds is not None else None),
        "round_graphs": {"plans": graph_stats["plans"], "recorded_rounds": graph_stats["captures"],
                         "replayed_rounds": graph_stats["replays"],
                         "note": ("refinement rounds launched kernel by kernel (--no-round-graphs)"
                                  if args.no_round_graphs else
                                  "refinement rounds replayed from hipGraphs recorded on the first partition")},
        "device_memory": {
            "peak_reserved_gib": round(alloc_after["peak_reserved_bytes"] / 2**30, 3),
            "peak_live_gib": round(alloc_after["peak_live_bytes"] / 2**30, 3),
            "allocations_in_timed_steps": alloc_after["segment_allocations"] - alloc_before["segment_allocations"],
            "allocation_seconds_in_timed_steps": round(alloc_after["segment_allocation_seconds"]
                                                       - alloc_before["segment_allocation_seconds"], 4),
            "allocations_total": alloc_after["segment_allocations"],
            "allocation_seconds_total": round(alloc_after["segment_allocation_seconds"], 4),
            "out_of_memory_releases": alloc_after["out_of_memory_releases"],
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if precision:
        line["fp32_vs_fp64"] = precision
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
