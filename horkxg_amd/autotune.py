"""Autotuner of the library's result-neutral tunables for gfx950.

Every tunable below changes how work is mapped onto the GPU, never the result:
the tuner partitions representative workloads with each candidate setting,
requires the labels of every candidate to be identical to the default's (a
mismatch aborts the run), and picks per tunable the candidate with the lowest
geometric mean time over the workloads, keeping the current default unless a
candidate beats it by more than `--margin` (2% by default, above run-to-run
noise).

Tunables (defaults compiled into the library from tuned_gfx950.hpp):
  rebuild_share, rebuild_min_entries  refinement rounds rebuild the affinity cache
                                      instead of updating it when the level has at
                                      least rebuild_min_entries entries and the
                                      movers hold more than 1/rebuild_share of them
  queue_own_blocks                    8-vertex blocks per workgroup of a team in the
                                      persistent bisection queue (team size rule)
  queue_single_max_task               largest task the single-graph driver gives the
                                      queue (0: per-depth launches only)
  batch_group                         graphs per group of partition_batch
  balance_blocks                      workgroups of the rebalance proposals launch at
                                      most (each flushes its bucket table once)

Run on the GPU from the repository root:
    python -m horkxg_amd.autotune --write
It prints a table per tunable and, with --write, rewrites
horkxg_amd/tuning/gfx950.json (every measurement) and
horkxg_amd/csrc/tuned_gfx950.hpp (the chosen defaults; rebuild the library after).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import statistics
import sys
import time
from pathlib import Path

import numpy as np

PACKAGE = Path(__file__).resolve().parent
JSON_PATH = PACKAGE / "tuning" / "gfx950.json"
HEADER_PATH = PACKAGE / "csrc" / "tuned_gfx950.hpp"

# The library's built-in defaults (the values tuned_gfx950.hpp held before tuning).
DEFAULTS = {"rebuild_share": 16, "rebuild_min_entries": 1 << 23, "queue_own_blocks": 8,
            "queue_single_max_task": 0, "batch_group": 16, "balance_blocks": 4096}


def say(*args):
    print(*args, file=sys.stderr, flush=True)


def digest(labels) -> str:
    return hashlib.sha256(np.ascontiguousarray(labels, dtype=np.int32).tobytes()).hexdigest()[:16]


class Workload:
    def __init__(self, hx, name, graph, k, mode):
        self.hx, self.name, self.graph, self.k, self.mode = hx, name, graph, k, mode

    def run(self, repeats):
        """(median algorithm seconds, labels digest) over `repeats` partitions after one warmup."""
        self.graph.partition(self.k, 1.03, 0, spectral_mode=self.mode, want_labels=False)
        times, digests = [], set()
        for _ in range(repeats):
            labels, rep = self.graph.partition(self.k, 1.03, 0, spectral_mode=self.mode)
            times.append(rep["algorithm_seconds"])
            digests.add(digest(labels))
        if len(digests) != 1:
            raise RuntimeError(f"{self.name}: labels differ between repeats")
        return statistics.median(times), digests.pop()


class MixWorkload:
    def __init__(self, hx, graphs, specs):
        self.hx, self.graphs, self.specs, self.name = hx, graphs, specs, f"mix{len(graphs)}"

    def run(self, repeats):
        args = ([s["k"] for s in self.specs], [s["ratio"] for s in self.specs], [s["seed"] for s in self.specs])
        self.hx.partition_batch(self.graphs, *args, spectral_mode="batched_fp32", want_labels=False)
        times, digests = [], set()
        for _ in range(repeats):
            labels, _, wall = self.hx.partition_batch(self.graphs, *args, spectral_mode="batched_fp32")
            times.append(wall)
            digests.add(digest(np.concatenate(labels)))
        if len(digests) != 1:
            raise RuntimeError("mix: labels differ between repeats")
        return statistics.median(times), digests.pop()


def apply(hx, setting):
    hx.set_refine_rebuild(setting["rebuild_min_entries"], setting["rebuild_share"])
    max_task = setting["queue_single_max_task"]
    hx.set_spectral_queue(max_task > 0, setting["queue_own_blocks"], max(1, max_task) if max_task else 4096)
    hx.set_batch_group(setting["batch_group"])
    hx.set_refine_balance_blocks(setting["balance_blocks"])


def sweep(hx, workloads, base, name, candidates, repeats, reference, margin):
    """Times every candidate value of one tunable (others at `base`)."""
    rows = []
    for value in candidates:
        setting = dict(base)
        if isinstance(name, tuple):
            for key, v in zip(name, value):
                setting[key] = v
        else:
            setting[name] = value
        apply(hx, setting)
        times = {}
        for w in workloads:
            t, d = w.run(repeats)
            if d != reference[w.name]:
                raise RuntimeError(f"{name}={value}: labels of {w.name} differ from the default's")
            times[w.name] = t
        gm = math.exp(sum(math.log(t) for t in times.values()) / len(times))
        rows.append({"value": value, "seconds": times, "geomean": gm})
        say(f"  {name}={value}: " + ", ".join(f"{k} {v * 1e3:.2f} ms" for k, v in times.items())
            + f" (geomean {gm * 1e3:.2f} ms)")
    current = next(r for r in rows if r["value"] == (tuple(base[k] for k in name) if isinstance(name, tuple)
                                                    else base[name]))
    best = min(rows, key=lambda r: r["geomean"])
    chosen = best if best["geomean"] < current["geomean"] * (1.0 - margin) else current
    return rows, chosen["value"]


def write_header(chosen, path=HEADER_PATH):
    text = f"""// tuned_gfx950.hpp -- defaults of the result-neutral tunables, chosen by
// horkxg_amd/autotune.py on an MI355X (measurements in horkxg_amd/tuning/gfx950.json).
// Generated: edit by re-running `python -m horkxg_amd.autotune --write` on the GPU.
#pragma once

#include <cstdint>

namespace gpart_tuned {{
constexpr std::int64_t kRebuildShare = {chosen['rebuild_share']};
constexpr std::int64_t kRebuildMinEntries = {chosen['rebuild_min_entries']};
constexpr int kQueueOwnBlocks = {chosen['queue_own_blocks']};
constexpr int kQueueSingleMaxTask = {chosen['queue_single_max_task']};  // 0: single graphs run depth by depth
constexpr int kBatchGroup = {chosen['batch_group']};
constexpr std::int64_t kBalanceBlocks = {chosen['balance_blocks']};
}}  // namespace gpart_tuned
"""
    path.write_text(text)


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--repeats", type=int, default=3)
    p.add_argument("--margin", type=float, default=0.02)
    p.add_argument("--quick", action="store_true", help="small workloads only (a smoke run of the tuner)")
    p.add_argument("--write", action="store_true", help="write the JSON record and the defaults header")
    p.add_argument("--out", default=None,
                   help="directory for the two files instead of their places in the package")
    args = p.parse_args(argv)

    import horkxg_amd as hx
    from horkxg_amd import workloads as W

    if hx.device_count() < 1:
        say("the autotuner needs a GPU")
        return 2
    hx.set_log(False)
    t0 = time.perf_counter()
    if args.quick:
        single = [Workload(hx, "grid128_k8", hx.Graph(*hx.generate_grid2d(128, 128)), 8, "batched_fp64"),
                  Workload(hx, "rmat14_k8", hx.Graph(*hx.generate_rmat(14, 16)), 8, "batched_fp32")]
        specs = W.mix_specs(8, min_n=2_000, max_n=20_000)
    else:
        single = [Workload(hx, "mesh128_k32", hx.Graph.generate_mesh3d(128, 128, 128), 32, "batched_fp64"),
                  Workload(hx, "rmat22_k8", hx.Graph.generate_rmat(22, 16), 8, "batched_fp32"),
                  Workload(hx, "grid256_k4", hx.Graph.generate_grid2d(256, 256), 4, "batched_fp64")]
        specs = W.mix_specs(48)
    mix = MixWorkload(hx, [W.device_graph(hx, s) for s in specs], specs)
    say(f"[autotune] workloads ready in {time.perf_counter() - t0:.1f} s")

    base = dict(DEFAULTS)
    apply(hx, base)
    reference = {}
    for w in single + [mix]:
        _, reference[w.name] = w.run(1)
    record = {"device": "gfx950", "defaults": dict(DEFAULTS), "workloads": [w.name for w in single + [mix]],
              "repeats": args.repeats, "margin": args.margin, "sweeps": {}}

    plan = [
        (("rebuild_share", "rebuild_min_entries"),
         [(16, 1 << 23), (8, 1 << 23), (32, 1 << 23), (16, 1 << 21), (16, 1 << 25), (0, 1 << 23)], single),
        ("queue_own_blocks", [8, 16, 32, 64, 128], single + [mix]),
        ("queue_single_max_task", [0, 256, 512, 1024, 4096], single),
        ("batch_group", [1, 4, 8, 16, 32], [mix]),
        ("balance_blocks", [4096, 512, 1024, 2048, 8192], single),
    ]
    for name, candidates, on in plan:
        say(f"[autotune] {name}")
        rows, value = sweep(hx, on, base, name, candidates, args.repeats, reference, args.margin)
        key = "+".join(name) if isinstance(name, tuple) else name
        record["sweeps"][key] = {"workloads": [w.name for w in on], "rows": rows, "chosen": value}
        if isinstance(name, tuple):
            for k, v in zip(name, value):
                base[k] = v
        else:
            base[name] = value
        say(f"[autotune] {key} -> {value}")
    record["chosen"] = base
    record["seconds"] = round(time.perf_counter() - t0, 1)
    print(json.dumps({"chosen": base, "seconds": record["seconds"]}))
    if args.write:
        json_path = Path(args.out) / JSON_PATH.name if args.out else JSON_PATH
        header_path = Path(args.out) / HEADER_PATH.name if args.out else HEADER_PATH
        json_path.parent.mkdir(parents=True, exist_ok=True)
        json_path.write_text(json.dumps(record, indent=1) + "\n")
        write_header(base, header_path)
        say(f"[autotune] wrote {json_path} and {header_path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
