"""The autotuner's committed output: the compiled-in defaults header and the
measurement record agree, and the header writer round-trips."""
from __future__ import annotations

import json
import re

from horkxg_amd import autotune


def header_values(text: str) -> dict:
    names = {"kRebuildShare": "rebuild_share", "kRebuildMinEntries": "rebuild_min_entries",
             "kQueueOwnBlocks": "queue_own_blocks", "kQueueSingleMaxTask": "queue_single_max_task",
             "kBatchGroup": "batch_group", "kBalanceBlocks": "balance_blocks"}
    out = {}
    for const, key in names.items():
        m = re.search(rf"constexpr [\w:]+ {const} = (-?\d+);", text)
        assert m, const
        out[key] = int(m.group(1))
    return out


def test_header_round_trip(tmp_path):
    chosen = {"rebuild_share": 8, "rebuild_min_entries": 1 << 21, "queue_own_blocks": 16,
              "queue_single_max_task": 512, "batch_group": 4, "balance_blocks": 1024}
    path = tmp_path / "tuned.hpp"
    autotune.write_header(chosen, path)
    assert header_values(path.read_text()) == chosen


def test_committed_header_matches_record():
    values = header_values(autotune.HEADER_PATH.read_text())
    assert set(values) == set(autotune.DEFAULTS)
    if autotune.JSON_PATH.exists():
        record = json.loads(autotune.JSON_PATH.read_text())
        assert record["chosen"] == values
        # every sweep kept labels identical (the tuner aborts otherwise) and
        # chose one of its own candidates
        for sweep in record["sweeps"].values():
            values_tried = [tuple(r["value"]) if isinstance(r["value"], list) else r["value"]
                            for r in sweep["rows"]]
            chosen = tuple(sweep["chosen"]) if isinstance(sweep["chosen"], list) else sweep["chosen"]
            assert chosen in values_tried
