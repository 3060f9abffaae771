// tuned_gfx950.hpp -- defaults of the result-neutral tunables, chosen by
// horkxg_amd/autotune.py on an MI355X (measurements in horkxg_amd/tuning/gfx950.json).
// Generated: edit by re-running `python -m horkxg_amd.autotune --write` on the GPU.
#pragma once

#include <cstdint>

namespace gpart_tuned {
constexpr std::int64_t kRebuildShare = 16;
constexpr std::int64_t kRebuildMinEntries = 8388608;
constexpr int kQueueOwnBlocks = 8;
constexpr int kQueueSingleMaxTask = 0;  // 0: single graphs run depth by depth
constexpr int kBatchGroup = 16;
constexpr std::int64_t kBalanceBlocks = 4096;
}  // namespace gpart_tuned
