"""horkxg_amd -- MI355X-native multilevel k-way graph partitioner.

Python host mirror of the partitioner's C ABIs:
  * include/gpart.h     (libgpart.so)     graphs, the full pipeline, coarsening,
                                          initial partition, refinement, hierarchy I/O,
                                          generators
  * include/gpart_hip.h (libgpart_hip.so) the gfx950 kernel families, device memory,
                                          live kernel timing

Both libraries are built in-tree (horkxg_amd/lib) by __graft_entry__.build(). There is
no fallback: importing this package on a machine without the built libraries raises.
"""
from __future__ import annotations

from ._native import (  # noqa: F401
    GpartError,
    Graph,
    Hierarchy,
    PROFILE_FAMILIES,
    SPECTRAL_MODES,
    device_count,
    generate_grid2d,
    generate_mesh3d,
    generate_rmat,
    hip_lib,
    host_lib,
    lib_dir,
    partition_batch,
    set_batch_group,
    batch_group,
    partition_capacity,
    profile_enable,
    profile_read,
    profile_reset,
    record_layout,
    cache_info,
    decimal_capacity,
    cache_stats,
    set_log,
    set_refine_rebuild,
    set_refine_engine,
    set_refine_graphs,
    set_refine_balance_blocks,
    set_contract_stash,
    refine_graph_stats,
    set_spectral_queue,
    spectral_queue_settings,
    refine_keep_ratio,
    trim,
)
from . import device  # noqa: F401
