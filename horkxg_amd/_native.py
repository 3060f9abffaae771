"""ctypes bindings of libgpart.so (include/gpart.h) and libgpart_hip.so (include/gpart_hip.h)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_LIB_DIR = Path(__file__).resolve().parent / "lib"

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

WIDTHS = {"standard": 0, "big": 1}
SPECTRAL_MODES = {"plain_fp64": 0, "batched_fp64": 1, "batched_fp32": 2}
PROFILE_FAMILIES = {
    "sclp_affinity": 0,
    "contract": 1,
    "refine_cache": 2,
    "refine_proposals": 3,
    "refine_filter": 4,
    "spectral": 5,
    "rebalance": 6,
    "cache_update": 7,
    "sclp_commit": 8,
}


class GpartError(RuntimeError):
    """A failing status from the C ABI. kind mirrors the reference's exception class."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.kind = {1: "invalid_argument", 2: "overflow_error", 3: "runtime_error"}.get(code, "error")


class PartitionReport(C.Structure):
    _fields_ = [("final_cut", C.c_uint64), ("max_part_weight", C.c_uint64), ("capacity", C.c_uint64),
                ("initial_cut", C.c_uint64), ("coarsest_vertices", C.c_int64), ("levels", C.c_int32),
                ("refined_levels", C.c_int32), ("coarsen_seconds", C.c_double),
                ("initial_seconds", C.c_double), ("uncoarsen_seconds", C.c_double),
                ("free_seconds", C.c_double), ("algorithm_seconds", C.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class VerifyReport(C.Structure):
    _fields_ = [("vertices", C.c_int64), ("undirected_edges", C.c_int64), ("cut", C.c_int64),
                ("cut_ratio", C.c_double), ("max_load", C.c_int64), ("max_load_ratio", C.c_double),
                ("capacity", C.c_int64), ("capacity_ok", C.c_int32), ("parts", C.c_int32)]


class QualityReport(C.Structure):
    _fields_ = [("parts", C.c_int32), ("pad", C.c_int32), ("vertices", C.c_int64),
                ("directed_edges", C.c_int64), ("edge_cut", C.c_uint64), ("weighted_edge_cut", C.c_uint64),
                ("edge_cut_ratio", C.c_double), ("avg_local_cut", C.c_double), ("max_local_cut", C.c_uint64),
                ("weighted_avg_local_cut", C.c_double), ("weighted_max_local_cut", C.c_uint64),
                ("vertex_avg", C.c_double), ("min_vertices", C.c_int64), ("max_vertices", C.c_int64),
                ("vertex_imbalance", C.c_double), ("vertex_max_deviation", C.c_double),
                ("edge_set_avg", C.c_double), ("min_edge_set", C.c_int64), ("max_edge_set", C.c_int64),
                ("edge_set_imbalance", C.c_double), ("edge_set_max_deviation", C.c_double)]


PER_PART_COUNTERS = ("vertices", "internal", "outgoing", "incoming", "weighted_internal", "weighted_outgoing",
                     "weighted_incoming")


class SpectralLevelError(C.Structure):
    _fields_ = [("depth", C.c_int32), ("tasks", C.c_int32), ("vertices", C.c_int64),
                ("max_abs_error", C.c_double), ("rms_error", C.c_double), ("splits_differ", C.c_int32),
                ("pad", C.c_int32), ("labels_differ", C.c_int64)]


class SpectralComparison(C.Structure):
    _fields_ = [("depths", C.c_int32), ("pad", C.c_int32), ("level", SpectralLevelError * 16),
                ("cut_fp64", C.c_uint64), ("cut_fp32", C.c_uint64), ("max_load_fp64", C.c_uint64),
                ("max_load_fp32", C.c_uint64), ("labels_differ", C.c_int64)]

    def as_dict(self) -> dict:
        levels = [{name: getattr(self.level[i], name) for name, _ in SpectralLevelError._fields_ if name != "pad"}
                  for i in range(self.depths)]
        out = {name: getattr(self, name) for name, _ in self._fields_ if name not in ("pad", "level", "depths")}
        out["depths"] = levels
        return out


class RefineReport(C.Structure):
    _fields_ = [("rounds", C.c_int32), ("admission_iterations", C.c_int32),
                ("admission_fallbacks", C.c_int32), ("pad", C.c_int32), ("proposals", C.c_uint64),
                ("accepted", C.c_uint64), ("initial_cut", C.c_uint64), ("final_cut", C.c_uint64),
                ("final_max_part_weight", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


_hip = None
_host = None


def lib_dir() -> Path:
    return _LIB_DIR


def hip_lib():
    """libgpart_hip.so: the kernel library."""
    global _hip
    if _hip is None:
        path = _LIB_DIR / "libgpart_hip.so"
        if not path.exists():
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        H = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
        H.gpart_last_error.restype = C.c_char_p
        H.gpart_device_count.argtypes = [C.POINTER(C.c_int)]
        H.gpart_record_layout.argtypes = [C.c_int64, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          C.POINTER(C.c_uint64)]
        H.gpart_record_layout.restype = None
        H.gpart_profile_enable.argtypes = [C.c_void_p, C.c_int, C.c_int]
        H.gpart_profile_read.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double)]
        H.gpart_profile_reset.argtypes = [C.c_void_p]
        H.gpart_profile_read_units.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double)]
        H.gpart_device_alloc.argtypes = [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]
        H.gpart_device_free.argtypes = [C.c_void_p, C.c_void_p]
        H.gpart_memcpy_htod.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        H.gpart_memcpy_dtoh.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        H.gpart_ctx_synchronize.argtypes = [C.c_void_p]
        H.gpart_cache_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        H.gpart_refine_set_rebuild.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
        H.gpart_cache_info.argtypes = [C.c_void_p, C.c_void_p]
        H.gpart_refine_set_engine.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int64]
        H.gpart_refine_keep_ratio.argtypes = [C.c_uint64, C.c_uint64]
        H.gpart_spectral_set_queue.argtypes = [C.c_int32]
        H.gpart_refine_set_graphs.argtypes = [C.c_void_p, C.c_int]
        H.gpart_refine_set_balance_blocks.argtypes = [C.c_void_p, C.c_int64]
        H.gpart_contract_set_stash.argtypes = [C.c_void_p, C.c_int64]
        H.gpart_refine_graph_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                               C.POINTER(C.c_int64)]
        H.gpart_spectral_queue_set_tuning.argtypes = [C.c_int32, C.c_int32]
        H.gpart_spectral_queue_tuning.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        H.gpart_spectral_queue_tuning.restype = None
        _hip = H
    return _hip


def host_lib():
    """libgpart.so: C++ scheduling and the high-level C ABI."""
    global _host
    if _host is None:
        hip_lib()
        path = _LIB_DIR / "libgpart.so"
        if not path.exists():
            raise ImportError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(path))
        L.gpart_partition_capacity_c.argtypes = [C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_uint64)]
        L.gpart_thread_context.restype = C.c_void_p
        L.gpart_graph_upload.argtypes = [C.c_int, C.c_int64, _i64p, C.c_int64, _i64p, C.c_void_p, C.c_void_p,
                                         C.c_int, C.POINTER(C.c_void_p)]
        L.gpart_graph_load.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]
        L.gpart_graph_free.argtypes = [C.c_void_p]
        L.gpart_graph_vertices.restype = C.c_int64
        L.gpart_graph_vertices.argtypes = [C.c_void_p]
        L.gpart_graph_edges.restype = C.c_int64
        L.gpart_graph_edges.argtypes = [C.c_void_p]
        L.gpart_graph_cut.argtypes = [C.c_void_p, _i32p, C.POINTER(C.c_uint64)]
        L.gpart_decimal_capacity.argtypes = [C.c_char_p, C.c_int64, C.c_int, C.POINTER(C.c_int64)]
        L.gpart_graph_verify.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_char_p, C.POINTER(VerifyReport),
                                         C.c_void_p]
        L.gpart_graph_analyze.argtypes = [C.c_void_p, _i32p, C.c_int, C.POINTER(QualityReport), C.c_void_p,
                                          C.c_int]
        L.gpart_graph_partition.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint32, C.c_int, C.c_int,
                                            C.c_void_p, C.POINTER(PartitionReport)]
        L.gpart_graph_generate_rmat.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double,
                                                C.c_uint64, C.POINTER(C.c_void_p)]
        L.gpart_graph_generate_lattice.argtypes = [C.c_int, C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]
        L.gpart_graph_arrays.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.gpart_graph_spectral_compare.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint32,
                                                   C.POINTER(SpectralComparison)]
        L.gpart_partition_batch.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                            C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int, C.c_int,
                                            C.POINTER(C.c_void_p), C.POINTER(PartitionReport),
                                            C.POINTER(C.c_double)]
        L.gpart_graph_coarsen.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.POINTER(C.c_void_p)]
        L.gpart_graph_aggregate.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, _i32p, _u64p,
                                            C.POINTER(C.c_int64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.gpart_hierarchy_read.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_void_p)]
        L.gpart_hierarchy_levels.argtypes = [C.c_void_p]
        L.gpart_hierarchy_level_size.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.gpart_hierarchy_level_copy.argtypes = [C.c_void_p, C.c_int, _i64p, _i32p, _u64p, _u64p]
        L.gpart_hierarchy_map_copy.argtypes = [C.c_void_p, C.c_int, _i32p]
        L.gpart_hierarchy_stop_reason.restype = C.c_char_p
        L.gpart_hierarchy_stop_reason.argtypes = [C.c_void_p]
        L.gpart_hierarchy_write.argtypes = [C.c_void_p, C.c_char_p]
        L.gpart_hierarchy_free.argtypes = [C.c_void_p]
        L.gpart_hierarchy_uncoarsen.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_double, C.c_int, C.c_double,
                                                C.c_int, _i32p, C.POINTER(C.c_uint64)]
        L.gpart_hierarchy_uncoarsen_host.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_double, C.c_int, C.c_double,
                                                     C.c_int, _i32p]
        L.gpart_graph_initial_partition.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint32, C.c_int, _i32p,
                                                    C.c_void_p]
        L.gpart_graph_refine.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(RefineReport)]
        for name in ("gpart_generate_grid2d", "gpart_generate_mesh3d", "gpart_generate_rmat"):
            getattr(L, name).restype = C.c_void_p
        L.gpart_generate_grid2d.argtypes = [C.c_int64, C.c_int64]
        L.gpart_generate_mesh3d.argtypes = [C.c_int64, C.c_int64, C.c_int64]
        L.gpart_generate_rmat.argtypes = [C.c_int, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64]
        L.gpart_host_csr_vertices.restype = C.c_int64
        L.gpart_host_csr_vertices.argtypes = [C.c_void_p]
        L.gpart_host_csr_entries.restype = C.c_int64
        L.gpart_host_csr_entries.argtypes = [C.c_void_p]
        L.gpart_host_csr_indptr.restype = C.POINTER(C.c_int64)
        L.gpart_host_csr_indptr.argtypes = [C.c_void_p]
        L.gpart_host_csr_indices.restype = C.POINTER(C.c_int64)
        L.gpart_host_csr_indices.argtypes = [C.c_void_p]
        L.gpart_host_csr_write.argtypes = [C.c_void_p, C.c_char_p, C.c_char_p]
        L.gpart_host_csr_free.argtypes = [C.c_void_p]
        L.gpart_trim.argtypes = []
        L.gpart_set_log.argtypes = [C.c_int]
        L.gpart_set_log.restype = None
        # Report lines stay off under Python unless HORKXG_LOG=1.
        L.gpart_set_log(1 if os.environ.get("HORKXG_LOG") == "1" else 0)
        _host = L
    return _host


def _check(rc: int) -> None:
    if rc:
        raise GpartError(rc, hip_lib().gpart_last_error().decode())


def device_count() -> int:
    n = C.c_int(0)
    rc = hip_lib().gpart_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def partition_capacity(total: int, parts: int, ratio: float) -> int:
    out = C.c_uint64()
    _check(host_lib().gpart_partition_capacity_c(total, parts, ratio, C.byref(out)))
    return out.value


def decimal_capacity(imbalance: str, n: int, parts: int) -> int:
    """floor(imbalance * n / parts) for a decimal string, exactly (host only)."""
    out = C.c_int64()
    _check(host_lib().gpart_decimal_capacity(str(imbalance).encode(), int(n), int(parts), C.byref(out)))
    return out.value


def record_layout(vertices: int, bound: int):
    vb, wb, mask = C.c_int(), C.c_int(), C.c_uint64()
    hip_lib().gpart_record_layout(vertices, bound, C.byref(vb), C.byref(wb), C.byref(mask))
    return vb.value, wb.value, mask.value


class _CacheInfo(C.Structure):
    _fields_ = [("reserved_bytes", C.c_uint64), ("live_bytes", C.c_uint64), ("free_bytes", C.c_uint64),
                ("peak_reserved_bytes", C.c_uint64), ("peak_live_bytes", C.c_uint64), ("segments", C.c_int64),
                ("live_blocks", C.c_int64), ("free_blocks", C.c_int64), ("segment_allocations", C.c_int64),
                ("segment_allocation_seconds", C.c_double), ("segment_releases", C.c_int64),
                ("out_of_memory_releases", C.c_int64)]


def cache_info() -> dict:
    """The calling thread's allocator counters: bytes reserved in segments, live and
    free, their peaks, and the device allocations and releases made so far."""
    info = _CacheInfo()
    _check(hip_lib().gpart_cache_info(_ctx(), C.byref(info)))
    return {name: getattr(info, name) for name, _ in _CacheInfo._fields_}


def cache_stats() -> dict:
    """The calling thread's allocator: cached blocks/bytes, live bytes and the last
    release's segment count and seconds (0, stream wait, freeing segments)."""
    blocks, cached, live, released = C.c_int64(), C.c_uint64(), C.c_uint64(), C.c_int64()
    secs = (C.c_double * 3)()
    _check(hip_lib().gpart_cache_stats(_ctx(), C.byref(blocks), C.byref(cached), C.byref(live),
                                       C.byref(released), secs))
    return {"cached_blocks": blocks.value, "cached_bytes": cached.value, "live_bytes": live.value,
            "released_blocks": released.value, "release_seconds": list(secs)}


def set_refine_rebuild(min_entries: int = 1 << 23, share: int = 16) -> None:
    """The calling thread's refinement tuning: rounds on levels of >= min_entries
    entries rebuild the affinity cache when their movers hold more than
    1/share of the entries (share <= 0: never). Results do not depend on it."""
    _check(hip_lib().gpart_refine_set_rebuild(_ctx(), int(min_entries), int(share)))


def set_refine_engine(max_vertices: int = 0, vertices_per_group: int = 512, max_grid: int = 128,
                      max_degree: int = 4096) -> None:
    """The calling thread's single-launch refinement engine: levels of at most
    max_vertices vertices whose rows have at most max_degree entries are refined
    in one cooperative launch of ceil(n / vertices_per_group) workgroups, at most
    max_grid (<= 128); max_vertices = 0 (the default) switches it off. Results do
    not depend on it."""
    _check(hip_lib().gpart_refine_set_engine(_ctx(), int(max_vertices), int(vertices_per_group), int(max_grid),
                                             int(max_degree)))


def set_spectral_queue(enabled: bool = True, own_blocks: int | None = None, max_task: int | None = None) -> None:
    """Single-graph batched bisection through the persistent queue, or depth by
    depth (the default); own_blocks: 8-vertex blocks per queue workgroup in the
    team-size rule (8..128); max_task: largest task the single-graph driver gives
    the queue (1..4096). Process-wide; results do not depend on them."""
    H = hip_lib()
    _check(H.gpart_spectral_set_queue(int(bool(enabled))))
    if own_blocks is not None or max_task is not None:
        cur = spectral_queue_settings()
        _check(H.gpart_spectral_queue_set_tuning(int(own_blocks if own_blocks is not None else cur["own_blocks"]),
                                                 int(max_task if max_task is not None else cur["max_task"])))


def spectral_queue_settings() -> dict:
    H = hip_lib()
    own, top = C.c_int32(), C.c_int32()
    H.gpart_spectral_queue_tuning(C.byref(own), C.byref(top))
    return {"enabled": bool(H.gpart_spectral_queue_enabled()), "own_blocks": own.value, "max_task": top.value}


def set_refine_graphs(enabled: bool = True) -> None:
    """The calling thread's refinement rounds recorded as hipGraphs and replayed
    (default) or launched kernel by kernel. Results do not depend on it."""
    _check(hip_lib().gpart_refine_set_graphs(_ctx(), int(bool(enabled))))


def set_refine_balance_blocks(blocks: int) -> None:
    """The calling thread's cap on the rebalance proposals grid (result-neutral)."""
    _check(hip_lib().gpart_refine_set_balance_blocks(_ctx(), int(blocks)))


def set_contract_stash(min_vertices: int) -> None:
    """The calling thread's contraction threshold: levels of at least min_vertices
    fine vertices keep every entry's coarse neighbour between the count and encode
    passes (result-neutral; default 2^24)."""
    _check(hip_lib().gpart_contract_set_stash(_ctx(), int(min_vertices)))


def refine_graph_stats() -> dict:
    plans, captures, replays = C.c_int64(), C.c_int64(), C.c_int64()
    _check(hip_lib().gpart_refine_graph_stats(_ctx(), C.byref(plans), C.byref(captures), C.byref(replays)))
    return {"plans": plans.value, "captures": captures.value, "replays": replays.value}


def refine_keep_ratio(current: int, best: int) -> bool:
    """The stale-round test (long double)current < 0.999L * best, as the engine
    evaluates it (host evaluation of the same code)."""
    return bool(hip_lib().gpart_refine_keep_ratio(C.c_uint64(current), C.c_uint64(best)))


def set_log(enabled: bool) -> None:
    """Switch the library's stdout report lines on or off (process-wide)."""
    host_lib().gpart_set_log(int(bool(enabled)))


def trim() -> None:
    _check(host_lib().gpart_trim())


def _ctx():
    ctx = host_lib().gpart_thread_context()
    if not ctx:
        raise GpartError(3, hip_lib().gpart_last_error().decode())
    return ctx


def profile_enable(family: str, enable: bool = True) -> None:
    _check(hip_lib().gpart_profile_enable(_ctx(), PROFILE_FAMILIES[family], int(enable)))


def profile_read(family: str) -> dict:
    ms, launches, nbytes = C.c_double(), C.c_int64(), C.c_double()
    _check(hip_lib().gpart_profile_read(_ctx(), PROFILE_FAMILIES[family], C.byref(ms), C.byref(launches),
                                        C.byref(nbytes)))
    units = C.c_double()
    _check(hip_lib().gpart_profile_read_units(_ctx(), PROFILE_FAMILIES[family], C.byref(units)))
    return {"total_ms": ms.value, "launches": launches.value, "bytes": nbytes.value, "atomics": units.value}


def profile_reset() -> None:
    _check(hip_lib().gpart_profile_reset(_ctx()))


def _take_csr(handle):
    L = host_lib()
    if not handle:
        raise GpartError(1, "graph generation rejected its arguments")
    try:
        n = L.gpart_host_csr_vertices(handle)
        m = L.gpart_host_csr_entries(handle)
        indptr = np.ctypeslib.as_array(L.gpart_host_csr_indptr(handle), shape=(n + 1,)).copy()
        indices = (np.ctypeslib.as_array(L.gpart_host_csr_indices(handle), shape=(m,)).copy()
                   if m else np.zeros(0, np.int64))
    finally:
        L.gpart_host_csr_free(handle)
    return indptr, indices


def generate_grid2d(rows: int, cols: int):
    return _take_csr(host_lib().gpart_generate_grid2d(rows, cols))


def generate_mesh3d(nx: int, ny: int, nz: int):
    return _take_csr(host_lib().gpart_generate_mesh3d(nx, ny, nz))


def generate_rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                  seed: int = 0):
    return _take_csr(host_lib().gpart_generate_rmat(scale, edge_factor, a, b, c, seed))


class Hierarchy:
    """Host copy of a coarsening hierarchy (levels, fine-to-coarse maps)."""

    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _host is not None:
            _host.gpart_hierarchy_free(self._h)
            self._h = None

    @classmethod
    def read(cls, path, width: str = "standard") -> "Hierarchy":
        h = C.c_void_p()
        _check(host_lib().gpart_hierarchy_read(WIDTHS[width], str(path).encode(), C.byref(h)))
        return cls(h.value)

    def __len__(self) -> int:
        return int(host_lib().gpart_hierarchy_levels(self._h))

    def level(self, i: int):
        n, m = C.c_int64(), C.c_int64()
        _check(host_lib().gpart_hierarchy_level_size(self._h, i, C.byref(n), C.byref(m)))
        off = np.empty(n.value + 1, np.int64)
        adj = np.empty(m.value, np.int32)
        ew = np.empty(m.value, np.uint64)
        vw = np.empty(n.value, np.uint64)
        _check(host_lib().gpart_hierarchy_level_copy(self._h, i, off, adj, ew, vw))
        return off, adj, ew, vw

    def map(self, i: int):
        n = C.c_int64()
        m = C.c_int64()
        _check(host_lib().gpart_hierarchy_level_size(self._h, i, C.byref(n), C.byref(m)))
        out = np.empty(n.value, np.int32)
        _check(host_lib().gpart_hierarchy_map_copy(self._h, i, out))
        return out

    @property
    def stop_reason(self) -> str:
        return host_lib().gpart_hierarchy_stop_reason(self._h).decode()

    def write(self, path) -> None:
        _check(host_lib().gpart_hierarchy_write(self._h, str(path).encode()))

    def uncoarsen(self, coarsest_labels, parts: int, ratio: float = 1.03, max_rounds: int = 32,
                  growth: float = 2.0, verify: bool = False, entry: str = "refine_hierarchy_device"):
        """Projection + refinement from labels of the coarsest level. entry
        "refine_hierarchy_device" returns (labels, final cut); entry "uncoarsen" goes
        through the host uncoarsen overload and returns (labels, None)."""
        lab = np.ascontiguousarray(coarsest_labels, dtype=np.int32)
        n0 = C.c_int64()
        m0 = C.c_int64()
        _check(host_lib().gpart_hierarchy_level_size(self._h, 0, C.byref(n0), C.byref(m0)))
        n_coarse = C.c_int64()
        _check(host_lib().gpart_hierarchy_level_size(self._h, len(self) - 1, C.byref(n_coarse), C.byref(m0)))
        if lab.shape != (n_coarse.value,):
            raise GpartError(1, "coarsest labels must have one entry per coarsest vertex")
        out = np.empty(n0.value, np.int32)
        if entry == "uncoarsen":
            _check(host_lib().gpart_hierarchy_uncoarsen_host(self._h, lab, parts, ratio, max_rounds, growth,
                                                             int(verify), out))
            return out, None
        if entry != "refine_hierarchy_device":
            raise ValueError(f"unknown entry {entry!r}")
        cut = C.c_uint64()
        _check(host_lib().gpart_hierarchy_uncoarsen(self._h, lab, parts, ratio, max_rounds, growth, int(verify),
                                                    out, C.byref(cut)))
        return out, cut.value


def set_batch_group(graphs: int = 16) -> None:
    """Graphs per group of partition_batch: a group's initial partitions run in
    one persistent launch (batched modes). 1: one graph at a time."""
    _check(host_lib().gpart_batch_set_group(int(graphs)))


def batch_group() -> int:
    return int(host_lib().gpart_batch_group())


def partition_batch(graphs, parts, ratios=None, seeds=None, spectral_mode: str = "batched_fp32",
                    workers: int = 1, want_labels: bool = True):
    """Partition many graphs, largest first, in the calling thread (workers must
    be 1: the concurrent form is switched off), in groups of batch_group() graphs
    of one width whose initial partitions share one launch. Returns (labels list or
    None, report dicts, wall seconds from the first coarsening start to the last
    labels on the device)."""
    count = len(graphs)
    parts = list(parts)
    ratios = [1.03] * count if ratios is None else list(ratios)
    seeds = list(range(count)) if seeds is None else list(seeds)
    if not (len(parts) == len(ratios) == len(seeds) == count):
        raise ValueError("parts, ratios and seeds need one entry per graph")
    handles = (C.c_void_p * max(count, 1))(*[g._h for g in graphs])
    labels = [np.empty(g.n, np.int32) for g in graphs] if want_labels else None
    label_ptrs = (C.c_void_p * max(count, 1))(*([a.ctypes.data for a in labels] if labels else [None] * count))
    reports = (PartitionReport * max(count, 1))()
    wall = C.c_double()
    _check(host_lib().gpart_partition_batch(count, handles, (C.c_int * max(count, 1))(*parts),
                                            (C.c_double * max(count, 1))(*ratios),
                                            (C.c_uint32 * max(count, 1))(*seeds), SPECTRAL_MODES[spectral_mode],
                                            workers, label_ptrs if labels else None, reports, C.byref(wall)))
    return labels, [reports[i].as_dict() for i in range(count)], wall.value


class Graph:
    """A validated input graph resident on the GPU (plus its host copy)."""

    def __init__(self, indptr, indices, width: str = "standard", edge_weights=None, vertex_weights=None,
                 validate: bool = False):
        L = host_lib()
        off = np.ascontiguousarray(indptr, dtype=np.int64)
        adj = np.ascontiguousarray(indices, dtype=np.int64)
        self._keep = []
        ew = vw = None
        if edge_weights is not None:
            ew = np.ascontiguousarray(edge_weights, dtype=np.uint64)
            self._keep.append(ew)
        if vertex_weights is not None:
            vw = np.ascontiguousarray(vertex_weights, dtype=np.uint64)
            self._keep.append(vw)
        h = C.c_void_p()
        _check(L.gpart_graph_upload(WIDTHS[width], len(off) - 1, off, len(adj), adj,
                                    ew.ctypes.data if ew is not None else None,
                                    vw.ctypes.data if vw is not None else None, int(validate), C.byref(h)))
        self._h = h.value
        self.width = width

    @classmethod
    def load(cls, indptr_path, indices_path, width: str = "standard", validate: bool = False) -> "Graph":
        g = cls.__new__(cls)
        h = C.c_void_p()
        _check(host_lib().gpart_graph_load(WIDTHS[width], str(indptr_path).encode(), str(indices_path).encode(),
                                           int(validate), C.byref(h)))
        g._h = h.value
        g._keep = []
        g.width = width
        return g

    @classmethod
    def _adopt(cls, handle, width: str) -> "Graph":
        g = cls.__new__(cls)
        g._h = handle
        g._keep = []
        g.width = width
        return g

    @classmethod
    def generate_rmat(cls, scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
                      seed: int = 0, width: str = "standard") -> "Graph":
        """The graph of generate_rmat(...), built on the GPU (unit weights)."""
        h = C.c_void_p()
        _check(host_lib().gpart_graph_generate_rmat(WIDTHS[width], scale, edge_factor, a, b, c, seed, C.byref(h)))
        return cls._adopt(h.value, width)

    @classmethod
    def generate_grid2d(cls, rows: int, cols: int, width: str = "standard") -> "Graph":
        h = C.c_void_p()
        _check(host_lib().gpart_graph_generate_lattice(WIDTHS[width], cols, rows, 1, C.byref(h)))
        return cls._adopt(h.value, width)

    @classmethod
    def generate_mesh3d(cls, nx: int, ny: int, nz: int, width: str = "standard") -> "Graph":
        h = C.c_void_p()
        _check(host_lib().gpart_graph_generate_lattice(WIDTHS[width], nx, ny, nz, C.byref(h)))
        return cls._adopt(h.value, width)

    def arrays(self):
        """(offsets int64, neighbors int32, edge weights u64, vertex weights u64) on the host."""
        n, m = self.n, self.m
        off = np.empty(n + 1, np.int64)
        nbr = np.empty(m, np.int32)
        ew = np.empty(m, np.uint64)
        vw = np.empty(n, np.uint64)
        _check(host_lib().gpart_graph_arrays(self._h, off.ctypes.data, nbr.ctypes.data, ew.ctypes.data,
                                             vw.ctypes.data))
        return off, nbr, ew, vw

    def __del__(self):
        if getattr(self, "_h", None) and _host is not None:
            _host.gpart_graph_free(self._h)
            self._h = None

    @property
    def n(self) -> int:
        return int(host_lib().gpart_graph_vertices(self._h))

    @property
    def m(self) -> int:
        return int(host_lib().gpart_graph_edges(self._h))

    def cut(self, labels) -> int:
        out = C.c_uint64()
        _check(host_lib().gpart_graph_cut(self._h, np.ascontiguousarray(labels, dtype=np.int32), C.byref(out)))
        return out.value

    def verify(self, labels, parts: int, imbalance: str = "1.03") -> dict:
        """tools/verify_partition.py:22-65 on the device: the same checks, error
        messages (GpartError, kind "invalid_argument") and result dictionary."""
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        if labels.size != self.n:
            raise GpartError(1, "partition label count does not match vertices")
        rep = VerifyReport()
        loads = np.zeros(max(int(parts), 1), np.int64)
        _check(host_lib().gpart_graph_verify(self._h, labels, int(parts), str(imbalance).encode(), C.byref(rep),
                                             loads.ctypes.data))
        return {
            "vertices": rep.vertices, "undirected_edges": rep.undirected_edges, "cut": rep.cut,
            "cut_ratio": rep.cut_ratio, "loads": [int(x) for x in loads[:parts]], "max_load": rep.max_load,
            "max_load_ratio": rep.max_load_ratio, "capacity": rep.capacity, "capacity_ok": bool(rep.capacity_ok),
        }

    def analyze(self, labels, parts: int = 0) -> dict:
        """graph_untils/analyze_partition.cpp:176-294 on the device; parts == 0 infers
        max label + 1. Weighted metrics use the graph's edge weights. per_part holds
        the per-part counters as int64 arrays."""
        labels = np.ascontiguousarray(labels, dtype=np.int32)
        if labels.size != self.n:
            raise GpartError(1, "Partition line count must equal CSR vertex count")
        cap = max(int(parts), int(labels.max()) + 1 if labels.size else 0, 1)
        per_part = np.zeros(7 * cap, np.uint64)
        rep = QualityReport()
        _check(host_lib().gpart_graph_analyze(self._h, labels, int(parts), C.byref(rep), per_part.ctypes.data, cap))
        out = {name: getattr(rep, name) for name, _ in QualityReport._fields_ if name != "pad"}
        k = rep.parts
        out["per_part"] = {name: per_part[i * k:(i + 1) * k].astype(np.int64)
                           for i, name in enumerate(PER_PART_COUNTERS)}
        return out

    def partition(self, parts: int, ratio: float = 1.03, seed: int = 0, verify: bool = False,
                  spectral_mode: str = "batched_fp64", want_labels: bool = True):
        labels = np.empty(self.n, np.int32) if want_labels else None
        rep = PartitionReport()
        _check(host_lib().gpart_graph_partition(self._h, parts, ratio, seed, int(verify),
                                                SPECTRAL_MODES[spectral_mode],
                                                labels.ctypes.data if labels is not None else None,
                                                C.byref(rep)))
        return labels, rep.as_dict()

    def spectral_compare(self, parts: int, ratio: float = 1.03, seed: int = 0) -> dict:
        """FP32 against FP64 diffusion of the batched bisection on this graph, per
        recursion depth, and the cuts of both complete recursions."""
        out = SpectralComparison()
        _check(host_lib().gpart_graph_spectral_compare(self._h, parts, ratio, seed, C.byref(out)))
        return out.as_dict()

    def coarsen(self, parts: int, seed: int = 0, stop_ratio: float = 0.90, max_levels: int = 24,
                cluster_size_factor: int = 256, pack_isolated: bool = False, verify: bool = False) -> Hierarchy:
        h = C.c_void_p()
        _check(host_lib().gpart_graph_coarsen(self._h, parts, seed, stop_ratio, max_levels, cluster_size_factor,
                                              int(pack_isolated), int(verify), C.byref(h)))
        return Hierarchy(h.value)

    def aggregate(self, parts: int, seed: int = 0, level: int = 0, cluster_size_factor: int = 256,
                  pack_isolated: bool = False):
        """One SCLP aggregation: (map, cluster weights, coarse count, capacity, max weight)."""
        mp = np.empty(self.n, np.int32)
        cw = np.empty(max(self.n, 1), np.uint64)
        coarse, cap, mx = C.c_int64(), C.c_uint64(), C.c_uint64()
        _check(host_lib().gpart_graph_aggregate(self._h, parts, seed, level, cluster_size_factor, int(pack_isolated),
                                                mp, cw, C.byref(coarse), C.byref(cap), C.byref(mx)))
        return mp, cw[: coarse.value].copy(), coarse.value, cap.value, mx.value

    def initial_partition(self, parts: int, ratio: float = 1.03, seed: int = 0,
                          spectral_mode: str = "batched_fp64", with_coords: bool = False):
        labels = np.empty(self.n, np.int32)
        coords = np.empty(self.n, np.float64) if with_coords else None
        _check(host_lib().gpart_graph_initial_partition(self._h, parts, ratio, seed, SPECTRAL_MODES[spectral_mode],
                                                        labels, coords.ctypes.data if coords is not None else None))
        return (labels, coords) if with_coords else labels

    def refine(self, labels, parts: int, ratio: float = 1.03, max_rounds: int = 32, stale_rounds: int = 3,
               verify: bool = False):
        lab = np.array(labels, dtype=np.int32, copy=True)
        rep = RefineReport()
        _check(host_lib().gpart_graph_refine(self._h, lab, parts, ratio, max_rounds, stale_rounds, int(verify),
                                             C.byref(rep)))
        return lab, rep.as_dict()
