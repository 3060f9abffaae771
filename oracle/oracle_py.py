"""ctypes binding of the CPU oracle (oracle/lib/libgpart_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is the parity checker and the CPU
baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "lib" / "libgpart_oracle.so"
_REF_PATH = _HERE / "_ref" / "libref_host.so"

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.kind = {1: "invalid_argument", 2: "overflow_error", 3: "runtime_error"}.get(code, "error")


class RefineStats(C.Structure):
    _fields_ = [("rounds", C.c_int), ("proposals", C.c_uint64), ("accepted", C.c_uint64),
                ("initial_cut", C.c_uint64), ("final_cut", C.c_uint64),
                ("final_max_part_weight", C.c_uint64), ("admission_iterations", C.c_int),
                ("admission_fallbacks", C.c_int)]


class PartitionResult(C.Structure):
    _fields_ = [("final_cut", C.c_uint64), ("max_part_weight", C.c_uint64),
                ("capacity", C.c_uint64), ("levels", C.c_int),
                ("coarsest_vertices", C.c_int64), ("initial_cut", C.c_uint64),
                ("coarsen_seconds", C.c_double), ("initial_seconds", C.c_double),
                ("uncoarsen_seconds", C.c_double), ("total_seconds", C.c_double)]


_lib = None


def available() -> bool:
    return _LIB_PATH.exists()


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise FileNotFoundError(f"oracle library not built: {_LIB_PATH} (run make -C oracle)")
    L = C.CDLL(str(_LIB_PATH))
    L.gpo_last_error.restype = C.c_char_p
    L.gpo_set_threads.argtypes = [C.c_int]
    L.gpo_get_threads.restype = C.c_int
    L.gpo_partition_capacity.argtypes = [C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_uint64)]
    L.gpo_graph_create.restype = C.c_void_p
    L.gpo_graph_create.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, C.c_void_p, C.c_void_p]
    L.gpo_graph_free.argtypes = [C.c_void_p]
    L.gpo_graph_n.restype = C.c_int64
    L.gpo_graph_n.argtypes = [C.c_void_p]
    L.gpo_graph_m.restype = C.c_int64
    L.gpo_graph_m.argtypes = [C.c_void_p]
    L.gpo_graph_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.gpo_aggregate.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                _i32p, _u64p, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                C.POINTER(C.c_int)]
    L.gpo_aggregate_trace.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _i32p, _i32p, _i32p, _u32p, C.POINTER(C.c_int)]
    L.gpo_sclp_proposals.argtypes = [C.c_void_p, _i32p, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                     C.c_uint32, _i32p, _u64p, _u32p]
    L.gpo_contract.restype = C.c_void_p
    L.gpo_contract.argtypes = [C.c_void_p, _i32p, C.c_int64, _u64p]
    L.gpo_coarsen.restype = C.c_void_p
    L.gpo_coarsen.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_double, C.c_int, C.c_int, C.c_int]
    L.gpo_hierarchy_levels.restype = C.c_int
    L.gpo_hierarchy_levels.argtypes = [C.c_void_p]
    L.gpo_hierarchy_level.restype = C.c_void_p
    L.gpo_hierarchy_level.argtypes = [C.c_void_p, C.c_int]
    L.gpo_hierarchy_map.restype = C.POINTER(C.c_int32)
    L.gpo_hierarchy_map.argtypes = [C.c_void_p, C.c_int]
    L.gpo_hierarchy_stop_reason.restype = C.c_char_p
    L.gpo_hierarchy_stop_reason.argtypes = [C.c_void_p]
    L.gpo_write_hierarchy.argtypes = [C.c_void_p, C.c_char_p]
    L.gpo_hierarchy_free.argtypes = [C.c_void_p]
    L.gpo_initial_partition_trace.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint32, _i32p,
                                              C.c_void_p, C.POINTER(C.c_int)]
    L.gpo_refine.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_double, C.c_int, C.c_int,
                             C.POINTER(RefineStats)]
    L.gpo_uncoarsen.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_double, C.c_int, C.c_double,
                                _i32p, C.POINTER(C.c_uint64)]
    L.gpo_partition.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_uint32, _i32p,
                                C.POINTER(PartitionResult)]
    L.gpo_cut.argtypes = [C.c_void_p, _i32p, C.POINTER(C.c_uint64)]
    L.gpo_part_weights.argtypes = [C.c_void_p, _i32p, C.c_int, _u64p]
    L.gpo_validate_csr.argtypes = [C.c_void_p, C.c_int, C.c_int]
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc:
        raise OracleError(rc, lib().gpo_last_error().decode())


def set_threads(threads: int) -> None:
    lib().gpo_set_threads(int(threads))


def threads() -> int:
    return int(lib().gpo_get_threads())


def partition_capacity(total: int, parts: int, ratio: float) -> int:
    out = C.c_uint64()
    _check(lib().gpo_partition_capacity(total, parts, ratio, C.byref(out)))
    return out.value


class Graph:
    """Oracle-side CSR graph (owned by the oracle library)."""

    def __init__(self, offsets, neighbors, edge_weights=None, vertex_weights=None, _handle=None):
        L = lib()
        if _handle is not None:
            self._h = _handle
            self._owned = False
            return
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        adj = np.ascontiguousarray(neighbors, dtype=np.int32)
        n, m = len(off) - 1, len(adj)
        self._keep = []
        ew = vw = None
        if edge_weights is not None:
            ew = np.ascontiguousarray(edge_weights, dtype=np.uint64)
            self._keep.append(ew)
        if vertex_weights is not None:
            vw = np.ascontiguousarray(vertex_weights, dtype=np.uint64)
            self._keep.append(vw)
        self._h = L.gpo_graph_create(n, m, off, adj,
                                     ew.ctypes.data if ew is not None else None,
                                     vw.ctypes.data if vw is not None else None)
        self._owned = True

    def __del__(self):
        if getattr(self, "_owned", False) and _lib is not None:
            _lib.gpo_graph_free(self._h)
            self._owned = False

    @property
    def n(self) -> int:
        return int(lib().gpo_graph_n(self._h))

    @property
    def m(self) -> int:
        return int(lib().gpo_graph_m(self._h))

    def arrays(self):
        n, m = self.n, self.m
        off = np.empty(n + 1, np.int64)
        adj = np.empty(m, np.int32)
        ew = np.empty(m, np.uint64)
        vw = np.empty(n, np.uint64)
        lib().gpo_graph_copy(self._h, off.ctypes.data, adj.ctypes.data, ew.ctypes.data, vw.ctypes.data)
        return off, adj, ew, vw


def aggregate(g: Graph, parts: int, seed: int, level: int, cluster_size_factor: int = 256,
              pack_isolated: bool = False):
    n = g.n
    mp = np.empty(n, np.int32)
    cw = np.empty(n, np.uint64)
    coarse, cap, rounds = C.c_int64(), C.c_uint64(), C.c_int()
    _check(lib().gpo_aggregate(g._h, parts, seed, level, cluster_size_factor, int(pack_isolated), mp, cw,
                               C.byref(coarse), C.byref(cap), C.byref(rounds)))
    return mp, cw[: coarse.value].copy(), coarse.value, cap.value, rounds.value


def aggregate_trace(g: Graph, parts: int, seed: int, level: int, cluster_size_factor: int = 256,
                    pack_isolated: bool = False):
    """One SCLP level step by step: a list of (kind, round, threshold, clusters)
    after every LP round ("lp"), isolated packing ("isolated") and two-hop
    pairing ("two_hop")."""
    n, cap = g.n, 8
    clusters = np.empty(cap * n, np.int32)
    kinds = np.empty(cap, np.int32)
    rounds = np.empty(cap, np.int32)
    thresholds = np.empty(cap, np.uint32)
    steps = C.c_int()
    _check(lib().gpo_aggregate_trace(g._h, parts, seed, level, cluster_size_factor, int(pack_isolated), cap,
                                     clusters, kinds, rounds, thresholds, C.byref(steps)))
    names = {0: "lp", 1: "isolated", 2: "two_hop"}
    return [(names[int(kinds[i])], int(rounds[i]), int(thresholds[i]), clusters[i * n:(i + 1) * n].copy())
            for i in range(steps.value)]


def sclp_proposals(g: Graph, clusters, seed: int, level: int, rnd: int, filter_roles: bool,
                   mover_threshold: int):
    n = g.n
    c = np.ascontiguousarray(clusters, dtype=np.int32)
    p = np.empty(n, np.int32)
    gn = np.empty(n, np.uint64)
    t = np.empty(n, np.uint32)
    _check(lib().gpo_sclp_proposals(g._h, c, seed, level, rnd, int(filter_roles), mover_threshold, p, gn, t))
    return p, gn, t


def contract(g: Graph, mapping, coarse: int, coarse_weights) -> Graph:
    mp = np.ascontiguousarray(mapping, dtype=np.int32)
    cw = np.ascontiguousarray(coarse_weights, dtype=np.uint64)
    h = lib().gpo_contract(g._h, mp, coarse, cw)
    if not h:
        raise OracleError(3, lib().gpo_last_error().decode())
    out = Graph(None, None, _handle=h)
    out._owned = True
    return out


class Hierarchy:
    def __init__(self, handle):
        self._h = handle

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.gpo_hierarchy_free(self._h)
            self._h = None

    def __len__(self):
        return int(lib().gpo_hierarchy_levels(self._h))

    def level(self, i: int) -> Graph:
        return Graph(None, None, _handle=lib().gpo_hierarchy_level(self._h, i))

    def map(self, i: int):
        n = self.level(i).n
        ptr = lib().gpo_hierarchy_map(self._h, i)
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    @property
    def stop_reason(self) -> str:
        return lib().gpo_hierarchy_stop_reason(self._h).decode()

    def write(self, path) -> None:
        _check(lib().gpo_write_hierarchy(self._h, str(path).encode()))


def coarsen(g: Graph, parts: int, seed: int = 0, stop_ratio: float = 0.90, max_levels: int = 24,
            cluster_size_factor: int = 256, pack_isolated: bool = False) -> Hierarchy:
    h = lib().gpo_coarsen(g._h, parts, seed, stop_ratio, max_levels, cluster_size_factor, int(pack_isolated))
    if not h:
        raise OracleError(3, lib().gpo_last_error().decode())
    return Hierarchy(h)


def initial_partition(g: Graph, parts: int, ratio: float = 1.03, seed: int = 0, with_coords=False):
    n = g.n
    labels = np.empty(n, np.int32)
    coords = np.empty(n, np.float64) if with_coords else None
    tasks = C.c_int()
    _check(lib().gpo_initial_partition_trace(g._h, parts, ratio, seed, labels,
                                             coords.ctypes.data if coords is not None else None,
                                             C.byref(tasks)))
    if with_coords:
        return labels, coords, tasks.value
    return labels


def refine(g: Graph, labels, parts: int, ratio: float = 1.03, max_rounds: int = 32, stale_rounds: int = 3):
    lab = np.array(labels, dtype=np.int32, copy=True)
    st = RefineStats()
    _check(lib().gpo_refine(g._h, lab, parts, ratio, max_rounds, stale_rounds, C.byref(st)))
    return lab, st


def uncoarsen(h: Hierarchy, coarsest_labels, parts: int, ratio: float = 1.03, max_rounds: int = 32,
              growth: float = 2.0):
    lab = np.ascontiguousarray(coarsest_labels, dtype=np.int32)
    out = np.empty(h.level(0).n, np.int32)
    cut = C.c_uint64()
    _check(lib().gpo_uncoarsen(h._h, lab, parts, ratio, max_rounds, growth, out, C.byref(cut)))
    return out, cut.value


def partition(g: Graph, parts: int, ratio: float = 1.03, seed: int = 0):
    labels = np.empty(g.n, np.int32)
    res = PartitionResult()
    _check(lib().gpo_partition(g._h, parts, ratio, seed, labels, C.byref(res)))
    return labels, res


def cut(g: Graph, labels) -> int:
    out = C.c_uint64()
    _check(lib().gpo_cut(g._h, np.ascontiguousarray(labels, dtype=np.int32), C.byref(out)))
    return out.value


def part_weights(g: Graph, labels, parts: int):
    out = np.empty(parts, np.uint64)
    _check(lib().gpo_part_weights(g._h, np.ascontiguousarray(labels, dtype=np.int32), parts, out))
    return out


def validate_csr(g: Graph, allow_self_loops: bool, strict: bool) -> None:
    _check(lib().gpo_validate_csr(g._h, int(allow_self_loops), int(strict)))


# ---- the reference's own host headers compiled into oracle/_ref ----

_ref = None


def reference_available() -> bool:
    return _REF_PATH.exists()


def ref_lib():
    global _ref
    if _ref is None:
        R = C.CDLL(str(_REF_PATH))
        R.ref_partition_capacity.argtypes = [C.c_uint64, C.c_int, C.c_double, C.POINTER(C.c_uint64)]
        R.ref_record_layout.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
        R.ref_record_encode.restype = C.c_uint64
        R.ref_record_encode.argtypes = [C.c_int, C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64]
        _ref = R
    return _ref


def ref_partition_capacity(total: int, parts: int, ratio: float):
    out = C.c_uint64()
    rc = ref_lib().ref_partition_capacity(total, parts, ratio, C.byref(out))
    return rc, out.value


def ref_record_layout(vertices: int, degree_per_weight: int, maximum_vertex_weight: int):
    vb, wb, mask = C.c_int(), C.c_int(), C.c_uint64()
    ref_lib().ref_record_layout(vertices, degree_per_weight, maximum_vertex_weight,
                                C.byref(vb), C.byref(wb), C.byref(mask))
    return vb.value, wb.value, mask.value
