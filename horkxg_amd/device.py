"""Kernel-family entry points of include/gpart_hip.h on host numpy arrays.

Each call uploads its inputs, runs one kernel family on the thread's context and
downloads the outputs. Used by the parity tests to compare single kernel
families (SCLP proposals of one round, contraction, primitives) with the oracle.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import WIDTHS, _check, _ctx, hip_lib


def _bind():
    H = hip_lib()
    vp, i64 = C.c_void_p, C.c_int64
    H.gpart_degree_classes_build.argtypes = [vp, vp, vp]
    H.gpart_sclp_proposals.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    H.gpart_contract.argtypes = [vp, vp, vp, vp, i64, C.c_uint64, vp, vp, vp, vp]
    H.gpart_sort_pairs_u64.argtypes = [vp, vp, vp, C.c_int, i64, C.c_int, C.c_int]
    H.gpart_exclusive_scan_u64.argtypes = [vp, vp, vp, i64]
    H.gpart_compact_flags.argtypes = [vp, vp, i64, vp, vp]
    H.gpart_contract_record_bounds.argtypes = [vp, i64, vp, vp, i64, vp, vp]
    H.gpart_cut.argtypes = [vp, vp, vp, vp]
    H.gpart_cut_noalloc.argtypes = [vp, vp, vp, vp]
    H.gpart_cut_noalloc_max_entries.restype = i64
    H.gpart_spectral_queue.argtypes = [vp, vp, C.c_int32, vp, vp, C.c_int32, C.c_int, vp]


_bind()


class DeviceArray:
    def __init__(self, host=None, dtype=None, count=None):
        if host is not None:
            host = np.ascontiguousarray(host, dtype=dtype)
            self.dtype, self.count = host.dtype, host.size
        else:
            self.dtype, self.count = np.dtype(dtype), int(count)
        self.nbytes = self.dtype.itemsize * self.count
        p = C.c_void_p()
        _check(hip_lib().gpart_device_alloc(_ctx(), max(self.nbytes, 8), C.byref(p)))
        self.ptr = p.value
        if host is not None and self.nbytes:
            _check(hip_lib().gpart_memcpy_htod(_ctx(), C.c_void_p(self.ptr), host.ctypes.data, self.nbytes))

    def get(self):
        out = np.empty(self.count, self.dtype)
        if self.nbytes:
            _check(hip_lib().gpart_memcpy_dtoh(_ctx(), out.ctypes.data, C.c_void_p(self.ptr), self.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                hip_lib().gpart_device_free(_ctx(), C.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None


class CsrView(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64), ("width", C.c_int32), ("offsets", C.c_void_p),
                ("neighbors", C.c_void_p), ("edge_weights", C.c_void_p), ("vertex_weights", C.c_void_p)]


class DegreeClasses(C.Structure):
    _fields_ = [("low", C.c_void_p), ("medium", C.c_void_p), ("high", C.c_void_p), ("low_count", C.c_int64),
                ("medium_count", C.c_int64), ("high_count", C.c_int64), ("high_edges", C.c_int64),
                ("max_degree", C.c_int64), ("isolated_count", C.c_int64)]


class SclpRound(C.Structure):
    _fields_ = [("role_salt", C.c_uint32), ("tie_salt", C.c_uint32), ("mover_threshold", C.c_uint32),
                ("filter_roles", C.c_int32), ("singleton", C.c_int32), ("value_bound", C.c_uint64)]


class DeviceGraph:
    """CSR arrays on the device in one width, with its degree classes."""

    def __init__(self, indptr, indices, edge_weights=None, vertex_weights=None, width: str = "standard"):
        off_t, w_t = (np.int32, np.uint32) if width == "standard" else (np.int64, np.uint64)
        n = len(indptr) - 1
        m = len(indices)
        ew = np.ones(m, np.uint64) if edge_weights is None else np.asarray(edge_weights, np.uint64)
        vw = np.ones(n, np.uint64) if vertex_weights is None else np.asarray(vertex_weights, np.uint64)
        self.off = DeviceArray(np.asarray(indptr), off_t)
        self.adj = DeviceArray(np.asarray(indices), np.int32)
        self.ew = DeviceArray(ew.astype(w_t), w_t)
        self.vw = DeviceArray(vw.astype(w_t), w_t)
        self.view = CsrView(n, m, WIDTHS[width], self.off.ptr, self.adj.ptr, self.ew.ptr, self.vw.ptr)
        self.n, self.m = n, m
        self.lists = [DeviceArray(dtype=np.int32, count=max(n, 1)) for _ in range(3)]
        self.classes = DegreeClasses(self.lists[0].ptr, self.lists[1].ptr, self.lists[2].ptr, 0, 0, 0, 0, 0, 0)
        _check(hip_lib().gpart_degree_classes_build(_ctx(), C.byref(self.view), C.byref(self.classes)))


def salts(seed: int, level: int, rnd: int):
    m = 0xFFFFFFFF
    role = (seed ^ ((level * 0x9e3779b9) & m) ^ ((rnd * 0x85ebca6b) & m) ^ 0x27d4eb2d) & m
    tie = (seed ^ ((level * 0xc2b2ae35) & m) ^ ((rnd * 0x165667b1) & m) ^ 0xd3a2646c) & m
    return role, tie


def sclp_proposals(g: DeviceGraph, clusters, seed: int, level: int, rnd: int, filter_roles: bool,
                   mover_threshold: int, singleton: bool = False, value_bound: int = 0):
    """value_bound: a bound on every row weight (0: unknown)."""
    role, tie = salts(seed, level, rnd)
    r = SclpRound(role, tie, mover_threshold, int(filter_roles), int(singleton), value_bound)
    cl = DeviceArray(clusters, np.int32)
    prop = DeviceArray(dtype=np.int32, count=g.n)
    gains = DeviceArray(dtype=np.uint64, count=g.n)
    ties = DeviceArray(dtype=np.uint32, count=g.n)
    _check(hip_lib().gpart_sclp_proposals(_ctx(), C.byref(g.view), C.c_void_p(cl.ptr), C.byref(g.classes),
                                          C.byref(r), C.c_void_p(prop.ptr), C.c_void_p(gains.ptr),
                                          C.c_void_p(ties.ptr)))
    return prop.get(), gains.get(), ties.get()


def contract(g: DeviceGraph, mapping, n_coarse: int, bound: int):
    """CSR contraction; returns (offsets int64, neighbors, edge weights) of the coarse graph."""
    mp = DeviceArray(mapping, np.int32)
    off, nbr, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
    m = C.c_int64()
    _check(hip_lib().gpart_contract(_ctx(), C.byref(g.view), C.byref(g.classes), C.c_void_p(mp.ptr), n_coarse,
                                    C.c_uint64(bound), C.byref(off), C.byref(nbr), C.byref(w), C.byref(m)))
    off_t, w_t = (np.int32, np.uint32) if g.view.width == 0 else (np.int64, np.uint64)
    out_off = np.empty(n_coarse + 1, off_t)
    out_nbr = np.empty(m.value, np.int32)
    out_w = np.empty(m.value, w_t)
    H = hip_lib()
    _check(H.gpart_memcpy_dtoh(_ctx(), out_off.ctypes.data, off, out_off.nbytes))
    if m.value:
        _check(H.gpart_memcpy_dtoh(_ctx(), out_nbr.ctypes.data, nbr, out_nbr.nbytes))
        _check(H.gpart_memcpy_dtoh(_ctx(), out_w.ctypes.data, w, out_w.nbytes))
    for p in (off, nbr, w):
        if p.value:
            H.gpart_device_free(_ctx(), p)
    return out_off.astype(np.int64), out_nbr, out_w.astype(np.uint64)


def contract_record_bounds(mapping, fine_bounds, n_coarse: int):
    """Saturating sums of fine row bounds per coarse vertex, and their maximum."""
    mp = DeviceArray(mapping, np.int32)
    fb = DeviceArray(fine_bounds, np.uint64)
    out = DeviceArray(dtype=np.uint64, count=max(n_coarse, 1))
    mx = C.c_uint64()
    _check(hip_lib().gpart_contract_record_bounds(_ctx(), mp.count, C.c_void_p(mp.ptr), C.c_void_p(fb.ptr), n_coarse,
                                                  C.c_void_p(out.ptr), C.byref(mx)))
    return out.get()[:n_coarse], mx.value


def sort_pairs(keys, values=None, begin_bit: int = 0, end_bit: int = 64):
    k = DeviceArray(keys, np.uint64)
    v = DeviceArray(values, values.dtype) if values is not None else None
    _check(hip_lib().gpart_sort_pairs_u64(_ctx(), C.c_void_p(k.ptr), C.c_void_p(v.ptr) if v else None,
                                          v.dtype.itemsize if v else 0, k.count, begin_bit, end_bit))
    return (k.get(), v.get()) if v is not None else k.get()


def exclusive_scan(values):
    a = DeviceArray(values, np.uint64)
    out = DeviceArray(dtype=np.uint64, count=a.count)
    _check(hip_lib().gpart_exclusive_scan_u64(_ctx(), C.c_void_p(a.ptr), C.c_void_p(out.ptr), a.count))
    return out.get()


def compact(flags):
    f = DeviceArray(flags, np.int32)
    out = DeviceArray(dtype=np.int32, count=max(f.count, 1))
    cnt = C.c_int64()
    _check(hip_lib().gpart_compact_flags(_ctx(), C.c_void_p(f.ptr), f.count, C.c_void_p(out.ptr), C.byref(cnt)))
    return out.get()[: cnt.value]


def cut(g: DeviceGraph, labels: DeviceArray, noalloc: bool = False) -> int:
    """Undirected cut of device labels: gpart_cut, or gpart_cut_noalloc (no
    allocation; scratch in a static device array)."""
    out = C.c_uint64()
    fn = hip_lib().gpart_cut_noalloc if noalloc else hip_lib().gpart_cut
    _check(fn(_ctx(), C.byref(g.view), C.c_void_p(labels.ptr), C.byref(out)))
    return out.value


def cut_noalloc_max_entries() -> int:
    return int(hip_lib().gpart_cut_noalloc_max_entries())


class BisectTask(C.Structure):
    _fields_ = [("begin", C.c_int32), ("count", C.c_int32), ("first_part", C.c_int32), ("parts", C.c_int32),
                ("weight", C.c_uint64)]


class QueueGraph(C.Structure):
    _fields_ = [("graph", CsrView), ("labels", C.c_void_p), ("order", C.c_void_p), ("vertex_weights", C.c_void_p),
                ("coords", C.c_void_p), ("capacity", C.c_uint64), ("seed", C.c_uint32), ("status", C.c_int32)]


class QueueReport(C.Structure):
    _fields_ = [("tasks", C.c_int64), ("slots", C.c_int64), ("entries", C.c_int64), ("vertices", C.c_int64),
                ("workgroups", C.c_int32), ("pad", C.c_int32)]


def spectral_queue(graphs, parts, capacities, seeds, precision: str = "fp64", vertex_weights=None):
    """gpart_spectral_queue: the recursive bisection of every graph (DeviceGraph)
    into parts[i] parts in one persistent launch, each graph from one root task
    over all its vertices. vertex_weights[i]: host weights (unit if None).
    Returns ([(labels, order, coords, status)] per graph, report dict)."""
    count = len(graphs)
    keep, qgs = [], (QueueGraph * max(count, 1))()
    roots = (BisectTask * max(count, 1))()
    root_graph = (C.c_int32 * max(count, 1))()
    for i, g in enumerate(graphs):
        w = np.ones(g.n, np.uint64) if vertex_weights is None or vertex_weights[i] is None \
            else np.asarray(vertex_weights[i], np.uint64)
        labels = DeviceArray(np.zeros(g.n, np.int32), np.int32)
        order = DeviceArray(np.arange(g.n, dtype=np.int32), np.int32)
        vw = DeviceArray(w, np.uint64)
        coords = DeviceArray(dtype=np.float64, count=g.n)
        keep.append((labels, order, vw, coords))
        qgs[i] = QueueGraph(g.view, labels.ptr, order.ptr, vw.ptr, coords.ptr, int(capacities[i]),
                            int(seeds[i]) & 0xFFFFFFFF, 0)
        roots[i] = BisectTask(0, g.n, 0, int(parts[i]), int(w.sum()))
        root_graph[i] = i
    rep = QueueReport()
    prec = {"fp64": 0, "fp32": 1}[precision]
    _check(hip_lib().gpart_spectral_queue(_ctx(), qgs, count, roots, root_graph, count, prec, C.byref(rep)))
    out = [(labels.get(), order.get(), coords.get(), int(qgs[i].status))
           for i, (labels, order, vw, coords) in enumerate(keep)]
    return out, {"tasks": rep.tasks, "slots": rep.slots, "entries": rep.entries, "vertices": rep.vertices,
                 "workgroups": rep.workgroups}
