"""CPU: the oracle pinned to the reference's own known-answer tests and closed forms.

Every check below comes from a reference test (cited) or from what the mathematics
fixes (closed forms, invariants). Where the reference's own host code compiles from
its headers (oracle/_ref), the oracle and the product's host code are compared with
it directly. No GPU is needed.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import kat_graphs as K


def ograph(oracle, g):
    return oracle.Graph(g["indptr"], g["indices"].astype(np.int32), g["edge_weights"], g["vertex_weights"])


def expect_error(oracle, code, fn, *args, **kw):
    with pytest.raises(oracle.OracleError) as info:
        fn(*args, **kw)
    assert info.value.code == code, info.value


# ---------------------------------------------------------------- capacity

CAPACITY_KATS = [((6, 2, 1.10), 3), ((100, 4, 1.03), 25), ((1000, 4, 1.03), 257), ((6, 2, 1.34), 4)]
CAPACITY_REJECTED = [(4, 3, 1.03), (100, 0, 1.03), (100, 33, 1.03), (100, 4, 0.99), (100, 4, math.nan)]


def test_capacity_kats(oracle, hx):
    """tests/test_partition_capacity.cpp:13-29."""
    for args, want in CAPACITY_KATS:
        assert oracle.partition_capacity(*args) == want
        assert hx.partition_capacity(*args) == want
    for args in CAPACITY_REJECTED:
        expect_error(oracle, 1, oracle.partition_capacity, *args)
        with pytest.raises(hx.GpartError) as info:
            hx.partition_capacity(*args)
        assert info.value.code == 1
    expect_error(oracle, 2, oracle.partition_capacity, 2**64 - 1, 2, 4.0)
    with pytest.raises(hx.GpartError) as info:
        hx.partition_capacity(2**64 - 1, 2, 4.0)
    assert info.value.code == 2


def test_capacity_matches_reference_header(oracle, hx):
    """partition_capacity of include/partition_capacity.hpp, compiled into oracle/_ref,
    against the oracle and the host library on random and edge inputs."""
    if not oracle.reference_available():
        pytest.skip("oracle/_ref not built (the reference tree is absent)")
    rng = np.random.default_rng(5)
    cases = [(int(t), int(p), float(r)) for t, p, r in zip(
        rng.integers(0, 2**40, 3000), rng.integers(0, 36, 3000), rng.uniform(0.9, 3.0, 3000))]
    cases += [(2**64 - 1, 2, 1.0), (2**63, 2, 1.5), (1, 1, 1.0), (0, 1, 1.0), (32, 32, 1.0), (31, 32, 1.03)]
    for total, parts, ratio in cases:
        rc, want = oracle.ref_partition_capacity(total, parts, ratio)
        try:
            got = oracle.partition_capacity(total, parts, ratio)
            got_rc = 0
        except oracle.OracleError as e:
            got, got_rc = None, e.code
        assert got_rc == rc, (total, parts, ratio)
        if rc == 0:
            assert got == want
        try:
            host = hx.partition_capacity(total, parts, ratio)
            host_rc = 0
        except hx.GpartError as e:
            host, host_rc = None, e.code
        assert host_rc == rc and (rc != 0 or host == want), (total, parts, ratio)


# ---------------------------------------------------------------- record layout

RECORD_VERTICES = [1, 2, 3, 4, 7, 8, 9, 1023, 1024, 4095, 4096, 2449029, 1295331, 2147483647]


def test_record_layout_matches_reference_header(oracle, hx):
    """SclpRecordLayout::checked (include/detail/sclp_record.hpp) and the properties
    tests/test_sclp_record.cpp:24-53 requires, against the product's layout choice
    (whose bound is degree_per_weight * maximum_vertex_weight)."""
    for n in RECORD_VERTICES:
        vb, wb, mask = hx.record_layout(n, 1)
        assert wb > 0, n  # small weights fit
        assert hx.record_layout(n, mask)[1] > 0  # exact bound accepted
        assert hx.record_layout(n, mask + 1)[1] == 0  # overflow refused
        assert hx.record_layout(n, (mask // 2 + 1) * 2)[1] == 0  # product overflow refused
        assert hx.record_layout(n, 2**64 - 1)[1] == 0  # saturated density refused
        assert vb + vb + wb == 64 and mask == (1 << wb) - 1
        if oracle.reference_available():
            for degree, vweight in [(1, 1), (mask, 1), (mask + 1, 1), (mask // 2 + 1, 2), (3, 0), (7, 5)]:
                want = oracle.ref_record_layout(n, degree, vweight)
                assert hx.record_layout(n, degree * vweight) == want, (n, degree, vweight)
            assert oracle.ref_record_layout(n, 2**64 - 1, 0)[1] == 0
    assert hx.record_layout(0, 0)[1] == 0  # empty id range
    assert hx.record_layout(2147483648, 1)[1] == 0  # id width overflow


# ---------------------------------------------------------------- cut, contraction, projection

@pytest.mark.parametrize("weight", [7, 1 << 40])
def test_cut_closed_form(oracle, weight):
    """tests/test_uncoarsen.cu:121-160: hub with repeated entries, isolated tail and
    (for 2^40) cut totals above 32 bits; refinement with zero rounds keeps labels."""
    g = K.hub259(weight)
    og = ograph(oracle, g)
    for interleaved in (False, True):
        labels = np.array([v % 4 if interleaved else v // 65 for v in range(259)], np.int32)
        expected = K.independent_cut(g, labels)
        assert oracle.cut(og, labels) == expected
        out, st = oracle.refine(og, labels, 4, 1.10, max_rounds=0)
        assert np.array_equal(out, labels)
        assert st.initial_cut == expected and st.final_cut == expected


@pytest.mark.parametrize("groups", [7, 1, 7])
@pytest.mark.parametrize("weight", [3, 1 << 40])
def test_contraction_closed_form(oracle, groups, weight):
    """tests/test_uncoarsen.cu:162-212: group cross entries by (map[v], map[u]), sum, sort."""
    g, mapping, cw, (e_ptr, e_idx, e_w) = K.contraction_case(groups, weight)
    out = oracle.contract(ograph(oracle, g), mapping, groups, cw)
    off, adj, ew, vw = out.arrays()
    assert np.array_equal(off, e_ptr)
    assert np.array_equal(adj, e_idx)
    assert np.array_equal(ew, e_w)
    assert np.array_equal(vw, cw)


@pytest.mark.parametrize("empty", [False, True])
def test_identity_contraction(oracle, empty):
    """tests/test_uncoarsen.cu:213-231: the identity map keeps singleton and empty rows."""
    fine, _, _ = K.path4()
    if empty:
        fine = dict(fine, indptr=np.zeros(5, np.int64), indices=np.zeros(0, np.int64),
                    edge_weights=np.zeros(0, np.uint64))
    out = oracle.contract(ograph(oracle, fine), np.arange(4, dtype=np.int32), 4, fine["vertex_weights"])
    off, adj, ew, vw = out.arrays()
    assert np.array_equal(off, fine["indptr"]) and np.array_equal(adj, fine["indices"])
    assert np.array_equal(ew, fine["edge_weights"]) and np.array_equal(vw, fine["vertex_weights"])


def test_projection_and_small_uncoarsen(oracle):
    """tests/test_uncoarsen.cu:341-400: projection keeps cut 5 and loads {2,2};
    an input over the ceil capacity is rejected; refinement of both levels keeps the
    cut at most 5 and every load at most 2."""
    fine, coarse, mapping = K.path4()
    coarse_labels = np.array([0, 1], np.int32)
    projected = coarse_labels[mapping]
    assert list(projected) == [0, 0, 1, 1]
    assert oracle.cut(ograph(oracle, coarse), coarse_labels) == 5
    assert oracle.cut(ograph(oracle, fine), projected) == 5
    assert list(oracle.part_weights(ograph(oracle, fine), projected, 2)) == [2, 2]
    expect_error(oracle, 1, oracle.refine, ograph(oracle, fine), np.array([0, 0, 0, 1], np.int32), 2, 1.10, 4)
    c_out, _ = oracle.refine(ograph(oracle, coarse), coarse_labels, 2, 1.10, 4)
    f_out, st = oracle.refine(ograph(oracle, fine), c_out[mapping], 2, 1.10, 4)
    assert st.final_cut <= 5 and K.independent_cut(fine, f_out) == st.final_cut
    assert max(K.loads(fine, f_out, 2)) <= 2


# ---------------------------------------------------------------- refinement KATs

def test_exchange_kat(oracle):
    """tests/test_uncoarsen.cu:454-471: cut 2 with three vertices in part 0."""
    g, labels = K.exchange()
    out, _ = oracle.refine(ograph(oracle, g), labels, 2, 1.03, 32)
    assert K.independent_cut(g, out) == 2
    assert int((out == 0).sum()) == 3


def test_repair_kat(oracle):
    """tests/test_uncoarsen.cu:472-484: cut 1, four vertices in part 0, rebalance used."""
    g, labels = K.repair()
    out, st = oracle.refine(ograph(oracle, g), labels, 2, 1.03, 32)
    assert K.independent_cut(g, out) == 1
    assert int((out == 0).sum()) == 4
    assert st.admission_iterations > 0


@pytest.mark.parametrize("k", [2, 4, 8, 32])
@pytest.mark.parametrize("scale", [1, 1 << 35])
def test_stress_kat(oracle, k, scale):
    """tests/test_uncoarsen.cu:485-518: cut never grows and loads stay under capacity."""
    g, labels = K.stress(k, scale)
    before = K.independent_cut(g, labels)
    out, st = oracle.refine(ograph(oracle, g), labels, k, 1.03, 32)
    assert K.independent_cut(g, out) <= before
    assert st.final_cut == K.independent_cut(g, out)
    cap = oracle.partition_capacity(int(g["vertex_weights"].sum()), k, 1.03)
    assert max(K.loads(g, out, k)) <= cap


def test_gain_overflow_rejected(oracle):
    """tests/test_uncoarsen.cu:287-310: edge weights INT64_MAX/2 exceed the signed gain range."""
    w = (2**63 - 1) // 2
    g = K.csr_from_rows([[(1, w)], [(0, w)]], [1, 1])
    expect_error(oracle, 2, oracle.refine, ograph(oracle, g), np.array([0, 1], np.int32), 2)


# ---------------------------------------------------------------- initial partition KATs

def _initial_ok(oracle, g, k, zero_cut):
    """tests/test_initial_partition_gpu.cu:13-33: repeatable, valid, under capacity."""
    og = ograph(oracle, g)
    a = oracle.initial_partition(og, k, 1.03, 0)
    b = oracle.initial_partition(og, k, 1.03, 0)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() < k
    cap = oracle.partition_capacity(int(g["vertex_weights"].astype(object).sum()), k, 1.03)
    assert max(K.loads(g, a, k)) <= cap
    if zero_cut:
        assert K.independent_cut(g, a) == 0
    return a


def test_initial_kats(oracle):
    """tests/test_initial_partition_gpu.cu:34-99."""
    _initial_ok(oracle, K.two_cliques(), 2, True)
    _initial_ok(oracle, K.two_cliques(1 << 40, 1 << 40), 2, True)
    _initial_ok(oracle, K.isolated(12), 3, True)
    hub = K.initial_hub259()
    for k in (2, 3, 4, 8):
        _initial_ok(oracle, hub, k, False)
    _initial_ok(oracle, K.isolated(256), 32, True)
    heavy = K.isolated(256)
    heavy["vertex_weights"][0] = 1000
    expect_error(oracle, 1, oracle.initial_partition, ograph(oracle, heavy), 3, 1.10)
    expect_error(oracle, 1, oracle.initial_partition, ograph(oracle, K.isolated(4)), 3, 1.03)
    expect_error(oracle, 3, oracle.initial_partition, ograph(oracle, K.isolated(3, 2)), 2, 1.03)
    big = K.csr_from_rows([[(1, 2**64 - 1)], [(0, 2**64 - 1)]], [1, 1])
    expect_error(oracle, 2, oracle.initial_partition, ograph(oracle, big), 2, 1.03)


# ---------------------------------------------------------------- coarsening invariants

def test_isolated_packing_kat(oracle):
    """tests/test_uncoarsen.cu:53-76: packing contracts the isolated vertices, keeps
    every cluster under capacity, never joins an isolated vertex to the connected
    pair, and the heavy pair stays split."""
    og = ograph(oracle, K.isolated_packing())
    _, _, before, _, _ = oracle.aggregate(og, 2, 0, 0, cluster_size_factor=2, pack_isolated=False)
    mp, cw, after, cap, _ = oracle.aggregate(og, 2, 0, 0, cluster_size_factor=2, pack_isolated=True)
    assert after < before
    assert cw.max() <= cap
    assert mp[0] != mp[1]
    assert all(mp[v] not in (mp[0], mp[1]) for v in range(2, 12))


def check_hierarchy(oracle, h, parts, factor, seed=17):
    """Invariants of tests/test_sclp_small.py:69-111 on every level: valid symmetric
    CSR, positive weights, no self loops above level 0, onto maps, summed vertex
    weights, cluster capacity, and the projected cut of random labels."""
    rng = np.random.default_rng(seed)
    levels = [h.level(i).arrays() for i in range(len(h))]
    total = int(levels[0][3].sum())
    for li, (off, adj, ew, vw) in enumerate(levels):
        n = len(off) - 1
        assert off[0] == 0 and off[-1] == len(adj) and np.all(np.diff(off) >= 0)
        assert len(adj) == 0 or (adj.min() >= 0 and adj.max() < n)
        assert np.all(ew > 0) and np.all(vw > 0)
        src = np.repeat(np.arange(n), np.diff(off))
        pairs = {}
        for a, b, w in zip(src.tolist(), adj.tolist(), ew.tolist()):
            pairs[(a, b)] = w
            if li > 0:
                assert a != b
        for (a, b), w in pairs.items():
            assert pairs.get((b, a)) == w
        if li > 0:
            mp = h.map(li - 1)
            fo, fa, fw, fv = levels[li - 1]
            assert len(mp) == len(fo) - 1 and mp.min() >= 0 and mp.max() < n
            assert len(np.unique(mp)) == n
            assert np.array_equal(np.bincount(mp, weights=fv.astype(np.float64), minlength=n).astype(np.uint64), vw)
            labels = rng.integers(0, 7, n).astype(np.int32)
            fine = dict(indptr=fo, indices=fa.astype(np.int64), edge_weights=fw)
            coarse = dict(indptr=off, indices=adj.astype(np.int64), edge_weights=ew)
            assert K.independent_cut(fine, labels[mp]) == K.independent_cut(coarse, labels)
            cap = max(int(fv.max()), -(-total // (factor * parts)))
            assert int(vw.max()) <= cap
    return levels


def test_sclp_small_hierarchy(oracle):
    """tests/test_sclp_small.py: gpart_coarsen defaults (factor 256, no packing),
    parts 4, seed 0, stop ratio 1.0, 24 levels; level 0 contracts to 45-60%."""
    g = K.sclp_small()
    h = oracle.coarsen(ograph(oracle, g), 4, 0, 1.0, 24)
    levels = check_hierarchy(oracle, h, 4, 256)
    assert len(levels) >= 2
    ratio = (len(levels[1][0]) - 1) / (len(levels[0][0]) - 1)
    assert 0.45 <= ratio <= 0.60


@pytest.mark.parametrize("n", [8, 4096])
def test_pairs_hierarchy(oracle, n):
    """tests/test_uncoarsen.cu:405-450: the no-contraction case (n = 8 stops at once)
    and a perfect matching, whose first-level clusters can only be a pair {v, v^1}
    or a singleton (no isolated packing, so no cluster spans two components)."""
    h = oracle.coarsen(ograph(oracle, K.pairs(n)), 4)
    levels = check_hierarchy(oracle, h, 4, 256)
    if n == 8:
        assert len(levels) == 1
        return
    assert len(levels) >= 2
    mp = h.map(0)
    n1 = len(levels[1][0]) - 1
    assert n // 2 <= n1 <= 0.90 * n
    members = [[] for _ in range(n1)]
    for v, c in enumerate(mp.tolist()):
        members[c].append(v)
    for group in members:
        assert len(group) == 1 or (len(group) == 2 and group[0] ^ 1 == group[1])


def test_hierarchy_file_format(oracle, hx, tmp_path):
    """The int32 exporter format of include/hierarchy.hpp:42-126, parsed independently
    as tests/test_sclp_small.py:69-84 does, equals the oracle's levels; the product's
    reader (host code) reads the same file back."""
    h = oracle.coarsen(ograph(oracle, K.sclp_small()), 4, 0, 1.0, 24)
    path = tmp_path / "h.bin"
    h.write(path)
    raw = np.fromfile(path, dtype="<i4")
    pos = 1
    parsed = []
    for li in range(int(raw[0])):
        n, m = int(raw[pos]), int(raw[pos + 1])
        pos += 2
        rows = raw[pos:pos + n + 1]; pos += n + 1
        cols = raw[pos:pos + m]; pos += m
        ew = raw[pos:pos + m]; pos += m
        vw = raw[pos:pos + n]; pos += n
        mp = None
        if li:
            fn = len(parsed[-1][0]) - 1
            mp = raw[pos:pos + fn]; pos += fn
        parsed.append((rows, cols, ew, vw, mp))
    assert pos == len(raw)
    assert len(parsed) == len(h)
    for li, (rows, cols, ew, vw, mp) in enumerate(parsed):
        off, adj, w, v = h.level(li).arrays()
        assert np.array_equal(rows, off) and np.array_equal(cols, adj)
        assert np.array_equal(ew, w) and np.array_equal(vw, v)
        if li:
            assert np.array_equal(mp, h.map(li - 1))
    back = hx.Hierarchy.read(path)
    assert len(back) == len(h)
    for li in range(len(h)):
        off, adj, w, v = back.level(li)
        o2, a2, w2, v2 = h.level(li).arrays()
        assert np.array_equal(off, o2) and np.array_equal(adj, a2)
        assert np.array_equal(w, w2) and np.array_equal(v, v2)
        if li:
            assert np.array_equal(back.map(li - 1), h.map(li - 1))


# ---------------------------------------------------------------- spectral recurrence

def lowbias32(x):
    x = np.asarray(x, np.uint64) & 0xffffffff
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xffffffff
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xffffffff
    return x ^ (x >> 16)


def diffusion_numpy(g, seed=0):
    """src/initial_partition_gpu.cu:81-142 for the first task (all vertices, part 0),
    in numpy's own summation order: x0 = mix32(v ^ seed)/2^32 - 0.5, then 64 lazy
    steps x <- (x + W x / d) / 2, each followed (and the start preceded) by
    normalisation to zero weighted mean and unit weighted variance, mass d (1 if d = 0)."""
    indptr, indices, w = g["indptr"], g["indices"], g["edge_weights"].astype(np.float64)
    n = len(indptr) - 1
    src = np.repeat(np.arange(n), np.diff(indptr))
    d = np.bincount(src, weights=w, minlength=n)
    mass = np.where(d > 0, d, 1.0)

    def normalise(a):
        mean = (mass * a).sum() / mass.sum()
        var = max(0.0, (mass * a * a).sum() / mass.sum() - mean * mean)
        scale = 1.0 / np.sqrt(var) if var > 1e-30 else 1.0
        return (a - mean) * scale

    x = lowbias32(np.arange(n, dtype=np.uint64) ^ seed).astype(np.float64) / 4294967296.0 - 0.5
    x = normalise(x)
    for _ in range(64):
        wx = np.bincount(src, weights=w * x[indices], minlength=n)
        x = np.where(d > 0, 0.5 * (x + wx / np.where(d > 0, d, 1.0)), x)
        x = normalise(x)
    return x


def split_labels(order_x, vw, capacity):
    """Split of src/initial_partition_gpu.cu:230-248 for two parts: stable sort by
    (x, id), smallest prefix reaching half the weight, clamped to the capacity."""
    ids = np.lexsort((np.arange(len(order_x)), order_x))
    prefix = np.cumsum(vw[ids].astype(np.uint64))
    total = int(prefix[-1])
    target = total // 2
    limit = total if capacity > total else capacity
    min_left = total - limit if total > limit else 0
    count = len(ids)
    boundary = int(np.searchsorted(prefix, target, side="left")) + 1
    boundary = max(1, min(count - 1, boundary))
    if int(prefix[boundary - 1]) > limit and boundary > 1:
        boundary -= 1
    if int(prefix[boundary - 1]) < min_left and boundary < count - 1:
        boundary += 1
    labels = np.empty(count, np.int32)
    labels[ids[:boundary]] = 0
    labels[ids[boundary:]] = 1
    return labels


@pytest.mark.parametrize("name", ["grid32", "two_cliques", "hub259"])
def test_spectral_recurrence(oracle, hx, name):
    """The oracle's FP64 diffusion vector agrees with an independent numpy evaluation
    of the recurrence to 1e-9 (summation orders differ, so bits need not), and its
    two-part labels follow the split rule applied to its own vector."""
    if name == "grid32":
        ip, ix = hx.generate_grid2d(32, 32)
        g = dict(indptr=ip, indices=ix, edge_weights=np.ones(len(ix), np.uint64),
                 vertex_weights=np.ones(len(ip) - 1, np.uint64))
    elif name == "two_cliques":
        g = K.two_cliques()
    else:
        g = K.initial_hub259()
    labels, coords, tasks = oracle.initial_partition(ograph(oracle, g), 2, 1.03, 0, with_coords=True)
    assert tasks == 1
    want = diffusion_numpy(g)
    assert np.max(np.abs(coords - want)) < 1e-9
    cap = oracle.partition_capacity(int(g["vertex_weights"].sum()), 2, 1.03)
    assert np.array_equal(labels, split_labels(coords, g["vertex_weights"], cap))
