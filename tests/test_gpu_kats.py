"""GPU: the reference's own known-answer tests through the HIP path, both widths.

The same graphs as tests/kat_graphs.py (each builder cites its reference test);
expected values are the reference's, or closed forms computed on the host here.
"""
from __future__ import annotations

import numpy as np
import pytest

import kat_graphs as K

pytestmark = pytest.mark.gpu

WIDTHS = ["standard", "big"]


def say(*args):
    print(*args, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")


def upload(hx, g, width="standard"):
    return hx.Graph(g["indptr"], g["indices"], width=width, edge_weights=g["edge_weights"],
                    vertex_weights=g["vertex_weights"])


def expect_code(hx, code, fn, *args, **kw):
    with pytest.raises(hx.GpartError) as info:
        fn(*args, **kw)
    assert info.value.code == code, info.value


@pytest.mark.parametrize("width", WIDTHS)
def test_cut_closed_form(hx, width):
    """tests/test_uncoarsen.cu:121-160 (2^40 weights in the big width)."""
    say(f"case kat cut width={width}")
    weight = 7 if width == "standard" else 1 << 40
    g = K.hub259(weight)
    dg = upload(hx, g, width)
    for interleaved in (False, True):
        labels = np.array([v % 4 if interleaved else v // 65 for v in range(259)], np.int32)
        expected = K.independent_cut(g, labels)
        assert dg.cut(labels) == expected
        out, rep = dg.refine(labels, 4, 1.10, max_rounds=0, verify=True)
        assert np.array_equal(out, labels)
        assert rep["initial_cut"] == expected and rep["final_cut"] == expected


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("groups", [7, 1])
def test_contraction_closed_form(hx, width, groups):
    """tests/test_uncoarsen.cu:162-212, packed and unpacked record layouts."""
    say(f"case kat contraction width={width} groups={groups}")
    weight = 3 if width == "standard" else 1 << 40
    g, mapping, cw, (e_ptr, e_idx, e_w) = K.contraction_case(groups, weight)
    dg = hx.device.DeviceGraph(g["indptr"], g["indices"], g["edge_weights"], g["vertex_weights"], width=width)
    total = int(g["edge_weights"].astype(object).sum())
    for bound in (total, 2**64 - 1):
        off, nbr, w = hx.device.contract(dg, mapping, groups, bound)
        assert np.array_equal(off.astype(np.int64), e_ptr)
        assert np.array_equal(nbr.astype(np.int64), e_idx)
        assert np.array_equal(w.astype(np.uint64), e_w)


@pytest.mark.parametrize("width", WIDTHS)
def test_identity_contraction(hx, width):
    """tests/test_uncoarsen.cu:213-231: identity map, with and without entries."""
    say(f"case kat identity contraction width={width}")
    fine, _, _ = K.path4()
    for empty in (False, True):
        f = fine if not empty else dict(fine, indptr=np.zeros(5, np.int64), indices=np.zeros(0, np.int64),
                                        edge_weights=np.zeros(0, np.uint64))
        dg = hx.device.DeviceGraph(f["indptr"], f["indices"], f["edge_weights"], f["vertex_weights"], width=width)
        off, nbr, w = hx.device.contract(dg, np.arange(4, dtype=np.int32), 4, 2**64 - 1)
        assert np.array_equal(off.astype(np.int64), f["indptr"])
        assert np.array_equal(nbr.astype(np.int64), f["indices"])
        assert np.array_equal(w.astype(np.uint64), f["edge_weights"])


@pytest.mark.parametrize("width", WIDTHS)
def test_isolated_packing(hx, width):
    """tests/test_uncoarsen.cu:53-76 through the GPU aggregate."""
    say(f"case kat isolated packing width={width}")
    dg = upload(hx, K.isolated_packing(), width)
    _, _, before, _, _ = dg.aggregate(2, 0, 0, cluster_size_factor=2, pack_isolated=False)
    mp, cw, after, cap, mx = dg.aggregate(2, 0, 0, cluster_size_factor=2, pack_isolated=True)
    assert after < before
    assert mx <= cap and cw.max() <= cap
    assert mp[0] != mp[1]
    assert all(mp[v] not in (mp[0], mp[1]) for v in range(2, 12))


@pytest.mark.parametrize("width", WIDTHS)
def test_overloaded_input_rejected(hx, width):
    """tests/test_uncoarsen.cu:362-368: loads {3,1} exceed floor(1.10*4/2) = 2."""
    say(f"case kat overloaded width={width}")
    fine, _, _ = K.path4()
    expect_code(hx, 1, upload(hx, fine, width).refine, np.array([0, 0, 0, 1], np.int32), 2, 1.10, 4)


@pytest.mark.parametrize("width", WIDTHS)
def test_exchange_and_repair(hx, width):
    """tests/test_uncoarsen.cu:454-484."""
    say(f"case kat exchange/repair width={width}")
    g, labels = K.exchange()
    out, _ = upload(hx, g, width).refine(labels, 2, 1.03, 32, verify=True)
    assert K.independent_cut(g, out) == 2 and int((out == 0).sum()) == 3
    g, labels = K.repair()
    out, rep = upload(hx, g, width).refine(labels, 2, 1.03, 32, verify=True)
    assert K.independent_cut(g, out) == 1 and int((out == 0).sum()) == 4
    assert rep["admission_iterations"] > 0


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("k", [2, 4, 8, 32])
def test_stress(hx, oracle, width, k):
    """tests/test_uncoarsen.cu:485-518 (scale 2^35 in the big width), strict
    verification on every round, labels equal to the oracle's."""
    say(f"case kat stress width={width} k={k}")
    g, labels = K.stress(k, 1 if width == "standard" else 1 << 35)
    before = K.independent_cut(g, labels)
    out, rep = upload(hx, g, width).refine(labels, k, 1.03, 32, verify=True)
    assert K.independent_cut(g, out) <= before and rep["final_cut"] == K.independent_cut(g, out)
    cap = hx.partition_capacity(int(g["vertex_weights"].sum()), k, 1.03)
    assert max(K.loads(g, out, k)) <= cap
    og = oracle.Graph(g["indptr"], g["indices"].astype(np.int32), g["edge_weights"], g["vertex_weights"])
    ol, _ = oracle.refine(og, labels, k, 1.03, 32)
    assert np.array_equal(out, ol)


def test_gain_overflow_rejected(hx):
    """tests/test_uncoarsen.cu:287-310 (big width)."""
    say("case kat gain overflow")
    w = (2**63 - 1) // 2
    g = K.csr_from_rows([[(1, w)], [(0, w)]], [1, 1])
    expect_code(hx, 2, upload(hx, g, "big").refine, np.array([0, 1], np.int32), 2)


@pytest.mark.parametrize("width", WIDTHS)
@pytest.mark.parametrize("mode", ["plain_fp64", "batched_fp64", "batched_fp32"])
def test_initial_kats(hx, width, mode):
    """tests/test_initial_partition_gpu.cu:13-99 for every bisection form."""
    say(f"case kat initial width={width} mode={mode}")

    def ok(g, k, zero_cut):
        dg = upload(hx, g, width)
        a = dg.initial_partition(k, 1.03, 0, mode)
        b = dg.initial_partition(k, 1.03, 0, mode)
        assert np.array_equal(a, b)
        assert a.min() >= 0 and a.max() < k
        cap = hx.partition_capacity(int(g["vertex_weights"].astype(object).sum()), k, 1.03)
        assert max(K.loads(g, a, k)) <= cap
        if zero_cut:
            assert K.independent_cut(g, a) == 0

    ok(K.two_cliques(), 2, True)
    if width == "big":
        ok(K.two_cliques(1 << 40, 1 << 40), 2, True)
    ok(K.isolated(12), 3, True)
    for k in (2, 3, 4, 8):
        ok(K.initial_hub259(), k, False)
    ok(K.isolated(256), 32, True)
    heavy = K.isolated(256)
    heavy["vertex_weights"][0] = 1000
    expect_code(hx, 1, upload(hx, heavy, width).initial_partition, 3, 1.10, 0, mode)
    expect_code(hx, 1, upload(hx, K.isolated(4), width).initial_partition, 3, 1.03, 0, mode)
    expect_code(hx, 3, upload(hx, K.isolated(3, 2), width).initial_partition, 2, 1.03, 0, mode)
    if width == "big":
        big = K.csr_from_rows([[(1, 2**64 - 1)], [(0, 2**64 - 1)]], [1, 1])
        expect_code(hx, 2, upload(hx, big, "big").initial_partition, 2, 1.03, 0, mode)


def test_sclp_small_hierarchy(hx, oracle):
    """tests/test_sclp_small.py through the GPU coarsening: equal to the oracle's
    hierarchy, whose invariants the CPU suite checks."""
    say("case kat sclp small hierarchy")
    g = K.sclp_small()
    gh = upload(hx, g).coarsen(4, 0, 1.0, 24, verify=True)
    og = oracle.Graph(g["indptr"], g["indices"].astype(np.int32), g["edge_weights"], g["vertex_weights"])
    oh = oracle.coarsen(og, 4, 0, 1.0, 24)
    assert len(gh) == len(oh)
    for i in range(len(gh)):
        for x, y in zip(gh.level(i), oh.level(i).arrays()):
            assert np.array_equal(x, y), i
        if i + 1 < len(gh):
            assert np.array_equal(gh.map(i), oh.map(i))


@pytest.mark.parametrize("verify", [False, True])
def test_edgeless_coarse_level(hx, oracle, verify):
    """A perfect matching contracts to levels without any cross entries (every
    edge internal); the hierarchy still equals the oracle's (regression: the
    empty coarse CSR once released its blocks twice)."""
    say(f"case kat edgeless coarse level verify={verify}")
    g = K.pairs(4096)
    gh = upload(hx, g).coarsen(4, 0, 1.0, 24, verify=verify)
    og = oracle.Graph(g["indptr"], g["indices"].astype(np.int32), g["edge_weights"], g["vertex_weights"])
    oh = oracle.coarsen(og, 4, 0, 1.0, 24)
    assert len(gh) == len(oh) and any(len(gh.level(i)[1]) == 0 for i in range(len(gh)))
    for i in range(len(gh)):
        for x, y in zip(gh.level(i), oh.level(i).arrays()):
            assert np.array_equal(x, y), i
        if i + 1 < len(gh):
            assert np.array_equal(gh.map(i), oh.map(i))
