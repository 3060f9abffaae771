"""GPU parity: every kernel family and the full pipeline against the CPU oracle.

Integer work (SCLP proposals, hierarchies, contraction, refinement labels) must
be bit-exact. The FP64 diffusion follows the same summation order as the
oracle and is compared bit-for-bit as well.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*args):
    print(*args, flush=True)


@pytest.fixture(scope="module", autouse=True)
def need_gpu(hx):
    if hx.device_count() < 1:
        pytest.skip("no GPU")


def hub_graph(n=6000, hub_degrees=(3000, 600), seed=1):
    """Ring + hubs: vertex i gets hub_degrees[i] random neighbours. Degrees above
    1024 take the grouped LDS SCLP path (in singleton rounds, chunks), above 4096 several items (and
    chunked cache rows in refinement)."""
    rng = np.random.default_rng(seed)
    rows = [set() for _ in range(n)]

    def add(a, b):
        if a != b:
            rows[a].add(b)
            rows[b].add(a)

    for v in range(n):
        add(v, (v + 1) % n)
    for h, degree in enumerate(hub_degrees):
        for u in rng.choice(np.arange(h + 1, n), degree, replace=False):
            add(h, int(u))
    indptr = [0]
    indices = []
    for r in rows:
        indices.extend(sorted(r))
        indptr.append(len(indices))
    return np.array(indptr, np.int64), np.array(indices, np.int64)


GRAPHS = {
    "grid64": lambda hx: hx.generate_grid2d(64, 64),
    "grid256": lambda hx: hx.generate_grid2d(256, 256),
    "rmat12": lambda hx: hx.generate_rmat(12, 16),
    # 3.6M cross entries after one aggregation: more than 1024 run-scan tiles
    "rmat17": lambda hx: hx.generate_rmat(17, 16),
    "hub": lambda hx: hub_graph(),
    "bighub": lambda hx: hub_graph(60000, (45000, 20000, 9000, 5000, 1500, 700), seed=3),
    # rows above 524288 entries: the per-hub global tables outside singleton rounds;
    # split rows of 16 to 128 groups
    "hugehub": lambda hx: hub_graph(700000, (600000, 400000, 300000, 100000, 70000), seed=5),
}


# sizes on both sides of the one-workgroup sort (at most 8192 keys) and of a tile
@pytest.mark.parametrize("n", [1, 2, 7, 2048, 2049, 5000, 8192, 8193, 100_000, 1_000_003])
def test_sort_pairs_stable(hx, n):
    say(f"case sort n={n}")
    rng = np.random.default_rng(n)
    for begin, end in [(0, 64), (0, 17), (20, 44), (60, 64)]:
        for spread in ("wide", "few"):
            # "few": a handful of distinct digits, so stability decides most of the order
            if spread == "wide" and n > 8:
                keys = rng.integers(0, 2**64, n, dtype=np.uint64)
            else:
                keys = rng.integers(0, 4, n).astype(np.uint64) << np.uint64(begin)
                keys |= rng.integers(0, 2**16, n).astype(np.uint64) << np.uint64(end % 64) if end < 64 else 0
            mask = np.uint64(((1 << (end - begin)) - 1) if end - begin < 64 else 2**64 - 1)
            digit = (keys >> np.uint64(begin)) & mask
            order = np.argsort(digit, kind="stable")
            vals = np.arange(n, dtype=np.uint32)
            k, v = hx.device.sort_pairs(keys, vals, begin, end)
            assert np.array_equal(v, vals[order]), (begin, end, spread)
            assert np.array_equal(k, keys[order]), (begin, end, spread)
            vals64 = rng.integers(0, 2**64, n, dtype=np.uint64)
            k, v = hx.device.sort_pairs(keys, vals64, begin, end)
            assert np.array_equal(v, vals64[order]), (begin, end, spread)
            assert np.array_equal(hx.device.sort_pairs(keys, None, begin, end), keys[order]), (begin, end, spread)


# one tile, tiles whose totals each workgroup reduces itself (up to 1024), and more
@pytest.mark.parametrize("n", [1, 5, 2048, 4097, 600_001, 2_097_152, 2_097_153, 3_000_001])
def test_scan_and_compact(hx, n):
    say(f"case scan n={n}")
    rng = np.random.default_rng(n)
    x = rng.integers(0, 1000, n, dtype=np.uint64)
    ex = hx.device.exclusive_scan(x)
    assert np.array_equal(ex, np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.uint64))
    flags = (rng.random(n) < 0.3).astype(np.int32)
    assert np.array_equal(hx.device.compact(flags), np.nonzero(flags)[0].astype(np.int32))


@pytest.mark.parametrize("name", list(GRAPHS))
def test_sclp_proposals_match_oracle(hx, oracle, name):
    say(f"case sclp_proposals graph={name}")
    indptr, indices = GRAPHS[name](hx)
    n = len(indptr) - 1
    rng = np.random.default_rng(7)
    ew = rng.integers(1, 5, len(indices)).astype(np.uint64)
    # symmetric weights: weight of (u,v) from the unordered pair
    src = np.repeat(np.arange(n), np.diff(indptr))
    lo, hi = np.minimum(src, indices), np.maximum(src, indices)
    ew = ((lo * 2654435761 + hi * 40503) % 5 + 1).astype(np.uint64)
    dg = hx.device.DeviceGraph(indptr, indices, ew)
    og = oracle.Graph(indptr, indices.astype(np.int32), ew)
    sums = np.concatenate([[0], np.cumsum(ew.astype(np.uint64))]).astype(np.uint64)
    row_bound = max(1, int((sums[indptr[1:]] - sums[indptr[:-1]]).max()))
    for clusters_kind in ("identity", "random", "blocks"):
        if clusters_kind == "identity":
            clusters = np.arange(n, dtype=np.int32)
        elif clusters_kind == "random":
            clusters = rng.integers(0, max(1, n // 3), n).astype(np.int32)
        else:
            clusters = (np.arange(n) // 5).astype(np.int32)
        for rnd, filt, thr in [(0, True, 2**31), (1, True, 1_500_000_000), (4, False, 0)]:
            single = clusters_kind == "identity" and rnd == 0
            op, ogn, ot = oracle.sclp_proposals(og, clusters, 3, 1, rnd, filt, thr)
            # unknown bound: 64-bit table sums; the row-weight bound: 32-bit sums
            for bound in (0, row_bound):
                p, g, t = hx.device.sclp_proposals(dg, clusters, 3, 1, rnd, filt, thr, singleton=single,
                                                   value_bound=bound)
                assert np.array_equal(p, op), (clusters_kind, rnd, bound)
                assert np.array_equal(g, ogn), (clusters_kind, rnd, bound)
                assert np.array_equal(t, ot), (clusters_kind, rnd, bound)


def mix32(x):
    """The lowbias32 integer hash (a bijection on 32-bit words)."""
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    return x ^ (x >> 16)


@pytest.mark.parametrize("bound", [0, 1 << 20])
def test_sclp_partition_overflow_splits(hx, oracle, bound):
    """A row of 16385 entries (8 hash partitions) whose neighbours sit in 16385
    distinct clusters that all hash into partition 0: the LDS pass overflows and
    is redone 2 then 4 ways finer; proposals still equal the oracle's."""
    say(f"case sclp partition overflow bound={bound}")
    d = 16385
    n = d + 1
    candidates = np.arange(1, 400_000, dtype=np.int64)
    ids = candidates[(mix32(candidates) & 7) == 0][:d]
    assert len(ids) == d
    cluster_count = int(ids.max()) + 1
    indptr = np.zeros(n + 1, np.int64)
    indptr[1] = d
    indptr[2:] = d + np.arange(1, n)
    indices = np.concatenate([np.arange(1, n), np.zeros(d, np.int64)])
    src = np.repeat(np.arange(n), np.diff(indptr))
    lo, hi = np.minimum(src, indices), np.maximum(src, indices)
    ew = ((lo * 2654435761 + hi * 40503) % 7 + 1).astype(np.uint64)
    clusters = np.zeros(n, np.int32)
    clusters[1:] = ids
    clusters[0] = cluster_count  # the hub alone
    dg = hx.device.DeviceGraph(indptr, indices, ew)
    og = oracle.Graph(indptr, indices.astype(np.int32), ew)
    for rnd, filt, thr in [(1, True, 3_000_000_000), (4, False, 0)]:
        op, ogn, ot = oracle.sclp_proposals(og, clusters, 3, 1, rnd, filt, thr)
        p, g, t = hx.device.sclp_proposals(dg, clusters, 3, 1, rnd, filt, thr, value_bound=bound)
        assert np.array_equal(p, op) and np.array_equal(g, ogn) and np.array_equal(t, ot), rnd


@pytest.mark.parametrize("name", ["grid64", "rmat12", "hub", "rmat17"])
@pytest.mark.parametrize("width", ["standard", "big"])
def test_contract_matches_oracle(hx, oracle, name, width):
    say(f"case contract graph={name} width={width}")
    indptr, indices = GRAPHS[name](hx)
    n = len(indptr) - 1
    og = oracle.Graph(indptr, indices.astype(np.int32))
    mp, cw, coarse, cap, _ = oracle.aggregate(og, 4, 0, 0)
    expect = oracle.contract(og, mp, coarse, cw).arrays()
    dg = hx.device.DeviceGraph(indptr, indices, width=width)
    # the coarse neighbours gathered twice (default at these sizes) and stashed
    # by the count pass for the encode pass (threshold 0)
    for stash in (1 << 24, 0):
        hx.set_contract_stash(stash)
        try:
            # bounds that fit the packed key and bounds that do not (rmat17: 30
            # weight bits); unit entry weights then still pack (the unpacked
            # layouts: test_contract_unpacked_heavy_weights)
            for bound in (64, 2**32 - 1, 2**64 - 1):
                off, nbr, w = hx.device.contract(dg, mp, coarse, bound)
                assert np.array_equal(off, expect[0]), (stash, bound)
                assert np.array_equal(nbr, expect[1]), (stash, bound)
                assert np.array_equal(w, expect[2]), (stash, bound)
        finally:
            hx.set_contract_stash(1 << 24)


@pytest.mark.parametrize("bound", [2**32 - 1, 2**64 - 1])
def test_contract_unpacked_heavy_weights(hx, oracle, bound):
    """Entry weights above what a packed key holds (rmat17's coarse level leaves
    30 weight bits) keep the unpacked layout: 4-byte values under a 32-bit row
    bound, 8-byte values otherwise; more than 1024 run-scan tiles."""
    say(f"case contract heavy weights bound={bound}")
    indptr, indices = GRAPHS["rmat17"](hx)
    n = len(indptr) - 1
    src = np.repeat(np.arange(n, dtype=np.uint64), np.diff(indptr))
    lo = np.minimum(src, indices.astype(np.uint64))
    hi = np.maximum(src, indices.astype(np.uint64))
    w = (np.uint64(1) << np.uint64(30)) + (lo * np.uint64(2654435761) + hi) % (np.uint64(1) << np.uint64(30))
    og = oracle.Graph(indptr, indices.astype(np.int32), edge_weights=w)
    mp, cw, coarse, cap, _ = oracle.aggregate(og, 4, 0, 0)
    expect = oracle.contract(og, mp, coarse, cw).arrays()
    assert int(w.max()) >= 1 << 30 and coarse >= 1 << 16  # no packed layout fits
    dg = hx.device.DeviceGraph(indptr, indices, edge_weights=w, width="big")
    off, nbr, wt = hx.device.contract(dg, mp, coarse, bound)
    assert np.array_equal(off, expect[0])
    assert np.array_equal(nbr, expect[1])
    assert np.array_equal(wt, expect[2])


@pytest.mark.parametrize("scale", ["small", "saturating"])
def test_contract_record_bounds(hx, scale):
    """Per-cluster saturating sums of row bounds (src/coarsen.cu:72-85), with zero
    bounds packed into one large cluster; closed form on Python integers."""
    say(f"case contract record bounds scale={scale}")
    rng = np.random.default_rng(11)
    n, coarse = 200_000, 5000
    mapping = rng.integers(0, coarse, n).astype(np.int32)
    mapping[: n // 4] = 0  # a packed cluster of isolated vertices
    if scale == "small":
        bounds = rng.integers(0, 1000, n).astype(np.uint64)
    else:
        bounds = rng.integers(2**58, 2**62, n).astype(np.uint64)
        bounds[rng.random(n) < 0.5] = 0
    bounds[: n // 4] = 0
    got, mx = hx.device.contract_record_bounds(mapping, bounds, coarse)
    sums = [0] * coarse
    for c, b in zip(mapping.tolist(), bounds.tolist()):
        sums[c] += b
    expect = np.array([min(x, 2**64 - 1) for x in sums], dtype=np.uint64)
    assert np.array_equal(got, expect)
    assert mx == int(expect.max())
    if scale == "saturating":
        assert (expect == np.uint64(2**64 - 1)).any()


def compare_hierarchy(gh, oh):
    assert len(gh) == len(oh)
    for level in range(len(gh)):
        a = gh.level(level)
        b = oh.level(level).arrays()
        for x, y, what in zip(a, b, ("offsets", "neighbors", "edge_weights", "vertex_weights")):
            assert np.array_equal(x, y), (level, what)
        if level + 1 < len(gh):
            assert np.array_equal(gh.map(level), oh.map(level)), level
    assert gh.stop_reason == oh.stop_reason


@pytest.mark.parametrize("name", ["grid256", "rmat12", "hub", "bighub"])
@pytest.mark.parametrize("production", [False, True])
def test_hierarchy_bit_exact(hx, oracle, name, production):
    say(f"case hierarchy graph={name} production={production}")
    indptr, indices = GRAPHS[name](hx)
    k = 4
    csf = max(64, (512 + k - 1) // k) if production else 256
    g = hx.Graph(indptr, indices)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    for seed in (0, 1):
        gh = g.coarsen(k, seed=seed, cluster_size_factor=csf, pack_isolated=production)
        oh = oracle.coarsen(og, k, seed=seed, cluster_size_factor=csf, pack_isolated=production)
        compare_hierarchy(gh, oh)


@pytest.mark.parametrize("mode", ["plain_fp64", "batched_fp64"])
def test_initial_partition_bit_exact(hx, oracle, mode):
    say(f"case initial mode={mode}")
    indptr, indices = hx.generate_grid2d(256, 256)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    oh = oracle.coarsen(og, 8, cluster_size_factor=64, pack_isolated=True)
    off, adj, ew, vw = oh.level(len(oh) - 1).arrays()
    coarse = hx.Graph(off, adj, edge_weights=ew, vertex_weights=vw)
    oc = oracle.Graph(off, adj, ew, vw)
    for k in (2, 3, 8):
        labels, coords = coarse.initial_partition(k, 1.03, 0, mode, with_coords=True)
        ol, ocoords, _ = oracle.initial_partition(oc, k, 1.03, 0, with_coords=True)
        assert np.array_equal(labels, ol), k
        assert np.array_equal(coords.view(np.uint64), ocoords.view(np.uint64)), k


def test_refine_kats(hx):
    say("case refine kats")
    # tests/test_uncoarsen.cu:454-484 of the reference
    ex = hx.Graph([0, 2, 5, 7, 10, 11, 12], [1, 3, 0, 2, 4, 1, 3, 0, 2, 5, 1, 3],
                  edge_weights=[1, 4, 1, 4, 100, 4, 1, 4, 1, 100, 100, 100], vertex_weights=[1] * 6)
    lab, rep = ex.refine([0, 0, 1, 1, 0, 1], 2, 1.03, 32, verify=True)
    assert ex.cut(lab) == 2 and (lab == 0).sum() == 3
    rp = hx.Graph([0, 2, 4, 5, 5, 7, 8, 8, 8], [1, 4, 0, 2, 1, 0, 5, 4],
                  edge_weights=[1, 4, 1, 100, 100, 4, 100, 100], vertex_weights=[1] * 8)
    lab, rep = rp.refine([0, 0, 0, 0, 1, 1, 1, 1], 2, 1.03, 32, verify=True)
    assert rp.cut(lab) == 1 and (lab == 0).sum() == 4 and rep["admission_iterations"] > 0


@pytest.mark.parametrize("k", [2, 4, 8, 32])
def test_refine_matches_oracle(hx, oracle, k):
    say(f"case refine oracle k={k}")
    indptr, indices = hx.generate_grid2d(64, 64)
    n = len(indptr) - 1
    labels = (np.arange(n) * 7919 % k).astype(np.int32)  # balanced, poor cut
    g = hx.Graph(indptr, indices)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    gl, grep = g.refine(labels, k, 1.03, 32)
    ol, orep = oracle.refine(og, labels, k, 1.03, 32)
    assert grep["final_cut"] == orep.final_cut
    assert np.array_equal(gl, ol)


@pytest.mark.parametrize("name,k", [("grid256", 4), ("grid256", 32), ("rmat12", 8), ("hub", 4), ("bighub", 8)])
def test_partition_matches_oracle(hx, oracle, name, k):
    say(f"case partition graph={name} k={k}")
    indptr, indices = GRAPHS[name](hx)
    g = hx.Graph(indptr, indices)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    labels, rep = g.partition(k, 1.03, 0)
    ol, orep = oracle.partition(og, k, 1.03, 0)
    cap = oracle.partition_capacity(len(indptr) - 1, k, 1.03)
    assert np.bincount(labels, minlength=k).max() <= cap
    assert oracle.cut(og, labels) == rep["final_cut"]
    assert rep["final_cut"] == orep.final_cut
    assert np.array_equal(labels, ol)


@pytest.mark.parametrize("name,k", [("rmat12", 8), ("bighub", 4), ("grid256", 32)])
def test_refine_rebuild_rounds_match_oracle(hx, oracle, name, k):
    """Rounds that rebuild the affinity cache instead of updating it (forced on
    every round here; by default only on levels of 2^23+ entries) give the same
    labels; strict verification compares the cache with a CSR rebuild each round."""
    say(f"case refine rebuild graph={name} k={k}")
    indptr, indices = GRAPHS[name](hx)
    n = len(indptr) - 1
    labels = (np.arange(n) * 7919 % k).astype(np.int32)
    og = oracle.Graph(indptr, indices.astype(np.int32))
    ol, orep = oracle.refine(og, labels, k, 1.03, 32)
    g = hx.Graph(indptr, indices)
    try:
        hx.set_refine_rebuild(0, 1 << 40)
        gl, grep = g.refine(labels, k, 1.03, 32, verify=True)
        pl, prep = g.partition(k, 1.03, 0)
    finally:
        hx.set_refine_rebuild()
    assert grep["final_cut"] == orep.final_cut
    assert np.array_equal(gl, ol)
    pol, porep = oracle.partition(og, k, 1.03, 0)
    assert prep["final_cut"] == porep.final_cut
    assert np.array_equal(pl, pol)


@pytest.mark.parametrize("name,k", [("grid256", 4), ("rmat12", 8), ("grid64", 32)])
def test_spectral_fp32_against_fp64(hx, oracle, name, k):
    """FP32 diffusion against FP64, per recursion depth, on the coarsest level of the
    production hierarchy: the FP64 side reproduces the plain initial partition, the
    FP32 side the batched FP32 one, and the per-depth vector error stays small."""
    say(f"case spectral fp32 vs fp64 graph={name} k={k}")
    indptr, indices = GRAPHS[name](hx)
    g = hx.Graph(indptr, indices)
    csf = max(64, (512 + k - 1) // k)
    h = g.coarsen(k, seed=0, cluster_size_factor=csf, pack_isolated=True)
    off, adj, ew, vw = h.level(len(h) - 1)
    coarse = hx.Graph(off, adj, edge_weights=ew, vertex_weights=vw)
    rep = coarse.spectral_compare(k, 1.03, 0)
    depths = rep["depths"]
    assert [d["depth"] for d in depths] == list(range(len(depths)))
    assert len(depths) == (k - 1).bit_length()
    assert [d["tasks"] for d in depths] == [2 ** i for i in range(len(depths))]
    for d in depths:
        assert d["vertices"] == len(off) - 1
        assert 0.0 <= d["rms_error"] <= d["max_abs_error"] < 1e-3, d
        assert d["labels_differ"] >= 0 and 0 <= d["splits_differ"] <= d["tasks"]
    lab64 = coarse.initial_partition(k, 1.03, 0, "batched_fp64")
    lab32 = coarse.initial_partition(k, 1.03, 0, "batched_fp32")
    assert rep["cut_fp64"] == coarse.cut(lab64)
    assert rep["cut_fp32"] == coarse.cut(lab32)
    assert rep["labels_differ"] == int((lab64 != lab32).sum())
    cap = hx.partition_capacity(int(vw.astype(object).sum()), k, 1.03)
    assert max(rep["max_load_fp64"], rep["max_load_fp32"]) <= cap
