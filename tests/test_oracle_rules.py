"""CPU: the oracle's rule choices, pinned independently of its own golden digests.

The reference ships no fixture of an SCLP map or of a refinement trajectory, so
these checks pin the rules that decide them:

1. Literal known answers on an 8-vertex weighted graph for the per-vertex SCLP
   proposal (src/coarsen.cu:218-229 role gate and candidate order, :306-330 tie
   key, :562-598 decode, :1152-1157 salts; mix32 is the public lowbias32 hash,
   :112-118). Each vector was worked out from those lines and typed here; the
   tie words (high 32 bits of the winning key) pin the tie salt of every
   (seed, level, round) used, the proposal sets under the role filter pin the
   role salt, and vertices 0, 3 and 6 of the first case choose between two
   neighbours of equal weight, so the direction of the tie comparison decides
   them.
2. Admission by brute force on small graphs: in every LP round of a level, each
   target admits exactly the longest prefix of its proposals ordered by (gain
   desc, tie desc, vertex asc) whose running weight fits the capacity left at the
   round's start (src/coarsen.cu:600-680, :1336-1340), and the mover threshold is
   (u32)(min(0.5, (clusters - ceil(n/2)) / clusters) * 2^32) (:1306-1315).
3. The reference's own diagnostic invariant (src/coarsen.cu:1380-1382,
   :1521-1524): no LP round and no two-hop pairing increases the cluster cut;
   run over the golden inputs at seeds 0-4.
4. Literal known answers for the refinement proposal filter
   (src/refine_incremental.cuh:27-46, :332-333): a negative-gain move is proposed
   iff current - best < filter * current, with filter 0.25 on unit-weight graphs and
   0.75 otherwise; together with the afterburner (:48-67) it decides which of two
   vertices move.

Rules that only a reading pins are listed where the design document covers the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import kat_graphs as K

INVALID = 0x7FFFFFFF

# ---------------------------------------------------------------- 1. SCLP proposals

# undirected (a, b, weight)
EDGES8 = [(0, 1, 2), (0, 2, 2), (0, 3, 1), (1, 4, 3), (2, 5, 3), (3, 6, 2), (3, 7, 2), (4, 5, 1), (6, 7, 1),
          (5, 6, 2)]


def graph_from_edges(oracle, edges, n, vertex_weights=None):
    rows = [[] for _ in range(n)]
    for a, b, w in edges:
        rows[a].append((b, w))
        rows[b].append((a, w))
    g = K.csr_from_rows([sorted(r) for r in rows], vertex_weights if vertex_weights is not None else [1] * n)
    return oracle.Graph(g["indptr"], g["indices"].astype(np.int32), g["edge_weights"], g["vertex_weights"]), g


SCLP_CASES = [
    # (clusters, seed, level, round, filter_roles, mover_threshold, proposals, gains, ties)
    ([0, 1, 2, 3, 4, 5, 6, 7], 0, 0, 0, False, 0,
     [1, 4, 5, 6, 1, 2, 3, 3], [2, 3, 3, 2, 3, 3, 2, 2],
     [0x9dfce202, 0x92b0a8c0, 0x57ff4c0d, 0xf2ab8eb8, 0xd45e485b, 0xe3d6c534, 0xab648f31, 0x98d28ab2]),
    ([0, 1, 2, 3, 4, 5, 6, 7], 0, 0, 1, True, 1 << 31,
     [1, INVALID, 5, 6, 1, INVALID, INVALID, 6], [2, 0, 3, 2, 3, 0, 0, 1],
     [0xbf6ddd24, 0, 0x985920e6, 0xbcfcc39f, 0xa99d1bb6, 0, 0, 0xe4651ba4]),
    ([0, 1, 2, 3, 4, 5, 6, 7], 7, 3, 2, True, 1 << 31,
     [1, INVALID, INVALID, INVALID, 1, 2, INVALID, 3], [2, 0, 0, 0, 3, 3, 0, 2],
     [0xa7951fbb, 0, 0, 0, 0xb43e6d65, 0x3f88b89b, 0, 0xe13ef81a]),
    ([0, 0, 2, 3, 4, 4, 6, 6], 1, 5, 3, True, 3 << 30,
     [INVALID, INVALID, 0, INVALID, 0, INVALID, 3, 3], [0, 0, 2, 0, 2, 0, 1, 1],
     [0, 0, 0x00ec57be, 0, 0x2afdb0d3, 0, 0x73e6965e, 0x1d2045a5]),
    # round 4 is the two-hop favourite pass: no role filter
    ([0, 0, 2, 3, 4, 4, 6, 6], 11, 2, 4, False, 0,
     [INVALID, 4, 4, 6, 0, 2, 4, 3], [0, 1, 3, 4, 2, 2, 1, 1],
     [0, 0x8c485e2c, 0xed04f7a2, 0x32f246f4, 0x57c66bc0, 0x48722d5b, 0x2cdde0e8, 0x51cd56a3]),
]


@pytest.mark.parametrize("case", range(len(SCLP_CASES)))
def test_sclp_proposals_known_answers(oracle, case):
    clusters, seed, level, rnd, filter_roles, threshold, want_p, want_g, want_t = SCLP_CASES[case]
    g, _ = graph_from_edges(oracle, EDGES8, 8)
    p, gains, ties = oracle.sclp_proposals(g, clusters, seed, level, rnd, filter_roles, threshold)
    assert p.tolist() == want_p
    assert gains.tolist() == want_g
    assert ties.tolist() == want_t


def test_sclp_equal_weight_choices_follow_the_tie_key(oracle):
    """Vertices 0, 3 and 6 of the first case have two neighbours of equal weight
    (0: 1 and 2; 3: 6 and 7; 6: 3 and 5); the larger 64-bit key wins, and the
    chosen target is recovered from its low word as 0xffffffff - low."""
    g, _ = graph_from_edges(oracle, EDGES8, 8)
    p, _, _ = oracle.sclp_proposals(g, list(range(8)), 0, 0, 0, False, 0)
    assert (p[0], p[3], p[6]) == (1, 6, 3)


# ---------------------------------------------------------------- 2. admission by brute force

def small_graphs(hx):
    rng = np.random.default_rng(5)
    out = []
    ip, ix = hx.generate_grid2d(12, 12)
    out.append(("grid12", ip, ix, None, None))
    n = len(ip) - 1
    out.append(("grid12_weighted", ip, ix, None, rng.integers(1, 6, n).astype(np.uint64)))
    ip, ix = hx.generate_rmat(8, 8)
    n = len(ip) - 1
    src = np.repeat(np.arange(n), np.diff(ip))
    a, b = np.minimum(src, ix), np.maximum(src, ix)
    w = (a * 131 + b * 71) % 5 + 1  # symmetric weights
    out.append(("rmat8_weighted", ip, ix, w.astype(np.uint64), rng.integers(1, 4, n).astype(np.uint64)))
    return out


def cluster_cut(g, clusters):
    return K.independent_cut(g, np.asarray(clusters))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sclp_admission_brute_force(hx, oracle, seed):
    for name, ip, ix, ew, vw in small_graphs(hx):
        og = oracle.Graph(ip, ix.astype(np.int32), ew, vw)
        n = og.n
        vwt = np.ones(n, np.uint64) if vw is None else vw
        parts, csf = 2, 8
        capacity = max(int(vwt.max()), -(-int(vwt.sum()) // (parts * csf)))
        _, _, _, cap, _ = oracle.aggregate(og, parts, seed, 0, csf)
        assert cap == capacity, name
        before = np.arange(n, dtype=np.int32)
        lp_steps = 0
        for kind, rnd, threshold, after in oracle.aggregate_trace(og, parts, seed, 0, csf):
            if kind != "lp":
                break
            assert rnd == lp_steps
            lp_steps += 1
            count = len(np.unique(before))
            fraction = min(0.5, max(0, count - (n + 1) // 2) / count)
            assert threshold == int(fraction * 4294967296.0), (name, rnd)
            p, gains, ties = oracle.sclp_proposals(og, before, seed, 0, rnd, True, threshold)
            weight = np.bincount(before, weights=vwt.astype(np.float64), minlength=n).astype(np.int64)
            expect = before.copy()
            for t in np.unique(p[p != INVALID]):
                members = [v for v in range(n) if p[v] == t]
                members.sort(key=lambda v: (-int(gains[v]), -int(ties[v]), v))
                room = capacity - int(weight[t])
                running = 0
                for v in members:
                    running += int(vwt[v])
                    if room < 0 or running > room:
                        break
                    expect[v] = t
            assert np.array_equal(after, expect), (name, seed, rnd)
            before = after
        assert lp_steps >= 1, name


# ---------------------------------------------------------------- 3. cut never increases

@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("name,k", [("grid256", 4), ("rmat12", 8)])
def test_sclp_steps_never_increase_the_cluster_cut(hx, oracle, name, k, seed):
    ip, ix = hx.generate_grid2d(256, 256) if name == "grid256" else hx.generate_rmat(12, 16)
    g = dict(indptr=ip, indices=ix, edge_weights=np.ones(len(ix), np.uint64))
    og = oracle.Graph(ip, ix.astype(np.int32))
    for csf, packed in ((max(64, -(-512 // k)), True), (256, False)):
        previous = cluster_cut(g, np.arange(og.n))
        steps = oracle.aggregate_trace(og, k, seed, 0, csf, packed)
        assert steps and steps[0][0] == "lp"
        for kind, rnd, _, clusters in steps:
            now = cluster_cut(g, clusters)
            assert now <= previous, (name, seed, kind, rnd)
            previous = now


# ---------------------------------------------------------------- 4. refinement filter

# Parts A = {0,1,2,3}, B = {4,5,6,7}. Vertex 0 gains by moving to B (conn 1 -> 2).
# Vertex 1 keeps conn 3 in A against 2 in B: gain -1, proposed only when
# 1 < filter * 3, i.e. with filter 0.75 (weighted graph) and not with 0.25 (unit
# weights). When proposed, the afterburner adds 2 * w(0,1) for vertex 0 (higher
# gain, moving from 1's part to 1's target), so 1 follows 0 into B.
REFINE_EDGES = [(0, 1), (0, 5), (0, 6), (1, 2), (1, 3), (1, 5), (1, 6), (2, 3), (4, 5), (4, 6), (4, 7), (5, 6),
                (5, 7), (6, 7)]


@pytest.mark.parametrize("scale,rounds,want,cut0,cut1", [
    (1, 1, [1, 0, 0, 0, 1, 1, 1, 1], 4, 3),   # filter 0.25: only vertex 0 moves
    (2, 1, [1, 1, 0, 0, 1, 1, 1, 1], 8, 4),   # filter 0.75: vertices 0 and 1 move
    (1, 32, [1, 1, 0, 0, 1, 1, 1, 1], 4, 2),  # unit weights reach the same labels in later rounds
])
def test_refine_filter_known_answers(oracle, scale, rounds, want, cut0, cut1):
    g, _ = graph_from_edges(oracle, [(a, b, scale) for a, b in REFINE_EDGES], 8)
    labels, st = oracle.refine(g, [0, 0, 0, 0, 1, 1, 1, 1], 2, 1.5, rounds, 3)
    assert labels.tolist() == want
    assert (st.initial_cut, st.final_cut) == (cut0, cut1)
