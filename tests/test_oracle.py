"""Pins the CPU oracle (oracle/usap_oracle.c) against independent checks:
numpy's Philox, high-precision logarithms, brute-force entropy differences,
exact enumeration of the samplers, hand-derived SCC and golden-section cases.
"""
import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from conftest import random_graph

FRAC = 1 << 48


# ---------------------------------------------------------------------------
# random stream
# ---------------------------------------------------------------------------

def test_philox_matches_numpy(orc, usap):
    rng = np.random.default_rng(0)
    for _ in range(50):
        ctr = rng.integers(0, 2**63, 4, dtype=np.uint64)
        key = rng.integers(0, 2**63, 2, dtype=np.uint64)
        # numpy's Philox increments its counter before producing a block
        prev = ctr.copy()
        prev[0] = prev[0] - np.uint64(1) if prev[0] > 0 else prev[0]
        if ctr[0] == 0:
            continue
        bg = np.random.Philox(key=key, counter=prev)
        want = bg.random_raw(4).astype(np.uint64)
        np.testing.assert_array_equal(orc.philox4x64(ctr, key), want)
        np.testing.assert_array_equal(usap.philox4x64(ctr, key), want)


def test_philox_known_answer(orc):
    # Random123 philox4x64-10 known-answer vector for counter 0, key 0
    out = orc.philox4x64([0, 0, 0, 0], [0, 0])
    assert [hex(int(v)) for v in out] == [
        "0x16554d9eca36314c", "0xdb20fe9d672d0fdc", "0xd7e772cee186176b", "0x7e68b68aec7ba23b"]


def test_pick_enumerates_weights(orc):
    rng = np.random.default_rng(1)
    for _ in range(20):
        w = rng.integers(0, 5, rng.integers(1, 9)).astype(np.int64)
        if w.sum() == 0:
            w[0] = 1
        counts = np.zeros(len(w), np.int64)
        for t in range(int(w.sum())):
            counts[orc.pick(w, t)] += 1
        np.testing.assert_array_equal(counts, w)


# ---------------------------------------------------------------------------
# fixed-point logarithm and entropy
# ---------------------------------------------------------------------------

def test_lambda_table_accuracy(orc, usap):
    getcontext().prec = 50
    lam = orc.lambda_table(5000)
    np.testing.assert_array_equal(lam, usap.lambda_table(5000))
    assert lam[0] == 0 and lam[1] == 0
    xs = list(range(1, 200)) + list(range(4000, 5001, 37))
    for x in xs:
        exact = Decimal(x).ln() * FRAC
        assert abs(Decimal(int(lam[x])) - exact) <= Decimal(2) * (1 + exact / Decimal(2**52) * 4)


def _data_S_textbook(M, dout, din):
    s = 0.0
    for a, b in zip(*np.nonzero(M)):
        m = M[a, b]
        s -= m * math.log(m / (dout[a] * din[b]))
    return s


def _model_S(N, E, B):
    x = B * B / E
    return E * (1 + x) * math.log(1 + x) - x * math.log(x) + N * math.log(B)


@pytest.mark.parametrize("init", ["scc", "identity", "random"])
def test_M_and_entropy(orc, usap, init):
    g = random_graph(300, 2500, seed=2, weighted=True, self_loops=True)
    o = orc.Oracle(g, threads=2, block_size=64)  # exercises the edge-list M path too
    if init == "scc":
        lab, B = o.scc()
    elif init == "identity":
        lab, B = np.arange(g.N, dtype=np.int32), g.N
    else:
        B = 17
        lab = np.random.default_rng(3).integers(0, B, g.N).astype(np.int32)
    o.set_labels(lab, B)
    src, dst = g.edges()
    M = np.zeros((B, B), np.int64)
    np.add.at(M, (lab[src], lab[dst]), g.out_w.astype(np.int64))
    np.testing.assert_array_equal(o.M_dense(), M)
    do, di = o.degrees()
    np.testing.assert_array_equal(do, M.sum(1))
    np.testing.assert_array_equal(di, M.sum(0))
    fx, S = o.entropy()
    want = _data_S_textbook(M, M.sum(1), M.sum(0))
    assert abs(fx / FRAC - want) <= 1e-9 * abs(want)
    assert abs(S - (want + _model_S(g.N, g.E, B))) <= 1e-9 * abs(S)
    assert abs(usap.model_S(g.N, g.E, B) - _model_S(g.N, g.E, B)) <= 1e-12 * abs(S)


# ---------------------------------------------------------------------------
# dS against brute force
# ---------------------------------------------------------------------------

def _fresh(orc, g, lab, B, **kw):
    o = orc.Oracle(g, threads=1, **kw)
    o.set_labels(lab, B)
    return o


@pytest.mark.parametrize("block_size", [1024, 8])
def test_merge_dS_equals_entropy_difference(orc, block_size):
    g = random_graph(120, 900, seed=4, weighted=True)
    rng = np.random.default_rng(5)
    B = 12
    lab = rng.integers(0, B, g.N).astype(np.int32)
    o = _fresh(orc, g, lab, B, block_size=block_size)
    base, _ = o.entropy()
    seen = 0
    for r in range(B):
        for idx in range(4):
            s, d = o.merge_propose_one(r, idx, 9)
            if s == r:
                assert d == 0
                continue
            merged = np.where(lab == r, s, lab)
            uniq, new = np.unique(merged, return_inverse=True)
            o2 = _fresh(orc, g, new.astype(np.int32), len(uniq))
            after, _ = o2.entropy()
            # the merged partition has B-1 blocks only in the model term; the
            # data term is what dS measures
            assert d == after - base
            seen += 1
    assert seen > 20


@pytest.mark.parametrize("block_size", [1024, 8])
def test_nodal_dS_exact_equals_entropy_difference(orc, block_size):
    g = random_graph(150, 1400, seed=6, weighted=True, self_loops=True)
    rng = np.random.default_rng(7)
    B = 9
    lab = rng.integers(0, B, g.N).astype(np.int32)
    o = _fresh(orc, g, lab, B, block_size=block_size, ds_mode=1)
    base, _ = o.entropy()
    checked = 0
    for n in range(g.N):
        for stream in range(3):
            s, d, inf, acc = o.nodal_propose_one(n, stream)
            r = lab[n]
            if s == r:
                assert d == 0 and acc == 0
                continue
            lab2 = lab.copy()
            lab2[n] = s
            after, _ = _fresh(orc, g, lab2, B).entropy()
            assert inf == 0
            assert d == after - base
            assert acc == (d < 0)
            checked += 1
    assert checked > 200


def _corner_T(M, r, s, dout, din, x_r, x_s, y_r, y_s):
    """sum of the 4 corner T terms (double) for entries x and degrees y."""
    tot = 0.0
    for m, x, y in ((M[0], y_r, x_r), (M[1], y_s, x_r), (M[2], y_r, x_s), (M[3], y_s, x_s)):
        if m:
            if x * y == 0:
                return math.inf
            tot += m * math.log(m / (x * y))
    return tot


def test_nodal_dS_ref_mode_quirk(orc):
    """ref mode (graph_parallel.hpp:1117-1145) differs from exact mode only in
    the four (r,s) corner entries of the new rows: they miss the updates of the
    in-edges from r and s.  Checked against an independent corner evaluation."""
    g = random_graph(150, 1400, seed=8)
    rng = np.random.default_rng(9)
    B = 7
    lab = rng.integers(0, B, g.N).astype(np.int32)
    oref = _fresh(orc, g, lab, B, ds_mode=0)
    oex = _fresh(orc, g, lab, B, ds_mode=1)
    M = oref.M_dense()
    dout, din = M.sum(1), M.sum(0)
    src, dst = g.edges()
    same, differ = 0, 0
    for n in range(g.N):
        s, dref, infr, _ = oref.nodal_propose_one(n, 4)
        s2, dex, infx, _ = oex.nodal_propose_one(n, 4)
        assert s == s2
        r = lab[n]
        if s == r:
            continue
        outs = dst[src == n]
        ins = src[dst == n]
        wo_r, wo_s = np.sum(lab[outs] == r), np.sum(lab[outs] == s)
        wi_r, wi_s = np.sum(lab[ins] == r), np.sum(lab[ins] == s)
        k_out, k_in = len(outs), len(ins)
        dor2, dos2 = dout[r] - k_out, dout[s] + k_out
        dir2, dis2 = din[r] - k_in, din[s] + k_in
        ref_c = (M[r, r] - wo_r, M[r, s] - wo_s, M[s, r] + wo_r, M[s, s] + wo_s)
        ex_c = (M[r, r] - wo_r - wi_r, M[r, s] - wo_s + wi_r, M[s, r] + wo_r - wi_s,
                M[s, s] + wo_s + wi_s)
        t_ref = _corner_T(ref_c, r, s, dout, din, dor2, dos2, dir2, dis2)
        t_ex = _corner_T(ex_c, r, s, dout, din, dor2, dos2, dir2, dis2)
        if math.isinf(t_ref):
            assert infr == -1
            differ += 1
            continue
        assert infr == 0 and infx == 0
        want = dex / FRAC + (t_ex - t_ref)
        assert abs(dref / FRAC - want) <= 1e-7 * (1 + abs(want))
        if wi_r == 0 and wi_s == 0:
            assert dref == dex
            same += 1
        else:
            differ += 1
    assert same > 10 and differ > 10


def test_nodal_proposal_distribution(orc):
    """Chi-square of the proposed block against the exact proposal law of
    graph_parallel.hpp:1006-1063 for one node."""
    g = random_graph(60, 500, seed=10)
    B = 6
    lab = (np.arange(g.N) % B).astype(np.int32)
    o = _fresh(orc, g, lab, B)
    M = o.M_dense()
    da = M.sum(1) + M.sum(0)
    n = int(np.argmax(np.diff(g.out_ptr) + np.diff(g.in_ptr)))
    nb = np.concatenate([g.out_idx[g.out_ptr[n]:g.out_ptr[n + 1]],
                         g.in_idx[g.in_ptr[n]:g.in_ptr[n + 1]]])
    p = np.zeros(B)
    for v in nb:
        u = lab[v]
        pu = B / (da[u] + B)
        p += (pu / B + (1 - pu) * (M[u] + M[:, u]) / da[u]) / len(nb)
    draws = 20000
    counts = np.zeros(B)
    for stream in range(draws):
        counts[o.nodal_propose_one(n, stream)[0]] += 1
    from scipy.stats import chisquare
    assert chisquare(counts, p * draws).pvalue > 1e-4


# ---------------------------------------------------------------------------
# carry-out, SCC, golden section
# ---------------------------------------------------------------------------

def _greedy_merges(best, dS, B_to_merge, lab):
    """Greedy merges in (dS, block) order as a union of classes."""
    rep = list(range(len(best)))

    def find(x):
        while rep[x] != x:
            x = rep[x]
        return x

    done = 0
    for r in sorted(range(len(best)), key=lambda i: (dS[i], i)):
        if done == B_to_merge:
            break
        t = find(best[r])
        if t != r:
            rep[r] = t
            done += 1
    roots = sorted({find(b) for b in set(lab.tolist())})
    rank = {b: i for i, b in enumerate(roots)}
    return np.array([rank[find(b)] for b in lab], np.int32)


def test_carry_out(orc):
    g = random_graph(200, 1500, seed=11)
    rng = np.random.default_rng(12)
    B = 40
    lab = rng.integers(0, B, g.N).astype(np.int32)
    o = _fresh(orc, g, lab, B)
    best, dS = o.block_merge(3)
    want = _greedy_merges(best, dS, 20, lab)
    assert o.carry_out(20) == 0
    np.testing.assert_array_equal(o.labels(), want)
    assert o.B == 20


def test_carry_out_infeasible(orc):
    g = random_graph(50, 300, seed=13)
    lab = (np.arange(g.N) % 10).astype(np.int32)
    o = _fresh(orc, g, lab, 10)
    o.block_merge(0)
    assert o.carry_out(9) in (0, -1)
    o2 = _fresh(orc, g, lab, 10)
    o2.block_merge(0)
    # at most B - (number of components of the best-merge graph) merges exist
    assert o2.carry_out(10) == -1


def _graph_from_edges(usap, edges, N):
    s = np.array([e[0] for e in edges], np.int32)
    d = np.array([e[1] for e in edges], np.int32)
    return usap.Graph.from_edges(s, d, N)


@pytest.mark.parametrize("case", ["cycle", "chain", "long_cycle", "two_sccs"])
def test_scc_hand_cases(orc, usap, case):
    if case == "cycle":
        g, want = _graph_from_edges(usap, [(0, 1), (1, 2), (2, 0)], 3), [0, 0, 0]
    elif case == "chain":
        g, want = _graph_from_edges(usap, [(0, 1), (1, 2)], 3), [0, 1, 2]
    elif case == "long_cycle":
        # finish order is 49..0; the transpose DFS from 0 walks 0,49,48,..,31
        # and stops at the depth cap (20 nodes); nodes 1..30 then each start a
        # DFS whose only transpose neighbour is already labelled
        g = _graph_from_edges(usap, [(i, (i + 1) % 50) for i in range(50)], 50)
        want = None
    else:  # {0,1} -> {2,3}
        g, want = _graph_from_edges(usap, [(0, 1), (1, 0), (1, 2), (2, 3), (3, 2)], 4), [0, 0, 1, 1]
    lab, B = orc.Oracle(g).scc()
    lab2, B2 = usap.scc_init(g)
    np.testing.assert_array_equal(lab, lab2)
    assert B == B2
    if want is not None:
        np.testing.assert_array_equal(lab, want)
    else:
        assert B == 31 and sorted(np.bincount(lab).tolist()) == [1] * 30 + [20]
        assert np.all(lab[[0] + list(range(31, 50))] == 0)


def test_scc_random_graphs_match(orc, usap):
    for seed in range(5):
        g = random_graph(400, 1200, seed=20 + seed)
        lab, B = orc.Oracle(g).scc()
        lab2, B2 = usap.scc_init(g)
        np.testing.assert_array_equal(lab, lab2)
        assert B == B2 and lab.max() == B - 1


INF = math.inf
EMPTY = (0, INF, -1)


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_golden_section_script(orc, usap, impl):
    step = orc.golden_update if impl == "oracle" else usap.golden_update
    slots = [EMPTY, EMPTY, EMPTY]
    rc, slots, nxt, btm = step(slots, (100, 50.0, 0))
    assert (rc, nxt[2], btm) == (0, 0, 50) and slots[1][2] == 0
    rc, slots, nxt, btm = step(slots, (50, 40.0, 1))
    assert (rc, nxt[2], btm) == (0, 1, 25) and slots[0][2] == 0 and slots[1][2] == 1
    rc, slots, nxt, btm = step(slots, (25, 45.0, 2))
    # bracket [100, 50, 25]: larger side first, next B = 50 + round(50*0.618) = 81
    assert (rc, nxt[2], btm) == (0, 0, 19) and [s[2] for s in slots] == [0, 1, 2]
    rc, slots, nxt, btm = step(slots, (81, 42.0, 3))
    # 81 worse than med (50) and above it -> large = 81; [81, 50, 25]: 31 >= 25
    # -> next = 50 + round(31*0.618) = 69 from large (81)
    assert (rc, nxt[2], btm) == (0, 3, 12) and [s[2] for s in slots] == [3, 1, 2]
    rc, slots, nxt, btm = step(slots, (69, 39.0, 4))
    # 69 better than med -> med.B (50) < 69 -> small = old med; med = 69:
    # [81, 69, 50]; 12 < 19 -> next = 50 + round(19*0.618) = 62 from med
    assert (rc, nxt[2], btm) == (0, 4, 7) and [s[2] for s in slots] == [3, 4, 1]
    # bracket of width 2 ends the search at med
    rc, slots, nxt, btm = step([(12, 5.0, 7), (11, 4.0, 8), (10, 6.0, 9)], (12, 4.5, 10))
    assert rc == 1 and nxt[2] == 8


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_golden_section_rise_on_second_halving(orc, usap, impl):
    # large is never set; the reference's size_t arithmetic is undefined here,
    # this build stops the search at med (the best partition so far)
    step = orc.golden_update if impl == "oracle" else usap.golden_update
    rc, slots, nxt, btm = step([EMPTY, EMPTY, EMPTY], (100, 50.0, 0))
    rc, slots, nxt, btm = step(slots, (50, 60.0, 1))  # worse on the first halving
    assert rc == 1 and nxt[2] == 0 and slots[2][2] == 1


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_golden_section_stops_at_one_block(orc, usap, impl):
    step = orc.golden_update if impl == "oracle" else usap.golden_update
    rc, slots, nxt, btm = step([EMPTY, EMPTY, EMPTY], (1, 5.0, 0))
    assert rc == 1 and btm == 0 and nxt[2] == 0


# ---------------------------------------------------------------------------
# end to end (oracle alone)
# ---------------------------------------------------------------------------

def test_oracle_partition_recovers_planted_c1(orc, usap):
    g = usap.config_graph("c1")
    o = orc.Oracle(g, threads=4, seed=1)
    o.set_labels(*o.scc())
    B, lab, st = o.partition()
    assert st["error"] == 0
    assert B == len(np.unique(g.truth))
    assert usap.nmi(g.truth, lab) > 0.99


def test_oracle_thread_count_invariance(orc, usap):
    g = usap.config_graph("c1")
    out = []
    for t in (1, 3, 8):
        o = orc.Oracle(g, threads=t, seed=5)
        o.set_labels(*o.scc())
        out.append(o.partition())
    for B, lab, st in out[1:]:
        assert B == out[0][0] and st == out[0][2]
        np.testing.assert_array_equal(lab, out[0][1])


# ---------------------------------------------------------------------------
# merge proposal law, graph_parallel.hpp:774-865
# ---------------------------------------------------------------------------

def _merge_law(M, r):
    """Exact law of the block a merge proposal of r draws
    (graph_parallel.hpp:774-865), from the dense M:
      k = sum_i M[r][i] + M[i][r]; k == 0: s uniform on [0, B) (r included);
      else u ~ (M[r][u] + M[u][r]) / k, then with probability B / (d_a[u] + B)
      a uniform draw whose hit on r moves to r + 1 (B wraps to 0), else
      s ~ (M[u][s] + M[s][u]) over s != r, or the same uniform-not-r draw when
      that weight is 0."""
    B = M.shape[0]
    w_r = (M[r] + M[:, r]).astype(float)
    k = w_r.sum()
    if k == 0:
        return np.full(B, 1.0 / B)
    not_r = np.full(B, 1.0 / B)
    not_r[(r + 1) % B] += not_r[r]
    not_r[r] = 0.0
    da = M.sum(1) + M.sum(0)
    p = np.zeros(B)
    for u in range(B):
        if w_r[u] == 0:
            continue
        pu = B / (da[u] + B)
        q = (M[u] + M[:, u]).astype(float)
        q[r] = 0.0
        branch = not_r if q.sum() == 0 else q / q.sum()
        p += (w_r[u] / k) * (pu * not_r + (1 - pu) * branch)
    return p


def _merge_draws(o, r, draws, stream=3):
    counts = None
    for idx in range(draws):
        s, _ = o.merge_propose_one(r, idx, stream)
        if counts is None:
            counts = np.zeros(o.B, np.int64)
        counts[s] += 1
    return counts


def _chi2_ok(counts, p):
    from scipy.stats import chisquare
    n = counts.sum()
    zero = p == 0
    assert np.all(counts[zero] == 0), (counts, p)
    return chisquare(counts[~zero], p[~zero] * n).pvalue > 1e-4


def _graph_of(usap, edges, N):
    s = np.array([e[0] for e in edges], np.int32)
    d = np.array([e[1] for e in edges], np.int32)
    return usap.Graph.from_edges(s, d, N)


@pytest.mark.parametrize("block_size", [1024, 4])
def test_merge_proposal_law_random(orc, usap, block_size):
    """Chi-square of the proposed merge target against the exact law on a
    random labelled graph (dense-M and edge-list paths)."""
    g = random_graph(90, 500, seed=14)
    B = 7
    lab = np.random.default_rng(15).integers(0, B, g.N).astype(np.int32)
    o = _fresh(orc, g, lab, B, block_size=block_size)
    M = o.M_dense()
    for r in (0, 3, 6):
        counts = _merge_draws(o, r, 12000)
        assert _chi2_ok(counts, _merge_law(M, r)), (r, counts)


def test_merge_proposal_law_edge_cases(orc, usap):
    """Hand-built blocks for the branches of the law:
      block 0: no edges at all (k = 0): s uniform on all B blocks, r included;
      block 1: self-loops plus in-edges from block 4;
      block 2: its only neighbour block is 3, whose only neighbour is 2: the
               multinomial over u = 3 has no weight outside r, so s falls back
               to uniform-not-r (r + 1 doubled);
      block 4 (the last): a uniform-not-r draw that hits r wraps to block 0."""
    # nodes: 0,1 -> block 0 (isolated); 2,3 -> block 1 (self-loops);
    # 4 -> block 2, 5 -> block 3 (2 <-> 3 only); 6,7 -> block 4 with 4 <-> 1
    edges = [(2, 2), (3, 3), (2, 2), (4, 5), (5, 4), (4, 5), (6, 7), (7, 2), (6, 3)]
    g = _graph_of(usap, edges, 8)
    lab = np.array([0, 0, 1, 1, 2, 3, 4, 4], np.int32)
    B = 5
    o = _fresh(orc, g, lab, B)
    M = o.M_dense()
    for r in range(B):
        counts = _merge_draws(o, r, 8000)
        assert _chi2_ok(counts, _merge_law(M, r)), (r, counts, _merge_law(M, r))
    # block 0 (k = 0): every block, r included, equally likely
    p0 = _merge_law(M, 0)
    np.testing.assert_allclose(p0, np.full(B, 0.2))
    # block 2: u = 3 always; 3's weight outside 2 is 0 -> uniform-not-2 (3 doubled)
    np.testing.assert_allclose(_merge_law(M, 2), [0.2, 0.2, 0.0, 0.4, 0.2])


def test_merge_proposal_law_pure_selfloops_and_two_blocks(orc, usap):
    """A block whose only edges are self-loops (the multinomial over u = r
    excludes r and has weight 0, so s is uniform-not-r: r + 1 twice as likely),
    and B = 2 (every proposal of r with k > 0 is the other block)."""
    g = _graph_of(usap, [(0, 0), (1, 1), (2, 3), (3, 4), (4, 2), (5, 5), (2, 5)], 6)
    lab = np.array([0, 0, 1, 1, 1, 2], np.int32)
    o = _fresh(orc, g, lab, 3)
    M = o.M_dense()
    p = _merge_law(M, 0)
    np.testing.assert_allclose(p, [0.0, 2 / 3, 1 / 3])
    assert _chi2_ok(_merge_draws(o, 0, 9000), p)
    # B = 2: block 0 = {0, 1} with self-loops, block 1 the rest
    lab2 = np.array([0, 0, 1, 1, 1, 1], np.int32)
    o2 = _fresh(orc, g, lab2, 2)
    for r in (0, 1):
        counts = _merge_draws(o2, r, 500)
        assert counts[r] == 0 and counts[1 - r] == 500


def test_merge_proposal_law_distinguishes_mistakes(orc, usap):
    """The law above rejects the plausible misreadings of :774-865 on a graph
    built to separate them.  Blocks 0 = {0}, 1 = {1}, 2 = {2, 3} with edges
    1 -> 0 and 2 <-> 3: a proposal of r = 1 always picks u = 0, takes the
    uniform branch with probability 3 / (1 + 3), and otherwise finds u's only
    weight on r, so it always ends in uniform-not-r: s = 2 (r + 1, doubled) with
    probability 2/3 and s = 0 with 1/3.  A redraw instead of the r + 1 wrap
    gives 1/2 each; a multinomial that keeps r would propose r itself."""
    g = _graph_of(usap, [(1, 0), (2, 3), (3, 2)], 4)
    lab = np.array([0, 1, 2, 2], np.int32)
    o = _fresh(orc, g, lab, 3)
    M = o.M_dense()
    np.testing.assert_allclose(_merge_law(M, 1), [1 / 3, 0.0, 2 / 3])
    counts = _merge_draws(o, 1, 6000)
    assert _chi2_ok(counts, _merge_law(M, 1))
    from scipy.stats import chisquare
    assert chisquare(counts[[0, 2]], [3000, 3000]).pvalue < 1e-10
    assert counts[1] == 0


# ---------------------------------------------------------------------------
# nodal-phase loop, graph_parallel.hpp:381-478 (hand-derived scripts)
# ---------------------------------------------------------------------------

ONE = 1 << 48  # 1.0 in the fixed point


def _script(N=8, batch=2, isf=0, S_start=None, sums=None, **kw):
    from oracle import oracle as O
    return O.nodal_phase_scripted(N, batch, isf, S_start, sums, **kw)


def test_nodal_loop_converges_on_third_sweep():
    """4 batches of -1 per sweep: itr_delta_entropy = -1, -2, -3, -4 within a
    sweep (cumulative), -4 at its end.  No test on sweeps 0 and 1; on sweep 2
    mean = (-4 + -4) / 2 = -4 and after its first batch S = 10000 - 1 = 9999:
    4 < 5e-4 * 9999 = 4.9995 fires.  3 sweeps, 9 batches, 18 proposals."""
    st, itr = _script(S_start=[10000.0] * 3, sums=[-ONE] * 12)
    assert st["sweeps"] == 3 and st["batches"] == 9 and st["converged"] == 1
    assert st["proposals"] == 18
    assert list(itr[:3]) == [-4.0, -4.0, -1.0]
    assert st["S_last"] == 9999.0


def test_nodal_loop_cumulative_S_and_threshold_switch():
    """Same script with the golden-section bracket complete (isf): threshold
    1e-4, so 4 < 0.9999 fails and the sweeps run on.  Sweep 2 moves nothing
    (sums 0): on sweep 3 mean = (-4 + 0) / 2 = -2, 2 < 1e-4 * S fails; on
    sweep 4 mean = 0 and 0 < 1e-4 * S fires after its first batch.  S within a
    sweep adds the cumulative sum after every batch (S_start 10000, sums -1 x 4
    -> 10000 - 1 - 2 - 3 - 4 = 9990 after sweep 0)."""
    sums = [-ONE] * 8 + [0] * 12
    st, itr = _script(isf=1, S_start=[10000.0] * 5, sums=sums)
    assert st["sweeps"] == 5 and st["batches"] == 17 and st["converged"] == 1
    assert list(itr[:5]) == [-4.0, -4.0, 0.0, 0.0, 0.0]
    st0, _ = _script(isf=1, S_start=[10000.0], sums=[-ONE] * 4, max_nodal_itr=1)
    assert st0["S_last"] == 9990.0 and st0["converged"] == 0 and st0["batches"] == 4
    # without the bracket (threshold 5e-4) the same script stops on sweep 2
    st1, _ = _script(isf=0, S_start=[10000.0] * 5, sums=sums)
    assert st1["sweeps"] == 3 and st1["batches"] == 9


def test_nodal_loop_window_is_the_two_previous_sweeps():
    """Batch sums -25 x 4 in sweep 0 (sweep total -100), -0.25 x 4 after (total
    -1), S 10000, threshold 5e-4 (-> about 5).  Sweep 2: mean = (-100 - 1) / 2
    = -50.5 does not fire; sweep 3: mean = (-1 - 1) / 2 = -1 fires after its
    first batch: 4 sweeps, 13 batches.  A window that took the running sweep's
    (still zero) entry would fire on sweep 2; one reaching back three sweeps
    would still hold the -100 on sweep 3."""
    sums = [-25 * ONE] * 4 + [-ONE // 4] * 12
    st, itr = _script(S_start=[10000.0] * 4, sums=sums)
    assert list(itr[:4]) == [-100.0, -1.0, -1.0, -0.25]
    assert st["sweeps"] == 4 and st["batches"] == 13 and st["converged"] == 1
    assert st["S_last"] == 10000.0 - 0.25


def test_nodal_loop_minus_infinity():
    """An accepted -inf dS (ref mode) makes the sweep's cumulative sum -inf:
    S = -inf fails the test for the rest of that sweep, and the mean stays
    -inf while the sweep is in the window.  -inf in sweep 1, zeros after:
    sweeps 2 and 3 cannot fire (mean -inf), sweep 4 (mean 0) fires after its
    first batch: 5 sweeps, 17 batches."""
    sums = [-ONE] * 4 + [0, 0, -math.inf, 0] + [0] * 12
    st, itr = _script(S_start=[10000.0] * 5, sums=sums)
    assert itr[1] == -math.inf and itr[0] == -4.0
    assert st["sweeps"] == 5 and st["batches"] == 17 and st["converged"] == 1


def test_nodal_loop_ragged_last_batch_and_cap():
    """N = 7, batch 2: batches [0,2) [2,4) [4,6) [6,7), 7 proposals a sweep;
    nothing converges with S = 0 (0 < 0 is false): max_nodal_itr sweeps."""
    st, _ = _script(N=7, batch=2, S_start=[0.0] * 6, sums=[-ONE] * 24, max_nodal_itr=6)
    assert st["sweeps"] == 6 and st["batches"] == 24 and st["proposals"] == 42
    assert st["converged"] == 0


def test_nodal_loop_no_test_before_window():
    """Nothing moves at all: the test may first run on sweep 2 (itr >=
    window - 1), where mean 0 < 5e-4 * S fires: 3 sweeps, 9 batches; with a
    window of 4 it first runs on sweep 3."""
    st, _ = _script(S_start=[100.0] * 4, sums=[0] * 16)
    assert st["sweeps"] == 3 and st["batches"] == 9
    st4, _ = _script(S_start=[100.0] * 4, sums=[0] * 16, avg_window=4)
    assert st4["sweeps"] == 4 and st4["batches"] == 13
