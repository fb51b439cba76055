"""CPU tests that pin the oracle (the CPU restatement of the reference path).

Each oracle function is checked against (a) the golden vectors written by the
compiled reference (tests/golden/make_golden.py), (b) hand-derived answers for
the reference's 18-vertex fixture, and (c) an independent textbook algorithm
on seeded graphs (Kosaraju vs Tarjan, BFS-labelling vs union-find, queue BFS
vs direction-optimising BFS, Dijkstra vs delta-stepping).
"""
import numpy as np
import pytest

import girhip
import oracle
from conftest import load_golden

# ---- hand-derived answers for data/example_graph (brute force on 18 vertices) ----
EX_WCC = [0] * 12 + [12] * 6
EX_SCC = [0, 0, 0, 3, 3, 5, 6, 6, 6, 0, 0, 0, 12, 12, 12, 12, 16, 17]
EX_DAG = [(0, 3, 3), (0, 5, 1), (3, 5, 1), (3, 6, 1), (5, 6, 1), (12, 16, 1), (16, 17, 1)]
EX_BFS = {
    9: [2, 3, 1, 2, 3, 1, 2, 3, 4, 0, 1, 2, -1, -1, -1, -1, -1, -1],
    0: [0, 1, 2, 1, 2, 2, 3, 2, 3, 3, 4, 5, -1, -1, -1, -1, -1, -1],
    12: [-1] * 12 + [0, 1, 2, 3, 2, 3],
}


def test_example_graph_hand_answers():
    g = load_golden("example_graph")
    fb, fa, bb, ba = g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"]
    lab, n = oracle.wcc(fb, fa)
    assert lab.tolist() == EX_WCC and n == 2
    s, ns = oracle.scc(fb, fa)
    assert s.tolist() == EX_SCC and ns == 7
    src, dst, mult = oracle.dag(fb, fa, s)
    assert list(zip(src.tolist(), dst.tolist(), mult.tolist())) == EX_DAG
    # 26 arcs - 17 intra-SCC arcs = 9 condensation arcs counted with multiplicity
    assert mult.sum() == 9
    for x, want in EX_BFS.items():
        assert oracle.bfs(fb, fa, bb, ba, x).tolist() == want


def test_oracle_matches_reference_golden(golden):
    name, g = golden
    fb, fa, bb, ba = g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"]
    assert (oracle.wcc(fb, fa)[0] == g["wcc"]).all()
    s, _ = oracle.scc(fb, fa)
    assert (s == g["scc"]).all()
    src, dst, mult = oracle.dag(fb, fa, s)
    assert (np.stack([src, dst], 1) == g["dag"]).all()
    for i, x in enumerate(g["bfs_src"]):
        assert (oracle.bfs(fb, fa, bb, ba, x) == g["bfs"][i]).all(), f"{name} bfs {x}"
    for i, x in enumerate(g["sssp_src"]):
        assert (oracle.sssp(fb, fa, None, x, 16) == g["sssp"][i]).all(), f"{name} sssp {x}"
    w = [oracle.edge_weight(u, v) for u, v in g["pairs"]]
    assert w == g["weight"].tolist()


def test_oracle_batch_matches_reference_golden(golden):
    name, g = golden
    fb, fa = g["fw_beg"], g["fw_adj"]
    ops = g["ops"]
    nfb, nfa, nbb, nba = oracle.apply_batch(fb, fa, ops[:, 0], ops[:, 1], ops[:, 2])
    assert (oracle.wcc(nfb, nfa)[0] == g["post_wcc"]).all(), name
    s, _ = oracle.scc(nfb, nfa)
    assert (s == g["post_scc"]).all(), name
    src, dst, _ = oracle.dag(nfb, nfa, s)
    assert (np.stack([src, dst], 1) == g["post_dag"]).all(), name


def _graphs():
    yield "rmat", girhip.rmat_graph(11, edge_factor=8, seed=21)
    yield "grid", girhip.grid_graph(30, 30, seed=22)
    V = 2000
    s, d = girhip.uniform_edges(V, 3000, seed=23)
    yield "er", (V,) + girhip.csr_from_edges(V, s, d)


@pytest.mark.parametrize("name,G", list(_graphs()))
def test_oracle_vs_textbook(name, G):
    V, fb, fa, bb, ba = G
    assert (oracle.wcc(fb, fa)[0] == oracle.wcc_bfs(fb, fa, bb, ba)).all()
    assert (oracle.scc(fb, fa)[0] == oracle.scc_kosaraju(fb, fa, bb, ba)).all()
    src = girhip.sources(fb, 6, 5)
    w = girhip.weights(len(fa), 1, 64, 9)
    for x in src:
        assert (oracle.bfs(fb, fa, bb, ba, x) == oracle.bfs_textbook(fb, fa, x)).all()
    for x in src[:3]:
        assert (oracle.sssp(fb, fa, w, x, 16) == oracle.dijkstra(fb, fa, w, x)).all()
        assert (oracle.sssp(fb, fa, None, x, 7) == oracle.dijkstra(fb, fa, None, x)).all()


@pytest.mark.parametrize("name,G", list(_graphs()))
def test_oracle_ranks_and_dag_invariants(name, G):
    V, fb, fa, bb, ba = G
    s, ns = oracle.scc(fb, fa)
    src, dst, mult = oracle.dag(fb, fa, s)
    r = oracle.scc_ranks(fb, fa)
    assert (r[src] < r[dst]).all()
    assert ((r >= 0) == (s == np.arange(V))).all()
    # multiplicities add up to the crossing arcs
    rows = np.repeat(np.arange(V), np.diff(fb))
    assert mult.sum() == int((s[rows] != s[fa]).sum())
    assert (src != dst).all()


def test_normalize_batch_last_op_wins():
    rng = np.random.default_rng(3)
    V = 12
    s, d = girhip.uniform_edges(V, 30, seed=2)
    fb, fa, bb, ba = girhip.csr_from_edges(V, s, d)
    present = {(u, int(v)) for u in range(V) for v in fa[fb[u]:fb[u + 1]]}
    for trial in range(50):
        n = int(rng.integers(0, 40))
        u = rng.integers(-1, V + 1, n).astype(np.int32)
        v = rng.integers(-1, V + 1, n).astype(np.int32)
        ins = rng.integers(0, 2, n).astype(np.uint8)
        last = {}
        for a, b, c in zip(u.tolist(), v.tolist(), ins.tolist()):
            if a == b or min(a, b) < 0 or max(a, b) >= V:
                continue
            last[(a, b)] = c
        want_ins = sorted(k for k, c in last.items() if c and k not in present)
        want_del = sorted(k for k, c in last.items() if not c and k in present)
        want_noop = sum(1 for k, c in last.items() if c and k in present)
        (iu, iv), (du, dv), nn = oracle.normalize_batch(fb, fa, u, v, ins)
        assert list(zip(iu.tolist(), iv.tolist())) == want_ins
        assert list(zip(du.tolist(), dv.tolist())) == want_del
        assert nn == want_noop
        nfb, nfa, nbb, nba = oracle.apply_batch(fb, fa, u, v, ins)
        after = (present | set(want_ins)) - set(want_del)
        got = {(x, int(y)) for x in range(V) for y in nfa[nfb[x]:nfb[x + 1]]}
        assert got == after
        gotb = {(int(y), x) for x in range(V) for y in nba[nbb[x]:nbb[x + 1]]}
        assert gotb == after


def test_normalize_csr_sorts_and_dedups():
    beg = np.array([0, 4, 4, 7], np.int32)
    adj = np.array([2, 1, 2, 0, 1, 1, 0], np.int32)
    b, a = oracle.normalize_csr(beg, adj)
    assert b.tolist() == [0, 3, 3, 5] and a.tolist() == [0, 1, 2, 0, 1]


def test_planted_mistake_is_caught():
    # A wrong component label planted in the oracle's output must fail the
    # comparison the GPU tests use (canonical labels compared element-wise).
    g = load_golden("rmat10")
    lab, _ = oracle.scc(g["fw_beg"], g["fw_adj"])
    bad = lab.copy()
    i = int(np.nonzero(bad != np.arange(len(bad)))[0][0])
    bad[i] = i
    assert not (bad == g["scc"]).all()


def _sym_csr(fb, fa):
    """Symmetric closure (both directions, no self loops) as CSR."""
    V = len(fb) - 1
    rows = np.repeat(np.arange(V, dtype=np.int64), np.diff(fb))
    cols = fa.astype(np.int64)
    keep = rows != cols
    s = np.concatenate([rows[keep], cols[keep]])
    d = np.concatenate([cols[keep], rows[keep]])
    k = np.unique(s * V + d)
    beg = np.zeros(V + 1, np.int32)
    np.add.at(beg, (k // V).astype(np.int64) + 1, 1)
    return np.cumsum(beg).astype(np.int32), (k % V).astype(np.int32)


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_bfs_component_pinned(seed):
    # ParallelBFS harness (gir_test.cpp:143-343) restated: pinned by
    # (i) a single label reduces it to the textbook BFS, (ii) inside an SCC
    # every member is reached (visited = |SCC|, levels only on members),
    # (iii) undirected + WCC labels equals textbook BFS on the symmetric
    # closure and reaches exactly the WCC.
    V, fb, fa, bb, ba = girhip.rmat_graph(11, edge_factor=4, seed=seed)
    src = girhip.sources(fb, 6, seed)
    one = np.zeros(V, np.int32)
    scc = oracle.scc_kosaraju(fb, fa, bb, ba)
    wcc = oracle.wcc_bfs(fb, fa, bb, ba)
    sb, sa = _sym_csr(fb, fa)
    for x in src:
        lv, vis = oracle.bfs_component(fb, fa, bb, ba, one, x, False)
        assert (lv == oracle.bfs_textbook(fb, fa, x)).all() and vis == (lv >= 0).sum()
        lv, vis = oracle.bfs_component(fb, fa, bb, ba, scc, x, False)
        members = scc == scc[x]
        assert vis == members.sum() and ((lv >= 0) == members).all()
        free = oracle.bfs_textbook(fb, fa, x)
        assert (lv[members] >= free[members]).all()  # a restriction can only lengthen paths
        lv, vis = oracle.bfs_component(fb, fa, bb, ba, wcc, x, True)
        assert (lv == oracle.bfs_textbook(sb, sa, x)).all()
        assert vis == (wcc == wcc[x]).sum()
    lv, vis = oracle.bfs_component(fb, fa, bb, ba, scc, -1, False)
    assert vis == 0 and (lv == -1).all()


def test_golden_reachable_matches_textbook_bfs(golden):
    # GIR::reachable (gir_query.cpp:789-828) answers from the compiled
    # reference against plain BFS reachability
    name, g = golden
    fb, fa = g["fw_beg"], g["fw_adj"]
    V = len(fb) - 1
    for (u, v), r in zip(g["pairs"], g["reach"]):
        want = u == v or (0 <= u < V and oracle.bfs_textbook(fb, fa, int(u))[v] >= 0)
        assert bool(r) == want, (name, u, v)


def test_oracle_reach_info_pinned(golden):
    # reachableSccs / reachInfo (gir_query.cpp:56-88) restated: the mask
    # equals the compiled reference's reachableSccsPublic, and (textbook) an
    # SCC is reachable iff its minimum vertex is; pending counts are the
    # reachable DAG predecessors counted over the golden arc list; vertices
    # and arcs summed over the reachable SCCs.
    name, g = golden
    fb, fa, scc = g["fw_beg"], g["fw_adj"], g["scc"]
    dag = g["dag"]
    outdeg = np.diff(fb)
    for i, s in enumerate(g["bfs_src"]):
        c = int(scc[s])
        mask, pend, (ns, nv, ne) = oracle.reach_info(fb, fa, scc, c)
        assert (mask[scc] == g["rsccs"][i].astype(bool)).all(), name
        reached = oracle.bfs_textbook(fb, fa, int(s)) >= 0
        ids = np.nonzero(scc == np.arange(len(scc)))[0]
        assert (mask[ids] == reached[ids]).all() and not mask[scc != np.arange(len(scc))].any()
        want = np.zeros(len(scc), np.int32)
        for a, b in dag:
            if mask[a] and mask[b]:
                want[b] += 1
        assert (pend == want).all(), name
        on = mask[scc]
        assert (ns, nv, ne) == (int(mask.sum()), int(on.sum()), int(outdeg[on].sum()))
    bad = np.nonzero(scc != np.arange(len(scc)))[0]
    with pytest.raises(ValueError):  # not an SCC id
        oracle.reach_info(fb, fa, scc, int(bad[0]) if len(bad) else -1)


def test_oracle_two_level_pinned(golden):
    # bfsTwoLevel / ssspTwoLevel (gir_query.cpp:304-350, 604-722) restated:
    # equal to the compiled reference's results, to a textbook BFS and to
    # Dijkstra (hash weights, and explicit weights at several bucket widths)
    name, g = golden
    fb, fa, scc = g["fw_beg"], g["fw_adj"], g["scc"]
    INF = np.uint64(2**64 - 1)
    for i, s in enumerate(g["bfs_src"]):
        d = oracle.two_level(fb, fa, None, scc, int(s), unit=True)
        lv = np.where(d == INF, -1, d.astype(np.int64)).astype(np.int32)
        assert (lv == g["bfs2"][i]).all() and (lv == oracle.bfs_textbook(fb, fa, int(s))).all(), name
    for i, s in enumerate(g["sssp_src"]):
        d = oracle.two_level(fb, fa, None, scc, int(s), delta=16)
        assert (d == g["sssp2"][i]).all() and (d == oracle.dijkstra(fb, fa, None, int(s))).all(), name
    w = np.random.default_rng(7).integers(1, 2000, len(fa)).astype(np.uint32)
    for s in g["sssp_src"][:2]:
        want = oracle.dijkstra(fb, fa, w, int(s))
        for delta in (1, 37, 5000):
            assert (oracle.two_level(fb, fa, w, scc, int(s), delta=delta) == want).all(), (name, delta)
    assert (oracle.two_level(fb, fa, None, scc, -1, unit=True) == INF).all()


def _bc_by_definition(fb, fa, sources):
    """Betweenness from its definition: sum over sources s and targets t of
    sigma_sv * sigma_vt / sigma_st over the vertices v != s, t on shortest
    s-t paths (all-pairs BFS path counts; tiny graphs only)."""
    V = len(fb) - 1
    dist = np.full((V, V), -1, np.int64)
    sig = np.zeros((V, V))
    for a in range(V):
        dist[a, a], sig[a, a] = 0, 1
        q = [a]
        for u in q:
            for w in fa[fb[u]:fb[u + 1]]:
                if dist[a, w] < 0:
                    dist[a, w] = dist[a, u] + 1
                    q.append(w)
                if dist[a, w] == dist[a, u] + 1:
                    sig[a, w] += sig[a, u]
    sc = np.zeros(V)
    for s in sources:
        if not 0 <= s < V:
            continue
        for t in range(V):
            if t == s or dist[s, t] < 0:
                continue
            on = (dist[s] >= 0) & (dist[:, t] >= 0) & (dist[s] + dist[:, t] == dist[s, t])
            on[[s, t]] = False
            sc[on] += sig[s, on] * sig[on, t] / sig[s, t]
    return sc


@pytest.mark.parametrize("seed", [5, 6])
def test_oracle_bc_matches_definition(seed):
    # GIR::bc (gir_query.cpp:728-778) restated; pinned by the definition of
    # betweenness on tiny graphs with many equal-length paths
    V, fb, fa, bb, ba = girhip.grid_graph(6, 7, p_keep=0.9, p_bidir=0.6, seed=seed)
    src = list(range(0, V, 3)) + [-1, 4, 4]  # invalid skipped, repeats count twice
    got = oracle.bc(fb, fa, bb, ba, src)
    want = _bc_by_definition(fb, fa, src)
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)
    assert want.max() > 0


def test_oracle_bc_matches_reference_golden(golden):
    name, g = golden
    got = oracle.bc(g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"], g["bfs_src"])
    assert np.allclose(got, g["bc"], rtol=1e-12, atol=1e-9), name


def _nx_bicc(fb, fa, scc):
    """Blocks of every SCC's undirected view by networkx (its
    biconnected_components; vertices in no block on their own)."""
    nx = pytest.importorskip("networkx")
    V = len(fb) - 1
    G = nx.Graph()
    G.add_nodes_from(range(V))
    src = np.repeat(np.arange(V), np.diff(fb))
    for u, w in zip(src.tolist(), fa.tolist()):
        if u != w and scc[u] == scc[w]:
            G.add_edge(u, w)
    blocks = [tuple(sorted(c)) for c in nx.biconnected_components(G)]
    covered = {x for b in blocks for x in b}
    blocks += [(x,) for x in range(V) if x not in covered]
    return sorted(blocks)


@pytest.mark.parametrize("kind", ["rmat", "lattice", "er"])
def test_oracle_bicc_matches_networkx(kind):
    # blocks of every SCC's undirected view (gir_test.cpp:448-465,
    # dynamic_repair.h:96-170) restated; pinned by networkx's biconnected
    # components on the same views
    if kind == "rmat":
        V, fb, fa, bb, ba = girhip.rmat_graph(11, edge_factor=6, seed=101)
    elif kind == "lattice":
        V, fb, fa, bb, ba = girhip.grid_graph(30, 30, p_keep=0.9, p_bidir=0.5, seed=102)
    else:
        s, d = girhip.uniform_edges(2000, 5000, seed=103)
        fb, fa, bb, ba = girhip.csr_from_edges(2000, s, d)
    scc, _ = oracle.scc(fb, fa)
    got = oracle.bicc(fb, fa, bb, ba, scc)
    want = _nx_bicc(fb, fa, scc)
    assert got == want
    assert any(len(b) > 2 for b in want)
    if kind != "er":  # bridges (two-vertex blocks) too
        assert any(len(b) == 2 for b in want)


def test_oracle_bicc_matches_reference_golden(golden):
    name, g = golden
    got = oracle.bicc(g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"], g["scc"])
    assert got == oracle.partition(g["bicc"][:, 0], g["bicc"][:, 1]), name
