"""Parity at the benchmark's own sizes, bit for bit against the oracle:

- the directed RMAT scale-22 graph (4,194,304 vertices, 2^26 samples,
  a/b/c/d = .57/.19/.19/.05, seed 2) with uint32 weights uniform [1, 64]
  (seed 12), the 64 BFS sources (seed 7) and 8 SSSP sources (seed 8) that
  bench.py times;
- the 4800 x 4800 percolated lattice (23,040,000 vertices, p_keep .9,
  p_bidir .8, seed 3) with weights uniform [1, 1000].

Checked: canonical WCC and SCC labels, the condensation DAG with arc
multiplicities, the rank invariant rank(a) < rank(b) on every DAG arc
(gir.cpp:833-836), BFS levels (gir_query.cpp:108-294) and SSSP distances
(gir_query.cpp:629-703, explicit weights and the GIR::edgeWeight hash).
"""
import numpy as np
import pytest

import girhip
import oracle

pytestmark = pytest.mark.gpu

SEED_W, SEED_BFS, SEED_SSSP = 12, 7, 8


def say(*a):
    print("[case]", *a, flush=True)


def layers_match(G, fb, fa, tag):
    lab, n = G.wcc()
    olab, on = oracle.wcc(fb, fa)
    assert n == on and np.array_equal(lab, olab), f"wcc {tag}"
    s, ns, nd = G.scc_dag()
    os_, ons = oracle.scc(fb, fa)
    assert ns == ons and np.array_equal(s, os_), f"scc {tag}"
    src, dst, mult = G.dag()
    osrc, odst, omult = oracle.dag(fb, fa, os_)
    assert nd == len(osrc), f"dag size {tag}"
    assert np.array_equal(src, osrc) and np.array_equal(dst, odst), f"dag arcs {tag}"
    assert np.array_equal(mult, omult), f"dag multiplicities {tag}"
    # multiplicities add up to the arcs between distinct SCCs
    rows = np.repeat(np.arange(len(fb) - 1, dtype=np.int32), np.diff(fb))
    assert int(mult.astype(np.int64).sum()) == int((s[rows] != s[fa]).sum()), f"dag total {tag}"
    r = G.ranks()
    reps = s == np.arange(len(s))
    assert np.array_equal(r >= 0, reps), f"rank reps {tag}"
    assert (r[src] < r[dst]).all(), f"rank order {tag}"
    assert len(np.unique(r[reps])) == int(reps.sum()), f"ranks distinct {tag}"


@pytest.fixture(scope="module")
def rmat22():
    say("generate rmat22 seed 2")
    V, fb, fa, bb, ba = girhip.rmat_graph(22, 16, 0.57, 0.19, 0.19, 2)
    w = girhip.weights(len(fa), 1, 64, SEED_W)
    G = girhip.Graph(fb, fa, bb, ba)
    yield dict(V=V, fb=fb, fa=fa, bb=bb, ba=ba, w=w, G=G)
    G.close()


@pytest.fixture(scope="module")
def grid4800():
    say("generate grid4800 seed 3")
    V, fb, fa, bb, ba = girhip.grid_graph(4800, 4800, 0.9, 0.8, 3)
    w = girhip.weights(len(fa), 1, 1000, SEED_W)
    G = girhip.Graph(fb, fa, bb, ba)
    yield dict(V=V, fb=fb, fa=fa, bb=bb, ba=ba, w=w, G=G)
    G.close()


def test_rmat22_layers(rmat22):
    g = rmat22
    say("rmat22 layers V", g["V"], "E", len(g["fa"]))
    assert g["V"] == 1 << 22
    layers_match(g["G"], g["fb"], g["fa"], "rmat22")


def test_rmat22_bfs_64_sources(rmat22):
    g = rmat22
    src = girhip.sources(g["fb"], 64, SEED_BFS)
    say("rmat22 bfs 64 sources")
    lv, m = g["G"].bfs(src, edges=True)
    deg = np.diff(g["fb"]).astype(np.int64)
    for i, x in enumerate(src):
        want = oracle.bfs(g["fb"], g["fa"], g["bb"], g["ba"], int(x))
        assert np.array_equal(lv[i], want), f"bfs source {x}"
        assert m[i] == int(deg[want >= 0].sum()), f"traversed arcs source {x}"


def test_rmat22_sssp_8_sources(rmat22):
    g = rmat22
    src = girhip.sources(g["fb"], 8, SEED_SSSP)
    G = g["G"]
    say("rmat22 sssp 8 sources weights [1,64] delta 16")
    G.set_weights(g["w"])
    try:
        d = G.sssp(src, 16)
        for i, x in enumerate(src):
            assert np.array_equal(d[i], oracle.sssp(g["fb"], g["fa"], g["w"], int(x), 16)), f"sssp source {x}"
        # one source against textbook Dijkstra as well
        assert np.array_equal(d[0], oracle.dijkstra(g["fb"], g["fa"], g["w"], int(src[0])))
    finally:
        G.set_weights(None)
    say("rmat22 sssp hash weights delta 16")
    d = G.sssp(src[:2], 16)
    for i, x in enumerate(src[:2]):
        assert np.array_equal(d[i], oracle.sssp(g["fb"], g["fa"], None, int(x), 16)), f"sssp hash source {x}"


def test_grid4800_layers(grid4800):
    g = grid4800
    say("grid4800 layers V", g["V"], "E", len(g["fa"]))
    assert g["V"] == 4800 * 4800
    layers_match(g["G"], g["fb"], g["fa"], "grid4800")


def test_grid4800_bfs_sssp(grid4800):
    g = grid4800
    G = g["G"]
    src = girhip.sources(g["fb"], 4, SEED_BFS)
    say("grid4800 bfs 4 sources")
    lv = G.bfs(src)
    for i, x in enumerate(src):
        assert np.array_equal(lv[i], oracle.bfs(g["fb"], g["fa"], g["bb"], g["ba"], int(x))), f"bfs source {x}"
    ss = girhip.sources(g["fb"], 2, SEED_SSSP)
    say("grid4800 sssp 2 sources weights [1,1000] delta 2048")
    G.set_weights(g["w"])
    try:
        d = G.sssp(ss, 2048)
        for i, x in enumerate(ss):
            assert np.array_equal(d[i], oracle.sssp(g["fb"], g["fa"], g["w"], int(x), 16)), f"sssp source {x}"
    finally:
        G.set_weights(None)
