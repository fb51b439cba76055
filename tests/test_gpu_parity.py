"""GPU parity tests: the gfx950 library through its C ABI against the oracle
and the reference's golden vectors.  Integer results must be bit-exact
(canonical component ids, BFS levels, SSSP distances, DAG arcs with
multiplicities)."""
import numpy as np
import pytest

import girhip
import oracle
from conftest import load_golden

pytestmark = pytest.mark.gpu
INF = np.iinfo(np.uint64).max


def say(*a):
    print("[case]", *a, flush=True)


def check_layers(G, fb, fa, bb, ba, tag="", rebuild=True):
    """WCC, SCC, DAG (+mult), ranks of G against the oracle on (fb, fa).
    rebuild=False reads the maintained layers (after updates) instead of
    building them again."""
    lab, n = G.wcc() if rebuild else G.get_wcc()
    olab, on = oracle.wcc(fb, fa)
    assert n == on and (lab == olab).all(), f"wcc {tag}"
    s, ns, nd = G.scc_dag() if rebuild else G.get_scc()
    os_, ons = oracle.scc(fb, fa)
    assert ns == ons and (s == os_).all(), f"scc {tag}"
    src, dst, mult = G.dag()
    osrc, odst, omult = oracle.dag(fb, fa, os_)
    assert nd == len(osrc)
    assert (src == osrc).all() and (dst == odst).all() and (mult == omult).all(), f"dag {tag}"
    r = G.ranks()
    reps = s == np.arange(len(s))
    assert ((r >= 0) == reps).all(), f"rank reps {tag}"
    assert (r[src] < r[dst]).all(), f"rank order {tag}"
    assert len(np.unique(r[reps])) == reps.sum()
    return s


def test_golden_cases(golden):
    name, g = golden
    fb, fa, bb, ba = g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"]
    say("golden", name, "V", len(fb) - 1, "E", len(fa))
    with girhip.Graph(fb, fa, bb, ba) as G:
        lab, _ = G.wcc()
        assert (lab == g["wcc"]).all()
        s, _, _ = G.scc_dag()
        assert (s == g["scc"]).all()
        src, dst, _ = G.dag()
        assert (np.stack([src, dst], 1) == g["dag"]).all()
        say("golden bfs", name)
        assert (G.bfs(g["bfs_src"]) == g["bfs"]).all()
        say("golden sssp", name)
        assert (G.sssp(g["sssp_src"], 16) == g["sssp"]).all()
        say("golden reachable", name)
        assert (G.reachable(g["pairs"][:, 0], g["pairs"][:, 1]) == g["reach"].astype(bool)).all()
        say("golden two-level", name)
        assert (G.bfs_two_level(g["bfs_src"]) == g["bfs2"]).all()
        assert (G.sssp_two_level(g["sssp_src"], 16) == g["sssp2"]).all()
        for i, x in enumerate(g["bfs_src"]):
            ri = G.reach_info(int(s[x]))
            assert (ri["mask"][s] == g["rsccs"][i].astype(bool)).all()
            m, p, cnt = oracle.reach_info(fb, fa, g["scc"], int(s[x]))
            assert (ri["mask"] == m).all() and (ri["pending"] == p).all()
            assert (ri["sccs"], ri["verts"], ri["edges"]) == cnt
        say("golden bc", name)
        assert np.allclose(G.bc(g["bfs_src"]), g["bc"], rtol=1e-9, atol=1e-9)
        say("golden batch", name)
        ops = g["ops"]
        G.apply_batch(ops[:, 0], ops[:, 1], ops[:, 2])
        assert (G.get_wcc()[0] == g["post_wcc"]).all()
        s, _, _ = G.get_scc()
        assert (s == g["post_scc"]).all()
        src, dst, _ = G.dag()
        assert (np.stack([src, dst], 1) == g["post_dag"]).all()


def test_example_graph_targeted_split():
    # gir_test.cpp:704-753: insert 3->12 merges the WCCs, deleting it splits
    # them again, deleting 9->2 splits SCC {0,1,2,9,10,11}.
    g = load_golden("example_graph")
    say("example targeted")
    with girhip.Graph(g["fw_beg"], g["fw_adj"]) as G:
        G.wcc()
        G.scc_dag()
        G.update_edge_directed(3, 12, True)
        assert G.get_wcc()[1] == 1
        G.update_edge_directed(3, 12, False)
        lab, n = G.get_wcc()
        assert n == 2 and lab.tolist() == [0] * 12 + [12] * 6
        G.update_edge_directed(9, 2, False)
        s, ns, _ = G.get_scc()
        assert s[[0, 1, 2]].tolist() == [0, 0, 0] and s[[9, 10, 11]].tolist() == [9, 9, 9]


@pytest.mark.parametrize("scale,ef,seed", [(10, 16, 31), (12, 8, 32), (16, 16, 1)])
def test_rmat_parity(scale, ef, seed):
    V, fb, fa, bb, ba = girhip.rmat_graph(scale, edge_factor=ef, seed=seed)
    say("rmat", scale, ef, seed, "E", len(fa))
    with girhip.Graph(fb, fa, bb, ba) as G:
        check_layers(G, fb, fa, bb, ba, f"rmat{scale}")
        src = girhip.sources(fb, 64 if scale == 16 else 16, seed + 100)
        say("rmat bfs", scale)
        lv, m = G.bfs(src, edges=True)
        for i, x in enumerate(src):
            want = oracle.bfs(fb, fa, bb, ba, x)
            assert (lv[i] == want).all(), f"bfs src {x}"
            assert m[i] == int(np.diff(fb)[want >= 0].sum())
        say("rmat bfs single-source engine", scale)
        G.set_param("msbfs", 0)
        assert (G.bfs(src[:4]) == lv[:4]).all()
        G.set_param("msbfs", 1)
        w = girhip.weights(len(fa), 1, 64, seed + 200)
        ss = src[:8]
        say("rmat sssp hash", scale)
        d = G.sssp(ss, 16)
        for i, x in enumerate(ss):
            assert (d[i] == oracle.sssp(fb, fa, None, x, 16)).all(), f"sssp hash {x}"
        say("rmat sssp weights", scale)
        G.set_weights(w)
        d, m = G.sssp(ss, 16, edges=True)
        for i, x in enumerate(ss):
            want = oracle.dijkstra(fb, fa, w, x)
            assert (d[i] == want).all(), f"sssp w {x}"
            assert m[i] == int(np.diff(fb)[want != INF].sum())


def test_grid_parity():
    V, fb, fa, bb, ba = girhip.grid_graph(120, 150, seed=33)
    say("grid 120x150 E", len(fa))
    with girhip.Graph(fb, fa, bb, ba) as G:
        check_layers(G, fb, fa, bb, ba, "grid")
        src = girhip.sources(fb, 4, 5)
        lv = G.bfs(src)
        for i, x in enumerate(src):
            assert (lv[i] == oracle.bfs_textbook(fb, fa, x)).all()
        w = girhip.weights(len(fa), 1, 1000, 6)
        G.set_weights(w)
        d = G.sssp(src, 0)
        for i, x in enumerate(src):
            assert (d[i] == oracle.dijkstra(fb, fa, w, x)).all()
        # the persistent cooperative kernels (reach levels of the SCC sweep,
        # SSSP rounds with and without a heavy phase) give the same results
        say("grid persistent kernels")
        G.set_param("persist", 1)
        s0 = G.get_scc()[0].copy()
        assert (G.scc_dag()[0] == s0).all()
        for delta in (0, 100, 2000):
            assert (G.sssp(src, delta) == d).all(), delta
        assert (G.bfs(src) == lv).all()


def test_sssp_one_workgroup_rounds():
    # the per-source rounds with small queues run one workgroup per source
    # and hand back to the grid kernels when a queue outgrows the workgroup
    # engine (and take over again when the queues shrink): a lattice (small
    # buckets throughout, stored and hash weights, 1..5 sources, deltas from
    # 1 to past the largest weight), a random graph whose middle rounds
    # outgrow it, and the SCC-restricted two-level rounds; against Dijkstra,
    # with the engine on and off
    V, fb, fa, bb, ba = girhip.grid_graph(90, 110, seed=64)
    src = girhip.sources(fb, 5, 65)
    w = girhip.weights(len(fa), 1, 1000, 66)
    want = [oracle.dijkstra(fb, fa, w, int(x)) for x in src]
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        G.set_param("sssp_interleave", 0)
        for tiny in (1, 0):
            G.set_param("sssp_tiny", tiny)
            for k, delta in ((1, 16), (4, 1), (5, 16), (3, 700), (2, 5000)):
                say("lattice sssp tiny", tiny, "k", k, "delta", delta)
                d = G.sssp(src[:k], delta)
                for i in range(k):
                    assert (d[i] == want[i]).all(), (tiny, k, delta, i)
        G.set_weights(None)
        hw = np.array([oracle.edge_weight(u, x) for u, x in zip(np.repeat(np.arange(V), np.diff(fb)), fa)],
                      np.uint32)
        G.set_param("sssp_tiny", 1)
        d = G.sssp(src[:4], 16)
        for i in range(4):
            assert (d[i] == oracle.dijkstra(fb, fa, hw, int(src[i]))).all(), ("hash", i)
        say("lattice two-level")
        G.scc_dag()
        assert (G.sssp_two_level(src[:2], 16) == d[:2]).all()
    # random graph: wide middle rounds
    u, v = girhip.uniform_edges(60000, 480000, seed=67)
    fb, fa, bb, ba = girhip.csr_from_edges(60000, u, v)
    w = girhip.weights(len(fa), 1, 64, 68)
    src = girhip.sources(fb, 3, 69)
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        G.set_param("sssp_interleave", 0)
        for tiny in (1, 0):
            G.set_param("sssp_tiny", tiny)
            say("uniform sssp tiny", tiny)
            d = G.sssp(src, 8)
            for i, x in enumerate(src):
                assert (d[i] == oracle.dijkstra(fb, fa, w, int(x))).all(), (tiny, i)


def test_bfs_one_workgroup_levels():
    # BFS on graphs without long rows walks each source's levels in one
    # workgroup; a frontier that outgrows it hands the source to the grid
    # engine: a lattice (thousands of levels), a random graph whose middle
    # levels are wide, invalid and repeated sources; against textbook BFS,
    # with the engine on and off
    V, fb, fa, bb, ba = girhip.grid_graph(70, 300, seed=71)
    src = np.concatenate([girhip.sources(fb, 5, 72), [0, 0, -2, V]]).astype(np.int32)
    u, v = girhip.uniform_edges(50000, 400000, seed=73)
    gb, ga, gbb, gba = girhip.csr_from_edges(50000, u, v)
    gsrc = girhip.sources(gb, 3, 74)
    for tiny in (1, 0):
        with girhip.Graph(fb, fa, bb, ba) as G:
            G.set_param("bfs_tiny", tiny)
            say("lattice bfs tiny", tiny)
            lv, m = G.bfs(src, edges=True)
            for i, x in enumerate(src):
                want = oracle.bfs_textbook(fb, fa, x) if 0 <= x < V else np.full(V, -1, np.int32)
                assert (lv[i] == want).all(), (tiny, i)
                assert m[i] == int(np.diff(fb)[want >= 0].sum())
            assert (G.bfs(src[:1]) == lv[:1]).all()
            assert (G.bfs(src[:2]) == lv[:2]).all()
        with girhip.Graph(gb, ga, gbb, gba) as G:
            G.set_param("bfs_tiny", tiny)
            say("uniform bfs tiny", tiny)
            lv = G.bfs(gsrc)
            for i, x in enumerate(gsrc):
                assert (lv[i] == oracle.bfs_textbook(gb, ga, x)).all(), (tiny, i)


def test_bfs_bit_parallel_dense_level_runs():
    # a ring of 12 layers, each layer's vertices all pointing into the next:
    # with sources spread over the layers every level's union frontier spans
    # most of V, so there are more dense levels than the finish keeps words
    # for (the rest write their levels at once); a second graph has dense
    # levels, then sparse ones, then dense ones again (only the first run is
    # kept).  Against textbook BFS per source.
    rng = np.random.default_rng(81)
    Lr, W = 12, 48
    layer = lambda i: np.arange(i * W, (i + 1) * W)
    u = np.concatenate([np.repeat(layer(i), W) for i in range(Lr)])
    v = np.concatenate([np.tile(layer((i + 1) % Lr), W) for i in range(Lr)])
    keep = rng.random(len(u)) < 0.5
    V = Lr * W
    fb, fa, bb, ba = girhip.csr_from_edges(V, u[keep], v[keep])
    src = np.concatenate([rng.integers(0, V, 40), [V - 1]]).astype(np.int32)
    # dense-sparse-dense: a wide fan, a long thin path, a second wide fan
    F, P = 600, 40
    hub2 = F + P
    u2 = np.concatenate([np.zeros(F, np.int64), np.arange(1, F + 1), np.arange(F + 1, hub2),
                         np.full(F, hub2), np.arange(hub2 + 1, hub2 + F + 1)])
    v2 = np.concatenate([np.arange(1, F + 1), np.full(F, F + 1), np.arange(F + 2, hub2 + 1),
                         np.arange(hub2 + 1, hub2 + F + 1), np.full(F, 0)])
    V2 = hub2 + F + 1
    gb, ga, gbb, gba = girhip.csr_from_edges(V2, u2, v2)
    src2 = np.array([0, 1, 2, F + 3, hub2, hub2 + 1], np.int32)
    for fbv, fav, bbv, bav, ss, n in ((fb, fa, bb, ba, src, V), (gb, ga, gbb, gba, src2, V2)):
        with girhip.Graph(fbv, fav, bbv, bav) as G:
            G.set_param("bfs_tiny", 0)
            say("dense-level runs V", n, "sources", len(ss))
            lv = G.bfs(ss)
            for i, x in enumerate(ss):
                assert (lv[i] == oracle.bfs_textbook(fbv, fav, x)).all(), (n, i)


def test_bfs_many_levels_bit_parallel():
    # long thin grid: hundreds of tiny levels -> sparse queue mode, chunked
    # levels and the counter ring wrapping; duplicate and invalid sources
    V, fb, fa, bb, ba = girhip.grid_graph(40, 700, seed=35)
    say("grid 40x700 msbfs E", len(fa))
    src = np.concatenate([girhip.sources(fb, 60, 7), [0, 0, V - 1, -3]]).astype(np.int32)
    with girhip.Graph(fb, fa, bb, ba) as G:
        lv, m = G.bfs(src, edges=True)
        for i, x in enumerate(src):
            want = oracle.bfs_textbook(fb, fa, x) if 0 <= x < V else np.full(V, -1, np.int32)
            assert (lv[i] == want).all(), f"bfs src {x}"
            assert m[i] == int(np.diff(fb)[want >= 0].sum())
        assert lv.max() > 300
        # same answers with the per-source engine and with pull forced early
        G.set_param("msbfs", 0)
        assert (G.bfs(src[:3]) == lv[:3]).all()
        G.set_param("msbfs", 1)
        # persistent cooperative kernel, and one level per readback
        G.set_param("persist", 1)
        assert (G.bfs(src) == lv).all()
        G.set_param("persist", 0)
        G.set_param("chunk", 1)
        assert (G.bfs(src) == lv).all()
        G.set_param("chunk", 16)
        G.set_param("bfs_alpha", 1)
        G.set_param("bfs_min_reach", 0)
        assert (G.bfs(src) == lv).all()


def test_bfs_sources_of_small_reach():
    # uniform random sources (many with no out-arcs or a tiny reach) beside
    # giant-component ones in one 64-source sweep: the pull levels stop
    # scanning once a vertex holds every source still expanding
    V, fb, fa, bb, ba = girhip.rmat_graph(14, seed=37)
    rng = np.random.default_rng(5)
    src = np.concatenate([rng.integers(0, V, 32).astype(np.int32), girhip.sources(fb, 32, 38)])
    say("bfs small-reach sources", V)
    with girhip.Graph(fb, fa, bb, ba) as G:
        lv = G.bfs(src)
        for i, x in enumerate(src):
            assert (lv[i] == oracle.bfs_textbook(fb, fa, int(x))).all(), f"bfs src {x}"


def test_sssp_split_views_across_deltas_weights_and_updates():
    # the per-delta split views (two kept, least recently used evicted) must
    # follow delta changes, weight changes and store changes
    V, fb, fa, bb, ba = girhip.rmat_graph(12, seed=39)
    src = girhip.sources(fb, 3, 40)
    w = girhip.weights(len(fa), 1, 64, 41)
    say("split views", V)
    with girhip.Graph(fb, fa, bb, ba) as G:
        want_hash = [oracle.sssp(fb, fa, None, int(x), 16) for x in src]
        for delta in (16, 32, 64, 16, 1, 100000):  # 64 evicts 16; 16 is rebuilt
            d = G.sssp(src, delta)
            for i in range(len(src)):
                assert (d[i] == want_hash[i]).all(), f"hash delta {delta}"
        G.set_weights(w)
        want_w = [oracle.dijkstra(fb, fa, w, int(x)) for x in src]
        for delta in (16, 8, 16):
            d = G.sssp(src, delta)
            for i in range(len(src)):
                assert (d[i] == want_w[i]).all(), f"weights delta {delta}"
        # a store change invalidates the views; inserted arcs carry the hash weight
        u = np.array([int(src[0])], np.int32)
        v = np.array([(int(src[0]) + 1) % V], np.int32)
        G.apply_batch(u, v, np.ones(1, np.uint8))
        fb2, fa2, _, _ = G.csr()
        rows = np.repeat(np.arange(V), np.diff(fb))
        wmap = {(int(a), int(b)): int(c) for a, b, c in zip(rows, fa, w)}
        rows2 = np.repeat(np.arange(V), np.diff(fb2))
        w2 = np.array([wmap.get((int(a), int(b)), oracle.edge_weight(int(a), int(b))) for a, b in zip(rows2, fa2)],
                      np.uint32)
        d = G.sssp(src, 16)
        for i, x in enumerate(src):
            assert (d[i] == oracle.dijkstra(fb2, fa2, w2, int(x))).all(), "after update"


def test_many_heavy_rows_in_one_wave():
    # 48 hubs with ids 0..47 (one wave of row owners) and 2,400-3,200 out- and
    # in-arcs each: every pass walks their rows through the heavy list, whose
    # slots and arc offsets a wave reserves with one packed atomic
    rng = np.random.default_rng(42)
    V = 20000
    s_bg = rng.integers(0, V, 60000)
    d_bg = rng.integers(0, V, 60000)
    hubs = np.arange(48)
    deg = rng.integers(2400, 3200, len(hubs))
    s_out = np.repeat(hubs, deg)
    d_out = rng.integers(0, V, deg.sum())
    s_in = rng.integers(0, V, deg.sum())
    d_in = np.repeat(hubs, deg)
    src = np.concatenate([s_bg, s_out, s_in]).astype(np.int32)
    dst = np.concatenate([d_bg, d_out, d_in]).astype(np.int32)
    fb, fa, bb, ba = girhip.csr_from_edges(V, src, dst)
    say("many heavy rows: V", V, "E", len(fa), "max row", int(np.diff(fb).max()))
    assert int(np.diff(fb)[:48].min()) > 2048
    with girhip.Graph(fb, fa, bb, ba) as G:
        check_layers(G, fb, fa, bb, ba, "heavy")
        srcs = girhip.sources(fb, 64, 43)
        lv = G.bfs(srcs)
        for i in range(0, 64, 9):
            assert (lv[i] == oracle.bfs_textbook(fb, fa, int(srcs[i]))).all(), f"bfs {srcs[i]}"
        w = girhip.weights(len(fa), 1, 64, 44)
        G.set_weights(w)
        d = G.sssp(srcs[:8], 16)
        for i in range(8):
            assert (d[i] == oracle.dijkstra(fb, fa, w, int(srcs[i]))).all(), f"sssp {srcs[i]}"


def test_tuning_knobs_do_not_change_results():
    V, fb, fa, bb, ba = girhip.rmat_graph(12, seed=34)
    say("knobs rmat12")
    src = girhip.sources(fb, 4, 1)
    with girhip.Graph(fb, fa, bb, ba) as G:
        base = (G.wcc()[0], G.scc_dag()[0], G.bfs(src), G.sssp(src, 16))
        for k, v in [("msbfs", 0), ("block", 64), ("block", 1024), ("bfs_alpha", 1), ("bfs_beta", 1000),
                     ("bfs_min_reach", 0),
                     ("wcc_rounds", 0), ("scc_trim_rounds", 0), ("scc_pivot", 0), ("sssp_delta", 1),
                     ("persist", 1), ("chunk", 1), ("sssp_dist32", 0), ("xcd", 1), ("split", 0), ("scc_block", 256), ("scc_block", 1024),
                     ("sssp_delta", 100000), ("sssp_interleave", 0), ("sssp_interleave", 1),
                     ("sssp_ms_batch", 3), ("sssp_ms_batch", 32), ("scc_trim2", 1), ("sssp_unroll", 1),
                     ("sssp_unroll", 2), ("bfs_alpha_auto", 0), ("bfs_alpha", 1000), ("update_bypass", 0),
                     ("scc_cert", 0), ("sssp_tiny", 0), ("bfs_tiny", 0), ("scc_pair", 0)]:
            say("knob", k, v)
            old = G.get_param(k)
            G.set_param(k, v)
            got = (G.wcc()[0], G.scc_dag()[0], G.bfs(src), G.sssp(src, 0 if k == "sssp_delta" else 16))
            for a, b in zip(base, got):
                assert (a == b).all(), (k, v)
            G.set_param(k, old)


def test_sssp_interleaved_batch_sizes():
    # the shared-bucket engine at every interleave width (k -> 1, 2, 4, 8,
    # 16, 32 distances per vertex; 33 and 40 sources split into even
    # batches), repeated and invalid sources, split views on and off, and
    # 64-bit distances, against Dijkstra
    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=38)
    w = girhip.weights(len(fa), 1, 64, 39)
    base = girhip.sources(fb, 40, 40)
    want = {int(x): oracle.dijkstra(fb, fa, w, int(x)) for x in base}
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        G.set_param("sssp_interleave", 1)
        G.set_param("sssp_ms_batch", 32)
        for k in (1, 2, 3, 5, 8, 13, 32, 33, 40):
            src = base[:k].copy()
            if k >= 3:
                src[1] = src[0]  # the same vertex twice in one batch
                src[2] = -1 if k % 2 else V  # out of range: all unreachable
            for split, unroll in ((1, 4), (0, 2), (1, 1)):
                say("interleaved sssp k", k, "split", split, "unroll", unroll)
                G.set_param("split", split)
                G.set_param("sssp_unroll", unroll)
                d = G.sssp(src, 7)
                for i, x in enumerate(src):
                    ref = want[int(x)] if 0 <= x < V else np.full(V, INF, np.uint64)
                    assert (d[i] == ref).all(), (k, split, i)
        G.set_param("split", 1)
        # weights near 2^31: 32-bit distances overflow, the batch is redone in 64 bits
        big = (w.astype(np.uint64) << 25).clip(1, 2**32 - 1).astype(np.uint32)
        G.set_weights(big)
        say("interleaved sssp 64-bit")
        d = G.sssp(base[:8], 1 << 30)
        for i, x in enumerate(base[:8]):
            assert (d[i] == oracle.dijkstra(fb, fa, big, int(x))).all()


def test_trim2_pairs():
    # trim-2 (trim_2_3.h:161-308): a chain of 2-cycles (pair i -> pair i+1)
    # is peeled one pair per trim round from its end (out-direction rule)
    # and, reversed, from its start (in-direction rule); mixed with a random
    # graph.  Labels equal the oracle's with and without trim-2.
    n = 300
    a = np.arange(n)
    src = np.concatenate([2 * a, 2 * a + 1, 2 * a[:-1] + 1])
    dst = np.concatenate([2 * a + 1, 2 * a, 2 * a[1:]])
    V0 = 2 * n
    rs, rd = girhip.uniform_edges(4000, 12000, seed=44)
    src = np.concatenate([src, rs + V0, np.array([V0 - 1])])
    dst = np.concatenate([dst, rd + V0, np.array([V0 + 5])])
    V = V0 + 4000
    for flip in (False, True):
        s_, d_ = (dst, src) if flip else (src, dst)
        fb, fa, bb, ba = girhip.csr_from_edges(V, s_, d_)
        want, nw = oracle.scc(fb, fa)
        say("trim2 chain flip", flip)
        with girhip.Graph(fb, fa, bb, ba) as G:
            for t2, rounds in ((0, 3), (1, 3), (1, 1000), (1, 0)):
                G.set_param("scc_trim2", t2)
                G.set_param("scc_trim_rounds", rounds)
                s, ns, _ = G.scc_dag()
                assert ns == nw and np.array_equal(s, want), (flip, t2, rounds)


@pytest.mark.parametrize("scale", [9, 18])
def test_unsorted_duplicate_input_is_normalised(scale):
    # (scale 18: 4.6 M arcs with duplicates, so the device sort runs over
    # thousands of tiles; the undirected closure then sorts twice as many)
    V, fb, fa, bb, ba = girhip.rmat_graph(scale, seed=35)
    rng = np.random.default_rng(0)
    # shuffle every row and duplicate some arcs (GIR normalises: gir.cpp:95-130)
    rows = np.repeat(np.arange(V), np.diff(fb))
    dup = rng.random(len(fa)) < 0.2
    r2 = np.concatenate([rows, rows[dup]])
    a2 = np.concatenate([fa, fa[dup]])
    perm = rng.permutation(len(r2))
    r2, a2 = r2[perm], a2[perm]
    order = np.argsort(r2, kind="stable")
    nfb = np.zeros(V + 1, np.int32)
    np.add.at(nfb, r2 + 1, 1)
    nfb = np.cumsum(nfb).astype(np.int32)
    nfa = a2[order].astype(np.int32)
    say("normalise rmat", scale)
    with girhip.Graph(nfb, nfa) as G:
        assert G.E == len(fa)
        gfb, gfa, gbb, gba = G.csr()
        assert (gfb == fb).all() and (gfa == fa).all() and (gbb == bb).all() and (gba == ba).all()
        if scale > 9:
            wl, _ = G.wcc()
            src = girhip.sources(fb, 2, 36)
            lv, vis = G.bfs_component(src, undirected=True)
            for i, x in enumerate(src):
                want, nv = oracle.bfs_component(fb, fa, bb, ba, wl, int(x), True)
                assert np.array_equal(lv[i], want) and vis[i] == nv, "undirected"


def test_bad_inputs_are_refused():
    say("bad inputs")
    with pytest.raises(girhip.GirError):
        girhip.Graph(np.array([0, 2, 1], np.int32), np.array([1, 0], np.int32))  # offsets not monotone
    with pytest.raises(girhip.GirError):
        girhip.Graph(np.array([0, 1, 2], np.int32), np.array([1, 7], np.int32))  # id out of range
    with pytest.raises(girhip.GirError):  # bw not the transpose
        girhip.Graph(np.array([0, 1, 1], np.int32), np.array([1], np.int32), np.array([0, 1, 1], np.int32),
                     np.array([1], np.int32))


def test_edge_cases():
    say("edge cases")
    # no arcs
    with girhip.Graph(np.zeros(6, np.int32), np.zeros(0, np.int32)) as G:
        assert G.wcc()[1] == 5 and G.scc_dag()[1] == 5 and len(G.dag()[0]) == 0
        assert (G.bfs([2]) == np.array([[-1, -1, 0, -1, -1]])).all()
        assert G.sssp([4])[0].tolist() == [INF] * 4 + [0]
    # self loops only (kept by normalisation, no effect on components)
    with girhip.Graph(np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32)) as G:
        assert G.wcc()[1] == 2 and G.scc_dag()[1] == 2 and len(G.dag()[0]) == 0
    # out-of-range sources -> unreachable rows (gir_query.cpp:112, 699)
    V, fb, fa, bb, ba = girhip.rmat_graph(6, seed=1)
    with girhip.Graph(fb, fa) as G:
        lv = G.bfs([-1, V, 0])
        assert (lv[0] == -1).all() and (lv[1] == -1).all() and lv[2][0] == 0
        d = G.sssp([V + 10])
        assert (d == INF).all()
        # invalid ops are dropped and counted
        st = G.apply_batch([1, 2, -1, V], [1, 3, 0, 0], [1, 1, 1, 1])
        assert st["invalid"] == 3


def _random_batch(rng, fb, fa, n, V):
    rows = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
    k = n // 2
    e = rng.integers(0, len(fa), k)
    u = np.concatenate([rows[e], rng.integers(0, V, n - k)]).astype(np.int32)
    v = np.concatenate([fa[e], rng.integers(0, V, n - k)]).astype(np.int32)
    ins = np.concatenate([np.zeros(k, np.uint8), np.ones(n - k, np.uint8)])
    p = rng.permutation(n)
    return u[p], v[p], ins[p]


def test_batch_sequence_matches_recompute():
    V, fb, fa, bb, ba = girhip.rmat_graph(12, seed=36)
    rng = np.random.default_rng(7)
    w = girhip.weights(len(fa), 1, 64, 8)
    src = girhip.sources(fb, 3, 9)
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        G.wcc()
        G.scc_dag()
        wmap = {(int(a), int(b)): int(c) for a, b, c in zip(np.repeat(np.arange(V), np.diff(fb)), fa, w)}
        for step, n in enumerate([1, 10, 1000, 5000, 300]):
            u, v, ins = _random_batch(rng, fb, fa, n, V)
            if step == 2:
                ins[:] = 1  # insert-only batch (incremental WCC hooking)
            say("batch step", step, "n", n)
            (iu, iv), (du, dv), nn = oracle.normalize_batch(fb, fa, u, v, ins)
            st = G.apply_batch(u, v, ins)
            assert st["inserted"] == len(iu) and st["deleted"] == len(du) and st["noop"] == nn
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            gfb, gfa, gbb, gba = G.csr()
            assert (gfb == fb).all() and (gfa == fa).all() and (gbb == bb).all() and (gba == ba).all()
            check_layers(G, fb, fa, bb, ba, f"step{step}", rebuild=False)
            # weights follow their arcs; inserted arcs get GIR::edgeWeight
            for a, b in zip(iu.tolist(), iv.tolist()):
                wmap[(a, b)] = oracle.edge_weight(a, b)
            wv = np.array([wmap[(int(a), int(b))] for a, b in zip(np.repeat(np.arange(V), np.diff(fb)), fa)],
                          np.uint32)
            d = G.sssp(src, 16)
            for i, x in enumerate(src):
                assert (d[i] == oracle.dijkstra(fb, fa, wv, x)).all()
            lv = G.bfs(src)
            for i, x in enumerate(src):
                assert (lv[i] == oracle.bfs_textbook(fb, fa, x)).all()


def test_incremental_scc_repair_paths():
    # the repair of the SCC layer after a batch (classified changes instead of
    # a rebuild): each path against the oracle, and the path the library took
    # ("batch_scc_path": 1 multiplicities only, 2 re-rank, 3 relabel the SCCs
    # that may change)
    g = load_golden("example_graph")
    fb, fa = g["fw_beg"], g["fw_adj"]
    V = len(fb) - 1
    with girhip.Graph(fb, fa) as G:
        G.set_param("update_bypass", 0)  # the paths below, not the bypass test first
        G.set_param("scc_cert", 0)       # nor the big SCC's certificate (own test below)
        G.wcc()
        s0, _, _ = G.scc_dag()
        r = G.ranks()
        rk = lambda x: r[s0[x]]
        present = set(zip(np.repeat(np.arange(V), np.diff(fb)).tolist(), fa.tolist()))
        # a forward pair between SCCs in different WCCs (no path either way)
        pairs = [(a, b) for a in range(V) for b in range(V) if s0[a] != s0[b] and (a, b) not in present
                 and ((a < 12) != (b < 12))]
        fwd = next((a, b) for a, b in pairs if rk(a) < rk(b))
        back = next((a, b) for a, b in pairs if rk(a) > rk(b))
        steps = [("insert inside an SCC", (0, 2, 1), 1), ("insert in rank order", fwd + (1,), 1),
                 ("delete between SCCs", (3, 5, 0), 1), ("insert against rank order, no cycle", back + (1,), 2),
                 ("insert closing a cycle", (6, 0, 1), 3), ("delete inside an SCC", (9, 2, 0), 3)]
        for what, (u, v, ins), path in steps:
            say("repair path:", what)
            G.apply_batch([u], [v], [ins])
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, [u], [v], [ins])
            check_layers(G, fb, fa, bb, ba, what, rebuild=False)
            assert G.last_ms("batch_scc_path") == path, (what, G.last_ms("batch_scc_path"))
    # random batches on a graph with a large SCC: same layers with the repair
    # switched off; inner deletes that keep the SCC together take path 3 too
    V, fb, fa, bb, ba = girhip.rmat_graph(12, seed=37)
    rng = np.random.default_rng(38)
    with girhip.Graph(fb, fa, bb, ba) as G, girhip.Graph(fb, fa, bb, ba) as H:
        H.set_param("scc_incremental", 0)
        for X in (G, H):
            X.wcc()
            X.scc_dag()
        for n in (1, 1, 5, 50, 2000):
            u, v, ins = _random_batch(rng, fb, fa, n, V)
            say("repair vs rebuild n", n)
            G.apply_batch(u, v, ins)
            H.apply_batch(u, v, ins)
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            check_layers(G, fb, fa, bb, ba, f"repair n={n}", rebuild=False)
            assert np.array_equal(G.get_scc()[0], H.get_scc()[0])
            assert H.last_ms("batch_scc_path") == 4


def _giant(s):
    lab, cnt = np.unique(s, return_counts=True)
    return int(lab[np.argmax(cnt)])


def test_big_scc_certificate():
    # the big SCC across batches through its spanning trees (no relabelling of
    # the giant; the build records them, and a batch rebuilds them when they
    # are absent): random mixed batches, deletes that strand members (every
    # out-arc or every in-arc of a member), inserts that pull outside SCCs
    # into it, and its smallest vertex leaving it (the canonical label moves);
    # every step against the oracle and against a rebuild
    V, fb, fa, bb, ba = girhip.rmat_graph(14, seed=71)
    rng = np.random.default_rng(72)
    paths = {}
    drops = joins = 0
    with girhip.Graph(fb, fa, bb, ba) as G, girhip.Graph(fb, fa, bb, ba) as H:
        H.set_param("scc_incremental", 0)
        for X in (G, H):
            X.wcc()
            X.scc_dag()

        def step(u, v, ins, tag):
            nonlocal fb, fa, bb, ba, drops, joins
            u, v, ins = (np.asarray(x) for x in (u, v, ins))
            say("certificate step", tag, "ops", len(u))
            G.apply_batch(u.astype(np.int32), v.astype(np.int32), ins.astype(np.uint8))
            H.apply_batch(u.astype(np.int32), v.astype(np.int32), ins.astype(np.uint8))
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            s = check_layers(G, fb, fa, bb, ba, tag, rebuild=False)
            assert np.array_equal(s, H.get_scc()[0]), tag
            c = int(G.last_ms("batch_scc_cert"))
            paths[c] = paths.get(c, 0) + 1
            if c:
                drops += int(G.last_ms("batch_scc_dropped"))
                joins += int(G.last_ms("batch_scc_joined"))
            return s

        s = G.get_scc()[0]
        for i in range(6):
            if i == 3:  # one batch without the trees: the next one builds them from the pivot
                G.set_param("scc_cert", 0)
            step(*_random_batch(rng, fb, fa, 200, V), f"random {i}")
            G.set_param("scc_cert", 1)
        for i in range(4):  # strand a member of the giant: all its out-arcs (or in-arcs) go
            s = G.get_scc()[0]
            gl = _giant(s)
            mem = np.nonzero(s == gl)[0]
            x = int(mem[rng.integers(0, len(mem))])
            if i % 2 == 0:
                nb = fa[fb[x]:fb[x + 1]]
                step(np.full(len(nb), x), nb, np.zeros(len(nb)), f"strand out {x}")
            else:
                nb = ba[bb[x]:bb[x + 1]]
                step(nb, np.full(len(nb), x), np.zeros(len(nb)), f"strand in {x}")
        for i in range(3):  # outside vertices pulled in: x reached from the giant, y reaching it
            s = G.get_scc()[0]
            gl = _giant(s)
            g0 = int(np.nonzero(s == gl)[0][0])
            fwd = oracle.bfs_textbook(fb, fa, g0) >= 0
            bwd = oracle.bfs_textbook(bb, ba, g0) >= 0
            xs = np.nonzero(fwd & ~bwd)[0]
            ys = np.nonzero(bwd & ~fwd)[0]
            if len(xs) == 0 or len(ys) == 0:
                continue
            k = min(5, len(xs), len(ys))
            step(xs[:k], ys[:k], np.ones(k), f"merge {k}")
        # a member cut off entirely (its WCC splits too: the WCC layer is rebuilt)
        s = G.get_scc()[0]
        mem = np.nonzero(s == _giant(s))[0]
        x = int(mem[len(mem) // 2])
        outs, ins_ = fa[fb[x]:fb[x + 1]], ba[bb[x]:bb[x + 1]]
        step(np.concatenate([np.full(len(outs), x), ins_]), np.concatenate([outs, np.full(len(ins_), x)]),
             np.zeros(len(outs) + len(ins_)), f"isolate {x}")
        assert G.last_ms("batch_wcc_unsafe") > 0
        # the giant's smallest vertex leaves it
        s = G.get_scc()[0]
        gl = _giant(s)
        nb = fa[fb[gl]:fb[gl + 1]]
        s = step(np.full(len(nb), gl), nb, np.zeros(len(nb)), f"label vertex {gl} leaves")
        assert s[gl] != _giant(s)
        for i in range(4):
            step(*_random_batch(rng, fb, fa, 300, V), f"random after {i}")
    assert paths.get(1, 0) > 0 and paths.get(2, 0) > 0, paths
    assert drops > 0 and joins > 0, (drops, joins)
    # a lattice: deep trees (hundreds of levels), strands and random batches
    V, fb, fa, bb, ba = girhip.grid_graph(120, 120, 0.9, 0.8, 5)
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.wcc()
        G.scc_dag()
        for i in range(5):
            u, v, ins = _random_batch(rng, fb, fa, 60, V)
            say("certificate lattice step", i)
            G.apply_batch(u, v, ins)
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            check_layers(G, fb, fa, bb, ba, f"lattice {i}", rebuild=False)


def test_delete_bypass():
    # deletes whose tail still reaches its head (a path of at most four arcs)
    # change no partition: the layers are repaired without relabelling; the
    # others fall back to the general repair.  Every step against the oracle.
    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=61)
    rng = np.random.default_rng(62)
    nby = 0
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.wcc()
        G.scc_dag()
        for step in range(24):
            rows = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
            k = 1 if step < 12 else int(rng.integers(2, 9))
            e = rng.choice(len(fa), k, replace=False)
            u, v = rows[e], fa[e]
            ins = np.zeros(k, np.uint8)
            if step >= 18:  # mixed with inserts
                u = np.concatenate([u, rng.integers(0, V, 5).astype(np.int32)])
                v = np.concatenate([v, rng.integers(0, V, 5).astype(np.int32)])
                ins = np.concatenate([ins, np.ones(5, np.uint8)])
            say("bypass step", step, "deletes", k)
            G.apply_batch(u, v, ins)
            nby += int(G.last_ms("batch_all_bypassed"))
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            check_layers(G, fb, fa, bb, ba, f"bypass step {step}", rebuild=False)
        assert nby > 0
    # a 2-cycle and an SCC {2, 3, 4} with a chord: deleting the chord 2 -> 4
    # keeps the path 2 -> 3 -> 4 (bypass); deleting 0 -> 1 splits the 2-cycle
    # (no bypass); WCCs checked as well
    fb2 = np.array([0, 1, 2, 4, 5, 6, 6], np.int32)  # 0->1, 1->0, 2->3, 2->4, 3->4, 4->2
    fa2 = np.array([1, 0, 3, 4, 4, 2], np.int32)
    with girhip.Graph(fb2, fa2) as G:
        G.wcc()
        G.scc_dag()
        f1, a1 = fb2, fa2
        for (u, v), want in (((2, 4), 1), ((0, 1), 0)):
            say("bypass cycle delete", u, v)
            G.apply_batch([u], [v], [0])
            assert G.last_ms("batch_all_bypassed") == want
            f1, a1, b1, c1 = oracle.apply_batch(f1, a1, [u], [v], [0])
            check_layers(G, f1, a1, b1, c1, f"cycle delete {u}->{v}", rebuild=False)


def test_ingest_then_sync():
    V, fb, fa, bb, ba = girhip.rmat_graph(11, seed=37)
    rng = np.random.default_rng(11)
    with girhip.Graph(fb, fa, bb, ba) as G:
        lab0, _ = G.wcc()
        s0, _, _ = G.scc_dag()
        u, v, ins = _random_batch(rng, fb, fa, 2000, V)
        say("ingest")
        G.ingest_batch(u, v, ins)
        assert not G.synced()
        nfb, nfa, nbb, nba = oracle.apply_batch(fb, fa, u, v, ins)
        # the store sees the change at once (BFS reads the store)
        assert (G.bfs([int(girhip.sources(nfb, 1, 2)[0])])[0] ==
                oracle.bfs_textbook(nfb, nfa, int(girhip.sources(nfb, 1, 2)[0]))).all()
        assert (G.get_wcc()[0] == lab0).all()  # layers lag until sync (gir.h:107-116)
        assert (G.get_scc()[0] == s0).all()
        G.sync()
        assert G.synced()
        check_layers(G, nfb, nfa, nbb, nba, "sync", rebuild=False)


def test_bfs_in_component_harness():
    # ParallelBFS::testParallelBFS (gir_test.cpp:143-343): BFS restricted to
    # the source's SCC (directed) / WCC (undirected) against the oracle
    for V, fb, fa, bb, ba in [girhip.rmat_graph(12, edge_factor=4, seed=36),
                              girhip.grid_graph(60, 90, seed=37)]:
        say("component bfs V", V)
        src = np.concatenate([girhip.sources(fb, 6, 2), [-1]]).astype(np.int32)
        with girhip.Graph(fb, fa, bb, ba) as G:
            with pytest.raises(girhip.GirError):
                G.bfs_component(src[:1])  # SCC layer not built yet
            scc, _, _ = G.scc_dag()
            wcc, _ = G.wcc()
            for und, lab in [(False, scc), (True, wcc)]:
                lv, vis = G.bfs_component(src, undirected=und)
                for i, x in enumerate(src):
                    want, wv = oracle.bfs_component(fb, fa, bb, ba, lab, x, und)
                    assert (lv[i] == want).all() and vis[i] == wv, (und, x)


def test_sssp_distances_beyond_32_bits():
    # weights near 2^30: distances pass 2^32 after a few hops, so the 32-bit
    # pass overflows and the batch is redone with 64-bit distances
    V, fb, fa, bb, ba = girhip.grid_graph(30, 40, seed=38)
    w = girhip.weights(len(fa), 1 << 29, (1 << 30) + 12345, 39)
    src = girhip.sources(fb, 3, 40)
    say("sssp 64-bit redo")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        d = G.sssp(src, 1 << 30)
        for i, x in enumerate(src):
            want = oracle.dijkstra(fb, fa, w, x)
            assert (d[i] == want).all()
        assert d[d != INF].max() > (1 << 32)
        # a small delta against these weights: the per-source rounds in one
        # workgroup per source, redone with 64-bit distances as well
        say("sssp 64-bit redo, one-workgroup rounds")
        assert (G.sssp(src, 1 << 26) == d).all()


def test_reachable_queries():
    # GIR::reachable (gir_query.cpp:789-828) on many pairs: trivial answers
    # (same SCC, rank order), DAG sweeps of 64 pairs, invalid ids, and the
    # sync of pending ingested changes before answering
    V, fb, fa, bb, ba = girhip.rmat_graph(13, edge_factor=3, seed=41)
    rng = np.random.default_rng(42)
    u = rng.integers(0, V, 700).astype(np.int32)
    v = rng.integers(0, V, 700).astype(np.int32)
    u[:3] = [-1, 5, V]
    v[:3] = [2, 5, 1]
    say("reachable rmat13")
    with girhip.Graph(fb, fa, bb, ba) as G:
        with pytest.raises(girhip.GirError):
            G.reachable(u[:1], v[:1])  # SCC layer not built yet
        G.scc_dag()
        got = G.reachable(u, v)
        lv = {}
        for i, (x, y) in enumerate(zip(u, v)):
            if not (0 <= x < V and 0 <= y < V):
                assert not got[i]
                continue
            if x not in lv:
                lv[x] = oracle.bfs_textbook(fb, fa, int(x))
            assert got[i] == (lv[x][y] >= 0), (x, y)
        assert 0 < got.sum() < len(got)
        # ingest an arc that makes a previously unreachable pair reachable;
        # reachable() syncs it first
        i = int(np.nonzero(~got[3:])[0][0]) + 3
        G.ingest_batch([u[i]], [v[i]], [1])
        assert G.reachable(u[i:i + 1], v[i:i + 1])[0]


@pytest.mark.gpu
def test_reachable_at_scale_both_paths():
    # GIR::reachable (gir_query.cpp:789-828) on a power-law graph whose giant
    # SCC has DAG rows far longer than a local search expands: answers of the
    # per-pair LDS search (with its fallback to sweeps) and of the sweeps
    # alone both equal reachability from the oracle's BFS
    V, fb, fa, bb, ba = girhip.rmat_graph(17, edge_factor=16, seed=43)
    rng = np.random.default_rng(44)
    srcs = rng.integers(0, V, 24).astype(np.int32)
    lv = {int(s): oracle.bfs_textbook(fb, fa, int(s)) for s in srcs}
    u, v = [], []
    for s in srcs:
        reached = np.nonzero(lv[int(s)] >= 0)[0]
        u.append(np.full(256, s, np.int32))
        v.append(np.concatenate([rng.integers(0, V, 128), rng.choice(reached, 128)]).astype(np.int32))
    u = np.concatenate(u)
    v = np.concatenate(v)
    perm = rng.permutation(len(u))
    u, v = u[perm], v[perm]
    want = np.array([lv[int(x)][y] >= 0 for x, y in zip(u, v)])
    say("reachable rmat17 ef16")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.scc_dag()
        ds, dd = G.dag()[:2]
        deg = max(np.bincount(ds).max(), np.bincount(dd).max())
        assert deg > 2048  # rows the local search sets aside
        got = G.reachable(u, v)
        assert np.array_equal(got, want)
        G.set_param("reach_local", 0)
        assert np.array_equal(G.reachable(u, v), want)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["rmat16", "lattice_oneway", "lattice_mixed"])
def test_two_level_queries_at_scale(kind):
    # bfsTwoLevel / ssspTwoLevel (gir_query.cpp:304-350, 604-722) and
    # reachInfo (56-88) on a power-law graph (giant SCC, DAG rows longer than
    # a wave's share) and on lattices with deep condensations (every wave
    # small): levels equal the oracle's textbook BFS, distances its Dijkstra
    # (hash weights, then explicit weights at two bucket widths)
    if kind == "rmat16":
        V, fb, fa, bb, ba = girhip.rmat_graph(16, edge_factor=16, seed=51)
    else:
        V, fb, fa, bb, ba = girhip.grid_graph(150, 150, p_keep=0.9, p_bidir=0.0 if kind == "lattice_oneway" else 0.3,
                                              seed=53)
    srcs = girhip.sources(fb, 3, 52)
    say("two-level", kind, "V", V, "E", len(fa))
    with girhip.Graph(fb, fa, bb, ba) as G:
        scc, _, _ = G.scc_dag()
        lv = G.bfs_two_level(srcs)
        d = G.sssp_two_level(srcs, 16)
        for i, x in enumerate(srcs):
            assert (lv[i] == oracle.bfs_textbook(fb, fa, int(x))).all(), (kind, x)
            assert (d[i] == oracle.dijkstra(fb, fa, None, int(x))).all(), (kind, x)
        ri = G.reach_info(int(scc[srcs[0]]))
        m, p, cnt = oracle.reach_info(fb, fa, scc, int(scc[srcs[0]]))
        assert (ri["mask"] == m).all() and (ri["pending"] == p).all()
        assert (ri["sccs"], ri["verts"], ri["edges"]) == cnt
        w = np.random.default_rng(54).integers(1, 5000, len(fa)).astype(np.uint32)
        G.set_weights(w)
        want = oracle.dijkstra(fb, fa, w, int(srcs[0]))
        for delta in (64, 4096):
            assert (G.sssp_two_level(srcs[:1], delta)[0] == want).all(), (kind, delta)


@pytest.mark.gpu
def test_two_level_edge_cases():
    # 64-bit redo when 32-bit distances would overflow, out-of-range
    # sources, calls before the SCC layer, and the sync of ingested changes
    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=55)
    say("two-level edge cases")
    with girhip.Graph(fb, fa, bb, ba) as G:
        with pytest.raises(girhip.GirError):
            G.bfs_two_level([0])
        scc, _, _ = G.scc_dag()
        with pytest.raises(girhip.GirError):  # not an SCC id
            G.reach_info(int(np.nonzero(scc != np.arange(V))[0][0]))
        x = int(girhip.sources(fb, 1, 56)[0])
        w = np.full(len(fa), 2**31 - 1, np.uint32)
        G.set_weights(w)
        d = G.sssp_two_level([x], 2**31)
        want = oracle.dijkstra(fb, fa, w, x)
        assert (d[0] == want).all()
        assert want[want != INF].max() > 2**32  # beyond 32 bits: the 64-bit redo ran
        G.set_weights(None)
        assert (G.bfs_two_level([-1, V])[:, :] == -1).all()
        assert (G.sssp_two_level([-1], 16) == INF).all()
        lv = G.bfs_two_level([x])[0]
        y = int(np.nonzero(lv < 0)[0][0])
        G.ingest_batch([x], [y], [1])
        fb2, fa2 = oracle.apply_batch(fb, fa, [x], [y], [1])[:2]
        assert (G.bfs_two_level([x])[0] == oracle.bfs_textbook(fb2, fa2, x)).all()


@pytest.mark.gpu
def test_sssp_graph_replay_matches_eager():
    # a lattice at delta 1 runs hundreds of chunks of rounds: with the
    # "graph" knob the full chunks replay captured HIP graphs (many replays,
    # readbacks between them, a second call re-captures); distances equal
    # the eager launches' and Dijkstra's
    V, fb, fa, bb, ba = girhip.grid_graph(120, 120, seed=61)
    srcs = girhip.sources(fb, 3, 62)
    say("sssp graph replay lattice 120x120")
    with girhip.Graph(fb, fa, bb, ba) as G:
        want = [oracle.dijkstra(fb, fa, None, int(x)) for x in srcs]
        for graph in (1, 0, 1):
            G.set_param("graph", graph)
            d = G.sssp(srcs, 1)
            for i in range(len(srcs)):
                assert (d[i] == want[i]).all(), (graph, i)
        G.set_param("graph", 1)
        G.scc_dag()
        d2 = G.sssp_two_level(srcs[:1], 1)
        assert (d2[0] == want[0]).all()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["rmat16", "lattice"])
def test_bc_matches_oracle(kind):
    # GIR::bc (gir_query.cpp:728-778): Brandes over more sources than one
    # batch holds, with an invalid and a repeated source, on a power-law
    # graph (hub rows through the heavy list) and a lattice (hundreds of
    # levels, many equal-length paths); fp64 sums in another order than the
    # oracle's, so within 1e-9 relative
    if kind == "rmat16":
        V, fb, fa, bb, ba = girhip.rmat_graph(16, edge_factor=16, seed=81)
    else:
        V, fb, fa, bb, ba = girhip.grid_graph(200, 200, p_keep=0.9, p_bidir=0.5, seed=82)
    src = list(girhip.sources(fb, 20, 83)) + [-1, V + 5]
    src += src[:2]
    say("bc", kind, "V", V, "sources", len(src))
    with girhip.Graph(fb, fa, bb, ba) as G:
        got = G.bc(src)
    want = oracle.bc(fb, fa, bb, ba, src)
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9), float(np.max(np.abs(got - want) / (1 + np.abs(want))))
    assert want.max() > 0


def _sp_arcs(fb, fa, d, wfun):
    """Arcs (u, v) on shortest paths (d[v] == d[u] + w(u, v)) - the arcs whose
    deletion can invalidate a dependence subtree."""
    V = len(fb) - 1
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(fb))
    ok = (d[src] != INF) & (d[fa] != INF)
    w = wfun(src, fa)
    on = ok & (d[fa].astype(np.float64) == d[src].astype(np.float64) + w)
    return src[on].astype(np.int32), fa[on].astype(np.int32)


@pytest.mark.gpu
@pytest.mark.parametrize("weighted", [0, 1])
def test_incremental_matches_from_scratch(weighted):
    # GIR::trackIncremental / updateIncremental / incDist
    # (gir_query.cpp:867-996): after every update the corrected distances
    # equal a from-scratch BFS / Dijkstra (the reference's own check,
    # gir_stream.cpp:54-67); batches delete shortest-path arcs (dependence
    # subtrees invalidated), insert random arcs, go through apply, ingest
    # and single updates, and updates may pile up before a correction
    V, fb, fa, bb, ba = girhip.rmat_graph(14, edge_factor=8, seed=91)
    rng = np.random.default_rng(92)
    x = int(girhip.sources(fb, 1, 93)[0])
    wf = (lambda u, v: np.array([oracle.edge_weight(int(a), int(b)) for a, b in zip(u, v)], np.float64)) \
        if weighted else (lambda u, v: np.ones(len(u)))
    ref = (lambda f, a: oracle.dijkstra(f, a, None, x)) if weighted else \
        (lambda f, a: np.where(oracle.bfs_textbook(f, a, x) < 0, INF, oracle.bfs_textbook(f, a, x)).astype(np.uint64))
    say("incremental", "sssp" if weighted else "bfs", "rmat14")
    with girhip.Graph(fb, fa, bb, ba) as G:
        with pytest.raises(girhip.GirError):
            G.update_incremental(weighted)
        G.track_incremental(x, weighted)
        assert (G.inc_dist(weighted) == ref(fb, fa)).all()
        f1, a1 = fb, fa
        for r in range(6):
            d = ref(f1, a1)
            su, sv = _sp_arcs(f1, a1, d, wf)
            pick = rng.choice(len(su), size=min(len(su), 300), replace=False)
            u = np.concatenate([su[pick], rng.integers(0, V, 300).astype(np.int32)])
            v = np.concatenate([sv[pick], rng.integers(0, V, 300).astype(np.int32)])
            ins = np.concatenate([np.zeros(len(pick), np.uint8), np.ones(300, np.uint8)])
            if r % 3 == 0:
                G.apply_batch(u, v, ins)
            elif r % 3 == 1:
                G.ingest_batch(u, v, ins)
            else:
                for i in range(0, 40):
                    G.update_edge_directed(int(u[i]), int(v[i]), bool(ins[i]))
                u, v, ins = u[:40], v[:40], ins[:40]
            f1, a1 = oracle.apply_batch(f1, a1, u, v, ins)[:2]
            if r == 3:
                continue  # two batches before one correction
            G.update_incremental(weighted)
            assert (G.inc_dist(weighted) == ref(f1, a1)).all(), r


@pytest.mark.gpu
def test_incremental_weight_changes():
    # A tracked SSSP follows weight changes: set_weights re-tracks it, and an
    # arc on a shortest path deleted in one batch and inserted again with a
    # larger explicit weight in the next, both before one correction, leaves
    # no head with its old, too small distance.
    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=96)
    x = int(girhip.sources(fb, 1, 97)[0])
    w = girhip.weights(len(fa), 1, 64, 98)
    rows = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
    say("incremental weight changes")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.track_incremental(x, 1)
        assert (G.inc_dist(1) == oracle.dijkstra(fb, fa, None, x)).all()
        G.set_weights(w)  # hash weights -> explicit ones
        d = oracle.dijkstra(fb, fa, w, x)
        assert (G.inc_dist(1) == d).all()
        # tree arcs: heads with exactly one shortest-path in-arc
        wmap = dict(zip(zip(rows.tolist(), fa.tolist()), w.tolist()))
        su, sv = _sp_arcs(fb, fa, d, lambda u, v: np.array([wmap[(int(a), int(b))] for a, b in zip(u, v)], float))
        heads, cnt = np.unique(sv, return_counts=True)
        only = np.isin(sv, heads[cnt == 1]) & (sv != x)
        su, sv = su[only][:50], sv[only][:50]
        assert len(su) > 0
        G.apply_batch(su, sv, np.zeros(len(su), np.uint8))
        G.apply_batch(su, sv, np.ones(len(su), np.uint8), w=np.full(len(su), 1000, np.uint32))
        for a, b in zip(su.tolist(), sv.tolist()):
            wmap[(a, b)] = 1000
        G.update_incremental(1)
        wv = np.array([wmap[(int(a), int(b))] for a, b in zip(rows, fa)], np.uint32)
        assert (G.inc_dist(1) == oracle.dijkstra(fb, fa, wv, x)).all()
        G.set_weights(None)  # back to the hash: tracked again
        assert (G.inc_dist(1) == oracle.dijkstra(fb, fa, None, x)).all()


@pytest.mark.gpu
def test_incremental_edge_cases():
    # out-of-range source (everything unreachable, updates keep it so), the
    # 32-bit packed distance limit, and both kinds tracked at once
    V, fb, fa, bb, ba = girhip.rmat_graph(11, edge_factor=6, seed=94)
    say("incremental edge cases")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.track_incremental(-1, 0)
        G.apply_batch([1, 2], [3, 4], [1, 1])
        G.update_incremental(0)
        assert (G.inc_dist(0) == INF).all()
        x = int(girhip.sources(fb, 1, 95)[0])
        G.track_incremental(x, 0)
        G.track_incremental(x, 1)
        G.apply_batch([x], [int(np.nonzero(oracle.bfs_textbook(fb, fa, x) < 0)[0][0])], [1])
        f1, a1 = oracle.apply_batch(fb, fa, [x], [int(np.nonzero(oracle.bfs_textbook(fb, fa, x) < 0)[0][0])], [1])[:2]
        G.update_incremental(0)
        G.update_incremental(1)
        lv = oracle.bfs_textbook(f1, a1, x)
        assert (G.inc_dist(0) == np.where(lv < 0, INF, lv).astype(np.uint64)).all()
        assert (G.inc_dist(1) == oracle.dijkstra(f1, a1, None, x)).all()
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(np.full(len(fa), 2**31 - 1, np.uint32))
        with pytest.raises(girhip.GirError):
            G.track_incremental(x, 1)  # distances beyond 32 bits


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["golden", "rmat16", "lattice"])
def test_bicc_matches_oracle(kind):
    # blocks of every SCC's undirected view (buildBCT / getBCTComponents /
    # articulationPoint, gir.cpp:2063-2110, 4140-4149): the same vertex-set
    # partition as the oracle (itself pinned to the compiled reference and
    # networkx), articulation points = vertices in two or more blocks
    cases = []
    if kind == "golden":
        for name in ("example_graph", "rmat10", "rmat12", "grid40", "er3000"):
            g = load_golden(name)
            cases.append((name, g["fw_beg"], g["fw_adj"], g["bw_beg"], g["bw_adj"],
                          oracle.partition(g["bicc"][:, 0], g["bicc"][:, 1])))
    elif kind == "rmat16":
        V, fb, fa, bb, ba = girhip.rmat_graph(16, edge_factor=16, seed=111)
        cases.append((kind, fb, fa, bb, ba, None))
    else:
        V, fb, fa, bb, ba = girhip.grid_graph(150, 150, p_keep=0.9, p_bidir=0.6, seed=112)
        cases.append((kind, fb, fa, bb, ba, None))
    for name, fb, fa, bb, ba, want in cases:
        say("bicc", name, "V", len(fb) - 1)
        with girhip.Graph(fb, fa, bb, ba) as G:
            with pytest.raises(girhip.GirError):
                G.bicc()  # SCC layer not built
            scc, _, _ = G.scc_dag()
            r = G.bicc()
        if want is None:
            want = oracle.bicc(fb, fa, bb, ba, scc)
        got = girhip.bicc_blocks(r["parent"], r["block"])
        assert got == want, name
        assert r["n_blocks"] == len(want)
        count = np.zeros(len(fb) - 1, np.int64)
        for b in want:
            count[list(b)] += 1
        assert (r["ap"] == (count >= 2)).all(), name
        p = r["parent"]
        assert ((p < 0) == (scc == np.arange(len(scc)))).all()  # roots = canonical ids
        assert (scc[p[p >= 0]] == scc[p >= 0]).all()
