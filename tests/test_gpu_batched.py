"""The batched form on the GPU: many BFS / SSSP queries in one call and
update batches of 10^5 and 10^6 operations, bit for bit against the oracle.

The reference answers a query batch one GIR::bfs / GIR::sssp call at a time
(bench_main.cpp:538-542) and repairs update batches with GIR::applyBatch
(gir.cpp:3634-4111); both are restated by the oracle."""
import numpy as np
import pytest

import girhip
import oracle
from test_gpu_parity import check_layers

pytestmark = pytest.mark.gpu
INF = np.iinfo(np.uint64).max


def say(*a):
    print("[case]", *a, flush=True)


def test_query_batch_matches_oracle():
    V, fb, fa, bb, ba = girhip.rmat_graph(14, edge_factor=8, seed=41)
    w = girhip.weights(len(fa), 1, 64, 42)
    rng = np.random.default_rng(43)
    deg = np.diff(fb)
    n = 220
    src = rng.integers(0, V, n).astype(np.int32)
    # zero-out-degree sources (only themselves reached) and invalid ones
    src[:6] = np.nonzero(deg == 0)[0][:6]
    src[6], src[7] = -1, V
    kind = (rng.random(n) < 0.4).astype(np.uint8)  # ~90 SSSP, ~130 BFS
    kind[:8] = [0, 1, 0, 1, 0, 1, 0, 1]
    assert (kind == 0).sum() > 64 and (kind == 1).sum() > 32
    say("query batch rmat14", n, "queries")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.set_weights(w)
        lv, d = G.query_batch(src, kind, 16)
        ib = np.nonzero(kind == 0)[0]
        isp = np.nonzero(kind == 1)[0]
        for r, i in enumerate(ib):
            x = int(src[i])
            want = oracle.bfs_textbook(fb, fa, x) if 0 <= x < V else np.full(V, -1, np.int32)
            assert np.array_equal(lv[r], want), f"bfs query {i} source {x}"
        for r, i in enumerate(isp):
            x = int(src[i])
            want = oracle.dijkstra(fb, fa, w, x) if 0 <= x < V else np.full(V, INF, np.uint64)
            assert np.array_equal(d[r], want), f"sssp query {i} source {x}"
        with pytest.raises(girhip.GirError) as e:
            G.query_batch([0, 1], [0, 2], 16)
        assert e.value.code == 1  # GIRHIP_EINVAL: kinds are 0 (BFS) or 1 (SSSP)


def _mixed_batch(rng, fb, fa, n):
    """The bench's recipe: half deletes of present arcs, half inserts of
    uniform pairs, shuffled."""
    V = len(fb) - 1
    rows = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
    k = n // 2
    e = rng.integers(0, len(fa), k)
    u = np.concatenate([rows[e], rng.integers(0, V, n - k)]).astype(np.int32)
    v = np.concatenate([fa[e], rng.integers(0, V, n - k)]).astype(np.int32)
    ins = np.concatenate([np.zeros(k, np.uint8), np.ones(n - k, np.uint8)])
    p = rng.permutation(n)
    return u[p], v[p], ins[p]


@pytest.mark.parametrize("nops", [100_000, 1_000_000])
def test_apply_large_batch(nops):
    V, fb, fa, bb, ba = girhip.rmat_graph(16, seed=1)
    rng = np.random.default_rng(nops)
    say("apply batch rmat16", nops, "ops")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.wcc()
        G.scc_dag()
        for step in range(2):
            u, v, ins = _mixed_batch(rng, fb, fa, nops)
            (iu, _), (du, _), nn = oracle.normalize_batch(fb, fa, u, v, ins)
            st = G.apply_batch(u, v, ins)
            assert (st["inserted"], st["deleted"], st["noop"]) == (len(iu), len(du), nn)
            fb, fa, bb, ba = oracle.apply_batch(fb, fa, u, v, ins)
            gfb, gfa, gbb, gba = G.csr()
            assert np.array_equal(gfb, fb) and np.array_equal(gfa, fa)
            assert np.array_equal(gbb, bb) and np.array_equal(gba, ba)
            check_layers(G, fb, fa, bb, ba, f"{nops} step {step}", rebuild=False)
            x = girhip.sources(fb, 2, step)
            lv = G.bfs(x)
            for i, s in enumerate(x):
                assert np.array_equal(lv[i], oracle.bfs_textbook(fb, fa, int(s)))


def test_run_batch_interleaved_queries_and_updates():
    # girhip_run_batch: queries and update runs in list order; every query
    # answers on the graph as the updates listed before it leave it, and the
    # layers end up as a from-scratch build on the final edge set
    V, fb, fa, bb, ba = girhip.rmat_graph(13, edge_factor=8, seed=51)
    rng = np.random.default_rng(52)
    deg = np.diff(fb)
    kinds, A, B = [], [], []
    plan = [("q", 70), ("u", 500), ("q", 40), ("u", 3000), ("q", 30), ("u", 7), ("q", 25)]
    f1, a1, g1, h1 = fb, fa, bb, ba
    expect = []  # per query: (kind, source, the graph it sees)
    for what, m in plan:
        if what == "q":
            src = rng.integers(0, V, m).astype(np.int32)
            src[:3] = [np.nonzero(deg == 0)[0][0], -1, V]  # no out-arc, out of range
            kd = (rng.random(m) < 0.45).astype(np.int32)
            kd[:3] = [0, 1, 0]
            for k_, s_ in zip(kd, src):
                expect.append((int(k_), int(s_), f1, a1))
            kinds += kd.tolist()
            A += src.tolist()
            B += [0] * m
        else:
            u, v, ins = _mixed_batch(rng, f1, a1, m)
            kinds += np.where(ins == 1, 2, 3).tolist()
            A += u.tolist()
            B += v.tolist()
            f1, a1, g1, h1 = oracle.apply_batch(f1, a1, u, v, ins)
    say("run_batch rmat13", len(kinds), "items")
    with girhip.Graph(fb, fa, bb, ba) as G:
        G.wcc()
        G.scc_dag()
        lv, d, st = G.run_batch(kinds, A, B, 16)
        ib = isp = 0
        for k_, s_, f_, a_ in expect:
            ok = 0 <= s_ < V
            if k_ == 0:
                want = oracle.bfs_textbook(f_, a_, s_) if ok else np.full(V, -1, np.int32)
                assert np.array_equal(lv[ib], want), ("bfs", ib, s_)
                ib += 1
            else:
                want = oracle.dijkstra(f_, a_, None, s_) if ok else np.full(V, INF, np.uint64)
                assert np.array_equal(d[isp], want), ("sssp", isp, s_)
                isp += 1
        assert st["inserted"] + st["deleted"] > 0
        check_layers(G, f1, a1, g1, h1, "after run_batch", rebuild=False)
        with pytest.raises(girhip.GirError) as e:
            G.run_batch([0, 5], [0, 0], [0, 0])
        assert e.value.code == 1
