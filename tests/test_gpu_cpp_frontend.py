"""The C++ drop-in GIR class (host/gir.h) on the GPU: runs the differential
self-test tests/cpp/gir_selftest (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "tests", "cpp", "gir_selftest")


def test_cpp_gir_selftest():
    assert os.path.exists(BIN), "build the C++ front end first (__graft_entry__.build())"
    print("[case] cpp gir_selftest 16 iterations", flush=True)
    r = subprocess.run([BIN, "16"], capture_output=True, text=True, timeout=110)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]


STREAM = os.path.join(ROOT, "oracle", "_ref", "gir_stream_gpu")


@pytest.mark.parametrize("mode", ["", "lazy", "inc", "lazy-inc"])
def test_reference_stream_caller_on_dropin(tmp_path, mode):
    # The reference's own streaming caller (experiments/baselines/gir/
    # gir_stream.cpp:21-88), compiled unchanged against host/gir.h and linked
    # to libgir.so, on the GPU: every result row is written, and in the
    # incremental modes its own check (corrected vs from-scratch BFS / SSSP,
    # gir_stream.cpp:61-73) finds no mismatch.  The same streams replayed
    # through the library then equal the oracle's layers and queries.
    import numpy as np

    import girhip
    import oracle

    if not os.path.exists(STREAM):
        pytest.skip("oracle/_ref/gir_stream_gpu is built only where the reference sources are present")
    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=73)
    d = str(tmp_path / "p")
    girhip.export_protocol(d, fb, fa, 3000, n_sources=2, seed=74)
    P = girhip.read_protocol(d)
    csv = str(tmp_path / "rows.csv")
    B, NB = 500, 4
    print(f"[case] reference gir_stream on the drop-in, rmat12 B={B} NB={NB} mode={mode or 'fmu'}", flush=True)
    r = subprocess.run([STREAM, d, "rmat12", csv, str(B), str(NB)] + ([mode] if mode else []),
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = [l.split(",") for l in open(csv).read().splitlines()]
    metrics = {x[2] for x in rows}
    vals = {x[2]: float(x[4]) for x in rows}
    # rows: dataset, system, metric, mode, value, unit, count, threads
    assert {x[0] for x in rows} == {"rmat12"}
    assert {x[1] for x in rows} == {"GIR" + ("-lazy" if mode.startswith("lazy") else "") +
                                    ("-inc" if mode.endswith("inc") else "")}
    for s in ("insert", "delete", "mixed"):
        assert {f"{s}_mutation_MEPS", f"bfs_{s}_recompute_ms", f"sssp_{s}_endtoend_ms"} <= metrics
        assert (f"sync_{s}_ms" in metrics) == mode.startswith("lazy")
        if mode.endswith("inc"):
            assert vals[f"inc_mismatch_{s}"] == 0
    lo = P["loaded"]
    fb0, fa0, _, _ = girhip.csr_from_edges(V, lo[:, 0], lo[:, 1])
    src = int(P["sources"][0])
    for s in ("insert", "delete", "mixed"):
        ins, arcs = P[s]
        f1, a1 = fb0, fa0
        with girhip.Graph(fb0, fa0) as G:
            G.wcc()
            G.scc_dag()
            for i in range(0, min(len(ins), B * NB), B):
                u, v, o = arcs[i:i + B, 0], arcs[i:i + B, 1], ins[i:i + B]
                if mode.startswith("lazy"):
                    G.ingest_batch(u, v, o)
                else:
                    G.apply_batch(u, v, o)
                f1, a1 = oracle.apply_batch(f1, a1, u, v, o)[:2]
            if mode.startswith("lazy"):
                G.sync()
            assert np.array_equal(G.get_scc()[0], oracle.scc(f1, a1)[0]), s
            assert np.array_equal(G.get_wcc()[0], oracle.wcc(f1, a1)[0]), s
            assert np.array_equal(G.bfs([src])[0], oracle.bfs_textbook(f1, a1, src)), s
            assert np.array_equal(G.sssp([src], 16)[0], oracle.dijkstra(f1, a1, None, src)), s


UPD = os.path.join(ROOT, "tests", "cpp", "gir_updates")


def test_single_edge_updates_through_dropin(tmp_path):
    # tests/cpp/gir_updates drives GIR::insertEdge / deleteEdge one call per
    # op (queued by the class, applied when read) and once more applying each
    # op at once; the layers after the ops equal the oracle's on the updated
    # store (canonical labels), including ops on present / absent arcs, u == v
    # and out-of-range ids, which change nothing.
    import numpy as np

    import girhip
    import oracle

    assert os.path.exists(UPD), "build the C++ front end first (__graft_entry__.build())"
    V, fb, fa, bb, ba = girhip.rmat_graph(11, edge_factor=8, seed=81)
    rng = np.random.default_rng(82)
    rows = np.repeat(np.arange(V, dtype=np.int32), np.diff(fb))
    e = rng.integers(0, len(fa), 60)
    ops = [(int(rows[i]), int(fa[i]), 0) for i in e]                              # deletes of present arcs
    ops += [(int(a), int(b), 1) for a, b in rng.integers(0, V, (60, 2))]         # inserts of random pairs
    ops += [(int(rows[e[0]]), int(fa[e[0]]), 1), (5, 5, 1), (V + 3, 1, 1), (1, -1, 0)]  # re-insert, invalid
    d = str(tmp_path / "g")
    girhip.write_csr(d, fb, fa, bb, ba)
    np.array(ops, np.int32).tofile(str(tmp_path / "ops.bin"))
    print("[case] cpp gir_updates rmat11", len(ops), "single ops", flush=True)
    r = subprocess.run([UPD, d + "/", str(tmp_path / "ops.bin"), str(tmp_path) + "/"], capture_output=True,
                       text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    rec = next(l for l in r.stdout.splitlines() if l.startswith("@@"))
    assert '"inserts": 63' in rec and '"deletes": 61' in rec, rec
    u, v, ins = (np.array([o[k] for o in ops], np.int64) for k in range(3))
    ok = (u >= 0) & (u < V) & (v >= 0) & (v < V) & (u != v)
    nfb, nfa, _, _ = oracle.apply_batch(fb, fa, u[ok].astype(np.int32), v[ok].astype(np.int32),
                                        ins[ok].astype(np.uint8))
    assert np.array_equal(np.fromfile(str(tmp_path / "post_scc.bin"), np.int32), oracle.scc(nfb, nfa)[0])
    assert np.array_equal(np.fromfile(str(tmp_path / "post_wcc.bin"), np.int32), oracle.wcc(nfb, nfa)[0])
