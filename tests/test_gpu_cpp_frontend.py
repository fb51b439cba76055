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


@pytest.mark.parametrize("mode", ["", "lazy", "inc", "lazy-inc"])
def test_stream_caller_matches_oracle(tmp_path, mode):
    # the streaming caller (the reference's gir_stream.cpp:21-88 command line)
    # on the GPU: after each stream the layers and the last BFS / SSSP equal
    # the oracle's on loaded.el plus the stream's first NB batches
    import numpy as np

    import girhip
    import oracle

    V, fb, fa, bb, ba = girhip.rmat_graph(12, edge_factor=8, seed=73)
    d = str(tmp_path / "p")
    girhip.export_protocol(d, fb, fa, 3000, n_sources=2, seed=74)
    P = girhip.read_protocol(d)
    dump = tmp_path / "dump"
    dump.mkdir()
    csv = str(tmp_path / "rows.csv")
    B, NB = 500, 4
    print(f"[case] gir_stream rmat12 B={B} NB={NB} mode={mode or 'fmu'}", flush=True)
    exe = os.path.join(ROOT, "graph-intermedia-representation_amd", "gir_stream")
    args = [exe, d, "rmat12", csv, str(B), str(NB)] + ([mode] if mode else []) + ["--dump", str(dump)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    rows = [l.split(",") for l in open(csv).read().splitlines()]
    metrics = {x[2] for x in rows}
    vals = {x[2]: float(x[4]) for x in rows}
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
        for i in range(0, min(len(ins), B * NB), B):
            f1, a1 = oracle.apply_batch(f1, a1, arcs[i:i + B, 0], arcs[i:i + B, 1], ins[i:i + B])[:2]
        rd = lambda f, t: np.fromfile(str(dump / f"{s}_{f}.bin"), t)
        assert (rd("scc", np.int32) == oracle.scc(f1, a1)[0]).all(), s
        assert (rd("wcc", np.int32) == oracle.wcc(f1, a1)[0]).all(), s
        assert (rd("bfs", np.int32) == oracle.bfs_textbook(f1, a1, src)).all(), s
        assert (rd("sssp", np.uint64) == oracle.dijkstra(f1, a1, None, src)).all(), s
