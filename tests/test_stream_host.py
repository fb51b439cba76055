"""The streaming protocol's inputs (bench_main.cpp:280-345 hold-out,
experiments/baselines/common/protocol.h formats) and the streaming caller's
command line, on the CPU."""
import os
import subprocess

import numpy as np

import girhip

PKG = os.path.dirname(girhip.__file__)


def _arcs(fb, fa):
    V = len(fb) - 1
    return np.stack([np.repeat(np.arange(V, dtype=np.int64), np.diff(fb)), fa.astype(np.int64)], 1)


def _keys(a):
    return set((a[:, 0].astype(np.int64) << 32 | a[:, 1].astype(np.int64)).tolist())


def test_protocol_hold_out_split(tmp_path):
    V, fb, fa, bb, ba = girhip.rmat_graph(11, edge_factor=8, seed=71)
    d = str(tmp_path / "p")
    girhip.export_protocol(d, fb, fa, 1000, n_sources=5, seed=9)
    P = girhip.read_protocol(d)
    E = len(fa)
    meta = dict(l.split() for l in open(os.path.join(d, "meta.txt")))
    assert int(meta["V"]) == V and int(meta["E"]) == E and int(meta["loaded"]) == E // 2
    k = min(1000, (E // 2) // 2)
    assert int(meta["ops_per_phase"]) == k
    full = _keys(_arcs(fb, fa))
    loaded = _keys(P["loaded"])
    assert len(P["loaded"]) == E // 2 and len(loaded) == E // 2 and loaded <= full
    ins, arcs = P["insert"]
    assert ins.all() and len(arcs) == k
    held = _keys(arcs)
    assert len(held) == k and not held & loaded and held <= full  # held-out arcs come back in
    dl, darcs = P["delete"]
    assert not dl.any() and len(darcs) == k and _keys(darcs) <= loaded
    mi, marcs = P["mixed"]
    assert len(mi) == 2 * (k // 2) and mi.sum() == k // 2
    mins, mdel = _keys(marcs[mi == 1]), _keys(marcs[mi == 0])
    assert mins <= full - loaded and not mins & held  # fresh held-out arcs
    assert mdel <= loaded and not mdel & _keys(darcs)  # fresh deletions
    src = P["sources"]
    assert len(src) == 5 and (np.diff(fb)[src] > 0).all()
    full_el = np.loadtxt(os.path.join(d, "full.el"), dtype=np.int64, ndmin=2)
    assert _keys(full_el) == full and len(full_el) == E


def test_reference_stream_caller_links_the_dropin():
    # oracle/_ref/gir_stream_gpu: the reference's gir_stream.cpp compiled
    # against host/gir.h; it must resolve libgir.so / libgirhip.so from this
    # tree, and with too few arguments it exits 2 (gir_stream.cpp:22).
    exe = os.path.join(os.path.dirname(PKG), "oracle", "_ref", "gir_stream_gpu")
    if not os.path.exists(exe):
        import pytest
        pytest.skip("built only where the reference sources are present")
    ldd = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "not found" not in ldd
    libs = {l.split()[0]: os.path.realpath(l.split()[2]) for l in ldd.splitlines() if "=>" in l and "/" in l}
    assert libs.get("libgir.so") == os.path.realpath(os.path.join(PKG, "libgir.so"))
    assert libs.get("libgirhip.so") == os.path.realpath(os.path.join(PKG, "libgirhip.so"))
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 2
