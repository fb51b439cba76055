"""ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference path used as
the checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline.
The product library never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_BIN = os.path.join(HERE, "_ref", "gir_ref")

_vp = ctypes.c_void_p
_i32, _i64, _u64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
_SIGS = {
    "orc_edge_weight": (ctypes.c_uint32, [_i32, _i32]),
    "orc_normalize_csr": (ctypes.c_int, [_i32, _vp, _vp, ctypes.POINTER(_i64)]),
    "orc_reach_info": (ctypes.c_int, [_i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp]),
    "orc_two_level": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _i32, _u64, _i32, _vp]),
    "orc_bc": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp]),
    "orc_bicc": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, ctypes.POINTER(_i64),
                                ctypes.POINTER(_i64)]),
    "orc_wcc": (ctypes.c_int, [_i32, _vp, _vp, _vp, ctypes.POINTER(_i32)]),
    "orc_scc": (ctypes.c_int, [_i32, _vp, _vp, _vp, ctypes.POINTER(_i32)]),
    "orc_dag": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64)]),
    "orc_scc_ranks": (ctypes.c_int, [_i32, _vp, _vp, _vp]),
    "orc_bfs": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _i32, _vp]),
    "orc_sssp": (ctypes.c_int, [_i32, _vp, _vp, _vp, _i32, _u64, _vp]),
    "orc_normalize_batch": (ctypes.c_int, [_i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64),
                                           _vp, _vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]),
    "orc_apply_batch": (ctypes.c_int, [_i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       ctypes.POINTER(_i64)]),
    "orc_bfs_textbook": (ctypes.c_int, [_i32, _vp, _vp, _i32, _vp]),
    "orc_dijkstra": (ctypes.c_int, [_i32, _vp, _vp, _vp, _i32, _vp]),
    "orc_bfs_component": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _vp, _vp]),
    "orc_wcc_bfs": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "orc_scc_kosaraju": (ctypes.c_int, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "orc_num_threads": (ctypes.c_int, []),
}
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _chk(rc, what):
    if rc != 0:
        raise ValueError(f"oracle {what} failed ({rc})")


def edge_weight(u, v):
    return lib().orc_edge_weight(int(u), int(v))


def normalize_csr(beg, adj):
    beg = beg.astype(np.int32).copy()
    adj = adj.astype(np.int32).copy()
    E = _i64()
    _chk(lib().orc_normalize_csr(len(beg) - 1, _p(beg), _p(adj), ctypes.byref(E)), "normalize_csr")
    return beg, adj[:E.value].copy()


def wcc(fb, fa):
    V = len(fb) - 1
    lab, n = np.empty(V, np.int32), _i32()
    _chk(lib().orc_wcc(V, _p(fb), _p(fa), _p(lab), ctypes.byref(n)), "wcc")
    return lab, n.value


def scc(fb, fa):
    V = len(fb) - 1
    lab, n = np.empty(V, np.int32), _i32()
    _chk(lib().orc_scc(V, _p(fb), _p(fa), _p(lab), ctypes.byref(n)), "scc")
    return lab, n.value


def dag(fb, fa, lab):
    V = len(fb) - 1
    n = _i64()
    _chk(lib().orc_dag(V, _p(fb), _p(fa), _p(lab), None, None, None, ctypes.byref(n)), "dag")
    s, d, m = (np.empty(max(n.value, 1), np.int32) for _ in range(3))
    _chk(lib().orc_dag(V, _p(fb), _p(fa), _p(lab), _p(s), _p(d), _p(m), ctypes.byref(n)), "dag")
    return s[:n.value], d[:n.value], m[:n.value]


def scc_ranks(fb, fa):
    V = len(fb) - 1
    r = np.empty(V, np.int64)
    _chk(lib().orc_scc_ranks(V, _p(fb), _p(fa), _p(r)), "scc_ranks")
    return r


def bfs(fb, fa, bb, ba, src):
    V = len(fb) - 1
    lv = np.empty(V, np.int32)
    _chk(lib().orc_bfs(V, _p(fb), _p(fa), _p(bb), _p(ba), int(src), _p(lv)), "bfs")
    return lv


def sssp(fb, fa, w, src, delta=16):
    V = len(fb) - 1
    d = np.empty(V, np.uint64)
    _chk(lib().orc_sssp(V, _p(fb), _p(fa), _p(w), int(src), int(delta), _p(d)), "sssp")
    return d


def normalize_batch(fb, fa, u, v, ins):
    V = len(fb) - 1
    u, v = np.ascontiguousarray(u, np.int32), np.ascontiguousarray(v, np.int32)
    ins = np.ascontiguousarray(ins, np.uint8)
    n = len(u)
    iu, iv, du, dv = (np.empty(max(n, 1), np.int32) for _ in range(4))
    ni, nd, nn = _i64(), _i64(), _i64()
    _chk(lib().orc_normalize_batch(V, _p(fb), _p(fa), n, _p(u), _p(v), _p(ins), _p(iu), _p(iv), ctypes.byref(ni),
                                   _p(du), _p(dv), ctypes.byref(nd), ctypes.byref(nn)), "normalize_batch")
    return (iu[:ni.value], iv[:ni.value]), (du[:nd.value], dv[:nd.value]), nn.value


def apply_batch(fb, fa, u, v, ins):
    V = len(fb) - 1
    u, v = np.ascontiguousarray(u, np.int32), np.ascontiguousarray(v, np.int32)
    ins = np.ascontiguousarray(ins, np.uint8)
    n = len(u)
    cap = len(fa) + n + 1
    nfb, nbb = np.empty(V + 1, np.int32), np.empty(V + 1, np.int32)
    nfa, nba = np.empty(cap, np.int32), np.empty(cap, np.int32)
    E = _i64()
    _chk(lib().orc_apply_batch(V, _p(fb), _p(fa), n, _p(u), _p(v), _p(ins), _p(nfb), _p(nfa), _p(nbb), _p(nba),
                               ctypes.byref(E)), "apply_batch")
    return nfb, nfa[:E.value].copy(), nbb, nba[:E.value].copy()


def bfs_textbook(fb, fa, src):
    lv = np.empty(len(fb) - 1, np.int32)
    _chk(lib().orc_bfs_textbook(len(fb) - 1, _p(fb), _p(fa), int(src), _p(lv)), "bfs_textbook")
    return lv


def bfs_component(fb, fa, bb, ba, label, src, undirected):
    """ParallelBFS::testParallelBFS (gir_test.cpp:143-343): levels inside
    src's component (label = SCC labels, directed; WCC labels, undirected)
    -> (levels, visited count)."""
    V = len(fb) - 1
    lv = np.empty(V, np.int32)
    vis = np.zeros(1, np.int64)
    label = np.ascontiguousarray(label, np.int32)
    _chk(lib().orc_bfs_component(V, _p(fb), _p(fa), _p(bb), _p(ba), _p(label), int(bool(undirected)), int(src),
                                 _p(lv), _p(vis)), "bfs_component")
    return lv, int(vis[0])


def reach_info(fb, fa, scc_labels, src_scc):
    """GIR::reachableSccs / reachInfo (gir_query.cpp:56-88) from the SCC with
    canonical id src_scc -> (mask bool[V] by SCC id, pending int32[V],
    (sccs, verts, edges))."""
    V = len(fb) - 1
    lab = np.ascontiguousarray(scc_labels, np.int32)
    mask = np.zeros(V, np.uint8)
    pend = np.zeros(V, np.int32)
    cnt = np.zeros(3, np.int64)
    _chk(lib().orc_reach_info(V, _p(fb), _p(fa), _p(lab), int(src_scc), _p(mask), _p(pend), _p(cnt)), "reach_info")
    return mask.astype(bool), pend, tuple(int(x) for x in cnt)


def two_level(fb, fa, w, scc_labels, src, delta=16, unit=False):
    """GIR::bfsTwoLevel (unit=True, levels as uint64) / ssspTwoLevel
    (gir_query.cpp:304-350, 604-722) -> uint64[V], 2^64-1 unreachable."""
    V = len(fb) - 1
    lab = np.ascontiguousarray(scc_labels, np.int32)
    d = np.empty(V, np.uint64)
    wa = None if w is None else np.ascontiguousarray(w, np.uint32)
    _chk(lib().orc_two_level(V, _p(fb), _p(fa), _p(wa), _p(lab), int(src), int(delta), int(bool(unit)), _p(d)),
         "two_level")
    return d


def bc(fb, fa, bb, ba, sources):
    """GIR::bc (gir_query.cpp:728-778) -> float64[V]."""
    V = len(fb) - 1
    src = np.ascontiguousarray(sources, np.int32)
    sc = np.empty(V, np.float64)
    _chk(lib().orc_bc(V, _p(fb), _p(fa), _p(bb), _p(ba), _p(src), len(src), _p(sc)), "bc")
    return sc


def bicc(fb, fa, bb, ba, scc_labels):
    """Blocks of every SCC's undirected view (gir_test.cpp:448-465,
    dynamic_repair.h:96-170) -> sorted list of sorted vertex tuples."""
    V = len(fb) - 1
    lab = np.ascontiguousarray(scc_labels, np.int32)
    cap = V + 2 * len(fa) + 1
    mb, mv = np.empty(cap, np.int32), np.empty(cap, np.int32)
    nm, nb = _i64(), _i64()
    _chk(lib().orc_bicc(V, _p(fb), _p(fa), _p(bb), _p(ba), _p(lab), _p(mb), _p(mv), cap, ctypes.byref(nm),
                        ctypes.byref(nb)), "bicc")
    return partition(mb[:nm.value], mv[:nm.value])


def partition(block_ids, vertices):
    """(block, vertex) membership pairs -> sorted list of sorted tuples."""
    groups = {}
    for b, v in zip(np.asarray(block_ids).tolist(), np.asarray(vertices).tolist()):
        groups.setdefault(b, set()).add(v)
    return sorted(tuple(sorted(g)) for g in groups.values())


def dijkstra(fb, fa, w, src):
    d = np.empty(len(fb) - 1, np.uint64)
    _chk(lib().orc_dijkstra(len(fb) - 1, _p(fb), _p(fa), _p(w), int(src), _p(d)), "dijkstra")
    return d


def wcc_bfs(fb, fa, bb, ba):
    lab = np.empty(len(fb) - 1, np.int32)
    _chk(lib().orc_wcc_bfs(len(fb) - 1, _p(fb), _p(fa), _p(bb), _p(ba), _p(lab)), "wcc_bfs")
    return lab


def scc_kosaraju(fb, fa, bb, ba):
    lab = np.empty(len(fb) - 1, np.int32)
    _chk(lib().orc_scc_kosaraju(len(fb) - 1, _p(fb), _p(fa), _p(bb), _p(ba), _p(lab)), "scc_kosaraju")
    return lab


def num_threads():
    return lib().orc_num_threads()
