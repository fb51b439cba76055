"""ctypes binding of the gfx950 GIR library (libgirhip.so) and the host
input-format library (libgirgen.so).

`Graph` wraps one girhip_graph handle (include/girhip.h): the forward and
backward CSR resident in HBM plus every layer built from it.  Arrays may be
numpy arrays (host) or torch tensors (host or cuda); device tensors are
written in place by the kernels.  There is no CPU fallback: if the library
or a GPU is missing, construction raises.

The method names follow the reference's GIR class (gir_constructor/gir.h:
buildWCC, buildSCC_DAG, bfs, sssp, applyBatch, ...); see each docstring for
the reference lines it stands in for.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgirhip.so")
GEN_PATH = os.path.join(_HERE, "libgirgen.so")

c_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ESTATE", 5: "ENODEV"}


class GirError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"girhip error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class BatchStats(ctypes.Structure):
    _fields_ = [("inserted", ctypes.c_int64), ("deleted", ctypes.c_int64), ("noop", ctypes.c_int64),
                ("invalid", ctypes.c_int64)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


_lib = None
_gen = None

# name -> (restype, argtypes)
_SIGS = {
    "girhip_last_error": (ctypes.c_char_p, []),
    "girhip_create": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(_vp)]),
    "girhip_destroy": (ctypes.c_int, [_vp]),
    "girhip_size": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "girhip_get_csr": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "girhip_set_weights": (ctypes.c_int, [_vp, _vp]),
    "girhip_wcc": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32)]),
    "girhip_scc_dag": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "girhip_get_wcc": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32)]),
    "girhip_get_scc": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64)]),
    "girhip_get_dag": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "girhip_get_ranks": (ctypes.c_int, [_vp, _vp]),
    "girhip_bfs": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp, _vp]),
    "girhip_bfs_noalloc": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "girhip_run_batch": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.c_uint64, _vp, _vp, _vp]),
    "girhip_bfs_component": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    "girhip_reachable": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int64, _vp]),
    "girhip_reach_info": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _vp, _vp, _vp, _vp]),
    "girhip_bfs_two_level": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp]),
    "girhip_sssp_two_level": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_uint64, _vp]),
    "girhip_bc": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp]),
    "girhip_bicc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "girhip_inc_track": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    "girhip_inc_update": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "girhip_inc_dist": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "girhip_sssp": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, ctypes.c_uint64, _vp, _vp]),
    "girhip_apply_batch": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.POINTER(BatchStats)]),
    "girhip_ingest_batch": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int64, ctypes.POINTER(BatchStats)]),
    "girhip_sync": (ctypes.c_int, [_vp]),
    "girhip_synced": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "girhip_has_edge": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "girhip_set_param": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_double]),
    "girhip_get_param": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    "girhip_last_ms": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    "girhip_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_double)]),
    "girhip_reset_kernel_stats": (ctypes.c_int, [_vp]),
    "girhip_kernel_stat_names": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]),
    "girhip_features": (ctypes.c_int, [_vp, ctypes.c_int32, _vp]),
    "girhip_autotune": (ctypes.c_int, [_vp, ctypes.c_char_p, _vp, ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_double), _vp]),
    "girhip_mlp_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "girhip_mlp_destroy": (ctypes.c_int, [_vp]),
    "girhip_mlp_num_params": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "girhip_mlp_get_params": (ctypes.c_int, [_vp, _vp]),
    "girhip_mlp_set_params": (ctypes.c_int, [_vp, _vp]),
    "girhip_mlp_forward": (ctypes.c_int, [_vp, _vp, ctypes.c_int32, _vp]),
    "girhip_mlp_grad": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, _vp, ctypes.POINTER(ctypes.c_float)]),
    "girhip_mlp_train": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                        ctypes.POINTER(ctypes.c_float)]),
    "girhip_learned_tune": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    "girhip_query_batch": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int32, ctypes.c_uint64, _vp, _vp]),
}

_GEN_SIGS = {
    "girgen_last_error": (ctypes.c_char_p, []),
    "girgen_rmat_edges": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_uint64, _vp, _vp]),
    "girgen_grid_edges": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_uint64, _vp, _vp, ctypes.POINTER(ctypes.c_int64)]),
    "girgen_uniform_edges": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, _vp, _vp]),
    "girgen_csr_from_edges": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp,
                                             ctypes.POINTER(ctypes.c_int64)]),
    "girgen_weights": (ctypes.c_int, [ctypes.c_int64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, _vp]),
    "girgen_sources": (ctypes.c_int, [ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_uint64, _vp]),
    "girgen_write_csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp]),
    "girgen_export_protocol": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, _vp, _vp, ctypes.c_int64,
                                              ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32]),
    "girgen_read_csr": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32),
                                       ctypes.POINTER(ctypes.c_int64), _vp, _vp, _vp, _vp]),
}


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib():
    """The loaded libgirhip.so (raises ImportError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        _lib = _bind(ctypes.CDLL(LIB_PATH), _SIGS)
    return _lib


def gen():
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise ImportError(f"{GEN_PATH} is not built (run __graft_entry__.build())")
        _gen = _bind(ctypes.CDLL(GEN_PATH), _GEN_SIGS)
    return _gen


def exported_symbols():
    return sorted(_SIGS)


def _check(rc):
    if rc != 0:
        raise GirError(rc, lib().girhip_last_error().decode())


def _gcheck(rc):
    if rc != 0:
        raise ValueError("girgen: " + gen().girgen_last_error().decode())


def _ptr(x, dtype, n=None, name="array"):
    """Pointer of a contiguous numpy array or torch tensor of the given dtype."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        if x.dtype != np.dtype(dtype):
            raise TypeError(f"{name}: expected {np.dtype(dtype)}, got {x.dtype}")
        if not x.flags["C_CONTIGUOUS"]:
            raise TypeError(f"{name}: not contiguous")
        if n is not None and x.size < n:
            raise ValueError(f"{name}: needs {n} entries, has {x.size}")
        return x.ctypes.data
    # torch tensor
    import torch
    tmap = {np.int32: torch.int32, np.uint32: torch.uint32 if hasattr(torch, "uint32") else torch.int32,
            np.int64: torch.int64, np.uint64: torch.uint64 if hasattr(torch, "uint64") else torch.int64,
            np.uint8: torch.uint8}
    want = tmap[np.dtype(dtype).type]
    if x.dtype != want and not (np.dtype(dtype) == np.uint32 and x.dtype == torch.int32) and \
            not (np.dtype(dtype) == np.uint64 and x.dtype == torch.int64):
        raise TypeError(f"{name}: expected {want}, got {x.dtype}")
    if not x.is_contiguous():
        raise TypeError(f"{name}: not contiguous")
    if n is not None and x.numel() < n:
        raise ValueError(f"{name}: needs {n} entries, has {x.numel()}")
    if x.is_cuda:
        # the library works on its own stream: work torch has queued on this
        # tensor (a pending clone, fill or read) must finish first
        torch.cuda.current_stream(x.device).synchronize()
    return x.data_ptr()


def _torch_int32():
    import torch
    return torch.int32


class Graph:
    """One device-resident graph (girhip_graph).

    Stands in for the reference's GIR object built by GIR(graph*) /
    GIR(path, path) (gir.cpp:38-63): rows are sorted and deduplicated on the
    device (normalizeBaseCsr, gir.cpp:95-130).
    """

    def __init__(self, fw_beg, fw_adj, bw_beg=None, bw_adj=None, V=None, E=None):
        L = lib()
        if V is None:
            V = (fw_beg.numel() if hasattr(fw_beg, "numel") else len(fw_beg)) - 1
        if E is None:
            E = fw_adj.numel() if hasattr(fw_adj, "numel") else len(fw_adj)
        h = _vp()
        _check(L.girhip_create(_ptr(fw_beg, np.int32, V + 1, "fw_beg"), _ptr(fw_adj, np.int32, E, "fw_adj"),
                               _ptr(bw_beg, np.int32, V + 1, "bw_beg"), _ptr(bw_adj, np.int32, E, "bw_adj"),
                               int(V), int(E), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().girhip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- sizes / store ----
    @property
    def V(self):
        v, e = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().girhip_size(self._h, ctypes.byref(v), ctypes.byref(e)))
        return v.value

    @property
    def E(self):
        v, e = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().girhip_size(self._h, ctypes.byref(v), ctypes.byref(e)))
        return e.value

    def csr(self):
        """Current canonical (fw_beg, fw_adj, bw_beg, bw_adj) as numpy arrays."""
        V, E = self.V, self.E
        fb, fa = np.empty(V + 1, np.int32), np.empty(max(E, 1), np.int32)
        bb, ba = np.empty(V + 1, np.int32), np.empty(max(E, 1), np.int32)
        _check(lib().girhip_get_csr(self._h, fb.ctypes.data, fa.ctypes.data, bb.ctypes.data, ba.ctypes.data))
        return fb, fa[:E], bb, ba[:E]

    def set_weights(self, w):
        """Per-arc uint32 weights aligned to the forward CSR, or None for the
        reference's GIR::edgeWeight hash (gir_query.cpp:50-54)."""
        _check(lib().girhip_set_weights(self._h, None if w is None else _ptr(w, np.uint32, self.E, "w")))

    def has_edge(self, u, v):
        """GIR::hasDirectedEdge (gir.h:160)."""
        r = ctypes.c_int32()
        _check(lib().girhip_has_edge(self._h, int(u), int(v), ctypes.byref(r)))
        return bool(r.value)

    # ---- layers ----
    def wcc(self, out=None):
        """GIR::buildWCC (gir.cpp:593-655) -> (labels, count); labels are the
        minimum vertex of each weakly connected component."""
        V = self.V
        lab = np.empty(V, np.int32) if out is None else out
        n = ctypes.c_int32()
        _check(lib().girhip_wcc(self._h, _ptr(lab, np.int32, V, "label"), ctypes.byref(n)))
        return lab, n.value

    def scc_dag(self, out=None):
        """GIR::buildSCC_DAG (gir.cpp:798-839) -> (scc labels, n_scc, n_dag)."""
        V = self.V
        lab = np.empty(V, np.int32) if out is None else out
        ns, nd = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().girhip_scc_dag(self._h, _ptr(lab, np.int32, V, "scc"), ctypes.byref(ns), ctypes.byref(nd)))
        self._n_dag = nd.value
        return lab, ns.value, nd.value

    def get_wcc(self):
        """Maintained WCC labels without rebuilding (GIR::getNodeToWcc,
        gir.cpp:4122-4126) -> (labels, count)."""
        lab = np.empty(self.V, np.int32)
        n = ctypes.c_int32()
        _check(lib().girhip_get_wcc(self._h, lab.ctypes.data, ctypes.byref(n)))
        return lab, n.value

    def get_scc(self):
        """Maintained SCC labels without rebuilding (GIR::getNodeToScc,
        gir.cpp:4117-4121) -> (labels, n_scc, n_dag)."""
        lab = np.empty(self.V, np.int32)
        ns, nd = ctypes.c_int32(), ctypes.c_int64()
        _check(lib().girhip_get_scc(self._h, lab.ctypes.data, ctypes.byref(ns), ctypes.byref(nd)))
        self._n_dag = nd.value
        return lab, ns.value, nd.value

    def dag(self):
        """Condensation arcs (src, dst, mult) sorted by (src, dst)
        (GIR::getDAG + the dag_out multiplicities, gir.cpp:818-831, 4127-4132)."""
        nd = getattr(self, "_n_dag", None)
        if nd is None:
            raise GirError(4, "dag: call scc_dag first")
        s, d, m = (np.empty(max(nd, 1), np.int32) for _ in range(3))
        _check(lib().girhip_get_dag(self._h, s.ctypes.data, d.ctypes.data, m.ctypes.data))
        return s[:nd], d[:nd], m[:nd]

    def ranks(self):
        """Topological ranks indexed by canonical SCC id (-1 elsewhere)."""
        r = np.empty(self.V, np.int64)
        _check(lib().girhip_get_ranks(self._h, r.ctypes.data))
        return r

    # ---- queries ----
    def bfs(self, sources, out=None, edges=False):
        """GIR::bfs (gir_query.cpp:108-294) for every source -> int32 (k, V)
        levels, -1 unreachable.  edges=True also returns m_s per source."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k, V = len(src), self.V
        lv = np.empty((k, V), np.int32) if out is None else out
        e = np.zeros(max(k, 1), np.int64)
        _check(lib().girhip_bfs(self._h, src.ctypes.data, k, _ptr(lv, np.int32, k * V, "levels"),
                                e.ctypes.data if edges else None))
        return (lv, e[:k]) if edges else lv

    def bfs_noalloc(self, source, out):
        """girhip_bfs_noalloc: GIR::bfs for one source with no device
        allocation.  out: a device int32 tensor of V entries, or the address
        (int) of V int32 of device memory."""
        if isinstance(out, int):
            ptr = out
        else:
            if not hasattr(out, "data_ptr") or not out.is_cuda:
                raise ValueError("bfs_noalloc: out must be device memory")
            if out.dtype != _torch_int32() or out.numel() < self.V or not out.is_contiguous():
                raise ValueError("bfs_noalloc: out must be a contiguous int32 tensor of V entries")
            ptr = out.data_ptr()
        _check(lib().girhip_bfs_noalloc(self._h, int(source), ptr))
        return out

    def sssp(self, sources, delta=0, out=None, edges=False):
        """GIR::sssp (gir_query.cpp:695-703) for every source -> uint64 (k, V)
        distances, 2^64-1 unreachable."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k, V = len(src), self.V
        d = np.empty((k, V), np.uint64) if out is None else out
        e = np.zeros(max(k, 1), np.int64)
        _check(lib().girhip_sssp(self._h, src.ctypes.data, k, int(delta), _ptr(d, np.uint64, k * V, "dist"),
                                 e.ctypes.data if edges else None))
        return (d, e[:k]) if edges else d

    # ---- updates ----
    def _batch(self, fn, u, v, ins, w):
        u = np.ascontiguousarray(np.asarray(u, dtype=np.int32))
        v = np.ascontiguousarray(np.asarray(v, dtype=np.int32))
        ins = np.ascontiguousarray(np.asarray(ins, dtype=np.uint8))
        if not (len(u) == len(v) == len(ins)):
            raise ValueError("batch: u, v, ins must have equal length")
        wp = None
        if w is not None:
            w = np.ascontiguousarray(np.asarray(w, dtype=np.uint32))
            wp = w.ctypes.data
        st = BatchStats()
        _check(fn(self._h, u.ctypes.data, v.ctypes.data, ins.ctypes.data, wp, len(u), ctypes.byref(st)))
        return st.as_dict()

    def apply_batch(self, u, v, ins, w=None):
        """GIR::applyBatch (gir.cpp:3634-4111): last op per arc wins; every
        built layer is repaired."""
        return self._batch(lib().girhip_apply_batch, u, v, ins, w)

    def ingest_batch(self, u, v, ins, w=None):
        """GIR::ingestBatch (gir.cpp:3588-3597): edge store only."""
        return self._batch(lib().girhip_ingest_batch, u, v, ins, w)

    def update_edge_directed(self, u, v, insert):
        """GIR::updateEdgeDirected (gir.cpp:3133)."""
        return self.apply_batch([u], [v], [1 if insert else 0])

    def sync(self):
        """GIR::sync (gir.cpp:3602-3632)."""
        _check(lib().girhip_sync(self._h))

    def synced(self):
        r = ctypes.c_int32()
        _check(lib().girhip_synced(self._h, ctypes.byref(r)))
        return bool(r.value)

    # ---- knobs / timing ----
    def set_param(self, name, value):
        _check(lib().girhip_set_param(self._h, name.encode(), float(value)))

    def get_param(self, name):
        r = ctypes.c_double()
        _check(lib().girhip_get_param(self._h, name.encode(), ctypes.byref(r)))
        return r.value

    def last_ms(self, kind):
        r = ctypes.c_double()
        _check(lib().girhip_last_ms(self._h, kind.encode(), ctypes.byref(r)))
        return r.value

    def kernel_stats(self, name):
        c, m = ctypes.c_int64(), ctypes.c_double()
        _check(lib().girhip_kernel_stats(self._h, name.encode(), ctypes.byref(c), ctypes.byref(m)))
        return c.value, m.value

    def reset_kernel_stats(self):
        _check(lib().girhip_reset_kernel_stats(self._h))

    def query_batch(self, sources, kinds, delta=0, levels=None, dist=None):
        """Mixed BFS (kind 0) / SSSP (kind 1) queries -> (levels rows of the
        BFS queries, dist rows of the SSSP queries), each in query order."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        kd = np.ascontiguousarray(np.asarray(kinds, dtype=np.uint8))
        nb, ns, V = int((kd == 0).sum()), int((kd != 0).sum()), self.V
        if levels is None and nb:
            levels = np.empty((nb, V), np.int32)
        if dist is None and ns:
            dist = np.empty((ns, V), np.uint64)
        _check(lib().girhip_query_batch(self._h, src.ctypes.data, kd.ctypes.data, len(src), int(delta),
                                        _ptr(levels, np.int32, nb * V, "levels") if nb else None,
                                        _ptr(dist, np.uint64, ns * V, "dist") if ns else None))
        return levels, dist

    def run_batch(self, kinds, a, b=None, delta=0, levels=None, dist=None):
        """Mixed queries and updates in list order (girhip_run_batch): kind
        0 BFS from a[i], 1 SSSP from a[i], 2 insert a[i] -> b[i], 3 delete
        it.  -> (levels rows of the BFS queries, dist rows of the SSSP
        queries, summed update stats)."""
        kd = np.ascontiguousarray(np.asarray(kinds, dtype=np.int32))
        av = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
        bv = None if b is None else np.ascontiguousarray(np.asarray(b, dtype=np.int32))
        if len(av) != len(kd) or (bv is not None and len(bv) != len(kd)):
            raise ValueError("run_batch: kinds, a, b must have equal length")
        nb, ns, V = int((kd == 0).sum()), int((kd == 1).sum()), self.V
        if levels is None and nb:
            levels = np.empty((nb, V), np.int32)
        if dist is None and ns:
            dist = np.empty((ns, V), np.uint64)
        st = BatchStats()
        _check(lib().girhip_run_batch(self._h, kd.ctypes.data, av.ctypes.data, None if bv is None else bv.ctypes.data,
                                      len(kd), int(delta),
                                      _ptr(levels, np.int32, nb * V, "levels") if nb else None,
                                      _ptr(dist, np.uint64, ns * V, "dist") if ns else None, ctypes.byref(st)))
        return levels, dist, st.as_dict()

    # ---- tuning ----
    def bfs_component(self, sources, undirected=False, out=None):
        """ParallelBFS harness (gir_test.cpp:143-343): BFS inside each
        source's SCC (directed) or WCC (undirected) -> (levels k x V, visited
        per source)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k = len(src)
        lv = out if out is not None else np.empty((k, self.V), np.int32)
        vis = np.zeros(k, np.int64)
        _check(lib().girhip_bfs_component(self._h, src.ctypes.data if k else None, k, int(bool(undirected)),
                                          _ptr(lv, np.int32, k * self.V, "levels"), vis.ctypes.data))
        return lv, vis

    def reachable(self, u, v):
        """GIR::reachable (gir_query.cpp:789-828) for pairs (u[i], v[i]) ->
        bool array."""
        u = np.ascontiguousarray(np.atleast_1d(np.asarray(u, dtype=np.int32)))
        v = np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.int32)))
        if len(u) != len(v):
            raise ValueError("reachable: u and v differ in length")
        out = np.zeros(len(u), np.uint8)
        _check(lib().girhip_reachable(self._h, u.ctypes.data if len(u) else None, v.ctypes.data if len(v) else None,
                                      len(u), out.ctypes.data))
        return out.astype(bool)

    def reach_info(self, scc_id):
        """GIR::reachableSccs / reachInfo (gir_query.cpp:56-88) from the SCC
        with canonical id scc_id -> dict: mask (bool per SCC id), pending
        (reachable DAG predecessors per reachable SCC id), sccs, verts,
        edges."""
        V = self.V
        mask = np.zeros(max(V, 1), np.uint8)
        pend = np.zeros(max(V, 1), np.int32)
        c = np.zeros(3, np.int64)
        _check(lib().girhip_reach_info(self._h, int(scc_id), mask.ctypes.data, pend.ctypes.data, c[0:].ctypes.data,
                                       c[1:].ctypes.data, c[2:].ctypes.data))
        return {"mask": mask[:V].astype(bool), "pending": pend[:V], "sccs": int(c[0]), "verts": int(c[1]),
                "edges": int(c[2])}

    def bfs_two_level(self, sources, out=None):
        """GIR::bfsTwoLevel (gir_query.cpp:604-620) for every source -> int32
        (k, V) levels, -1 unreachable (equal to bfs)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k, V = len(src), self.V
        lv = np.empty((k, V), np.int32) if out is None else out
        _check(lib().girhip_bfs_two_level(self._h, src.ctypes.data if k else None, k,
                                          _ptr(lv, np.int32, k * V, "levels")))
        return lv

    def sssp_two_level(self, sources, delta=0, out=None):
        """GIR::ssspTwoLevel (gir_query.cpp:705-722) for every source ->
        uint64 (k, V) distances, 2^64-1 unreachable (equal to sssp)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k, V = len(src), self.V
        d = np.empty((k, V), np.uint64) if out is None else out
        _check(lib().girhip_sssp_two_level(self._h, src.ctypes.data if k else None, k, int(delta),
                                           _ptr(d, np.uint64, k * V, "dist")))
        return d

    def bc(self, sources, out=None):
        """GIR::bc (gir_query.cpp:728-778): Brandes betweenness over the
        sources -> float64 (V,)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        k = len(src)
        sc = np.empty(max(self.V, 1), np.float64) if out is None else out
        if out is not None and not isinstance(out, np.ndarray):
            ptr = out.data_ptr()
        else:
            if sc.dtype != np.float64 or not sc.flags["C_CONTIGUOUS"]:
                raise TypeError("bc: out must be contiguous float64")
            ptr = sc.ctypes.data
        _check(lib().girhip_bc(self._h, src.ctypes.data if k else None, k, ptr))
        return sc[:self.V] if out is None else out

    def bicc(self):
        """Blocks of every SCC's undirected view (GIR::buildBCT /
        getBCTComponents / articulationPoint, gir.cpp:2063-2110, 4140-4149)
        -> dict: parent, block (int32 per vertex, see girhip_bicc), ap (bool
        per vertex), n_blocks."""
        V = max(self.V, 1)
        par, blk = np.empty(V, np.int32), np.empty(V, np.int32)
        ap = np.empty(V, np.uint8)
        nb = ctypes.c_int64()
        _check(lib().girhip_bicc(self._h, par.ctypes.data, blk.ctypes.data, ap.ctypes.data, ctypes.byref(nb)))
        return {"parent": par[:self.V], "block": blk[:self.V], "ap": ap[:self.V].astype(bool), "n_blocks": nb.value}

    def track_incremental(self, source, weighted):
        """GIR::trackIncremental (gir_query.cpp:867-891)."""
        _check(lib().girhip_inc_track(self._h, int(source), int(bool(weighted))))

    def update_incremental(self, weighted):
        """GIR::updateIncremental (gir_query.cpp:893-996)."""
        _check(lib().girhip_inc_update(self._h, int(bool(weighted))))

    def inc_dist(self, weighted):
        """GIR::incDist for every vertex (gir.h:229-233) -> uint64 (V,),
        2^64-1 unreachable (BFS: levels)."""
        d = np.empty(max(self.V, 1), np.uint64)
        _check(lib().girhip_inc_dist(self._h, int(bool(weighted)), d.ctypes.data))
        return d[:self.V]

    def features(self, k):
        """16 graph/query features (see csrc/tune.hip) for k query sources."""
        f = np.empty(16, np.float32)
        _check(lib().girhip_features(self._h, int(k), f.ctypes.data))
        return f

    def autotune(self, what, sources, candidates, reps=1):
        """Time every candidate value of knob `what`, keep the fastest ->
        (best value, per-candidate ms)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        cand = np.ascontiguousarray(np.asarray(candidates, dtype=np.float64))
        times = np.empty(len(cand), np.float64)
        best = ctypes.c_double()
        _check(lib().girhip_autotune(self._h, what.encode(), src.ctypes.data if len(src) else None, len(src),
                                     cand.ctypes.data, len(cand), int(reps), ctypes.byref(best), times.ctypes.data))
        return best.value, times

    def learned_tune(self, model, norm, sources, judge=0):
        """Set delta and the SCC schedule (pivot rule, trim rounds) from the
        learned model (defaults as fallback / judge) -> dict(delta, trim,
        raw_delta, used, pivot, schedule_used)."""
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.int32))
        norm = np.ascontiguousarray(np.asarray(norm, dtype=np.float32))
        out = np.zeros(6, np.float64)
        _check(lib().girhip_learned_tune(self._h, model._m, norm.ctypes.data, src.ctypes.data if len(src) else None,
                                         len(src), int(judge), out.ctypes.data))
        return {"delta": out[0], "trim": out[1], "raw_delta": out[2], "used": bool(out[3]), "pivot": out[4],
                "schedule_used": bool(out[5])}

    def all_kernel_stats(self):
        """{kernel name: (launches, ms)} recorded in profile mode."""
        need = ctypes.c_int64()
        _check(lib().girhip_kernel_stat_names(self._h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(lib().girhip_kernel_stat_names(self._h, buf, need.value, None))
        names = [n for n in buf.value.decode().split("\n") if n]
        return {n: self.kernel_stats(n) for n in names}


class Mlp:
    """The learned-tuning MLP (16 -> 64 -> 64 -> 2) on the GPU (girhip_mlp)."""

    NI, NH, NO = 16, 64, 2

    def __init__(self, seed=1):
        m = _vp()
        _check(lib().girhip_mlp_create(int(seed), ctypes.byref(m)))
        self._m = m
        n = ctypes.c_int32()
        _check(lib().girhip_mlp_num_params(ctypes.byref(n)))
        self.num_params = n.value

    def close(self):
        if getattr(self, "_m", None):
            lib().girhip_mlp_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.empty(self.num_params, np.float32)
        _check(lib().girhip_mlp_get_params(self._m, p.ctypes.data))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().girhip_mlp_set_params(self._m, p.ctypes.data))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.NI)
        y = np.empty((len(x), self.NO), np.float32)
        _check(lib().girhip_mlp_forward(self._m, x.ctypes.data, len(x), y.ctypes.data))
        return y

    def grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.NI)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.NO)
        g = np.empty(self.num_params, np.float32)
        loss = ctypes.c_float()
        _check(lib().girhip_mlp_grad(self._m, x.ctypes.data, t.ctypes.data, len(x), g.ctypes.data, ctypes.byref(loss)))
        return loss.value, g

    def train(self, x, t, epochs, lr=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, self.NI)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, self.NO)
        loss = ctypes.c_float()
        _check(lib().girhip_mlp_train(self._m, x.ctypes.data, t.ctypes.data, len(x), int(epochs), float(lr),
                                      ctypes.byref(loss)))
        return loss.value


# ---------------------------------------------------------------------------
# Input formats (libgirgen.so)

def rmat_edges(scale, n_samples, a=0.57, b=0.19, c=0.19, seed=1):
    src = np.empty(n_samples, np.int32)
    dst = np.empty(n_samples, np.int32)
    _gcheck(gen().girgen_rmat_edges(scale, n_samples, a, b, c, seed, src.ctypes.data, dst.ctypes.data))
    return src, dst


def grid_edges(rows, cols, p_keep=0.9, p_bidir=0.8, seed=3):
    n = ctypes.c_int64()
    _gcheck(gen().girgen_grid_edges(rows, cols, p_keep, p_bidir, seed, None, None, ctypes.byref(n)))
    src = np.empty(max(n.value, 1), np.int32)
    dst = np.empty(max(n.value, 1), np.int32)
    _gcheck(gen().girgen_grid_edges(rows, cols, p_keep, p_bidir, seed, src.ctypes.data, dst.ctypes.data,
                                    ctypes.byref(n)))
    return src[:n.value], dst[:n.value]


def uniform_edges(V, n_arcs, seed=4):
    src = np.empty(n_arcs, np.int32)
    dst = np.empty(n_arcs, np.int32)
    _gcheck(gen().girgen_uniform_edges(V, n_arcs, seed, src.ctypes.data, dst.ctypes.data))
    return src, dst


def csr_from_edges(V, src, dst):
    """Canonical fw/bw CSR (self loops and duplicates dropped, rows sorted)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    n = len(src)
    fb, bb = np.empty(V + 1, np.int32), np.empty(V + 1, np.int32)
    fa, ba = np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.int32)
    E = ctypes.c_int64()
    _gcheck(gen().girgen_csr_from_edges(V, n, src.ctypes.data, dst.ctypes.data, fb.ctypes.data, fa.ctypes.data,
                                        bb.ctypes.data, ba.ctypes.data, ctypes.byref(E)))
    return fb, fa[:E.value].copy(), bb, ba[:E.value].copy()


def weights(E, lo, hi, seed):
    w = np.empty(max(E, 1), np.uint32)
    _gcheck(gen().girgen_weights(E, lo, hi, seed, w.ctypes.data))
    return w[:E]


def sources(fw_beg, k, seed):
    fw_beg = np.ascontiguousarray(fw_beg, dtype=np.int32)
    out = np.empty(max(k, 1), np.int32)
    _gcheck(gen().girgen_sources(len(fw_beg) - 1, fw_beg.ctypes.data, k, seed, out.ctypes.data))
    return out[:k]


def write_csr(dirpath, fb, fa, bb, ba):
    os.makedirs(dirpath, exist_ok=True)
    _gcheck(gen().girgen_write_csr(dirpath.encode(), len(fb) - 1, len(fa), fb.ctypes.data, fa.ctypes.data,
                                   bb.ctypes.data, ba.ctypes.data))


def bicc_blocks(parent, block):
    """Blocks as vertex sets from girhip_bicc's forest arrays: { v, parent[v]
    : block[v] = b } per block, plus every root without children (a
    single-vertex SCC) on its own -> sorted list of sorted tuples."""
    parent = np.asarray(parent)
    block = np.asarray(block)
    nonroot = parent >= 0
    v = np.nonzero(nonroot)[0]
    groups = {}
    for b, x, p in zip(block[v].tolist(), v.tolist(), parent[v].tolist()):
        g = groups.setdefault(b, set())
        g.add(x)
        g.add(p)
    has_child = np.zeros(len(parent), bool)
    has_child[parent[nonroot]] = True
    out = [tuple(sorted(g)) for g in groups.values()]
    out += [(int(x),) for x in np.nonzero(~nonroot & ~has_child)[0]]
    return sorted(out)


def export_protocol(dirpath, fb, fa, k_ops, n_sources=8, seed=2027, full=True):
    """Streaming-protocol inputs (loaded.el, full.el, insert/delete/mixed.ops,
    sources.txt, meta.txt) of the reference's baseline drivers
    (bench_main.cpp:280-345, experiments/baselines/common/protocol.h) for the
    canonical CSR (fb, fa): a seeded 50% hold-out (full.el only if full)."""
    os.makedirs(dirpath, exist_ok=True)
    fb = np.ascontiguousarray(fb, np.int32)
    fa = np.ascontiguousarray(fa, np.int32)
    _gcheck(gen().girgen_export_protocol(dirpath.encode(), len(fb) - 1, fb.ctypes.data, fa.ctypes.data, int(k_ops),
                                         int(n_sources), int(seed), int(bool(full))))


def read_protocol(dirpath):
    """The protocol's inputs back as arrays: dict(V, loaded (n, 2), insert /
    delete / mixed: (ops (n,) uint8 1 = insert, arcs (n, 2)), sources)."""
    meta = dict(l.split() for l in open(os.path.join(dirpath, "meta.txt")) if l.strip())
    V = int(meta["V"])

    def el(f):
        a = np.loadtxt(os.path.join(dirpath, f), dtype=np.int64, ndmin=2)
        return a.reshape(-1, 2).astype(np.int32)

    def ops(f):
        rows = [l.split() for l in open(os.path.join(dirpath, f)) if l.strip()]
        ins = np.array([r[0] == "a" for r in rows], np.uint8)
        arcs = np.array([[int(r[1]), int(r[2])] for r in rows], np.int32).reshape(-1, 2)
        return ins, arcs

    return {"V": V, "loaded": el("loaded.el"), "insert": ops("insert.ops"), "delete": ops("delete.ops"),
            "mixed": ops("mixed.ops"),
            "sources": np.loadtxt(os.path.join(dirpath, "sources.txt"), dtype=np.int32, ndmin=1)}


def read_csr(dirpath):
    V, E = ctypes.c_int32(), ctypes.c_int64()
    _gcheck(gen().girgen_read_csr(dirpath.encode(), ctypes.byref(V), ctypes.byref(E), None, None, None, None))
    fb, bb = np.empty(V.value + 1, np.int32), np.empty(V.value + 1, np.int32)
    fa, ba = np.empty(max(E.value, 1), np.int32), np.empty(max(E.value, 1), np.int32)
    _gcheck(gen().girgen_read_csr(dirpath.encode(), ctypes.byref(V), ctypes.byref(E), fb.ctypes.data,
                                  fa.ctypes.data, bb.ctypes.data, ba.ctypes.data))
    return fb, fa[:E.value], bb, ba[:E.value]


# Workloads named by what they are (the seeds and sizes of the benchmark
# configurations).
def rmat_graph(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    V = 1 << scale
    s, d = rmat_edges(scale, edge_factor * V, a, b, c, seed)
    return (V,) + csr_from_edges(V, s, d)


def grid_graph(rows, cols, p_keep=0.9, p_bidir=0.8, seed=3):
    V = rows * cols
    s, d = grid_edges(rows, cols, p_keep, p_bidir, seed)
    return (V,) + csr_from_edges(V, s, d)
