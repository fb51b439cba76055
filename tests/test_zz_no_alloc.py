"""The allocation-free entry point girhip_bfs_noalloc, through the C ABI:
the same levels as girhip_bfs (k = 1) on two calls in a row, the second on
a smaller graph, no device memory allocated by the calls, host outputs and
sizes beyond the scratch array refused.  Device buffers come from the HIP
runtime the library itself uses (libamdhip64.so.7, through ctypes)."""
import ctypes

import numpy as np
import pytest

import girhip

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def hip():
    girhip.lib()  # loads the runtime; the name below resolves to that copy
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    h.hipDeviceSynchronize.argtypes = []
    return h


def dmalloc(h, nbytes):
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), nbytes) == 0
    return p.value


def free_bytes(h):
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def test_bfs_noalloc_matches_bfs():
    H = hip()
    cases = [girhip.rmat_graph(16, seed=61), girhip.grid_graph(50, 70, seed=62)]  # second: smaller
    for ci, (V, fb, fa, bb, ba) in enumerate(cases):
        say("bfs_noalloc case", ci, "V", V, "E", len(fa))
        with girhip.Graph(fb, fa, bb, ba) as G:
            srcs = [int(x) for x in girhip.sources(fb, 3, 63 + ci)] + [-1, V]
            want = G.bfs(srcs)  # the existing entry point (allocates its scratch)
            outs = [dmalloc(H, 4 * V) for _ in srcs]
            try:
                # one call before the measurement: the runtime loads the entry
                # point's code object on its first launch in a process
                G.bfs_noalloc(srcs[0], outs[0])
                free0 = free_bytes(H)
                for x, o in zip(srcs, outs):
                    G.bfs_noalloc(x, o)
                assert free_bytes(H) == free0, "bfs_noalloc allocated device memory"
                for i, (x, o) in enumerate(zip(srcs, outs)):
                    got = np.empty(V, np.int32)
                    assert H.hipMemcpy(got.ctypes.data, o, 4 * V, 2) == 0  # device to host
                    assert np.array_equal(got, want[i]), (ci, x)
                host = np.empty(V, np.int32)
                with pytest.raises(girhip.GirError) as e:
                    G.bfs_noalloc(srcs[0], host.ctypes.data)  # host output refused
                assert e.value.code == 1
            finally:
                for o in outs:
                    H.hipFree(o)


def test_bfs_noalloc_refuses_graphs_beyond_its_scratch():
    H = hip()
    V = (1 << 22) + 1
    fb = np.zeros(V + 1, np.int32)
    fa = np.zeros(0, np.int32)
    say("bfs_noalloc size limit V", V)
    with girhip.Graph(fb, fa) as G:
        out = dmalloc(H, 4 * V)
        try:
            with pytest.raises(girhip.GirError) as e:
                G.bfs_noalloc(0, out)
            assert e.value.code == 1  # GIRHIP_EINVAL
        finally:
            H.hipFree(out)
