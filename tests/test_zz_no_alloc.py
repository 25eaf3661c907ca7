"""dpf_hm_forward_noalloc, the allocation-free form of dpf_hm_forward: its
staging is a fixed device array shared by every call on the GPU.  Two calls
in a row (the second on a smaller mesh, so the array holds the first call's
longer contents) must give K_e bit for bit equal to dpf_hm_forward's on the
same device inputs; host memory and oversized meshes are refused.

Device memory and the stream come from the HIP runtime the library links
(ctypes), not from torch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_matches_hm_forward(pkg):
    from dpfehm_jl_amd import _lib, api
    L = _lib.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")
    vp = ctypes.c_void_p

    def ok(code):
        assert code == 0, code

    stream = vp()
    ok(hip.hipStreamCreateWithFlags(ctypes.byref(stream), 1))
    bufs = []

    def dalloc(nbytes):
        p = vp()
        ok(hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)))
        bufs.append(p)
        return p

    try:
        for ns in ((96, 96, 48), (40, 36, 20)):
            print("noalloc", ns, flush=True)
            w = pkg.workloads.steady_2m(ns=ns, maxs=(0.5 * ns[0], 0.5 * ns[1], 0.5 * ns[2]))
            topo = api.get_topology(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes)
            lk = np.ascontiguousarray(w.logK)
            d_lk, d_a, d_b = dalloc(8 * w.n), dalloc(8 * topo.ne), dalloc(8 * topo.ne)
            ok(hip.hipMemcpy(d_lk, lk.ctypes.data_as(vp), ctypes.c_size_t(8 * w.n), 1))
            o = api._opts({}, memory=_lib.DPF_DEVICE, stream=stream)
            _lib.check(L.dpf_hm_forward_noalloc(topo.handle, d_lk, d_a, ctypes.byref(o)))
            ok(hip.hipStreamSynchronize(stream))
            _lib.check(L.dpf_hm_forward(topo.handle, d_lk, d_b, ctypes.byref(o)))
            ok(hip.hipStreamSynchronize(stream))
            a, b = np.empty(topo.ne), np.empty(topo.ne)
            ok(hip.hipMemcpy(a.ctypes.data_as(vp), d_a, ctypes.c_size_t(8 * topo.ne), 2))
            ok(hip.hipMemcpy(b.ctypes.data_as(vp), d_b, ctypes.c_size_t(8 * topo.ne), 2))
            assert np.array_equal(a, b)
            assert np.all(np.isfinite(a)) and np.all(a > 0)
            # host memory is refused (the staging of host inputs would allocate)
            oh = api._opts({})
            assert L.dpf_hm_forward_noalloc(topo.handle, lk.ctypes.data_as(vp),
                                            a.ctypes.data_as(vp), ctypes.byref(oh)) == \
                _lib.DPF_EINVAL
    finally:
        for p in bufs:
            hip.hipFree(p)
        hip.hipStreamDestroy(stream)


def test_noalloc_refuses_oversized_mesh(pkg):
    """A mesh of more cells than the staging holds is refused before any
    launch, with a message naming the limit."""
    from dpfehm_jl_amd import _lib, api
    L = _lib.lib()
    n = 4194304 + 64
    # a chain of cells, one Dirichlet end: cheap to set up at this size
    nb = np.stack([np.arange(1, n), np.arange(2, n + 1)], 1).astype(np.int64)
    topo = api.Topology(n, nb, np.ones(n - 1), np.array([1], dtype=np.int64))
    hip = ctypes.CDLL("libamdhip64.so.7")
    d_lk, d_Ke = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_lk), ctypes.c_size_t(8 * n)) == 0
    assert hip.hipMalloc(ctypes.byref(d_Ke), ctypes.c_size_t(8 * (n - 1))) == 0
    try:
        o = api._opts({}, memory=_lib.DPF_DEVICE, stream=None)
        code = L.dpf_hm_forward_noalloc(topo.handle, d_lk, d_Ke, ctypes.byref(o))
        assert code == _lib.DPF_EINVAL
        assert b"4194304" in L.dpf_last_error()
    finally:
        hip.hipFree(d_lk)
        hip.hipFree(d_Ke)
        topo.close()
