"""Multi-core FP64 CPU baseline -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/_build/libamg_cpu.so`` (``amg_cpu.cpp``): one
forward + adjoint evaluation of the steady groundwater inverse problem with
smoothed-aggregation AMG-PCG and OpenMP on every host core.  Only the
``cpu_baseline`` leg of ``bench.py`` and ``tests/`` use it; the product never
does.  It is a port of the same mathematics, not the reference (pure Julia,
not installed here).  Pinned against the scalar oracle (``gw_oracle``) by
``tests/test_oracle.py::test_cpu_amg_baseline_matches_oracle``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libamg_cpu.so")
_lib_handle = None


def _lib():
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(_SO):
            from . import gw_oracle
            gw_oracle.build()
        L = ctypes.CDLL(_SO)
        vp = ctypes.c_void_p
        L.acpu_create.restype = vp
        L.acpu_create.argtypes = [ctypes.c_int64, ctypes.c_int64, vp, vp, ctypes.c_int64, vp,
                                  ctypes.c_int]
        L.acpu_destroy.argtypes = [vp]
        L.acpu_info.argtypes = [vp, vp]
        L.acpu_eval.restype = ctypes.c_int
        L.acpu_eval.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_double, vp, vp,
                                vp, vp]
        _lib_handle = L
    return _lib_handle


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class CpuAmg:
    """Mesh set-up once (aggregates and sparsity patterns), then ``eval``."""

    def __init__(self, n, neighbors, areasoverlengths, dirichletnodes, threads=0):
        nb = np.ascontiguousarray(np.asarray(neighbors, dtype=np.int64).reshape(-1, 2))
        aol = np.ascontiguousarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
        dn = np.ascontiguousarray(dirichletnodes, dtype=np.int64)
        self.n = int(n)
        self._h = _lib().acpu_create(self.n, nb.shape[0], _p(nb), _p(np.ascontiguousarray(aol)),
                                     dn.size, _p(dn), int(threads))
        if not self._h:
            raise ValueError("amg_cpu: singular mesh or coarsest level too large")
        info = np.zeros(4)
        _lib().acpu_info(self._h, _p(info))
        self.levels, self.nnz, self.operator_complexity, self.threads = info
        self.threads = int(self.threads)

    def eval(self, logK, dirichleths, Qs, obs, hobs, reltol=1e-10, want_h=False):
        """J, dJ/dlogK (and h) of J = sum over obs (1-based) of (h - hobs)^2;
        info = [forward iterations, adjoint iterations, set-up ms, forward
        solve ms, adjoint solve + gradient ms]."""
        lk = np.ascontiguousarray(logK, dtype=np.float64)
        dh = np.ascontiguousarray(dirichleths, dtype=np.float64)
        Q = np.ascontiguousarray(Qs, dtype=np.float64)
        ob = np.ascontiguousarray(obs, dtype=np.int64)
        ho = np.ascontiguousarray(hobs, dtype=np.float64)
        J = np.zeros(1)
        g = np.empty(self.n)
        h = np.empty(self.n) if want_h else None
        info = np.zeros(5)
        rc = _lib().acpu_eval(self._h, _p(lk), _p(dh), _p(Q), _p(ob), _p(ho), ob.size,
                              float(reltol), _p(J), _p(g), _p(h) if want_h else None, _p(info))
        if rc:
            raise RuntimeError("amg_cpu: solve failed")
        return (float(J[0]), g, info, h) if want_h else (float(J[0]), g, info)

    def close(self):
        if self._h:
            _lib().acpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
