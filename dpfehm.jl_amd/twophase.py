"""DPFEHM's two-phase (IMPES) interface on the GPU (host side, Python).

Same names, argument meaning and order as src/twophase.jl:

* ``solvepressure`` / ``rrule_solvepressure`` -- src/twophase.jl:48-80: the
  transmissivity system (tx = 2 aol) for edge permeabilities, and its
  pullback (dKs_neighbors, ddirichleths, dQs)
* ``tpfa``          -- src/twophase.jl:189-198: arithmetic-mean K, pressure,
  Darcy fluxes Vn
* ``upstream``      -- src/twophase.jl:208-235: CFL-sub-stepped upwind
  saturation transport
* ``solvetwophase`` -- src/twophase.jl:238-264: the IMPES time loop
* ``rrule_solvetwophase`` -- its reverse mode (dK, dS0), the gradient the
  reference's test/twophase.jl:192-205 takes with Zygote

``fluid`` is any object with ``vw``, ``vo``, ``swc``, ``sor`` attributes (the
reference's ``Fluid``), or the 4-sequence ``(vw, vo, swc, sor)``.

Every call runs the library's HIP kernels.  A topology without Dirichlet cells
(the reference's no-flow examples) makes the pressure system singular; the
calls then pin the last cell at head 0 and return the pressure shifted to mean
zero.  The reference's AMG-CG iterate carries some other constant; fluxes and
saturations do not depend on it.  Pressure solves default to reltol 1e-12 (the
reference asks its AMG for 1e-14, which an FP64 Krylov solve of a large
system does not reliably reach).
"""
import ctypes

import numpy as np

from . import _lib
from .api import (NoTangent, State, _f64, _opts, _ptr, get_topology,
                  rrule_groundwater_steadystate)

_REL = 1e-12


def _fluid(fluid):
    if all(hasattr(fluid, a) for a in ("vw", "vo", "swc", "sor")):
        v = (fluid.vw, fluid.vo, fluid.swc, fluid.sor)
    else:
        v = tuple(fluid)
        if len(v) != 4:
            raise ValueError("fluid must have vw, vo, swc, sor")
    return np.ascontiguousarray(v, dtype=np.float64)


def _pinned(n, dirichletnodes, dirichleths):
    """(dirichletnodes, dirichleths, zero_mean) with the last cell pinned when
    there is no Dirichlet cell."""
    dn = np.asarray(dirichletnodes, dtype=np.int64).ravel()
    dh = _f64(np.asarray(dirichleths, dtype=np.float64)).copy()
    if dh.shape[0] != n:
        raise ValueError("dirichleths must have one entry per cell")
    if dn.size:
        return dn, dh, 0
    dh[n - 1] = 0.0
    return np.array([n], dtype=np.int64), dh, 1


def _kw(kwargs):
    kw = dict(kwargs)
    kw.setdefault("reltol", _REL)
    return kw


def rrule_solvepressure(Ks_neighbors, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                        Qs, **kwargs):
    """src/twophase.jl:59-80.  Returns (h, pullback); pullback(delta) gives
    (NoTangent, dKs_neighbors, NoTangent, NoTangent, NoTangent, ddirichleths,
    dQs).  The transmissivity operator with K_e is the groundwater operator
    with 2 K_e (tx = 2 aol, src/twophase.jl:26), so the solve and its adjoint
    run on the groundwater path, and dK_e = 2 dK'_e."""
    Qs = _f64(Qs)
    n = Qs.shape[0]
    dn, dh, zm = _pinned(n, dirichletnodes, dirichleths)
    K2 = 2.0 * _f64(Ks_neighbors)
    h, pb = rrule_groundwater_steadystate(K2, neighbors, areasoverlengths, dn, dh, Qs,
                                          **_kw(kwargs))
    shift = h.sum() / n if zm else 0.0

    def pullback(delta):
        # the mean shift is linear: delta -> delta - mean(delta) on every cell
        d = _f64(delta)
        if zm:
            d = d - d.sum() / n
        t = pb(d)
        ddh = t[5]
        if zm:
            ddh = ddh.copy()
            ddh[n - 1] = 0.0  # the pinned head is not an input
        return (NoTangent, 2.0 * t[1], NoTangent, NoTangent, NoTangent, ddh, t[6])

    return h - shift, pullback


def solvepressure(Ks_neighbors, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                  **kwargs):
    """src/twophase.jl:48-56.  Returns the heads (n)."""
    h, _ = rrule_solvepressure(Ks_neighbors, neighbors, areasoverlengths, dirichletnodes,
                               dirichleths, Qs, **kwargs)
    return h


def tpfa(Ks, dirichleths, dirichletnodes, Qs, areasoverlengths, neighbors, stats=None,
         **kwargs):
    """src/twophase.jl:189-198 (argument order as there).  Returns (P, Vn)."""
    Qs = _f64(Qs)
    n = Qs.shape[0]
    dn, dh, zm = _pinned(n, dirichletnodes, dirichleths)
    topo = get_topology(n, neighbors, areasoverlengths, dn)
    K = _f64(Ks)
    if K.shape[0] != n:
        raise ValueError("Ks must have one entry per cell")
    P = np.empty(n)
    Vn = np.empty(topo.ne)
    o = _opts(_kw(kwargs))
    st = _lib.Stats()
    code = _lib.lib().dpf_twophase_tpfa(topo.handle, _ptr(K), _ptr(dh), _ptr(Qs), _ptr(P),
                                        _ptr(Vn), zm, ctypes.byref(o), ctypes.byref(st))
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return P, Vn


def upstream(S, fluid, Qs, T, P, Vn, neighbors, volumes, dirichletnodes=(),
             areasoverlengths=None):
    """src/twophase.jl:208-235 (argument order as there).  Returns the new
    saturation.  ``dirichletnodes`` / ``areasoverlengths`` only select the
    cached topology (the transport itself uses neither)."""
    Qs = _f64(Qs)
    n = Qs.shape[0]
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = np.ones(nb.shape[0]) if areasoverlengths is None else areasoverlengths
    dn, _, _ = _pinned(n, dirichletnodes, np.zeros(n))
    topo = get_topology(n, nb, aol, dn)
    out = _f64(S).copy()
    Vn = _f64(Vn)
    if out.shape[0] != n or _f64(P).shape[0] != n or Vn.shape[0] != topo.ne:
        raise ValueError("S and P need one entry per cell, Vn one per edge")
    nsub = ctypes.c_int64()
    o = _opts({})
    _lib.check(_lib.lib().dpf_twophase_upstream(topo.handle, _ptr(out), _ptr(_fluid(fluid)),
                                                _ptr(Qs), float(T), _ptr(_f64(P)), _ptr(Vn),
                                                _ptr(_f64(volumes)), ctypes.byref(nsub),
                                                ctypes.byref(o)))
    upstream.last_substeps = int(nsub.value)
    return out


def _solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes, areasoverlengths,
                   fluid, dt, neighbors, nt, everyStep, stats, keep, kwargs):
    Qs = _f64(Qs)
    n = Qs.shape[0]
    nt = int(nt)
    dn, dh, zm = _pinned(n, dirichletnodes, dirichleths)
    topo = get_topology(n, neighbors, areasoverlengths, dn)
    P, S = np.empty(n), np.empty(n)
    Pall = np.empty((nt, n)) if everyStep else None
    Sall = np.empty((nt, n)) if everyStep else None
    nsub = ctypes.c_int64()
    o = _opts(_kw(kwargs))
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    code = _lib.lib().dpf_solvetwophase(
        topo.handle, _ptr(_f64(h0)), _ptr(_f64(S0)), _ptr(_f64(K)), _ptr(dh), _ptr(Qs),
        _ptr(_f64(volumes)), _ptr(_fluid(fluid)), float(dt), nt, zm, _ptr(P), _ptr(S),
        _ptr(Pall) if everyStep else None, _ptr(Sall) if everyStep else None,
        ctypes.byref(nsub), ctypes.byref(sh) if keep else None, ctypes.byref(o),
        ctypes.byref(st))
    if stats is not None:
        stats.update(st.as_dict())
        stats["substeps"] = int(nsub.value)
    _lib.check(code)
    out = (list(Pall), list(Sall)) if everyStep else (P, S)
    return out, (State(topo, sh) if keep else None), o


def solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes, areasoverlengths, fluid,
                  dt, neighbors, nt, everyStep=False, stats=None, **kwargs):
    """src/twophase.jl:238-264 (the reference's ``args`` tuple, in order).
    Returns (P, S) of the last step, or the lists of every step's P and S
    when ``everyStep``."""
    out, _, _ = _solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes,
                               areasoverlengths, fluid, dt, neighbors, nt, everyStep, stats,
                               False, kwargs)
    return out


def rrule_solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes, areasoverlengths,
                        fluid, dt, neighbors, nt, stats=None, **kwargs):
    """Forward ``solvetwophase`` (last step) and its pullback: the gradient
    test/twophase.jl:192-205 takes with Zygote.  ``pullback(dP, dS)`` (either
    may be None) returns (dh0, dS0, dK): dh0 is zero for nt >= 1 (every step
    recomputes the pressure).  Tangents in dirichleths, Qs, volumes,
    areasoverlengths, fluid and dt are not formed."""
    (P, S), state, o = _solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes,
                                      areasoverlengths, fluid, dt, neighbors, nt, False, stats,
                                      True, kwargs)
    n = P.shape[0]

    def pullback(dP=None, dS=None, stats=None):
        if int(nt) == 0:
            return (_f64(dP) if dP is not None else np.zeros(n),
                    _f64(dS) if dS is not None else np.zeros(n), np.zeros(n))
        dK, dS0 = np.empty(n), np.empty(n)
        st = _lib.Stats()
        code = _lib.lib().dpf_solvetwophase_pullback(
            state._h, _ptr(_f64(dP)) if dP is not None else None,
            _ptr(_f64(dS)) if dS is not None else None, _ptr(dK), _ptr(dS0), ctypes.byref(o),
            ctypes.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        _lib.check(code)
        return np.zeros(n), dS0, dK

    pullback.state = state
    return (P, S), pullback
