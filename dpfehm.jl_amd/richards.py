"""DPFEHM's steady Richards interface on the GPU (host side, Python).

Same names, argument meaning and order as src/richards.jl:

* ``richards_residuals``         -- src/richards.jl:97-117
* ``richards_steadystate``       -- src/richards.jl:156-179
* ``rrule_richards_steadystate`` -- src/richards.jl:181-205: returns
  ``(psi, pullback)``; ``pullback(delta)`` gives the 11-tuple (NoTangent,
  dpsi0 = 0, dKs, NoTangent x 3, ddirichletpsis, NoTangent, dalphas, dNs, dQs)

``coords`` is the reference's (3, n) matrix; only its third row (the
elevation) enters the equations.  ``alphas`` and ``Ns`` are per edge, as in
the reference.  The reference solves with NLsolve to its default residual
tolerance; here Newton's method runs on the GPU until max|r| falls below
``reltol`` (default 1e-10) times its initial value, each Newton system solved
by BiCGStab with the AMG V-cycle of the Picard operator as preconditioner
(``precision="mixed_pc"`` runs that V-cycle in FP32; everything else stays
FP64).  Non-convergence raises instead of returning.
"""
import ctypes

import numpy as np

from . import _lib
from .api import NoTangent, State, _f64, _opts, _ptr, get_topology


def _z(coords, n):
    c = np.asarray(coords, dtype=np.float64)
    if c.ndim != 2 or c.shape[0] != 3 or c.shape[1] != n:
        raise ValueError("coords must be a 3 x n matrix")
    return np.ascontiguousarray(c[2])


def _inputs(Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis, coords, alphas, Ns,
            Qs):
    Qs = _f64(Qs)
    n = Qs.shape[0]
    topo = get_topology(n, neighbors, areasoverlengths, dirichletnodes)
    arrs = dict(Ks=_f64(Ks), dirichletpsis=_f64(dirichletpsis), z=_z(coords, n),
                alphas=_f64(np.broadcast_to(np.asarray(alphas, dtype=np.float64), (topo.ne,))),
                Ns=_f64(np.broadcast_to(np.asarray(Ns, dtype=np.float64), (topo.ne,))), Qs=Qs)
    if arrs["Ks"].shape[0] != topo.ne or arrs["dirichletpsis"].shape[0] != n:
        raise ValueError("Ks needs one entry per edge, dirichletpsis one per cell")
    return topo, n, arrs


def richards_residuals(psi, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                       coords, alphas, Ns, Qs, specificstorage, volumes):
    """src/richards.jl:97-117.  Returns the residuals at the free cells."""
    topo, n, a = _inputs(Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis, coords,
                         alphas, Ns, Qs)
    psi = _f64(psi)
    if psi.shape[0] != topo.nfree:
        raise ValueError("psi must have one entry per free cell")
    r = np.empty(topo.nfree)
    o = _opts({})
    _lib.check(_lib.lib().dpf_richards_residuals(
        topo.handle, _ptr(psi), _ptr(a["Ks"]), _ptr(a["dirichletpsis"]), _ptr(a["z"]),
        _ptr(a["alphas"]), _ptr(a["Ns"]), _ptr(a["Qs"]), _ptr(_f64(specificstorage)),
        _ptr(_f64(volumes)), _ptr(r), ctypes.byref(o)))
    return r


def _steady(psi0, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis, coords,
            alphas, Ns, Qs, keep, stats, kwargs):
    topo, n, a = _inputs(Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis, coords,
                         alphas, Ns, Qs)
    kw = dict(kwargs)
    kw.pop("callback", None)
    o = _opts(kw)
    psi = np.empty(n)
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    code = _lib.lib().dpf_richards_steady(
        topo.handle, _ptr(_f64(psi0)), _ptr(a["Ks"]), _ptr(a["dirichletpsis"]), _ptr(a["z"]),
        _ptr(a["alphas"]), _ptr(a["Ns"]), _ptr(a["Qs"]), _ptr(psi), ctypes.byref(o),
        ctypes.byref(st), ctypes.byref(sh) if keep else None)
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return psi, (State(topo, sh) if keep else None), o, topo


def richards_steadystate(psi0, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                         coords, alphas, Ns, Qs, stats=None, **kwargs):
    """src/richards.jl:156-179.  Returns psi (n)."""
    psi, _, _, _ = _steady(psi0, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                           coords, alphas, Ns, Qs, False, stats, kwargs)
    return psi


def rrule_richards_steadystate(psi0, Ks, neighbors, areasoverlengths, dirichletnodes,
                               dirichletpsis, coords, alphas, Ns, Qs, stats=None, **kwargs):
    """src/richards.jl:181-205.  Returns (psi, pullback)."""
    psi, state, o, topo = _steady(psi0, Ks, neighbors, areasoverlengths, dirichletnodes,
                                  dirichletpsis, coords, alphas, Ns, Qs, True, stats, kwargs)
    n, ne = psi.shape[0], topo.ne

    def pullback(delta, stats=None):
        delta = _f64(delta)
        if delta.shape[0] != n:
            raise ValueError("delta must have one entry per cell")
        dK, dal, dN = np.empty(ne), np.empty(ne), np.empty(ne)
        ddp, dQ = np.empty(n), np.empty(n)
        st = _lib.Stats()
        code = _lib.lib().dpf_richards_pullback(state._h, _ptr(delta), _ptr(dK), _ptr(ddp),
                                                _ptr(dal), _ptr(dN), _ptr(dQ), ctypes.byref(o),
                                                ctypes.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        _lib.check(code)
        return (NoTangent, np.zeros(n), dK, NoTangent, NoTangent, NoTangent, ddp, NoTangent, dal,
                dN, dQ)

    pullback.state = state
    return psi, pullback
