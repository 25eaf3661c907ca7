"""DPFEHM's advection-dispersion interface on the GPU (host side, Python).

* ``transport_residuals`` -- src/transport.jl:2-33 (argument order as there)
* ``transport_steps`` / ``rrule_transport_steps`` -- fixed-step backward
  Euler of du/dt = transport_residuals(u) and its gradient, the role
  DifferentiableBackwardEuler.steps plays in test/transport.jl:30-70 (which
  adapts its steps to 1e-4; here the step count is given)

``coords`` is the reference's (3, n) matrix; ``vxs``, ``vys``, ``vzs`` and
``Ds`` are per edge.  Every call runs the library's HIP kernels.
"""
import ctypes

import numpy as np

from . import _lib
from .api import State, _f64, _opts, _ptr, get_topology


def _args(vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes, dirichletus, Qs,
          volumes, coords):
    Qs = _f64(Qs)
    n = Qs.shape[0]
    topo = get_topology(n, neighbors, areasoverlengths, dirichletnodes)
    c = np.asarray(coords, dtype=np.float64)
    if c.ndim != 2 or c.shape[0] != 3 or c.shape[1] != n:
        raise ValueError("coords must be a 3 x n matrix (the reference asserts 3-D)")
    ne = topo.ne
    a = [_f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (ne,))) for v in
         (vxs, vys, vzs, Ds)]
    a += [_f64(dirichletus), Qs, _f64(volumes)]
    a += [np.ascontiguousarray(c[k]) for k in range(3)]
    return topo, n, a


def transport_residuals(u, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                        dirichletus, Qs, volumes, coords, use_upwind=True):
    """src/transport.jl:2-33: residuals at every cell."""
    topo, n, a = _args(vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                       dirichletus, Qs, volumes, coords)
    r = np.empty(n)
    o = _opts({})
    _lib.check(_lib.lib().dpf_transport_residuals(topo.handle, _ptr(_f64(u)),
                                                  *[_ptr(x) for x in a], int(bool(use_upwind)),
                                                  _ptr(r), ctypes.byref(o)))
    return r


def _steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes, dirichletus, Qs,
           volumes, coords, t0, t1, nsteps, use_upwind, return_all, keep, stats, kwargs):
    topo, n, a = _args(vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                       dirichletus, Qs, volumes, coords)
    o = _opts(kwargs)
    u = np.empty(n)
    allu = np.empty((int(nsteps) + 1, n)) if return_all else None
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    code = _lib.lib().dpf_transport_steps(
        topo.handle, _ptr(_f64(u0)), *[_ptr(x) for x in a], int(bool(use_upwind)), float(t0),
        float(t1), int(nsteps), _ptr(u), _ptr(allu) if return_all else None, ctypes.byref(o),
        ctypes.byref(st), ctypes.byref(sh) if keep else None)
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return (allu.T if return_all else u), (State(topo, sh) if keep else None), o, topo


def transport_steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                    dirichletus, Qs, volumes, coords, t0, t1, nsteps, use_upwind=True,
                    return_all=False, stats=None, **kwargs):
    """u at t1 (n), or the n x (nsteps + 1) matrix of every state (the
    reference's ``steps`` output) when ``return_all``."""
    out, _, _, _ = _steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                          dirichletus, Qs, volumes, coords, t0, t1, nsteps, use_upwind,
                          return_all, False, stats, kwargs)
    return out


def rrule_transport_steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                          dirichletus, Qs, volumes, coords, t0, t1, nsteps, use_upwind=True,
                          stats=None, **kwargs):
    """Returns (u_final, pullback); pullback(dJ_du) gives
    dict(vxs, vys, vzs, Ds, Qs)."""
    u, state, o, topo = _steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths,
                               dirichletnodes, dirichletus, Qs, volumes, coords, t0, t1, nsteps,
                               use_upwind, False, True, stats, kwargs)
    n, ne = u.shape[0], topo.ne

    def pullback(g, stats=None):
        outs = {k: np.empty(ne) for k in ("vxs", "vys", "vzs", "Ds")}
        outs["Qs"] = np.empty(n)
        st = _lib.Stats()
        code = _lib.lib().dpf_transport_pullback(
            state._h, _ptr(_f64(g)), _ptr(outs["vxs"]), _ptr(outs["vys"]), _ptr(outs["vzs"]),
            _ptr(outs["Ds"]), _ptr(outs["Qs"]), ctypes.byref(o), ctypes.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        _lib.check(code)
        return outs

    pullback.state = state
    return u, pullback
