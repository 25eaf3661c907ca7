"""DPFEHM's groundwater interface on the GPU (host side, Python).

Same names, argument meaning and error behaviour as the reference's Julia
functions, with 1-based ``neighbors`` (an ``(E, 2)`` array of pairs) and
``dirichletnodes``:

* ``groundwater_steadystate``  -- src/groundwater.jl:75-83
* ``rrule_groundwater_steadystate`` -- the ChainRulesCore.rrule of
  src/groundwater.jl:85-109: returns ``(h, pullback)``, ``pullback(delta)``
  returns the 7-tuple ``(NoTangent, dKs, NoTangent, NoTangent, NoTangent,
  ddirichleths, dQs)``
* ``groundwater_residuals``    -- src/groundwater.jl:1-38
* ``hm``, ``logKs2Ks_harmonic`` and its pullback -- DPFEHM.hm
  (src/richards.jl:50-52) applied per edge as the configs' logK -> K map

Every call goes through libdpfgpu.so (HIP kernels on the GPU).  The topology
(free-cell maps, Jacobian pattern, AMG symbolic setup) is cached per
``(n, neighbors, areasoverlengths, dirichletnodes)`` content, as the reference
recomputes it in every call.

Keyword arguments: ``reltol`` (default 1e-10), ``abstol``, ``maxiter``,
``precision`` ("fp64", "mixed" or "mixed_pc"), ``inner_reltol``, ``max_outer``,
``cheb_degree``, ``graph``, ``fp32_storage`` (the FP32 V-cycle reads FP32
instead of bf16 operator values), ``autotune`` (the topology's first solve in
a precision times smoother degrees and coarse-level lanes per row and keeps
the fastest; without it the defaults or pinned values apply, and results do
not depend on timing).  ``linear_solver`` selects amg_solver (the
default: AMG-PCG to ``reltol``) or cholesky_solver (the same AMG-PCG to a
relative residual of 1e-14, a direct solve's accuracy; tolerance keywords
ignored), as a function or by name; the reference's rrule always uses AMG
(src/groundwater.jl:90), and here the rrule honours the choice.  Unknown
keywords raise TypeError.  A steady system with a component that no edge of nonzero
conductance connects to a Dirichlet cell (e.g. ``dirichletnodes=[]``) raises
``SingularSystem``.  Non-convergence raises ``NotConverged`` instead of returning
silently.
"""
import ctypes
import threading

import numpy as np

from . import _lib
from ._lib import DpfError, NotConverged, SingularSystem  # noqa: F401

try:
    import xxhash

    def _digest(*arrays):
        h = xxhash.xxh3_128()
        for a in arrays:
            h.update(np.ascontiguousarray(a).view(np.uint8))
        return h.hexdigest()
except ImportError:  # pragma: no cover
    import hashlib

    def _digest(*arrays):
        h = hashlib.blake2b(digest_size=16)
        for a in arrays:
            h.update(np.ascontiguousarray(a).view(np.uint8))
        return h.hexdigest()


class _NoTangent:
    def __repr__(self):
        return "NoTangent()"


NoTangent = _NoTangent()

_OPT_KEYS = {"reltol", "abstol", "maxiter", "precision", "inner_reltol", "max_outer",
             "cheb_degree"}


def amg_solver(A=None, b=None, **kwargs):
    """Marker for ``linear_solver=`` (src/groundwater.jl:45-49), the default:
    AMG-preconditioned CG to ``reltol`` on the GPU.  Not callable on its own."""
    raise TypeError("amg_solver selects the GPU solve through linear_solver=; it is not "
                    "called directly")


def cholesky_solver(A=None, b=None, **kwargs):
    """Marker for ``linear_solver=`` (src/groundwater.jl:56-60).  There is no
    sparse Cholesky on the GPU: the system is solved by the same AMG-PCG to a
    relative residual of 1e-14, the accuracy a direct solve delivers, and
    tolerance keyword arguments are ignored as the reference's Cholesky
    ignores them.  Not callable on its own."""
    raise TypeError("cholesky_solver selects the GPU solve through linear_solver=; it is not "
                    "called directly")


def _opts(kwargs, memory=_lib.DPF_HOST, stream=None):
    """dpf_opts from keyword arguments; ``graph=True`` sets DPF_FLAG_GRAPH
    (each PCG iteration replayed as one captured hipGraph), ``fp32_storage=True``
    DPF_FLAG_FP32_STORAGE, ``autotune=True`` DPF_FLAG_AUTOTUNE; ``linear_solver``
    is amg_solver (default) or cholesky_solver, as a function or by name."""
    kw = dict(kwargs)
    ls = kw.pop("linear_solver", None)
    name = getattr(ls, "__name__", ls)
    if name not in (None, "amg_solver", "cholesky_solver"):
        raise TypeError(f"linear_solver must be amg_solver or cholesky_solver, not {ls!r}")
    if name == "cholesky_solver":
        kw["reltol"], kw["abstol"] = 1e-14, 0.0
    graph = bool(kw.pop("graph", False))
    fp32_storage = bool(kw.pop("fp32_storage", False))
    autotune = bool(kw.pop("autotune", False))
    bad = set(kw) - _OPT_KEYS
    if bad:
        raise TypeError(f"unexpected keyword argument(s): {sorted(bad)}")
    o = _lib.default_opts(**kw)
    if graph:
        o.flags |= _lib.DPF_FLAG_GRAPH
    if fp32_storage:
        o.flags |= _lib.DPF_FLAG_FP32_STORAGE
    if autotune:
        o.flags |= _lib.DPF_FLAG_AUTOTUNE
    o.memory = memory
    o.stream = stream
    return o


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Topology:
    """A mesh + Dirichlet set cached on the GPU (dpf_topology)."""

    def __init__(self, n, neighbors, areasoverlengths, dirichletnodes):
        nb = np.ascontiguousarray(np.asarray(neighbors, dtype=np.int64).reshape(-1, 2))
        ne = nb.shape[0]
        aol = _f64(np.asarray(areasoverlengths, dtype=np.float64)[:ne])
        if aol.shape[0] != ne:
            raise ValueError("areasoverlengths is shorter than neighbors")
        dn = np.ascontiguousarray(np.asarray(dirichletnodes, dtype=np.int64).ravel())
        self.n, self.ne = int(n), int(ne)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().dpf_topology_create(self.n, self.ne, _ptr(nb), _ptr(aol),
                                                  dn.size, _ptr(dn), ctypes.byref(h)))
        self._h = h
        info = self.info()
        self.nfree = int(info[2])
        self.nnz = int(info[3])
        self.levels = int(info[4])
        self.coarse_rows = int(info[5])
        self.operator_complexity = info[6] / 1000.0
        self.symbolic_setup_ms = info[7] / 1000.0
        isfree = np.ones(self.n, dtype=bool)
        isfree[dn - 1] = False
        self.isfree = isfree
        self.lock = threading.Lock()

    def info(self):
        info = np.zeros(21, dtype=np.int64)
        _lib.lib().dpf_topology_info(self._h, _ptr(info), 21)
        return info

    def layout(self):
        """{'fine_offset_slices': k}: how many of the fine operator's 64-row
        slices store column offsets instead of column indices."""
        return {"fine_offset_slices": int(self.info()[12])}

    def tuned_degrees(self):
        """{precision: (finest degree, second-level degree, deeper levels'
        degree, interval ratio)} in use (defaults, pinned or autotuned)."""
        i = self.info()
        return {"fp64": (int(i[8]), int(i[9]), int(i[18]), int(i[15])),
                "mixed": (int(i[10]), int(i[11]), int(i[19]), int(i[16])),
                "mixed_pc": (int(i[13]), int(i[14]), int(i[20]), int(i[17]))}

    def lanes(self):
        """(lanes exponent per level, {(level, exponent): sweep us}): each
        level's split copy has 2^lg lanes a row (0: no split copy); the
        autotuner's sweep timings where it ran."""
        L = self.levels
        lg = np.zeros(L, dtype=np.int32)
        us = np.zeros(7 * L)
        _lib.check(_lib.lib().dpf_topology_lanes(self._h, _ptr(lg), _ptr(us), L))
        trials = {(l, k): float(us[7 * l + k]) for l in range(L) for k in range(7)
                  if us[7 * l + k] > 0}
        return [int(v) for v in lg], trials

    def set_lanes(self, level, lg):
        """Pin level ``level``'s split copy at 2^lg lanes a row."""
        _lib.check(_lib.lib().dpf_topology_set_lanes(self._h, int(level), int(lg)))

    def tuning(self, precision="fp64"):
        """The autotuner's trials for ``precision``: [(ms, iterations)] for
        the degree triples (finest, second, deeper levels) (1,1,2), (1,1,3),
        (1,1,4), (1,2,2), (1,2,3) at interval ratio 15, then the best triple at
        ratios 30 and 8 (empty before tuning)."""
        ms = np.zeros(7)
        its = np.zeros(7, dtype=np.int32)
        _lib.check(_lib.lib().dpf_topology_tuning(self._h, _lib.precision_code(precision),
                                                  _ptr(ms), _ptr(its), 7))
        return [(float(a), int(b)) for a, b in zip(ms, its)] if its.any() else []

    def set_tuning(self, precision, fine, coarse, ratio):
        """Fix the Chebyshev degrees (finest level, every coarser level) and
        the interval ratio of ``precision`` instead of autotuning them."""
        _lib.check(_lib.lib().dpf_topology_set_tuning(self._h, _lib.precision_code(precision),
                                                      int(fine), int(coarse), float(ratio)))

    def set_damping(self, fine, coarse):
        """Damping of the prolongator smoothing on the finest level and on
        the coarser levels (default 1.5 and 2; 4/3 on both can suit fields
        that are rough from cell to cell)."""
        _lib.check(_lib.lib().dpf_topology_set_damping(self._h, float(fine), float(coarse)))

    def set_smoother(self, precision, fine, second, deeper, ratio):
        """Fix the Chebyshev degrees (finest level, second level, every level
        below it) and the interval ratio of ``precision``."""
        _lib.check(_lib.lib().dpf_topology_set_smoother(
            self._h, _lib.precision_code(precision), int(fine), int(second), int(deeper),
            float(ratio)))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().dpf_topology_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_topo_cache = {}
_topo_lock = threading.Lock()


def get_topology(n, neighbors, areasoverlengths, dirichletnodes):
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = _f64(np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]])
    dn = np.asarray(dirichletnodes, dtype=np.int64).ravel()
    key = (int(n), nb.shape[0], dn.size, _digest(nb, aol, dn))
    with _topo_lock:
        t = _topo_cache.get(key)
        if t is None:
            if len(_topo_cache) >= 8:
                _topo_cache.pop(next(iter(_topo_cache))).close()
            t = Topology(n, nb, aol, dn)
            _topo_cache[key] = t
        return t


def clear_cache():
    with _topo_lock:
        for t in _topo_cache.values():
            t.close()
        _topo_cache.clear()


class State:
    """What a pullback needs from its forward solve (dpf_state)."""

    def __init__(self, topo, handle):
        self.topo = topo
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().dpf_state_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def groundwater_steadystate(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                            linear_solver=None, stats=None, **kwargs):
    """src/groundwater.jl:75-83.  Returns h (n,).  ``linear_solver``:
    amg_solver (default) or cholesky_solver."""
    h, _ = _steady(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs, False,
                   stats, dict(kwargs, linear_solver=linear_solver))
    return h


def _steady(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs, keep, stats,
            kwargs):
    o = _opts(kwargs)
    Qs = _f64(Qs)
    n = Qs.shape[0]
    topo = get_topology(n, neighbors, areasoverlengths, dirichletnodes)
    Ks, dh = _f64(Ks), _f64(dirichleths)
    if Ks.shape[0] != topo.ne or dh.shape[0] != n:
        raise ValueError("Ks must have one entry per edge and dirichleths one per cell")
    h = np.empty(n)
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    code = _lib.lib().dpf_steady_forward(topo.handle, _ptr(Ks), _ptr(dh), _ptr(Qs), _ptr(h),
                                         ctypes.byref(o), ctypes.byref(st),
                                         ctypes.byref(sh) if keep else None)
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return h, (State(topo, sh) if keep else None)


def rrule_groundwater_steadystate(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                                  Qs, stats=None, **kwargs):
    """src/groundwater.jl:85-109.  Returns (h, pullback)."""
    h, state = _steady(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs, True,
                       stats, kwargs)
    o = _opts(kwargs)
    n = h.shape[0]
    ne = state.topo.ne

    def pullback(delta, stats=None):
        delta = _f64(delta)
        if delta.shape[0] != n:
            raise ValueError("delta must have one entry per cell")
        dK = np.empty(ne)
        ddh = np.empty(n)
        dQ = np.empty(n)
        st = _lib.Stats()
        code = _lib.lib().dpf_steady_pullback(state._h, _ptr(delta), _ptr(dK), _ptr(ddh),
                                              _ptr(dQ), ctypes.byref(o), ctypes.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        _lib.check(code)
        return (NoTangent, dK, NoTangent, NoTangent, NoTangent, ddh, dQ)

    pullback.state = state
    return h, pullback


def groundwater_residuals(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                          specificstorage, volumes):
    """src/groundwater.jl:1-38.  Returns the residuals at the free cells."""
    Qs = _f64(Qs)
    n = Qs.shape[0]
    topo = get_topology(n, neighbors, areasoverlengths, dirichletnodes)
    h = _f64(h)
    if h.shape[0] != topo.nfree:
        raise ValueError("h must have one entry per free cell")
    r = np.empty(topo.nfree)
    o = _opts({})
    _lib.check(_lib.lib().dpf_residuals(topo.handle, _ptr(h), _ptr(_f64(Ks)),
                                        _ptr(_f64(dirichleths)), _ptr(Qs),
                                        _ptr(_f64(specificstorage)), _ptr(_f64(volumes)),
                                        _ptr(r), ctypes.byref(o)))
    return r


def _jacobian(which, hfree, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
              specificstorage, volumes):
    import scipy.sparse as sp
    topo = get_topology(len(Qs), neighbors, areasoverlengths, dirichletnodes)
    L = _lib.lib()
    nnz = ctypes.c_int64()
    _lib.check(L.dpf_jacobian_pattern(topo.handle, which, ctypes.byref(nnz), None, None))
    rp = np.empty(topo.nfree + 1, dtype=np.int64)
    cols = np.empty(max(nnz.value, 1), dtype=np.int32)
    _lib.check(L.dpf_jacobian_pattern(topo.handle, which, ctypes.byref(nnz), _ptr(rp),
                                      _ptr(cols)))
    vals = np.empty(max(nnz.value, 1))
    arrs = [None if a is None else _f64(a) for a in (hfree, Ks, dirichleths, specificstorage,
                                                    volumes)]
    o = _opts({})
    _lib.check(L.dpf_jacobian_values(topo.handle, which,
                                     *[None if a is None else _ptr(a) for a in arrs],
                                     _ptr(vals), ctypes.byref(o)))
    ncol = {_lib.DPF_JAC_H: topo.nfree, _lib.DPF_JAC_KS: topo.ne}.get(which, topo.n)
    m = nnz.value
    return sp.csr_matrix((vals[:m], cols[:m], rp), shape=(topo.nfree, ncol))


def groundwater_h(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                  specificstorage, volumes):
    """d groundwater_residuals / dh (free cells x free cells) as a scipy CSR
    matrix, the generated groundwater_h of src/groundwater.jl:1 (f_gw_u in
    examples/gw_transient_3d/inputdeck.jl:82-85).  ``h`` is not read (the
    residual is affine in h)."""
    return _jacobian(_lib.DPF_JAC_H, None, Ks, neighbors, areasoverlengths, dirichletnodes,
                     dirichleths, Qs, specificstorage, volumes)


def groundwater_Ks(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                   specificstorage, volumes):
    """d groundwater_residuals / dKs (free cells x edges), scipy CSR: the
    generated groundwater_Ks (f_gw_p's Ks block, inputdeck.jl:86-91)."""
    return _jacobian(_lib.DPF_JAC_KS, h, None, neighbors, areasoverlengths, dirichletnodes,
                     dirichleths, Qs, specificstorage, volumes)


def groundwater_dirichleths(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                            Qs, specificstorage, volumes):
    """d groundwater_residuals / d dirichleths (free cells x cells), scipy
    CSR: the generated groundwater_dirichleths (f_gw_p's dirichleths block)."""
    return _jacobian(_lib.DPF_JAC_DIRICHLETHS, None, Ks, neighbors, areasoverlengths,
                     dirichletnodes, dirichleths, Qs, specificstorage, volumes)


def groundwater_Qs(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                   specificstorage, volumes):
    """d groundwater_residuals / dQs (free cells x cells), scipy CSR: the
    generated groundwater_Qs."""
    return _jacobian(_lib.DPF_JAC_QS, None, None, neighbors, areasoverlengths, dirichletnodes,
                     dirichleths, Qs, specificstorage, volumes)


def hm(x, y):
    """DPFEHM.hm (src/richards.jl:50-52), elementwise on host arrays."""
    return 2.0 / (1.0 / np.asarray(x) + 1.0 / np.asarray(y))


def logKs2Ks_harmonic(logKs, neighbors, dirichletnodes=(), areasoverlengths=None):
    """K_e = hm(exp(logK_a), exp(logK_b)) computed on the GPU."""
    logKs = _f64(logKs)
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = np.ones(nb.shape[0]) if areasoverlengths is None else areasoverlengths
    topo = get_topology(logKs.shape[0], nb, aol, dirichletnodes)
    Ke = np.empty(topo.ne)
    o = _opts({})
    _lib.check(_lib.lib().dpf_hm_forward(topo.handle, _ptr(logKs), _ptr(Ke), ctypes.byref(o)))
    return Ke


def logKs2Ks_harmonic_pullback(logKs, Ke, dKe, neighbors, dirichletnodes=(), areasoverlengths=None):
    logKs = _f64(logKs)
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = np.ones(nb.shape[0]) if areasoverlengths is None else areasoverlengths
    topo = get_topology(logKs.shape[0], nb, aol, dirichletnodes)
    out = np.empty(logKs.shape[0])
    o = _opts({})
    _lib.check(_lib.lib().dpf_hm_pullback(topo.handle, _ptr(logKs), _ptr(_f64(Ke)),
                                          _ptr(_f64(dKe)), _ptr(out), ctypes.byref(o)))
    return out


def logKs2Ks_geometric(logKs, neighbors, dirichletnodes=(), areasoverlengths=None):
    """K_e = exp(0.5 (logK_a + logK_b)), the map of
    examples/gw_steadystate_3d/inputdeck.jl:51, computed on the GPU."""
    logKs = _f64(logKs)
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = np.ones(nb.shape[0]) if areasoverlengths is None else areasoverlengths
    topo = get_topology(logKs.shape[0], nb, aol, dirichletnodes)
    Ke = np.empty(topo.ne)
    o = _opts({})
    _lib.check(_lib.lib().dpf_geo_forward(topo.handle, _ptr(logKs), _ptr(Ke), ctypes.byref(o)))
    return Ke


def logKs2Ks_geometric_pullback(logKs, Ke, dKe, neighbors, dirichletnodes=(),
                                areasoverlengths=None):
    logKs = _f64(logKs)
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2)
    aol = np.ones(nb.shape[0]) if areasoverlengths is None else areasoverlengths
    topo = get_topology(logKs.shape[0], nb, aol, dirichletnodes)
    out = np.empty(logKs.shape[0])
    o = _opts({})
    _lib.check(_lib.lib().dpf_geo_pullback(topo.handle, _ptr(_f64(Ke)), _ptr(_f64(dKe)),
                                           _ptr(out), ctypes.byref(o)))
    return out


def _transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
               specificstorage, volumes, u0, nsteps, dt, qscale, checkpoint_every, keep,
               return_states, stats, kwargs):
    o = _opts(kwargs)
    Qs = _f64(Qs)
    n = Qs.shape[0]
    topo = get_topology(n, neighbors, areasoverlengths, dirichletnodes)
    u0 = _f64(u0)
    if u0.shape[0] != topo.nfree:
        raise ValueError("u0 must have one entry per free cell")
    qscale = _f64(np.ones(nsteps + 1) if qscale is None else qscale)
    if qscale.shape[0] != nsteps + 1:
        raise ValueError("qscale must have nsteps + 1 entries")
    dts = None
    if np.ndim(dt) > 0:
        dts = _f64(dt)
        if dts.shape != (int(nsteps),):
            raise ValueError("dt must be a number or hold one length per step")
    uN = np.empty(topo.nfree)
    states = np.empty((nsteps + 1, topo.nfree)) if return_states else None
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    head = (topo.handle, _ptr(_f64(Ks)), _ptr(_f64(dirichleths)), _ptr(Qs),
            _ptr(_f64(specificstorage)), _ptr(_f64(volumes)), _ptr(u0), int(nsteps))
    tail = (_ptr(qscale), int(checkpoint_every), _ptr(uN),
            _ptr(states) if return_states else None, ctypes.byref(o), ctypes.byref(st),
            ctypes.byref(sh) if keep else None)
    if dts is None:
        code = _lib.lib().dpf_transient_forward(*head, float(dt), *tail)
    else:
        code = _lib.lib().dpf_transient_forward_dts(*head, _ptr(dts), *tail)
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return uN, states, (State(topo, sh) if keep else None), o


def groundwater_transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                          specificstorage, volumes, u0, nsteps, dt, qscale=None,
                          return_states=False, stats=None, **kwargs):
    """Backward Euler of du/dt = -groundwater_residuals(u, ...), the system
    DifferentiableBackwardEuler.steps integrates in
    examples/gw_transient_3d/inputdeck.jl:60-92.  ``u0`` and the result are on
    the free cells; ``dt`` is one step length or a sequence of ``nsteps``
    lengths (e.g. the steps an adaptive integrator took); ``qscale[n]``
    scales Qs at t_n.  Returns u_N, or the (nsteps + 1, nfree) matrix of
    states with ``return_states`` (row k = u_k, the reference's columns)."""
    uN, states, _, _ = _transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                                  Qs, specificstorage, volumes, u0, nsteps, dt, qscale, 0, False,
                                  return_states, stats, kwargs)
    return states if return_states else uN


def rrule_groundwater_transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                                specificstorage, volumes, u0, nsteps, dt, qscale=None,
                                checkpoint_every=20, stats=None, **kwargs):
    """Returns (u_N, pullback).  ``pullback(dJdu)`` takes dJ/du_N (nfree,) or
    the seeds of every state, an (nsteps + 1, nfree) array with row k =
    dJ/du_k (rows of zeros are skipped), and returns
    ``(dKs, ddirichleths, dQs, du0)``: the gradient in p = [Ks; dirichleths]
    that examples/gw_transient_3d/inputdeck.jl:63,103-104 takes through
    DifferentiableBackwardEuler.steps, plus the tangents in Qs and u0.  The
    states between checkpoints are recomputed during the reverse sweep."""
    uN, _, state, o = _transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                                 Qs, specificstorage, volumes, u0, nsteps, dt, qscale,
                                 checkpoint_every, True, False, stats, kwargs)
    ne, n, nf = state.topo.ne, state.topo.n, state.topo.nfree

    def pullback(dJdu, stats=None):
        g = _f64(dJdu)
        if g.shape == (nf,):
            steps = np.array([nsteps], dtype=np.int32)
            seeds = g
        elif g.shape == (nsteps + 1, nf):
            steps = np.nonzero(np.any(g != 0.0, axis=1))[0].astype(np.int32)
            seeds = np.ascontiguousarray(g[steps])
        else:
            raise ValueError("dJdu must be (nfree,) or (nsteps + 1, nfree)")
        steps = np.ascontiguousarray(steps)
        dK = np.empty(ne)
        ddh = np.empty(n)
        dQ = np.empty(n)
        du0 = np.empty(nf)
        st = _lib.Stats()
        code = _lib.lib().dpf_transient_pullback_seeds(
            state._h, int(steps.size), _ptr(steps), _ptr(seeds), _ptr(dK), _ptr(ddh), _ptr(dQ),
            _ptr(du0), ctypes.byref(o), ctypes.byref(st))
        if stats is not None:
            stats.update(st.as_dict())
        _lib.check(code)
        return dK, ddh, dQ, du0

    pullback.state = state
    return uN, pullback


class BatchTopology(Topology):
    """Several meshes packed into one block-diagonal topology
    (dpf_batch_topology_create).  ``members`` is a list of
    ``(n, neighbors, areasoverlengths, dirichletnodes)`` with member-local
    1-based indices.  The packed arrays (cells, edges, free cells) are the
    members' arrays concatenated; ``cell0``/``edge0``/``free0`` give the
    offsets."""

    def __init__(self, members):  # noqa: D107 - Topology.__init__ is not reused
        ns = np.array([int(m[0]) for m in members], dtype=np.int64)
        nbs = [np.asarray(m[1], dtype=np.int64).reshape(-1, 2) for m in members]
        es = np.array([nb.shape[0] for nb in nbs], dtype=np.int64)
        aols = [np.asarray(m[2], dtype=np.float64)[:e] for m, e in zip(members, es)]
        dns = [np.asarray(m[3], dtype=np.int64).ravel() for m in members]
        nds = np.array([d.size for d in dns], dtype=np.int64)
        nb = np.ascontiguousarray(np.concatenate(nbs))
        aol = _f64(np.concatenate(aols))
        dn = np.ascontiguousarray(np.concatenate(dns)) if nds.sum() else np.zeros(0, np.int64)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().dpf_batch_topology_create(
            len(members), _ptr(ns), _ptr(es), _ptr(nb), _ptr(aol), _ptr(nds), _ptr(dn),
            ctypes.byref(h)))
        self._h = h
        self.nmembers = len(members)
        self.cell0 = np.concatenate([[0], np.cumsum(ns)])
        self.edge0 = np.concatenate([[0], np.cumsum(es)])
        self.n, self.ne = int(self.cell0[-1]), int(self.edge0[-1])
        isfree = np.ones(self.n, dtype=bool)
        isfree[(np.concatenate([d - 1 + c for d, c in zip(dns, self.cell0[:-1])])
                if nds.sum() else np.zeros(0, np.int64))] = False
        self.isfree = isfree
        self.free0 = np.concatenate([[0], np.cumsum([isfree[a:b].sum() for a, b in
                                                     zip(self.cell0[:-1], self.cell0[1:])])])
        info = self.info()
        self.nfree = int(info[2])
        self.nnz = int(info[3])
        self.levels = int(info[4])
        self.coarse_rows = int(info[5])
        self.operator_complexity = info[6] / 1000.0
        self.symbolic_setup_ms = info[7] / 1000.0
        self.lock = threading.Lock()

    def member_relres(self):
        out = np.empty(self.nmembers)
        _lib.check(_lib.lib().dpf_member_stats(self._h, _ptr(out), self.nmembers))
        return out


def batch_steadystate(topo, Ks, dirichleths, Qs, stats=None, **kwargs):
    """Steady solves of every member of a BatchTopology in one set of launches.
    Ks, dirichleths, Qs: the members' arrays concatenated.  Returns h (packed)."""
    o = _opts(kwargs)
    Ks, dh, Qs = _f64(Ks), _f64(dirichleths), _f64(Qs)
    if Ks.shape[0] != topo.ne or dh.shape[0] != topo.n or Qs.shape[0] != topo.n:
        raise ValueError("packed arrays do not match the batch topology")
    h = np.empty(topo.n)
    st = _lib.Stats()
    code = _lib.lib().dpf_steady_forward(topo.handle, _ptr(Ks), _ptr(dh), _ptr(Qs), _ptr(h),
                                         ctypes.byref(o), ctypes.byref(st), None)
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    return h


def rrule_batch_steadystate(topo, Ks, dirichleths, Qs, stats=None, **kwargs):
    """Batched forward + pullback: pullback(delta_packed) -> (dKs, ddirichleths, dQs)."""
    o = _opts(kwargs)
    Ks, dh, Qs = _f64(Ks), _f64(dirichleths), _f64(Qs)
    h = np.empty(topo.n)
    st = _lib.Stats()
    sh = ctypes.c_void_p()
    code = _lib.lib().dpf_steady_forward(topo.handle, _ptr(Ks), _ptr(dh), _ptr(Qs), _ptr(h),
                                         ctypes.byref(o), ctypes.byref(st), ctypes.byref(sh))
    if stats is not None:
        stats.update(st.as_dict())
    _lib.check(code)
    state = State(topo, sh)

    def pullback(delta, stats=None):
        delta = _f64(delta)
        dK = np.empty(topo.ne)
        ddh = np.empty(topo.n)
        dQ = np.empty(topo.n)
        st2 = _lib.Stats()
        code = _lib.lib().dpf_steady_pullback(state._h, _ptr(delta), _ptr(dK), _ptr(ddh), _ptr(dQ),
                                              ctypes.byref(o), ctypes.byref(st2))
        if stats is not None:
            stats.update(st2.as_dict())
        _lib.check(code)
        return dK, ddh, dQ

    pullback.state = state
    return h, pullback


def hm_forward(topo, logK):
    """K_e = hm(exp(logK_a), exp(logK_b)) on a Topology / BatchTopology."""
    logK = _f64(logK)
    Ke = np.empty(topo.ne)
    o = _opts({})
    _lib.check(_lib.lib().dpf_hm_forward(topo.handle, _ptr(logK), _ptr(Ke), ctypes.byref(o)))
    return Ke


def hm_pullback(topo, logK, Ke, dKe):
    out = np.empty(topo.n)
    o = _opts({})
    _lib.check(_lib.lib().dpf_hm_pullback(topo.handle, _ptr(_f64(logK)), _ptr(_f64(Ke)),
                                          _ptr(_f64(dKe)), _ptr(out), ctypes.byref(o)))
    return out
