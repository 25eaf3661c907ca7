"""FP64 CPU oracle for DPFEHM's groundwater path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product
(``dpfehm.jl_amd``) never loads it.

The arithmetic lives in ``gw_oracle.c`` (plain C, built by ``oracle/Makefile``
into ``oracle/_build/libgw_oracle.so``); this file is a ctypes wrapper that
speaks the reference's conventions (1-based ``neighbors`` pairs and
``dirichletnodes``).

The reference (DPFEHM.jl) is pure Julia and its hot-path arithmetic lives in
packages that are not on disk (NonlinearEquations.jl, AlgebraicMultigrid.jl
0.5.0, DifferentiableBackwardEuler.jl; Manifest.toml:19-23,229-240,629-635).
There is no Julia in this image, so there is no ``oracle/_ref`` build and no
reference golden vector; the oracle is pinned by the checks listed below.

Function (restates) -- pinned by (tests/test_oracle.py unless noted)

* ``getfreenodes``            src/utilities.jl:10-24 and src/twophase.jl:15-20
  -- brute-force restatement of both reference forms.
* ``groundwater_residuals``   src/groundwater.jl:1-21
  -- linearity r(h) = A h - b at random h; hand-computed 1-D/2-D cases;
  invariance to edge order; a duplicated edge equals one edge of summed K;
  the numpy closed form in tests/golden/make_golden.py.
* ``groundwater_h``           generated Jacobian of src/groundwater.jl:1
  -- central finite differences of ``groundwater_residuals``; symmetry;
  zero row sums away from Dirichlet cells.
* ``groundwater_steadystate`` src/groundwater.jl:75-83
  -- exact linear-head solution (examples/fracture_network_solver_scaling/
  porous_media_comparison.jl:24-46); Thiem drawdown within 10 % (test/thiem.jl:
  10-12,67); scipy sparse direct solve; dense solve on tiny grids.
* ``steadystate_rrule``       src/groundwater.jl:85-109
  -- central finite differences of the forward solve for every tangent
  (Ks, dirichleths, Qs); dense inverse on tiny grids; reciprocity
  dh_k/dQ_j = dh_j/dQ_k; the one-sided FD gate of test/thiem.jl:14-25,68-71.
* ``hm_forward``/``hm_pullback`` src/richards.jl:50-52 (hm) with the
  logK -> K_e map of examples/gw_steadystate_3d/inputdeck.jl:51
  -- central finite differences; closed form at equal arguments.
* ``geo_forward``/``geo_pullback`` the geometric-mean map K_e =
  exp(0.5 (logK_a + logK_b)) of examples/gw_steadystate_3d/inputdeck.jl:51
  -- central finite differences; closed form sqrt(k_a k_b).
* ``transient``/``transient_adjoint`` examples/gw_transient_3d/inputdeck.jl:
  60-92 (f = -groundwater_residuals, parameters p = [Ks; dirichleths] at
  :63,86-91) with backward Euler over a given step sequence in place of
  DifferentiableBackwardEuler.steps
  -- a dense restatement in the reference's (I - dt f_u) form, uniform and
  non-uniform steps; the Theis drawdown within 0.1 m (test/theis.jl:12-21,
  102); the steady limit; central finite differences of objectives of one
  and of several states for every tangent (Ks, dirichleths, Qs, u0); the
  tangent in dirichleths against the dense product of the reference's
  f_gw_p block -groundwater_dirichleths.
  The reference integrates with adaptive steps inside DiffEq; matching its
  exact step sequence would need that package: parity unpinned for the
  adaptive sequence, pinned for fixed steps by the checks above.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libgw_oracle.so")
_lib_handle = None

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib():
    global _lib_handle
    if _lib_handle is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.orc_freenodes.restype = ctypes.c_int64
        L.orc_jacobian_h.restype = ctypes.c_int64
        for f in ("orc_pcg", "orc_steady", "orc_steady_pullback", "orc_transient",
                  "orc_transient_adjoint"):
            getattr(L, f).restype = ctypes.c_int
        _lib_handle = L
    return _lib_handle


def _p(a, kind):
    return a.ctypes.data_as({"i": _i64p, "f": _f64p, "u": _u8p}[kind])


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _edges0(neighbors):
    return _i64(np.asarray(neighbors, dtype=np.int64).reshape(-1, 2) - 1)


class OracleError(RuntimeError):
    pass


def getfreenodes(n, dirichletnodes):
    """Returns (isfreenode, nodei2freenodei, freenodei2nodei), 1-based, -1 on
    Dirichlet cells."""
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    isfree = np.empty(n, dtype=np.uint8)
    n2f = np.empty(n, dtype=np.int64)
    f2n = np.empty(n, dtype=np.int64)
    nf = _lib().orc_freenodes(ctypes.c_int64(n), ctypes.c_int64(d.size), _p(d, "i"),
                              _p(isfree, "u"), _p(n2f, "i"), _p(f2n, "i"))
    n2f1 = np.where(n2f >= 0, n2f + 1, -1)
    return isfree.astype(bool), n2f1, f2n[:nf] + 1


def _maps(n, dirichletnodes):
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    isfree = np.empty(n, dtype=np.uint8)
    n2f = np.empty(n, dtype=np.int64)
    f2n = np.empty(n, dtype=np.int64)
    nf = _lib().orc_freenodes(ctypes.c_int64(n), ctypes.c_int64(d.size), _p(d, "i"),
                              _p(isfree, "u"), _p(n2f, "i"), _p(f2n, "i"))
    return d, isfree, n2f, f2n[:nf], int(nf)


def groundwater_residuals(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                          Qs, specificstorage, volumes):
    n = len(Qs)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    _, isfree, n2f, _, nf = _maps(n, dirichletnodes)
    h, K, aol = _f64(h), _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    dh, Q, Ss, V = _f64(dirichleths), _f64(Qs), _f64(specificstorage), _f64(volumes)
    r = np.empty(nf)
    _lib().orc_residuals(ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"),
                         _p(aol, "f"), _p(isfree, "u"), _p(n2f, "i"), _p(h, "f"),
                         _p(dh, "f"), _p(Q, "f"), _p(Ss, "f"), _p(V, "f"),
                         ctypes.c_int64(nf), _p(r, "f"))
    return r


def groundwater_h(h, Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                  specificstorage, volumes):
    """dr/dh as a scipy CSR matrix (sorted columns, duplicates summed)."""
    import scipy.sparse as sp
    n = len(Qs)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    _, isfree, n2f, _, nf = _maps(n, dirichletnodes)
    K, aol = _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    Ss, V = _f64(specificstorage), _f64(volumes)
    rp = np.empty(nf + 1, dtype=np.int64)
    args = (ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"), _p(aol, "f"),
            _p(isfree, "u"), _p(n2f, "i"), _p(Ss, "f"), _p(V, "f"), ctypes.c_int64(nf),
            _p(rp, "i"))
    nnz = _lib().orc_jacobian_h(*args, None, None)
    ci = np.empty(max(nnz, 1), dtype=np.int64)
    v = np.empty(max(nnz, 1))
    _lib().orc_jacobian_h(*args, _p(ci, "i"), _p(v, "f"))
    return sp.csr_matrix((v[:nnz], ci[:nnz], rp), shape=(nf, nf))


def pcg(A, b, x0=None, reltol=1e-13, maxiter=20000):
    """IC(0)-preconditioned CG on a scipy CSR matrix.  Returns (x, iters, relres)."""
    A = A.tocsr()
    A.sort_indices()
    n = A.shape[0]
    rp, ci, v = _i64(A.indptr), _i64(A.indices), _f64(A.data)
    b = _f64(b)
    x = np.zeros(n) if x0 is None else _f64(x0).copy()
    info = np.zeros(2)
    st = _lib().orc_pcg(ctypes.c_int64(n), _p(rp, "i"), _p(ci, "i"), _p(v, "f"), _p(b, "f"),
                        _p(x, "f"), ctypes.c_double(reltol), ctypes.c_int64(maxiter),
                        _p(info, "f"))
    if st:
        raise OracleError(f"oracle PCG status {st} after {int(info[0])} iterations, "
                          f"relres {info[1]:.3e}")
    return x, int(info[0]), float(info[1])


def groundwater_steadystate(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                            reltol=1e-13, maxiter=20000, return_info=False):
    n = len(Qs)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    K, aol = _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    dh, Q = _f64(dirichleths), _f64(Qs)
    h = np.empty(n)
    hfree = np.empty(n - np.unique(d).size + 1)
    info = np.zeros(2)
    st = _lib().orc_steady(ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"),
                           _p(aol, "f"), ctypes.c_int64(d.size), _p(d, "i"), _p(dh, "f"),
                           _p(Q, "f"), ctypes.c_double(reltol), ctypes.c_int64(maxiter),
                           _p(h, "f"), _p(hfree, "f"), _p(info, "f"))
    if st:
        raise OracleError(f"oracle steady solve status {st}: relres {info[1]:.3e}")
    if return_info:
        return h, hfree[:-1], dict(iterations=int(info[0]), relres=float(info[1]))
    return h


def steadystate_rrule(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
                      reltol=1e-13, maxiter=20000):
    """Returns (h, pullback); pullback(delta) gives the 7-tuple of
    src/groundwater.jl:100-106 with None for NoTangent."""
    h, hfree, _ = groundwater_steadystate(Ks, neighbors, areasoverlengths, dirichletnodes,
                                          dirichleths, Qs, reltol, maxiter, return_info=True)
    n = len(Qs)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    K, aol = _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    dh = _f64(dirichleths)
    hf = _f64(hfree)

    def pullback(delta):
        delta = _f64(delta)
        dK = np.empty(ne)
        ddh = np.empty(n)
        dQ = np.empty(n)
        lam = np.empty(hf.size + 1)
        info = np.zeros(2)
        st = _lib().orc_steady_pullback(
            ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"), _p(aol, "f"),
            ctypes.c_int64(d.size), _p(d, "i"), _p(dh, "f"), _p(hf, "f"), _p(delta, "f"),
            ctypes.c_double(reltol), ctypes.c_int64(maxiter), _p(dK, "f"), _p(ddh, "f"),
            _p(dQ, "f"), _p(lam, "f"), _p(info, "f"))
        if st:
            raise OracleError(f"oracle adjoint solve status {st}: relres {info[1]:.3e}")
        pullback.lam = lam[:-1]
        return (None, dK, None, None, None, ddh, dQ)

    return h, pullback


def hm_forward(logK, neighbors):
    nb = _edges0(neighbors)
    lk = _f64(logK)
    Ke = np.empty(nb.shape[0])
    _lib().orc_hm_forward(ctypes.c_int64(nb.shape[0]), _p(nb, "i"), _p(lk, "f"), _p(Ke, "f"))
    return Ke


def hm_pullback(logK, neighbors, dKe):
    nb = _edges0(neighbors)
    lk, dk = _f64(logK), _f64(dKe)
    out = np.empty(lk.size)
    _lib().orc_hm_pullback(ctypes.c_int64(lk.size), ctypes.c_int64(nb.shape[0]), _p(nb, "i"),
                           _p(lk, "f"), _p(dk, "f"), _p(out, "f"))
    return out


def geo_forward(logK, neighbors):
    nb = _edges0(neighbors)
    lk = _f64(logK)
    Ke = np.empty(nb.shape[0])
    _lib().orc_geo_forward(ctypes.c_int64(nb.shape[0]), _p(nb, "i"), _p(lk, "f"), _p(Ke, "f"))
    return Ke


def geo_pullback(logK, neighbors, dKe):
    nb = _edges0(neighbors)
    lk, dk = _f64(logK), _f64(dKe)
    out = np.empty(lk.size)
    _lib().orc_geo_pullback(ctypes.c_int64(lk.size), ctypes.c_int64(nb.shape[0]), _p(nb, "i"),
                            _p(lk, "f"), _p(dk, "f"), _p(out, "f"))
    return out

def _dts(dt, nsteps):
    """Per-step lengths from a scalar dt or a sequence of nsteps lengths."""
    d = np.asarray(dt, dtype=np.float64)
    d = np.full(int(nsteps), float(d)) if d.ndim == 0 else np.ascontiguousarray(d)
    if d.shape != (int(nsteps),):
        raise ValueError("dt must be a scalar or have nsteps entries")
    return d


def transient(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths, Qs,
              specificstorage, volumes, u0, nsteps, dt, qscale, reltol=1e-13, maxiter=20000):
    """Backward Euler with step lengths dt (a scalar, or one per step);
    returns states (nsteps+1, nf) on free cells."""
    n = len(Qs)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    nf = n - np.unique(d).size
    K, aol = _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    dh, Q, Ss, V = _f64(dirichleths), _f64(Qs), _f64(specificstorage), _f64(volumes)
    u0 = _f64(u0)
    qs = _f64(qscale)
    dts = _dts(dt, nsteps)
    states = np.empty((nsteps + 1, nf))
    info = np.zeros(2)
    st = _lib().orc_transient(ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"),
                              _p(aol, "f"), ctypes.c_int64(d.size), _p(d, "i"), _p(dh, "f"),
                              _p(Q, "f"), _p(Ss, "f"), _p(V, "f"), _p(u0, "f"),
                              ctypes.c_int64(nsteps), _p(dts, "f"), _p(qs, "f"),
                              ctypes.c_double(reltol), ctypes.c_int64(maxiter),
                              _p(states, "f"), _p(info, "f"))
    if st:
        raise OracleError(f"oracle transient status {st}")
    return states


def transient_adjoint(Ks, neighbors, areasoverlengths, dirichletnodes, dirichleths,
                      specificstorage, volumes, nsteps, dt, qscale, states, dJdu,
                      reltol=1e-13, maxiter=20000, full=False):
    """Reverse mode of ``transient`` for an objective of the states; dJdu is
    (nsteps+1, nf), row n = dJ/du_n.  Returns (dJ/dKs (E,), dJ/dQs (n,)), or
    with ``full`` (dJ/dKs, dJ/ddirichleths (n,), dJ/dQs, dJ/du0 (nf,))."""
    n = len(volumes)
    nb = _edges0(neighbors)
    ne = nb.shape[0]
    d = _i64(np.asarray(dirichletnodes, dtype=np.int64) - 1)
    K, aol = _f64(Ks), _f64(np.asarray(areasoverlengths)[:ne])
    dh, Ss, V = _f64(dirichleths), _f64(specificstorage), _f64(volumes)
    qs, S, G = _f64(qscale), _f64(states), _f64(dJdu)
    dts = _dts(dt, nsteps)
    dK = np.empty(ne)
    dQ = np.empty(n)
    ddh = np.empty(n)
    du0 = np.empty(S.shape[1])
    info = np.zeros(2)
    st = _lib().orc_transient_adjoint(
        ctypes.c_int64(n), ctypes.c_int64(ne), _p(nb, "i"), _p(K, "f"), _p(aol, "f"),
        ctypes.c_int64(d.size), _p(d, "i"), _p(dh, "f"), _p(Ss, "f"), _p(V, "f"),
        ctypes.c_int64(nsteps), _p(dts, "f"), _p(qs, "f"), _p(S, "f"), _p(G, "f"),
        ctypes.c_double(reltol), ctypes.c_int64(maxiter), _p(dK, "f"), _p(dQ, "f"),
        _p(ddh, "f"), _p(du0, "f"), _p(info, "f"))
    if st:
        raise OracleError(f"oracle transient adjoint status {st}")
    return (dK, ddh, dQ, du0) if full else (dK, dQ)
