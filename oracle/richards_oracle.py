"""FP64 CPU oracle for DPFEHM's steady Richards path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module, and only as the checker.  The product
(``dpfehm.jl_amd``) never loads it.  numpy / scipy.

Function (restates) -- pinned by (tests/test_oracle_richards.py)

* ``kr`` / ``kr_partials``     src/richards.jl:18-27 (van Genuchten) and the
  ForwardDiff gradient the generated Jacobian takes of it (:36-43)
  -- kr = 1 for psi >= 0; central differences of the complex-step partials;
  the closed form at N = 2.
* ``richards_residuals``       src/richards.jl:97-117 (macro equations;
  coords[3] is the elevation, Ss V per cell)
  -- with every head saturated (psi >= 0, kr = 1) it is the groundwater
  residual of h = psi + z (src/groundwater.jl:1-21, gw_oracle) times
  1 / (Ss V) with the sign flipped; edge-order invariance.
* ``richards_jacobian``        the generated richards_psi
  -- central differences of ``richards_residuals``.
* ``richards_steadystate``     src/richards.jl:156-179 (NLsolve on the
  residual; here Newton with backtracking and a sparse direct solve, to a
  residual far below NLsolve's default ftol)
  -- the residual at the solution; the saturated case against the
  groundwater oracle's heads.
* ``richards_rrule``           src/richards.jl:181-205
  -- central differences of the forward solve in Ks, dirichletpsis, alphas,
  Ns and Qs.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def kr(psi, alpha, N):
    """src/richards.jl:18-27, elementwise."""
    psi = np.asarray(psi, dtype=np.float64)
    alpha = np.broadcast_to(np.asarray(alpha, dtype=np.float64), psi.shape)
    N = np.broadcast_to(np.asarray(N, dtype=np.float64), psi.shape)
    out = np.ones(psi.shape)
    neg = psi < 0
    if np.any(neg):
        p, a, n = psi[neg], alpha[neg], N[neg]
        m = (n - 1) / n
        s = np.abs(a * p)
        denom = 1 + s ** n
        numer = 1 - s ** (n - 1) * denom ** (-m)
        out[neg] = numer ** 2 / denom ** (m / 2)
    return out


def _kr_complex(psi, alpha, N):
    """kr for psi < 0 in complex arithmetic (|alpha psi| written as
    -alpha psi with alpha > 0, psi < 0, so that it is analytic)."""
    m = (N - 1) / N
    s = -alpha * psi
    denom = 1 + s ** N
    numer = 1 - s ** (N - 1) * denom ** (-m)
    return numer ** 2 / denom ** (m / 2)


def kr_partials(psi, alpha, N, h=1e-30):
    """(kr, dkr/dpsi, dkr/dalpha, dkr/dN) by complex steps (exact to
    rounding); all derivatives are 0 where psi >= 0."""
    psi = np.asarray(psi, dtype=np.float64)
    alpha = np.broadcast_to(np.asarray(alpha, dtype=np.float64), psi.shape).astype(np.float64)
    N = np.broadcast_to(np.asarray(N, dtype=np.float64), psi.shape).astype(np.float64)
    k = kr(psi, alpha, N)
    d = [np.zeros(psi.shape) for _ in range(3)]
    neg = psi < 0
    if np.any(neg):
        p, a, n = psi[neg], alpha[neg], N[neg]
        d[0][neg] = _kr_complex(p + 1j * h, a, n).imag / h
        d[1][neg] = _kr_complex(p, a + 1j * h, n).imag / h
        d[2][neg] = _kr_complex(p, a, n + 1j * h).imag / h
    return k, d[0], d[1], d[2]


def hm(x, y):
    return 2 / (1 / x + 1 / y)


def _setup(n, neighbors, dirichletnodes):
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2) - 1
    dn = np.asarray(dirichletnodes, dtype=np.int64).ravel() - 1
    isfree = np.ones(n, dtype=bool)
    isfree[dn] = False
    n2f = -np.ones(n, dtype=np.int64)
    n2f[isfree] = np.arange(isfree.sum())
    return nb, isfree, n2f


def _edge_terms(psi_full, Ks, nb, aol, z, alphas, Ns):
    """Per edge: kr at both ends with the edge's (alpha, N), hm, its partials,
    and the head difference (psi + z)_b - (psi + z)_a."""
    a, b = nb[:, 0], nb[:, 1]
    ka, dka, dkaa, dkaN = kr_partials(psi_full[a], alphas, Ns)
    kb, dkb, dkba, dkbN = kr_partials(psi_full[b], alphas, Ns)
    H = hm(ka, kb)
    Ha = 2 * kb * kb / (ka + kb) ** 2  # dhm/dka
    Hb = 2 * ka * ka / (ka + kb) ** 2
    dlt = psi_full[b] + z[b] - psi_full[a] - z[a]
    return dict(ka=ka, kb=kb, dka=dka, dkb=dkb, dkaa=dkaa, dkba=dkba, dkaN=dkaN, dkbN=dkbN,
                H=H, Ha=Ha, Hb=Hb, dlt=dlt)


def _full(psi_free, isfree, dirichletpsis):
    psi = np.asarray(dirichletpsis, dtype=np.float64).copy()
    psi[isfree] = psi_free
    return psi


def richards_residuals(psi_free, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                       coords, alphas, Ns, Qs, specificstorage, volumes):
    """src/richards.jl:97-117: residuals at the free cells, the Q term then
    per edge in list order."""
    Q = np.asarray(Qs, dtype=np.float64)
    n = Q.size
    nb, isfree, n2f = _setup(n, neighbors, dirichletnodes)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    z = np.asarray(coords, dtype=np.float64)[2]
    sv = np.asarray(specificstorage, dtype=np.float64) * np.asarray(volumes, dtype=np.float64)
    psi = _full(psi_free, isfree, dirichletpsis)
    T = _edge_terms(psi, np.asarray(Ks, dtype=np.float64), nb, aol, z, alphas, Ns)
    flux = T["H"] * np.asarray(Ks, dtype=np.float64) * T["dlt"] * aol  # from a towards b's head
    r_full = Q / sv
    a, b = nb[:, 0], nb[:, 1]
    both = isfree[a] & isfree[b]
    ok_a = isfree[a] & (a != b)
    ok_b = isfree[b] & (a != b)
    idx = np.stack([a, b], axis=1).ravel()
    val = np.stack([np.where(ok_a, flux / sv[a], 0.0), np.where(ok_b, -flux / sv[b], 0.0)],
                   axis=1).ravel()
    del both
    np.add.at(r_full, idx, val)
    return r_full[isfree]


def richards_jacobian(psi_free, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                      coords, alphas, Ns, Qs, specificstorage, volumes):
    """dr/dpsi at the free cells (scipy CSR, nf x nf)."""
    Q = np.asarray(Qs, dtype=np.float64)
    n = Q.size
    nb, isfree, n2f = _setup(n, neighbors, dirichletnodes)
    K = np.asarray(Ks, dtype=np.float64)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    z = np.asarray(coords, dtype=np.float64)[2]
    sv = np.asarray(specificstorage, dtype=np.float64) * np.asarray(volumes, dtype=np.float64)
    psi = _full(psi_free, isfree, dirichletpsis)
    T = _edge_terms(psi, K, nb, aol, z, alphas, Ns)
    a, b = nb[:, 0], nb[:, 1]
    # flux F = H K dlt aol; dF/dpsi_a, dF/dpsi_b
    dFa = (T["Ha"] * T["dka"] * T["dlt"] - T["H"]) * K * aol
    dFb = (T["Hb"] * T["dkb"] * T["dlt"] + T["H"]) * K * aol
    rows, cols, vals = [], [], []
    keep = a != b
    for (r, s_, sign) in ((a, b, 1.0), (b, a, -1.0)):
        # equation at r gets sign * F / sv[r]
        for (c, dF) in ((a, dFa), (b, dFb)):
            m = keep & isfree[r] & isfree[c]
            rows.append(n2f[r[m]])
            cols.append(n2f[c[m]])
            vals.append(sign * dF[m] / sv[r[m]])
    nf = int(isfree.sum())
    J = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nf, nf))
    J.sum_duplicates()
    return J


def richards_steadystate(psi0, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                         coords, alphas, Ns, Qs, tol=1e-13, maxiter=100, return_info=False):
    """src/richards.jl:156-179 (Ss = V = 1).  Newton with backtracking on
    ||r||_2 until ||r||_inf <= tol * max(1, ||r(psi0)||_inf)."""
    Q = np.asarray(Qs, dtype=np.float64)
    n = Q.size
    nb, isfree, n2f = _setup(n, neighbors, dirichletnodes)
    one = np.ones(n)
    args = (Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis, coords, alphas, Ns,
            Qs, one, one)
    x = np.asarray(psi0, dtype=np.float64)[isfree].copy()
    r = richards_residuals(x, *args)
    r0 = max(1.0, np.abs(r).max())
    it = 0
    while np.abs(r).max() > tol * r0:
        if it >= maxiter:
            raise RuntimeError("oracle Richards Newton did not converge")
        J = richards_jacobian(x, *args)
        dx = spla.spsolve(J.tocsc(), -r)
        lam, nr0 = 1.0, np.linalg.norm(r)
        while True:
            xn = x + lam * dx
            rn = richards_residuals(xn, *args)
            if np.linalg.norm(rn) < (1 - 1e-4 * lam) * nr0 or lam < 1e-6:
                break
            lam *= 0.5
        x, r = xn, rn
        it += 1
    psi = _full(x, isfree, dirichletpsis)
    if return_info:
        return psi, dict(iterations=it, resid=float(np.abs(r).max()))
    return psi


def richards_rrule(psi0, Ks, neighbors, areasoverlengths, dirichletnodes, dirichletpsis,
                   coords, alphas, Ns, Qs, tol=1e-13):
    """src/richards.jl:181-205.  Returns (psi, pullback); pullback(delta) ->
    dict(Ks, dirichletpsis, alphas, Ns, Qs, psi0=zeros)."""
    psi = richards_steadystate(psi0, Ks, neighbors, areasoverlengths, dirichletnodes,
                               dirichletpsis, coords, alphas, Ns, Qs, tol=tol)
    Q = np.asarray(Qs, dtype=np.float64)
    n = Q.size
    nb, isfree, n2f = _setup(n, neighbors, dirichletnodes)
    K = np.asarray(Ks, dtype=np.float64)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    z = np.asarray(coords, dtype=np.float64)[2]
    one = np.ones(n)
    J = richards_jacobian(psi[isfree], K, neighbors, aol, dirichletnodes, dirichletpsis, coords,
                          alphas, Ns, Qs, one, one)
    T = _edge_terms(psi, K, nb, aol, z, alphas, Ns)
    a, b = nb[:, 0], nb[:, 1]

    def pullback(delta):
        delta = np.asarray(delta, dtype=np.float64)
        lam = spla.spsolve(J.T.tocsc(), delta[isfree])
        lf = np.zeros(n)
        lf[isfree] = lam
        # equation at a: +F, at b: -F (only where the equation's cell is free)
        wa = np.where(isfree[a] & (a != b), lf[a], 0.0)
        wb = np.where(isfree[b] & (a != b), lf[b], 0.0)
        g = wa - wb  # dJ/dF_e through the residuals, before the minus of the rrule
        F_over_K = T["H"] * T["dlt"] * aol
        dK = -g * F_over_K
        dHa_dalpha = T["Ha"] * T["dkaa"] + T["Hb"] * T["dkba"]
        dHa_dN = T["Ha"] * T["dkaN"] + T["Hb"] * T["dkbN"]
        dal = -g * dHa_dalpha * K * T["dlt"] * aol
        dNs = -g * dHa_dN * K * T["dlt"] * aol
        # Dirichlet psis: F depends on psi at a Dirichlet end
        dFa = (T["Ha"] * T["dka"] * T["dlt"] - T["H"]) * K * aol
        dFb = (T["Hb"] * T["dkb"] * T["dlt"] + T["H"]) * K * aol
        ddp = np.zeros(n)
        np.add.at(ddp, a, np.where(~isfree[a], -g * dFa, 0.0))
        np.add.at(ddp, b, np.where(~isfree[b], -g * dFb, 0.0))
        ddp[~isfree] += delta[~isfree]
        dQ = np.zeros(n)
        dQ[isfree] = -lam
        return dict(psi0=np.zeros(n), Ks=dK, dirichletpsis=ddp, alphas=dal, Ns=dNs, Qs=dQ)

    return psi, pullback
