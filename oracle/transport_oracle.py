"""FP64 CPU oracle for DPFEHM's advection-dispersion path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module, and only as the checker.  The product
(``dpfehm.jl_amd``) never loads it.  numpy / scipy.

Function (restates) -- pinned by (tests/test_oracle_transport.py)

* ``transport_residuals``  src/transport.jl:2-33 (upwind, or downwind with
  use_upwind=False) -- affine in u (r(u) = J u + r(0)); a hand-computed
  two-cell case; edge-order invariance.
* ``transport_jacobian``   the generated transport_u -- equals r(e_j) - r(0).
* ``transport_steps``      fixed-step backward Euler of du/dt = r(u) (the
  reference integrates with DifferentiableBackwardEuler.steps, adaptive steps
  to 1e-4, test/transport.jl:30-33: parity of the step sequence unpinned) --
  the reference test's physics gates (test/transport.jl:37-41: mass conserved,
  mean displaced by c t, variance within 10 %).
* ``transport_adjoint``    gradient of g . u_N in vxs, vys, vzs, Ds (per edge)
  and Qs -- central differences of ``transport_steps`` (and the reference
  test's one-sided check, test/transport.jl:58-69).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _geom(neighbors, coords, vxs, vys, vzs, use_upwind):
    nb = np.asarray(neighbors, dtype=np.int64).reshape(-1, 2) - 1
    c = np.asarray(coords, dtype=np.float64)
    a, b = nb[:, 0], nb[:, 1]
    d = c[:, a] - c[:, b]  # node1 = a, node2 = b
    vd = np.asarray(vxs) * d[0] + np.asarray(vys) * d[1] + np.asarray(vzs) * d[2]
    # direction (a, b): term at a active when (vd > 0) == use_upwind;
    # direction (b, a): delta flips, vd -> -vd
    act_a = (vd > 0) if use_upwind else ~(vd > 0)
    act_b = (-vd > 0) if use_upwind else ~(-vd > 0)
    return nb, d, vd, act_a, act_b


def transport_jacobian(neighbors, areasoverlengths, dirichletnodes, volumes, coords, vxs, vys,
                       vzs, Ds, use_upwind=True):
    """dr/du (n x n CSR); r is affine in u."""
    V = np.asarray(volumes, dtype=np.float64)
    n = V.size
    nb, d, vd, act_a, act_b = _geom(neighbors, coords, vxs, vys, vzs, use_upwind)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    D = np.asarray(Ds, dtype=np.float64)
    isd = np.zeros(n, dtype=bool)
    isd[np.asarray(dirichletnodes, dtype=np.int64).ravel() - 1] = True
    rows, cols, vals = [], [], []
    rows.append(np.nonzero(isd)[0])
    cols.append(np.nonzero(isd)[0])
    vals.append(np.ones(isd.sum()))
    for (i1, i2, w, act) in ((nb[:, 0], nb[:, 1], vd, act_a), (nb[:, 1], nb[:, 0], -vd, act_b)):
        m1 = ~isd[i1]
        # advection -(v.delta)(u1 - u2) aol / V1 where active; u2 free or Dirichlet
        adv = np.where(act & m1, -w * aol / V[i1], 0.0)
        dif = np.where(m1, D * aol / V[i1], 0.0)
        rows += [i1, i1]
        cols += [i1, i1]
        vals += [adv, -dif]
        m2 = m1 & ~isd[i2]
        rows += [i1, i1]
        cols += [i2, i2]
        vals += [np.where(m2, -adv, 0.0), np.where(m2, dif, 0.0)]
    J = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    J.sum_duplicates()
    return J


def transport_residuals(u, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                        dirichletus, Qs, volumes, coords, use_upwind=True):
    """src/transport.jl:2-33."""
    u = np.asarray(u, dtype=np.float64)
    V = np.asarray(volumes, dtype=np.float64)
    n = u.size
    nb, d, vd, act_a, act_b = _geom(neighbors, coords, vxs, vys, vzs, use_upwind)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    D = np.asarray(Ds, dtype=np.float64)
    ud = np.asarray(dirichletus, dtype=np.float64)
    isd = np.zeros(n, dtype=bool)
    isd[np.asarray(dirichletnodes, dtype=np.int64).ravel() - 1] = True
    r = np.where(isd, u - ud, np.asarray(Qs, dtype=np.float64) / V)
    ut = np.where(isd, ud, u)  # the value a neighbour contributes
    for (i1, i2, w, act) in ((nb[:, 0], nb[:, 1], vd, act_a), (nb[:, 1], nb[:, 0], -vd, act_b)):
        m1 = ~isd[i1]
        adv = np.where(act & m1, -w * (u[i1] - ut[i2]) * aol / V[i1], 0.0)
        dif = np.where(m1, D * (ut[i2] - u[i1]) * aol / V[i1], 0.0)
        np.add.at(r, i1, adv + dif)
    return r


def transport_steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                    dirichletus, Qs, volumes, coords, t0, t1, nsteps, use_upwind=True,
                    return_all=False):
    """Fixed-step backward Euler of du/dt = r(u) from t0 to t1."""
    args = (vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes, dirichletus, Qs,
            volumes, coords, use_upwind)
    dt = (t1 - t0) / nsteps
    J = transport_jacobian(neighbors, areasoverlengths, dirichletnodes, volumes, coords, vxs,
                           vys, vzs, Ds, use_upwind)
    n = len(u0)
    r0 = transport_residuals(np.zeros(n), *args)
    A = (sp.identity(n, format="csc") - dt * J).tocsc()
    lu = spla.splu(A)
    u = np.asarray(u0, dtype=np.float64).copy()
    us = [u.copy()]
    for _ in range(nsteps):
        u = lu.solve(u + dt * r0)
        us.append(u.copy())
    return np.array(us) if return_all else u


def transport_adjoint(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                      dirichletus, Qs, volumes, coords, t0, t1, nsteps, g, use_upwind=True):
    """Gradient of g . u_N: dict(vxs, vys, vzs, Ds, Qs)."""
    us = transport_steps(u0, vxs, vys, vzs, Ds, neighbors, areasoverlengths, dirichletnodes,
                         dirichletus, Qs, volumes, coords, t0, t1, nsteps, use_upwind,
                         return_all=True)
    dt = (t1 - t0) / nsteps
    V = np.asarray(volumes, dtype=np.float64)
    n = V.size
    J = transport_jacobian(neighbors, areasoverlengths, dirichletnodes, volumes, coords, vxs,
                           vys, vzs, Ds, use_upwind)
    A = (sp.identity(n, format="csc") - dt * J).T.tocsc()
    lu = spla.splu(A)
    nb, d, vd, act_a, act_b = _geom(neighbors, coords, vxs, vys, vzs, use_upwind)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    ud = np.asarray(dirichletus, dtype=np.float64)
    isd = np.zeros(n, dtype=bool)
    isd[np.asarray(dirichletnodes, dtype=np.int64).ravel() - 1] = True
    out = {k: np.zeros(nb.shape[0]) for k in ("vxs", "vys", "vzs", "Ds")}
    out["Qs"] = np.zeros(n)
    lam = np.asarray(g, dtype=np.float64).copy()
    a, b = nb[:, 0], nb[:, 1]
    for k in range(nsteps, 0, -1):
        mu = lu.solve(lam)
        u = us[k]
        ut = np.where(isd, ud, u)
        # d r_i1 / d v (component c) = -delta_c (u1 - u2) aol / V1 where active
        for (i1, i2, sgn, act) in ((a, b, 1.0, act_a), (b, a, -1.0, act_b)):
            m1 = ~isd[i1]
            base = np.where(act & m1, -(u[i1] - ut[i2]) * aol / V[i1], 0.0) * mu[i1] * dt
            for c, key in enumerate(("vxs", "vys", "vzs")):
                out[key] += base * sgn * d[c]
            out["Ds"] += np.where(m1, (ut[i2] - u[i1]) * aol / V[i1], 0.0) * mu[i1] * dt
        out["Qs"] += np.where(isd, 0.0, mu / V) * dt
        lam = mu
    return us[-1], out
