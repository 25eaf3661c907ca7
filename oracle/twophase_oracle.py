"""FP64 CPU oracle for DPFEHM's two-phase IMPES path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` imports this module, and only as the checker.  The product
(``dpfehm.jl_amd``) never loads it.  Plain numpy; the pressure solve goes
through ``gw_oracle.groundwater_steadystate`` (IC(0)-PCG in C).

Function (restates) -- pinned by (tests/test_oracle_twophase.py)

* ``relativeperm``            src/twophase.jl:201-206
  -- closed form at S = swc, 1 - sor and the midpoint.
* ``tpfa``                    src/twophase.jl:189-198 (arithmetic-mean K,
  solvepressure :48-56 on the transmissivity equations :23-44, tx = 2 aol)
  -- the textbook TPFA of test/twophase.jl:54-88 (Aarnes, Gimse and Lie's
  MATLAB code, assembled from grid arrays instead of the edge list) on the
  reference test's 16x16 quarter five-spot; flux antisymmetry and zero net
  flux out of every cell without a source.
* ``saturation_residual``     src/twophase.jl:85-112 (saturationcalc, every
  cell an unknown) -- the textbook upstream matrix GenA of
  test/twophase.jl:132-148 applied to f; the edge terms sum to zero (what
  leaves one cell enters the other).
* ``upstream``                src/twophase.jl:208-235 -- the textbook Upstream
  of test/twophase.jl:109-130 (no clamping there; none is active in that
  case); the discrete water balance sum(pv dS) = tau (sum max(Q,0) + sum
  min(Q,0) f) per unclamped sub-step.
* ``solvetwophase``           src/twophase.jl:238-264 -- the reference test's
  comparison with the textbook solution (test/twophase.jl:150-163,198-206,
  total_norm < 1e-5, here with the pressure constant removed) and its
  Buckley-Leverett check (test/twophase.jl:208-249, RMS error < 1e-2).
* ``solvetwophase_grad``      reverse mode of the IMPES loop (the gradient
  test/twophase.jl:192-205 takes with Zygote through the rrules of
  solvepressure and saturationcalc, src/twophase.jl:59-80,117-183) --
  central differences in K of c_P . P + c_S . S on a box, and of the
  reference test's FindGrad objective (P at the centre cell after two steps)
  on its singular quarter five-spot.

Singular pressure systems (no Dirichlet cell): the reference's AMG-CG picks
some constant; here one cell is pinned and P is shifted to mean zero, and the
tests compare pressures up to that constant (fluxes and saturations do not
depend on it).
"""
import numpy as np

from . import gw_oracle as G


def relativeperm(s, fluid):
    """src/twophase.jl:201-206.  fluid = (vw, vo, swc, sor)."""
    vw, vo, swc, sor = fluid
    S = (np.asarray(s, dtype=np.float64) - swc) / ((1.0 - swc) - sor)
    mw = S * S / vw
    o = 1.0 - S
    mo = o * o / vo
    return mw, mo


def fracflow(s, fluid):
    mw, mo = relativeperm(s, fluid)
    return mw / (mw + mo)


def _nb0(neighbors):
    return np.asarray(neighbors, dtype=np.int64).reshape(-1, 2) - 1


def pressure_solve(Ksum, neighbors, aol, dirichletnodes, dirichleths, Qs, reltol=1e-14):
    """The transmissivity system with K_e = Ksum (transmissibility aol Ksum =
    2 aol 0.5 Ksum); singular systems (no Dirichlet cell) are pinned at the
    last cell and shifted to mean zero."""
    n = len(Qs)
    dn = np.asarray(dirichletnodes, dtype=np.int64).ravel()
    dh = np.asarray(dirichleths, dtype=np.float64).copy()
    pinned = dn.size == 0
    if pinned:
        dn = np.array([n], dtype=np.int64)
        dh[n - 1] = 0.0
    P = G.groundwater_steadystate(Ksum, neighbors, aol, dn, dh, Qs, reltol=reltol)
    if pinned:
        P = P - P.sum() / n
    return P


def tpfa(K, dirichleths, dirichletnodes, Qs, areasoverlengths, neighbors, reltol=1e-14):
    """src/twophase.jl:189-198: returns (P, Vn)."""
    nb = _nb0(neighbors)
    K = np.asarray(K, dtype=np.float64)
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    Ksum = K[nb[:, 0]] + K[nb[:, 1]]
    P = pressure_solve(Ksum, neighbors, aol, dirichletnodes, dirichleths, Qs, reltol)
    Vn = (P[nb[:, 0]] - P[nb[:, 1]]) * (aol * Ksum)
    return P, Vn


def saturation_residual(f, Qs, neighbors, P, Vn):
    """src/twophase.jl:85-112 with every cell an unknown: the mass term, then
    per edge in list order the (a, b) and (b, a) upwind terms."""
    nb = _nb0(neighbors)
    f = np.asarray(f, dtype=np.float64)
    Q = np.asarray(Qs, dtype=np.float64)
    R = np.minimum(Q, 0.0) * f
    a, b = nb[:, 0], nb[:, 1]
    keep = a != b
    a, b, v = a[keep], b[keep], np.asarray(Vn, dtype=np.float64)[keep]

    def term(j1, j2):
        up = (P[j2] - P[j1]) >= 0
        t_up = np.where(v > 0, f[j2] * v, -f[j2] * v)
        t_dn = np.where(v > 0, -f[j1] * v, f[j1] * v)
        return np.where(up, t_up, t_dn)

    idx = np.stack([a, b], axis=1).ravel()
    val = np.stack([term(a, b), term(b, a)], axis=1).ravel()
    np.add.at(R, idx, val)  # sequential: per cell in edge order
    return R


def upstream(S, fluid, Qs, T, P, Vn, neighbors, volumes):
    """src/twophase.jl:208-235.  Returns (S, Nts)."""
    vw, vo, swc, sor = fluid
    nb = _nb0(neighbors)
    S = np.asarray(S, dtype=np.float64).copy()
    Q = np.asarray(Qs, dtype=np.float64)
    Vn = np.asarray(Vn, dtype=np.float64)
    pv = np.asarray(volumes, dtype=np.float64) * 1.0
    fi = np.maximum(Q, 0.0)
    Vi = np.zeros(S.size)
    keep = nb[:, 0] != nb[:, 1]
    a, b, v = nb[keep, 0], nb[keep, 1], Vn[keep]
    neg = v < 0
    idx = np.where(neg, a, b)
    val = np.where(neg, -v, v)
    np.add.at(Vi, idx, val)
    with np.errstate(divide="ignore", invalid="ignore"):
        pm = np.min(pv / (Vi + fi))
    cfl = (((1.0 - swc) - sor) / 3.0) * pm
    nts = int(np.ceil(T / cfl))
    if nts <= 0:
        return S, 0
    dtx = (T / nts) / pv
    fi = np.maximum(Q, 0.0) * dtx
    for _ in range(nts):
        f = fracflow(S, fluid)
        S = S + (saturation_residual(f, Q, neighbors, P, Vn) * dtx + fi)
        S = np.where(S > 1.0 - sor, 1.0 - sor, np.where(S < swc, swc, S))
    return S, nts


def solvetwophase(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes, areasoverlengths,
                  fluid, dt, neighbors, nt, everystep=False, reltol=1e-14):
    """src/twophase.jl:238-264.  Returns (P, S), or the lists of every step's
    (P, S) when everystep."""
    S = np.asarray(S0, dtype=np.float64).copy()
    P = np.asarray(h0, dtype=np.float64).copy()
    K = np.asarray(K, dtype=np.float64)
    Ps, Ss = [], []
    for _ in range(nt):
        mw, mo = relativeperm(S, fluid)
        Km = (mw + mo) * K
        P, Vn = tpfa(Km, dirichleths, dirichletnodes, Qs, areasoverlengths, neighbors, reltol)
        S, _ = upstream(S, fluid, Qs, dt, P, Vn, neighbors, volumes)
        if everystep:
            Ps.append(P.copy())
            Ss.append(S.copy())
    return (Ps, Ss) if everystep else (P, S)


# ---------------------------------------------------------------- textbook restatement

def textbook_quarter_five_spot(N=16, nt=5, dt=0.7 / 25, fluid=(1.0, 1.0, 0.0, 0.0)):
    """The MATLAB-style computation of test/twophase.jl:39-163 (Aarnes, Gimse
    and Lie), on an N x N x 1 grid of unit square, hz = 1, unit K and porosity,
    q = +1 at the first cell and -1 at the last.  Cells are numbered x fastest
    as there; returns every step's (P, S) in that numbering, P with the
    constant removed (pinned last cell, then mean zero)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    Nx, Ny = N, N
    n = Nx * Ny
    hx = hy = 1.0 / N
    hz = 1.0
    vol = hx * hy * hz
    q = np.zeros(n)
    q[0], q[n - 1] = 1.0, -1.0
    vw, vo, swc, sor = fluid

    def rel(s):
        S = (s - swc) / ((1.0 - swc) - sor)
        return S * S / vw, (1.0 - S) ** 2 / vo

    S = np.zeros(n)
    Ps, Ss = [], []
    for _ in range(nt):
        mw, mo = rel(S)
        L = (mw + mo).reshape(Ny, Nx)  # [j, i], x fastest
        tx = 2 * hy * hz / hx
        ty = 2 * hx * hz / hy
        TX = np.zeros((Ny, Nx + 1))
        TY = np.zeros((Ny + 1, Nx))
        TX[:, 1:Nx] = tx * (0.5 * (L[:, :Nx - 1] + L[:, 1:]))
        TY[1:Ny, :] = ty * (0.5 * (L[:Ny - 1, :] + L[1:, :]))
        x1, x2 = TX[:, :Nx].ravel(), TX[:, 1:].ravel()
        y1, y2 = TY[:Ny, :].ravel(), TY[1:, :].ravel()
        A = sp.diags([-y2[:n - Nx], -x2[:n - 1], x1 + x2 + y1 + y2, -x1[1:], -y1[Nx:]],
                     [-Nx, -1, 0, 1, Nx], shape=(n, n)).tocsr()
        # singular (pure Neumann): pin the last cell, then remove the mean
        u = np.zeros(n)
        u[:n - 1] = spla.spsolve(A[:n - 1, :n - 1].tocsc(), q[:n - 1])
        u -= u.mean()
        Pg = u.reshape(Ny, Nx)
        Vx = np.zeros((Ny, Nx + 1))
        Vy = np.zeros((Ny + 1, Nx))
        Vx[:, 1:Nx] = (Pg[:, :Nx - 1] - Pg[:, 1:]) * TX[:, 1:Nx]
        Vy[1:Ny, :] = (Pg[:Ny - 1, :] - Pg[1:, :]) * TY[1:Ny, :]
        # Upstream (test/twophase.jl:109-130) with GenA (:132-148)
        pv = np.full(n, vol)
        fi = np.maximum(q, 0.0)
        XP, XN = np.maximum(Vx, 0), np.minimum(Vx, 0)
        YP, YN = np.maximum(Vy, 0), np.minimum(Vy, 0)
        Vi = (XP[:, :Nx] + YP[:Ny, :] - XN[:, 1:] - YN[1:, :]).ravel()
        pm = np.min(pv / (Vi + fi))
        cfl = (((1.0 - swc) - sor) / 3.0) * pm
        nts = int(np.ceil(dt / cfl))
        dtx = (dt / nts) / pv
        fp = np.minimum(q, 0.0)
        gx1, gy1 = XN[:, :Nx].ravel(), YN[:Ny, :].ravel()
        gx2, gy2 = XP[:, 1:].ravel(), YP[1:, :].ravel()
        GA = sp.diags([gy2[:n - Nx], gx2[:n - 1], fp + gx1 - gx2 + gy1 - gy2, -gx1[1:], -gy1[Nx:]],
                      [-Nx, -1, 0, 1, Nx], shape=(n, n)).tocsr()
        GA = sp.diags(dtx) @ GA
        fi = np.maximum(q, 0.0) * dtx
        for _ in range(nts):
            mw, mo = rel(S)
            S = S + (GA @ (mw / (mw + mo)) + fi)
        Ps.append(u.copy())
        Ss.append(S.copy())
    return Ps, Ss


def buckley_leverett(x, t):
    """Analytical Buckley-Leverett saturation for unit mobilities ratio
    (vw = vo = 1, swc = sor = 0): f(S) = S^2 / (S^2 + (1-S)^2); the shock
    saturation 1/sqrt(2); behind it S solves f'(S) = x/t on [S*, 1]
    (test/twophase.jl:218-241), here by bisection."""
    def fprime(s):
        return 2 * s * (1 - s) / (2 * s * s - 2 * s + 1) ** 2
    s_star = 1.0 / np.sqrt(2.0)
    v_shock = fprime(s_star)
    out = np.zeros(len(x))
    for i, xv in enumerate(x):
        xi = xv / t
        if xi > v_shock:
            continue
        lo, hi = s_star, 1.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            # f' decreases on [S*, 1]: f'(S*) = v_shock >= xi >= f'(1) = 0
            if fprime(mid) > xi:
                lo = mid
            else:
                hi = mid
        out[i] = 0.5 * (lo + hi)
    return out


# ---------------------------------------------------------------- reverse mode

def _mobility_and_slope(s, fluid):
    """Mt = Mw + Mo and dMt/ds, f = Mw / Mt and df/ds."""
    vw, vo, swc, sor = fluid
    span = (1.0 - swc) - sor
    S = (np.asarray(s, dtype=np.float64) - swc) / span
    mw, mo = S * S / vw, (1.0 - S) ** 2 / vo
    dmw, dmo = 2.0 * S / vw / span, -2.0 * (1.0 - S) / vo / span
    mt = mw + mo
    return mt, dmw + dmo, mw / mt, (dmw * mo - mw * dmo) / (mt * mt)


def solvetwophase_grad(h0, S0, K, dirichleths, dirichletnodes, Qs, volumes, areasoverlengths,
                       fluid, dt, neighbors, nt, seedP, seedS, reltol=1e-14):
    """Reverse mode of ``solvetwophase`` for J with dJ/dP_nt = seedP and
    dJ/dS_nt = seedS.  Returns (P, S, dJ/dK, dJ/dS0).  Sub-step counts, upwind
    choices and clamp activity are held fixed (they are piecewise constant in
    K); the clamp passes the tangent where lo <= u <= hi."""
    vw, vo, swc, sor = fluid
    lo, hi = swc, 1.0 - sor
    nb = _nb0(neighbors)
    a, b = nb[:, 0], nb[:, 1]
    aol = np.asarray(areasoverlengths, dtype=np.float64)[:nb.shape[0]]
    K = np.asarray(K, dtype=np.float64)
    Q = np.asarray(Qs, dtype=np.float64)
    pv = np.asarray(volumes, dtype=np.float64)
    n = K.size
    dn = np.asarray(dirichletnodes, dtype=np.int64).ravel()
    pinned = dn.size == 0
    dn_eff = np.array([n], dtype=np.int64) if pinned else dn
    dh_eff = np.asarray(dirichleths, dtype=np.float64).copy()
    if pinned:
        dh_eff[n - 1] = 0.0
    qneg = np.minimum(Q, 0.0)
    tape = []
    S = np.asarray(S0, dtype=np.float64).copy()
    P = np.asarray(h0, dtype=np.float64).copy()
    for _ in range(nt):
        mt = relativeperm(S, fluid)
        Km = (mt[0] + mt[1]) * K
        Ksum = Km[a] + Km[b]
        P, pb = G.steadystate_rrule(Ksum, neighbors, aol, dn_eff, dh_eff, Q, reltol=reltol)
        if pinned:
            P = P - P.sum() / n
        Vn = (P[a] - P[b]) * (aol * Ksum)
        Snew, nts = upstream(S, fluid, Q, dt, P, Vn, neighbors, volumes)
        tape.append((S, Ksum, P, Vn, pb, nts))
        S = Snew
    Pfin, Sfin = P, S
    dS = np.asarray(seedS, dtype=np.float64).copy()
    dK = np.zeros(n)
    for t in range(nt - 1, -1, -1):
        S_prev, Ksum, P, Vn, pb, nts = tape[t]
        av = np.abs(Vn)
        up_a = (P[b] - P[a]) >= 0  # the term at a takes f_b
        up_b = (P[a] - P[b]) >= 0
        dabs = np.zeros(Vn.size)
        dVn = np.zeros(Vn.size)
        if nts > 0:
            dtx = (dt / nts) / pv
            fi = np.maximum(Q, 0.0) * dtx
            states = [S_prev]
            for _ in range(nts - 1):
                s_ = states[-1]
                f = fracflow(s_, fluid)
                u = s_ + (saturation_residual(f, Q, neighbors, P, Vn) * dtx + fi)
                states.append(np.where(u > hi, hi, np.where(u < lo, lo, u)))
            for k in range(nts - 1, -1, -1):
                s_ = states[k]
                _, _, f, dfds = _mobility_and_slope(s_, fluid)
                u = s_ + (saturation_residual(f, Q, neighbors, P, Vn) * dtx + fi)
                du = dS * ((u >= lo) & (u <= hi))
                dR = du * dtx
                df = qneg * dR
                # term at a: up_a ? f_b |v| : -f_a |v|; at b: up_b ? f_a |v| : -f_b |v|
                np.add.at(df, b, np.where(up_a, av * dR[a], 0.0))
                np.add.at(df, a, np.where(up_a, 0.0, -av * dR[a]))
                np.add.at(df, a, np.where(up_b, av * dR[b], 0.0))
                np.add.at(df, b, np.where(up_b, 0.0, -av * dR[b]))
                dabs += np.where(up_a, f[b], -f[a]) * dR[a] + np.where(up_b, f[a], -f[b]) * dR[b]
                dS = du + dfds * df
            dVn = np.sign(Vn) * dabs
        # Vn = (P_a - P_b) aol Ksum
        dP = np.zeros(n)
        np.add.at(dP, a, dVn * aol * Ksum)
        np.add.at(dP, b, -dVn * aol * Ksum)
        if t == nt - 1:
            dP += np.asarray(seedP, dtype=np.float64)
        dKsum = dVn * (P[a] - P[b]) * aol
        if pinned:
            dP = dP - dP.sum() / n
        dKsum = dKsum + pb(dP)[1]
        dKm = np.zeros(n)
        np.add.at(dKm, a, dKsum)
        np.add.at(dKm, b, dKsum)
        mt, dmt, _, _ = _mobility_and_slope(S_prev, fluid)
        dK += mt * dKm
        dS = dS + K * dmt * dKm
    return Pfin, Sfin, dK, dS
