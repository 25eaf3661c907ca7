"""Pins the two-phase oracle (oracle/twophase_oracle.py) against the textbook
computation and the Buckley-Leverett solution of the reference's
test/twophase.jl, closed forms and conservation identities."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def tpo():
    from oracle import twophase_oracle
    return twophase_oracle


FLUID = (1.0, 1.0, 0.0, 0.0)


def quarter_five_spot(pkg, N=16):
    """test/twophase.jl:182-191: regulargrid2d on [0, 1 - 1/N]^2, dz = 1, no
    Dirichlet cell, unit K, q = +1 / -1 at the first / last cell."""
    coords, nb, aol, vol = pkg.grid.regulargrid2d((0.0, 0.0), (1 - 1 / N, 1 - 1 / N), (N, N), 1.0)
    n = coords.shape[1]
    q = np.zeros(n)
    q[0], q[n - 1] = 1.0, -1.0
    return nb, aol, vol, q, n


def buckley_leverett_column(pkg, ns=100):
    """test/twophase.jl:208-230: a 1-D column of unit length, head 0 at x = 1,
    injection 0.03 into the first cell."""
    coords, nb, aol, vol = pkg.grid.regulargrid1d(0.0, 1.0, ns, 1.0, 1.0)
    n = coords.shape[1]
    dn = np.array([i + 1 for i in range(n) if coords[0, i] == 1.0], dtype=np.int64)
    Q = np.zeros(n)
    Q[0] = 0.03
    return coords, nb, aol, vol, dn, Q, n


def _to_grid_order(v, N):
    """x-fastest (textbook) numbering -> regulargrid2d's y-fastest numbering."""
    return np.asarray(v).reshape(N, N).T.ravel()


def test_relativeperm_closed_form(tpo):
    fl = (2.0, 0.5, 0.1, 0.2)
    s = np.array([0.1, 0.8, 0.45])
    mw, mo = tpo.relativeperm(s, fl)
    assert np.allclose(mw, [0.0, 1.0 / 2.0, 0.25 / 2.0], rtol=1e-15, atol=0)
    assert np.allclose(mo, [1.0 / 0.5, 0.0, 0.25 / 0.5], rtol=1e-15, atol=1e-30)


def test_quarter_five_spot_matches_textbook(pkg, tpo):
    """The reference test's comparison (test/twophase.jl:150-163,198-206):
    DPFEHM's edge-list IMPES against the grid-array textbook computation, every
    step, pressures with the constant removed."""
    N = 16
    nb, aol, vol, q, n = quarter_five_spot(pkg, N)
    Ps, Ss = tpo.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), np.zeros(n), [], q, vol,
                               aol, FLUID, 0.7 / 25, nb, 5, everystep=True)
    Pt, St = tpo.textbook_quarter_five_spot(N, 5)
    for k in range(5):
        pt, st = _to_grid_order(Pt[k], N), _to_grid_order(St[k], N)
        assert np.linalg.norm(Ps[k] - pt) / np.linalg.norm(pt) < 1e-9
        assert np.abs(Ss[k] - st).max() < 1e-11
    total = np.linalg.norm(Ps[-1] - _to_grid_order(Pt[-1], N)) + \
        np.linalg.norm(Ss[-1] - _to_grid_order(St[-1], N))
    assert total < 1e-5  # the reference's own gate
    assert Ss[-1].max() > 0.3 and Ss[-1].min() == 0.0  # water has entered, not everywhere


def test_buckley_leverett(pkg, tpo):
    """test/twophase.jl:208-249: the reference's gate |mean(S_BL - S)| < 1e-2,
    plus a looser RMS bound."""
    coords, nb, aol, vol, dn, Q, n = buckley_leverett_column(pkg)
    dh = np.zeros(n)
    P, S = tpo.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), dh, dn, Q, vol, aol, FLUID,
                             0.7 / 25, nb, 25)
    x = coords[0]
    S_BL = tpo.buckley_leverett(x, Q[0] * 25 * (0.7 / 25) / 1.0)
    assert abs(np.mean(S_BL - S)) < 1e-2
    assert np.sqrt(np.mean((S_BL - S) ** 2)) < 0.1
    assert S[0] > 0.5 and S[-1] == 0.0


def test_buckley_leverett_closed_form(tpo):
    """f'(S) = x/t behind the shock: check the bisection on the closed form."""
    t = 0.5
    x = np.array([0.0, 0.1, 0.3, 0.6, 0.7])
    s = tpo.buckley_leverett(x, t)

    def fprime(v):
        return 2 * v * (1 - v) / (2 * v * v - 2 * v + 1) ** 2
    v_shock = fprime(1 / np.sqrt(2))
    for xi, si in zip(x, s):
        if xi / t > v_shock:
            assert si == 0.0
        else:
            assert abs(fprime(si) - xi / t) < 1e-12 and si >= 1 / np.sqrt(2) - 1e-15
    assert s[0] == pytest.approx(1.0)


def _random_box(pkg, seed=3, ns=(6, 5, 4)):
    rng = np.random.default_rng(seed)
    coords, nb, aol, vol = pkg.grid.regulargrid3d((0, 0, 0), (5.0, 4.0, 3.0), ns)
    n = coords.shape[1]
    K = np.exp(rng.normal(size=n))
    Q = np.zeros(n)
    Q[0], Q[n - 1] = 0.5, -0.5
    Q[n // 2] = 0.2
    S = 0.3 * rng.random(n)
    return nb, aol, vol, K, Q, S, n, rng


def test_flux_antisymmetry_and_balance(pkg, tpo):
    nb, aol, vol, K, Q, S, n, _ = _random_box(pkg)
    dn = np.array([n // 3], dtype=np.int64)
    dh = np.zeros(n)
    P, Vn = tpo.tpfa(K, dh, dn, Q, aol, nb)
    # Vn_e = (P_a - P_b) tx K_e with K_e the arithmetic mean
    Ke = 0.5 * (K[nb[:, 0] - 1] + K[nb[:, 1] - 1])
    assert np.allclose(Vn, (P[nb[:, 0] - 1] - P[nb[:, 1] - 1]) * 2 * aol * Ke, rtol=1e-14)
    # net outflow of every free cell equals its source
    out = np.zeros(n)
    np.add.at(out, nb[:, 0] - 1, Vn)
    np.add.at(out, nb[:, 1] - 1, -Vn)
    free = np.setdiff1d(np.arange(n), dn - 1)
    assert np.allclose(out[free], Q[free], atol=1e-10 * np.abs(Q).max())


def test_saturation_residual_edge_terms_cancel(pkg, tpo):
    nb, aol, vol, K, Q, S, n, rng = _random_box(pkg)
    P = rng.normal(size=n)
    Vn = (P[nb[:, 0] - 1] - P[nb[:, 1] - 1]) * aol
    f = tpo.fracflow(S, FLUID)
    R = tpo.saturation_residual(f, Q, nb, P, Vn)
    assert abs(R.sum() - (np.minimum(Q, 0) * f).sum()) < 1e-13 * np.abs(Vn).sum()
    # two cells: flow from the high-pressure cell carries its own f
    R2 = tpo.saturation_residual(np.array([0.3, 0.9]), np.zeros(2), np.array([[1, 2]]),
                                 np.array([2.0, 1.0]), np.array([5.0]))
    assert np.allclose(R2, [-0.3 * 5.0, 0.3 * 5.0], rtol=0, atol=0)


def test_upstream_water_balance_one_substep(pkg, tpo):
    """One unclamped sub-step: sum(pv dS) = tau (sum max(Q,0) + sum min(Q,0) f)."""
    nb, aol, vol, K, Q, S, n, _ = _random_box(pkg)
    dn = np.array([n // 3], dtype=np.int64)
    P, Vn = tpo.tpfa(K, np.zeros(n), dn, Q, aol, nb)
    # cfl from the same formula, then T = cfl / 2 gives exactly one sub-step
    Vi = np.zeros(n)
    a, b = nb[:, 0] - 1, nb[:, 1] - 1
    np.add.at(Vi, np.where(Vn < 0, a, b), np.abs(Vn))
    cfl = (1.0 / 3.0) * np.min(vol / (Vi + np.maximum(Q, 0)))
    S1, nts = tpo.upstream(S, FLUID, Q, 0.5 * cfl, P, Vn, nb, vol)
    assert nts == 1 and 0 <= S1.min() and S1.max() <= 1
    lhs = (vol * (S1 - S)).sum()
    rhs = 0.5 * cfl * (np.maximum(Q, 0).sum() + (np.minimum(Q, 0) * tpo.fracflow(S, FLUID)).sum())
    assert abs(lhs - rhs) < 1e-13 * abs(rhs)


def test_solvetwophase_gradient_fd(pkg, tpo):
    """Reverse mode of the oracle's IMPES loop against central differences in
    K, for J = c_P . P_nt + c_S . S_nt, on a heterogeneous box with a Dirichlet
    face (and on the reference test's singular quarter five-spot with its
    FindGrad objective, P at the centre cell, test/twophase.jl:192-204)."""
    nb, aol, vol, K, Q, S, n, rng = _random_box(pkg, ns=(6, 5, 4))
    dn = np.arange(4, n + 1, 4, dtype=np.int64)
    dh = np.zeros(n)
    fluid = (1.0, 3.0, 0.05, 0.1)
    S0 = np.full(n, 0.05)
    cP, cS = rng.normal(size=n), rng.normal(size=n)
    args = (np.zeros(n), S0, None, dh, dn, Q, vol, aol, fluid, 0.8, nb, 3)

    def J(Kv):
        a = list(args)
        a[2] = Kv
        P, S = tpo.solvetwophase(*a)
        return cP @ P + cS @ S
    a = list(args)
    a[2] = K
    P, S, dK, _ = tpo.solvetwophase_grad(*a, cP, cS)
    for i in np.argsort(-np.abs(dK))[:5]:
        eps = 1e-6 * K[i]
        Kp, Km = K.copy(), K.copy()
        Kp[i] += eps
        Km[i] -= eps
        fd = (J(Kp) - J(Km)) / (2 * eps)
        assert abs(fd - dK[i]) < 1e-5 * abs(dK[i]) + 1e-9 * np.abs(dK).max()
    # the reference test's objective on its singular 16x16 case, two steps
    N = 16
    nb, aol, vol, q, n = quarter_five_spot(pkg, N)
    crit = int(np.ceil(N * N / 2)) - 1
    seed = np.zeros(n)
    seed[crit] = 1.0
    Kq = np.ones(n)
    P, S, dK, _ = tpo.solvetwophase_grad(np.zeros(n), np.zeros(n), Kq, np.zeros(n), [], q, vol,
                                         aol, FLUID, 0.7 / 25, nb, 2, seed, np.zeros(n))
    for i in np.argsort(-np.abs(dK))[:3]:
        Kp, Km = Kq.copy(), Kq.copy()
        Kp[i] += 1e-6
        Km[i] -= 1e-6
        fp = tpo.solvetwophase(np.zeros(n), np.zeros(n), Kp, np.zeros(n), [], q, vol, aol, FLUID,
                               0.7 / 25, nb, 2)[0][crit]
        fm = tpo.solvetwophase(np.zeros(n), np.zeros(n), Km, np.zeros(n), [], q, vol, aol, FLUID,
                               0.7 / 25, nb, 2)[0][crit]
        assert abs((fp - fm) / 2e-6 - dK[i]) < 1e-4 * abs(dK[i])
