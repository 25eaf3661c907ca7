"""Pins the FP64 oracle (oracle/gw_oracle.c) against closed forms, brute force,
finite differences and the known-answer problems of the reference's tests."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _tiny(pkg, ns=(4, 3, 5), nd=4, seed=0, hetero=True):
    rng = np.random.default_rng(seed)
    coords, nb, aol, vol = pkg.grid.regulargrid3d((0, 0, 0), (3.0, 2.0, 4.0), ns)
    n = coords.shape[1]
    K = np.exp(rng.normal(size=nb.shape[0])) if hetero else np.ones(nb.shape[0])
    dn = np.sort(rng.choice(np.arange(1, n + 1), size=nd, replace=False))
    dh = rng.normal(size=n)
    Q = rng.normal(size=n) * 0.1
    return nb, aol, vol, dn, dh, Q, K, n


# ---------------------------------------------------------------- getfreenodes

def test_freenodes_matches_both_reference_forms(oracle):
    rng = np.random.default_rng(1)
    n = 50
    dn = np.sort(rng.choice(np.arange(1, n + 1), size=7, replace=False))
    isfree, n2f, f2n = oracle.getfreenodes(n, dn)
    # src/twophase.jl:15-20
    ref_free = np.array([i not in set(dn) for i in range(1, n + 1)])
    ref_n2f = np.array([ref_free[:i].sum() if ref_free[i - 1] else -1 for i in range(1, n + 1)])
    ref_f2n = np.array([i for i in range(1, n + 1) if ref_free[i - 1]])
    assert np.array_equal(isfree, ref_free)
    assert np.array_equal(n2f, ref_n2f)
    assert np.array_equal(f2n, ref_f2n)


# ---------------------------------------------------------------- residual / Jacobian

def test_residual_is_affine_with_jacobian(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg)
    rng = np.random.default_rng(2)
    Ss = rng.uniform(0.5, 2.0, n)
    V = rng.uniform(0.5, 2.0, n)
    nf = n - dn.size
    h = rng.normal(size=nf)
    args = (K, nb, aol, dn, dh, Q, Ss, V)
    r0 = oracle.groundwater_residuals(np.zeros(nf), *args)
    r = oracle.groundwater_residuals(h, *args)
    J = oracle.groundwater_h(h, *args)
    assert np.allclose(r - r0, J @ h, rtol=1e-13, atol=1e-13)


def test_jacobian_matches_central_differences(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, ns=(3, 3, 3), nd=3, seed=3)
    rng = np.random.default_rng(4)
    Ss, V = rng.uniform(0.5, 2.0, n), rng.uniform(0.5, 2.0, n)
    nf = n - dn.size
    h = rng.normal(size=nf)
    args = (K, nb, aol, dn, dh, Q, Ss, V)
    J = oracle.groundwater_h(h, *args).toarray()
    eps = 1e-6
    Jfd = np.empty_like(J)
    for k in range(nf):
        hp, hm = h.copy(), h.copy()
        hp[k] += eps
        hm[k] -= eps
        Jfd[:, k] = (oracle.groundwater_residuals(hp, *args) - oracle.groundwater_residuals(hm, *args)) / (2 * eps)
    assert np.allclose(J, Jfd, rtol=1e-7, atol=1e-9 * np.abs(J).max())


def test_jacobian_symmetric_and_row_sums(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=5)
    nf = n - dn.size
    ones = np.ones(n)
    A = oracle.groundwater_h(np.zeros(nf), K, nb, aol, dn, dh, Q, ones, ones)
    assert abs(A - A.T).max() < 1e-15
    isfree, n2f, f2n = oracle.getfreenodes(n, dn)
    touches_d = np.zeros(n, dtype=bool)
    for a, b in nb:
        if not isfree[a - 1] or not isfree[b - 1]:
            touches_d[a - 1] = touches_d[b - 1] = True
    rows = np.array([not touches_d[c - 1] for c in f2n])
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert np.all(np.abs(rs[rows]) < 1e-13)
    assert np.all(rs[~rows] > 0)


def test_residual_hand_computed_1d(oracle):
    # cells 1-2-3, edges (1,2) and (2,3); Dirichlet h=2 at cell 3; Q=0.5 at cell 1
    nb = np.array([[1, 2], [2, 3]])
    K = np.array([2.0, 3.0])
    aol = np.array([1.5, 0.5])
    dn = np.array([3])
    dh = np.array([0.0, 0.0, 2.0])
    Q = np.array([0.5, 0.0, 0.0])
    one = np.ones(3)
    h = np.array([1.0, -1.0])
    r = oracle.groundwater_residuals(h, K, nb, aol, dn, dh, Q, one, one)
    t1, t2 = 2.0 * 1.5, 3.0 * 0.5
    assert np.allclose(r, [-0.5 + t1 * (1.0 + 1.0), t1 * (-1.0 - 1.0) + t2 * (-1.0 - 2.0)], rtol=1e-15)
    h_sol = oracle.groundwater_steadystate(K, nb, aol, dn, dh, Q)
    # flux balance: Q enters at 1, leaves through 2 -> 3
    assert np.isclose(h_sol[2], 2.0)
    assert np.isclose(h_sol[1], 2.0 + 0.5 / t2)
    assert np.isclose(h_sol[0], h_sol[1] + 0.5 / t1)


def test_residual_edge_order_and_duplicate_edges(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=6)
    rng = np.random.default_rng(7)
    nf = n - dn.size
    h = rng.normal(size=nf)
    one = np.ones(n)
    r = oracle.groundwater_residuals(h, K, nb, aol, dn, dh, Q, one, one)
    perm = rng.permutation(nb.shape[0])
    rp = oracle.groundwater_residuals(h, K[perm], nb[perm][:, ::-1], aol[perm], dn, dh, Q, one, one)
    assert np.allclose(r, rp, rtol=1e-12, atol=1e-12)
    # splitting every edge into two parallel edges of half conductivity
    nb2 = np.concatenate([nb, nb])
    K2 = np.concatenate([K * 0.25, K * 0.75])
    aol2 = np.concatenate([aol, aol])
    r2 = oracle.groundwater_residuals(h, K2, nb2, aol2, dn, dh, Q, one, one)
    assert np.allclose(r, r2, rtol=1e-12, atol=1e-12)
    A = oracle.groundwater_h(h, K, nb, aol, dn, dh, Q, one, one)
    A2 = oracle.groundwater_h(h, K2, nb2, aol2, dn, dh, Q, one, one)
    assert A.nnz == A2.nnz and abs(A - A2).max() < 1e-12


# ---------------------------------------------------------------- steady solve

@pytest.mark.parametrize("dim", [1, 2, 3])
def test_linear_head_exact(pkg, oracle, dim):
    """examples/fracture_network_solver_scaling/porous_media_comparison.jl:24-46:
    K = 1, h = 1 at x = min, h = 0 at x = max, Q = 0 -> h = (max - x)/(max - min)."""
    if dim == 1:
        coords, nb, aol, _ = pkg.grid.regulargrid1d([0.0], [100.0], [9], 1.0, 1.0)
    elif dim == 2:
        coords, nb, aol, _ = pkg.grid.regulargrid2d([0, 0], [100.0, 100.0], [9, 7], 1.0)
    else:
        coords, nb, aol, _ = pkg.grid.regulargrid3d([0, 0, 0], [100.0, 100.0, 100.0], [8, 6, 5])
    n = coords.shape[1]
    x = coords[0]
    dn = np.array([i + 1 for i in range(n) if x[i] in (0.0, 100.0)])
    dh = np.where(x == 0.0, 1.0, 0.0)
    h = oracle.groundwater_steadystate(np.ones(nb.shape[0]), nb, aol, dn, dh, np.zeros(n))
    assert np.allclose(h, (100.0 - x) / 100.0, rtol=0, atol=1e-12)


def test_thiem_drawdown(pkg, oracle):
    """test/thiem.jl:27-67 (10 % gate)."""
    side, thick, n = 50.0, 10.0, 101
    coords, nb, aol, _ = pkg.grid.regulargrid2d([-side, -side], [side, side], [n, n], thick)
    k, Q = 1e-5, 1e-3
    Ks = np.full(nb.shape[0], k)
    Qs = np.zeros(coords.shape[1])
    Qs[n * ((n + 1) // 2 - 1) + (n + 1) // 2 - 1] = -Q
    r = np.hypot(coords[0], coords[1])
    dn = np.nonzero(r >= side)[0] + 1
    dh = np.where(r >= side, 1e3, 0.0)
    h = oracle.groundwater_steadystate(Ks, nb, aol, dn, dh, Qs, reltol=1e-12)
    good = np.nonzero((coords[1] == 0) & (coords[0] > 0.1))[0]
    rs = coords[0, good]
    thiem = Q * np.log(side / rs) / (2 * np.pi * thick * k)
    assert np.allclose(thiem, 1e3 - h[good], rtol=1e-1)


def test_steady_matches_direct_solve(pkg, oracle):
    w = pkg.workloads.steady_2m(sigma=2.0, ns=(24, 22, 14), maxs=(12.0, 11.0, 7.0))
    K = w.Ks()
    h = oracle.groundwater_steadystate(K, w.neighbors, w.areasoverlengths, w.dirichletnodes,
                                       w.dirichleths, w.Qs)
    n = w.n
    one = np.ones(n)
    nf = n - w.dirichletnodes.size
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, one, one)
    b = -oracle.groundwater_residuals(np.zeros(nf), *args)
    A = oracle.groundwater_h(np.zeros(nf), *args).tocsc()
    hf = spla.spsolve(A, b)
    isfree, n2f, f2n = oracle.getfreenodes(n, w.dirichletnodes)
    assert np.linalg.norm(h[f2n - 1] - hf) / np.linalg.norm(hf) < 1e-11


# ---------------------------------------------------------------- adjoint

def _steady_fixture(pkg, ns=(4, 4, 3), seed=8):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, ns=ns, nd=3, seed=seed)
    return nb, aol, dn, dh, Q, K, n


def test_pullback_matches_central_fd(pkg, oracle):
    nb, aol, dn, dh, Q, K, n = _steady_fixture(pkg)
    rng = np.random.default_rng(9)
    w = rng.normal(size=n)
    J = lambda K_, dh_, Q_: w @ oracle.groundwater_steadystate(K_, nb, aol, dn, dh_, Q_)
    h, pb = oracle.steadystate_rrule(K, nb, aol, dn, dh, Q)
    _, dK, _, _, _, ddh, dQ = pb(w)
    eps = 1e-6
    for e in range(nb.shape[0]):
        Kp, Km = K.copy(), K.copy()
        Kp[e] *= 1 + eps
        Km[e] *= 1 - eps
        fd = (J(Kp, dh, Q) - J(Km, dh, Q)) / (2 * eps * K[e])
        assert np.isclose(dK[e], fd, rtol=1e-6, atol=1e-9 * np.abs(dK).max())
    for i in range(n):
        dp, dm = dh.copy(), dh.copy()
        dp[i] += eps
        dm[i] -= eps
        assert np.isclose(ddh[i], (J(K, dp, Q) - J(K, dm, Q)) / (2 * eps), rtol=1e-6, atol=1e-8)
        qp, qm = Q.copy(), Q.copy()
        qp[i] += eps
        qm[i] -= eps
        assert np.isclose(dQ[i], (J(K, dh, qp) - J(K, dh, qm)) / (2 * eps), rtol=1e-6, atol=1e-8)


def test_pullback_dense_inverse_and_reciprocity(pkg, oracle):
    nb, aol, dn, dh, Q, K, n = _steady_fixture(pkg, seed=10)
    one = np.ones(n)
    isfree, n2f, f2n = oracle.getfreenodes(n, dn)
    nf = f2n.size
    A = oracle.groundwater_h(np.zeros(nf), K, nb, aol, dn, dh, Q, one, one).toarray()
    Ainv = np.linalg.inv(A)
    rng = np.random.default_rng(11)
    delta = rng.normal(size=n)
    _, pb = oracle.steadystate_rrule(K, nb, aol, dn, dh, Q)
    _, dK, _, _, _, ddh, dQ = pb(delta)
    lam = Ainv.T @ delta[f2n - 1]
    assert np.allclose(dQ[f2n - 1], lam, rtol=1e-10, atol=1e-12)
    # reciprocity: dh_k/dQ_j == dh_j/dQ_k  (A symmetric)
    assert np.allclose(Ainv, Ainv.T, rtol=1e-10, atol=1e-14)
    # dK closed form (Appendix: -aol (h~a - h~b)(lam~a - lam~b))
    h = oracle.groundwater_steadystate(K, nb, aol, dn, dh, Q)
    lt = np.zeros(n)
    lt[f2n - 1] = lam
    a, b = nb[:, 0] - 1, nb[:, 1] - 1
    assert np.allclose(dK, -aol[:nb.shape[0]] * (h[a] - h[b]) * (lt[a] - lt[b]), rtol=1e-9, atol=1e-12)


def test_reference_one_sided_fd_gate(pkg, oracle):
    """test/thiem.jl:14-25,68-71 on a small 2-D well problem."""
    side, n = 10.0, 21
    coords, nb, aol, _ = pkg.grid.regulargrid2d([-side, -side], [side, side], [n, n], 2.0)
    Ks = np.full(nb.shape[0], 1e-5)
    Qs = np.zeros(coords.shape[1])
    Qs[n * ((n + 1) // 2 - 1) + (n + 1) // 2 - 1] = -1e-4
    r = np.hypot(coords[0], coords[1])
    dn = np.nonzero(r >= side)[0] + 1
    dh = np.where(r >= side, 1e3, 0.0)
    node = np.nonzero((coords[1] == 0) & (coords[0] > 0.1))[0][2]
    g = lambda K: oracle.groundwater_steadystate(K, nb, aol, dn, dh, Qs)[node]
    delta = np.zeros(coords.shape[1])
    delta[node] = 1.0
    _, pb = oracle.steadystate_rrule(Ks, nb, aol, dn, dh, Qs)
    grad = pb(delta)[1]
    f0 = g(Ks)
    for i in np.argsort(-np.abs(grad))[:3]:
        x = Ks.copy()
        x[i] += 1e-8
        assert np.isclose(grad[i], (g(x) - f0) / 1e-8, rtol=1e-1)


def test_hm_chain_rule(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=12)
    rng = np.random.default_rng(13)
    logK = rng.normal(size=n) - 9.0
    Ke = oracle.hm_forward(logK, nb)
    k = np.exp(logK)
    assert np.allclose(Ke, 2 * k[nb[:, 0] - 1] * k[nb[:, 1] - 1] / (k[nb[:, 0] - 1] + k[nb[:, 1] - 1]),
                       rtol=1e-14)
    same = np.full(n, -3.0)
    assert np.allclose(oracle.hm_forward(same, nb), np.exp(-3.0), rtol=1e-15)
    w = rng.normal(size=nb.shape[0]) * 1e4
    g = oracle.hm_pullback(logK, nb, w)
    eps = 1e-6
    for c in range(0, n, 7):
        lp, lm = logK.copy(), logK.copy()
        lp[c] += eps
        lm[c] -= eps
        fd = (w @ oracle.hm_forward(lp, nb) - w @ oracle.hm_forward(lm, nb)) / (2 * eps)
        assert np.isclose(g[c], fd, rtol=1e-6, atol=1e-12)



def test_geo_chain_rule(pkg, oracle):
    """Geometric-mean map of examples/gw_steadystate_3d/inputdeck.jl:51: the
    closed form sqrt(k_a k_b), and its pullback against central differences."""
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=14)
    rng = np.random.default_rng(15)
    logK = rng.normal(size=n) - 9.0
    Ke = oracle.geo_forward(logK, nb)
    k = np.exp(logK)
    assert np.allclose(Ke, np.sqrt(k[nb[:, 0] - 1] * k[nb[:, 1] - 1]), rtol=1e-14)
    w = rng.normal(size=nb.shape[0]) * 1e4
    g = oracle.geo_pullback(logK, nb, w)
    eps = 1e-6
    for c in range(0, n, 7):
        lp, lm = logK.copy(), logK.copy()
        lp[c] += eps
        lm[c] -= eps
        fd = (w @ oracle.geo_forward(lp, nb) - w @ oracle.geo_forward(lm, nb)) / (2 * eps)
        assert np.isclose(g[c], fd, rtol=1e-6, atol=1e-12)

# ---------------------------------------------------------------- transient

def _dbe_dense(oracle, K, nb, aol, dn, dh, Q, Ss, V, u0, nsteps, dt, qscale):
    """Backward Euler in the reference's form: u_n = u_{n-1} + dt_n f(u_n) with
    f = -groundwater_residuals, solved densely as (I - dt_n f_u) u_n = u_{n-1}
    + dt_n (f(0) at t_n); dt is a scalar or one length per step."""
    nf = u0.size
    dts = np.broadcast_to(np.asarray(dt, dtype=float), (nsteps,))
    us = [u0]
    for t in range(1, nsteps + 1):
        f0 = -oracle.groundwater_residuals(np.zeros(nf), K, nb, aol, dn, dh, Q * qscale[t], Ss, V)
        fu = -oracle.groundwater_h(np.zeros(nf), K, nb, aol, dn, dh, Q, Ss, V).toarray()
        us.append(np.linalg.solve(np.eye(nf) - dts[t - 1] * fu, us[-1] + dts[t - 1] * f0))
    return np.array(us)


@pytest.mark.parametrize("dt", [0.7, [0.2, 1.5, 0.05, 0.7, 3.0]])
def test_transient_matches_dense_reference_form(pkg, oracle, dt):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=14)
    rng = np.random.default_rng(15)
    Ss = rng.uniform(0.1, 1.0, n)
    nf = n - dn.size
    u0 = rng.normal(size=nf)
    qs = np.linspace(0, 1, 6)
    st = oracle.transient(K, nb, aol, dn, dh, Q, Ss, vol, u0, 5, dt, qs)
    ref = _dbe_dense(oracle, K, nb, aol, dn, dh, Q, Ss, vol, u0, 5, dt, qs)
    assert np.allclose(st, ref, rtol=1e-11, atol=1e-11)


def test_transient_steady_limit(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, seed=16)
    nf = n - dn.size
    st = oracle.transient(K, nb, aol, dn, dh, Q, np.full(n, 1e-3), vol, np.zeros(nf), 3, 1e9,
                          np.ones(4))
    hs, hf, _ = oracle.groundwater_steadystate(K, nb, aol, dn, dh, Q, return_info=True)
    assert np.allclose(st[-1], hf, rtol=1e-8, atol=1e-10)


def test_transient_adjoint_fd(pkg, oracle):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, ns=(3, 3, 3), seed=17)
    rng = np.random.default_rng(18)
    Ss = rng.uniform(0.1, 1.0, n)
    nf = n - dn.size
    u0 = rng.normal(size=nf)
    qs = np.linspace(0, 1, 9)
    w = rng.normal(size=nf)
    N, dt = 8, 0.3

    def J(K_, Q_):
        return w @ oracle.transient(K_, nb, aol, dn, dh, Q_, Ss, vol, u0, N, dt, qs)[-1]

    st = oracle.transient(K, nb, aol, dn, dh, Q, Ss, vol, u0, N, dt, qs)
    G = np.zeros_like(st)
    G[-1] = w
    dK, dQ = oracle.transient_adjoint(K, nb, aol, dn, dh, Ss, vol, N, dt, qs, st, G)
    eps = 1e-6
    for e in range(0, nb.shape[0], 3):
        Kp, Km = K.copy(), K.copy()
        Kp[e] *= 1 + eps
        Km[e] *= 1 - eps
        assert np.isclose(dK[e], (J(Kp, Q) - J(Km, Q)) / (2 * eps * K[e]), rtol=1e-6, atol=1e-9)
    for i in range(0, n, 4):
        qp, qm = Q.copy(), Q.copy()
        qp[i] += eps
        qm[i] -= eps
        assert np.isclose(dQ[i], (J(K, qp) - J(K, qm)) / (2 * eps), rtol=1e-6, atol=1e-9)


def _transient_case(pkg, seed):
    nb, aol, vol, dn, dh, Q, K, n = _tiny(pkg, ns=(3, 3, 3), seed=seed)
    rng = np.random.default_rng(seed + 1)
    Ss = rng.uniform(0.1, 1.0, n)
    nf = n - dn.size
    u0 = rng.normal(size=nf)
    N = 7
    dts = rng.uniform(0.05, 0.8, N)
    qs = np.linspace(0, 1, N + 1)
    # objective of several states, u_0 included: J = sum_n W[n] . u_n
    W = np.zeros((N + 1, nf))
    for k in (0, 2, 5, 7):
        W[k] = rng.normal(size=nf)
    return nb, aol, vol, dn, dh, Q, K, n, Ss, nf, u0, N, dts, qs, W


def test_transient_adjoint_fd_every_tangent(pkg, oracle):
    """Reverse mode of a non-uniform step sequence for an objective of several
    states: every tangent (Ks, dirichleths, Qs, u0) against central finite
    differences of the forward run."""
    nb, aol, vol, dn, dh, Q, K, n, Ss, nf, u0, N, dts, qs, W = _transient_case(pkg, 21)

    def J(K_=K, dh_=dh, Q_=Q, u0_=u0):
        return np.sum(W * oracle.transient(K_, nb, aol, dn, dh_, Q_, Ss, vol, u0_, N, dts, qs))

    st = oracle.transient(K, nb, aol, dn, dh, Q, Ss, vol, u0, N, dts, qs)
    dK, ddh, dQ, du0 = oracle.transient_adjoint(K, nb, aol, dn, dh, Ss, vol, N, dts, qs, st, W,
                                                full=True)
    eps = 1e-6
    for e in range(0, nb.shape[0], 2):
        Kp, Km = K.copy(), K.copy()
        Kp[e] *= 1 + eps
        Km[e] *= 1 - eps
        assert np.isclose(dK[e], (J(K_=Kp) - J(K_=Km)) / (2 * eps * K[e]), rtol=1e-6, atol=1e-9)
    free = np.ones(n, dtype=bool)
    free[dn - 1] = False
    for i in range(n):
        hp, hm = dh.copy(), dh.copy()
        hp[i] += eps
        hm[i] -= eps
        fd = (J(dh_=hp) - J(dh_=hm)) / (2 * eps)
        assert np.isclose(ddh[i], fd, rtol=1e-6, atol=1e-9), i
        if free[i]:
            assert ddh[i] == 0.0
        qp, qm = Q.copy(), Q.copy()
        qp[i] += eps
        qm[i] -= eps
        assert np.isclose(dQ[i], (J(Q_=qp) - J(Q_=qm)) / (2 * eps), rtol=1e-6, atol=1e-9)
    for j in range(nf):
        up, um = u0.copy(), u0.copy()
        up[j] += eps
        um[j] -= eps
        assert np.isclose(du0[j], (J(u0_=up) - J(u0_=um)) / (2 * eps), rtol=1e-6, atol=1e-9)
    assert np.any(ddh != 0.0)


def test_transient_ddh_dense_f_gw_p(pkg, oracle):
    """The dirichleths tangent as the reference's reverse sweep forms it: the
    dense adjoint recursion with the parameter block of f_gw_p,
    -groundwater_dirichleths (examples/gw_transient_3d/inputdeck.jl:86-91),
    built column by column from groundwater_residuals (affine in dh)."""
    nb, aol, vol, dn, dh, Q, K, n, Ss, nf, u0, N, dts, qs, W = _transient_case(pkg, 23)
    st = oracle.transient(K, nb, aol, dn, dh, Q, Ss, vol, u0, N, dts, qs)
    _, ddh, _, du0 = oracle.transient_adjoint(K, nb, aol, dn, dh, Ss, vol, N, dts, qs, st, W,
                                              full=True)
    r0 = oracle.groundwater_residuals(np.zeros(nf), K, nb, aol, dn, dh, Q, Ss, vol)
    f_p = np.empty((nf, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        f_p[:, i] = -(oracle.groundwater_residuals(np.zeros(nf), K, nb, aol, dn, dh + e, Q, Ss,
                                                   vol) - r0)
    f_u = -oracle.groundwater_h(np.zeros(nf), K, nb, aol, dn, dh, Q, Ss, vol).toarray()
    lam = np.zeros(nf)
    g = np.zeros(n)
    for t in range(N, 0, -1):
        # u_t = u_{t-1} + dt f(u_t): (I - dt f_u)' lam_t = W_t + lam_{t+1}
        lam = np.linalg.solve((np.eye(nf) - dts[t - 1] * f_u).T, W[t] + lam)
        g += dts[t - 1] * f_p.T @ lam
    assert np.allclose(ddh, g, rtol=1e-9, atol=1e-12)
    assert np.allclose(du0, W[0] + lam, rtol=1e-9, atol=1e-12)


def test_theis_drawdown(pkg, oracle):
    """test/theis.jl:27-102 (atol 0.1 m) with 200 fixed backward-Euler steps
    (geometric) in place of the adaptive DiffEq integrator."""
    side, thick, n = 50.0, 10.0, 101
    coords, nb, aol, vol = pkg.grid.regulargrid2d([-side, -side], [side, side], [n, n], thick)
    k, Q, Ss = 1e-5, 1e-3, 0.1
    T, S = thick * k, Ss * thick
    Ks = np.full(nb.shape[0], k)
    Qs = np.zeros(coords.shape[1])
    Qs[n * ((n + 1) // 2 - 1) + (n + 1) // 2 - 1] = -Q
    edge = (np.abs(coords[0]) == side) | (np.abs(coords[1]) == side)
    dn = np.nonzero(edge)[0] + 1
    dh = np.where(edge, 1e3, 0.0)
    nf = coords.shape[1] - dn.size
    tfinal = 86400.0 * 10
    # fixed steps: integrate in 100 equal steps
    st = oracle.transient(Ks, nb, aol, dn, dh, Qs, np.full(coords.shape[1], Ss), vol,
                          np.full(nf, 1e3), 100, tfinal / 100, np.ones(101), reltol=1e-12)
    isfree, n2f, f2n = oracle.getfreenodes(coords.shape[1], dn)
    h = dh.copy()
    h[f2n - 1] = st[-1]
    good = np.nonzero((coords[1] == 0) & (coords[0] > 0.1) & (coords[0] <= side / 2))[0]
    rs = coords[0, good]

    def W(u):
        if u <= 1:
            return (-np.log(u) - 0.57721566 + 0.99999193 * u - 0.24991055 * u ** 2
                    + 0.05519968 * u ** 3 - 0.00976004 * u ** 4 + 0.00107857 * u ** 5)
        return (u ** 2 + 2.334733 * u + 0.250621) / (u ** 2 + 3.330657 * u + 1.681534) * np.exp(-u) / u

    theis = np.array([Q * W(r ** 2 * S / (4 * T * tfinal)) / (4 * np.pi * T) for r in rs])
    assert np.allclose(theis, 1e3 - h[good], atol=1e-1)


# ---------------------------------------------------------------- multi-core CPU baseline

def test_cpu_amg_baseline_matches_oracle(pkg, oracle):
    """The multi-core SA-AMG-PCG baseline (oracle/amg_cpu.cpp) evaluates the
    same objective, heads and logK gradient as the scalar oracle."""
    from oracle import amg_cpu
    w = pkg.workloads.steady_2m(sigma=2.0, ns=(40, 36, 20), maxs=(20.0, 18.0, 10.0))
    rng = np.random.default_rng(5)
    hobs = rng.normal(scale=1e-2, size=w.obs.size)
    cpu = amg_cpu.CpuAmg(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes, threads=4)
    assert cpu.levels >= 3
    J, g, info, h = cpu.eval(w.logK, w.dirichleths, w.Qs, w.obs, hobs, reltol=1e-13,
                             want_h=True)
    K = oracle.hm_forward(w.logK, w.neighbors)
    h_ref, pb = oracle.steadystate_rrule(K, w.neighbors, w.areasoverlengths, w.dirichletnodes,
                                         w.dirichleths, w.Qs, reltol=1e-14)
    delta = np.zeros(w.n)
    delta[w.obs - 1] = 2.0 * (h_ref[w.obs - 1] - hobs)
    g_ref = oracle.hm_pullback(w.logK, w.neighbors, pb(delta)[1])
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)  # noqa: E731
    assert rel(h, h_ref) < 1e-11
    assert np.isclose(J, np.sum((h_ref[w.obs - 1] - hobs) ** 2), rtol=1e-10)
    assert rel(g, g_ref) < 1e-9
    assert 0 < info[0] < 60 and 0 < info[1] < 60
    cpu.close()
