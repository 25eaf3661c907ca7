"""Two-phase IMPES on the GPU (dpf_twophase_tpfa / _upstream / dpf_solvetwophase)
against the oracle, the textbook computation and the Buckley-Leverett gate of
the reference's test/twophase.jl."""
import numpy as np
import pytest

from test_oracle_twophase import (FLUID, _random_box, _to_grid_order, buckley_leverett_column,
                                  quarter_five_spot)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tpo():
    from oracle import twophase_oracle
    return twophase_oracle


@pytest.fixture(scope="module")
def tp(pkg):
    import importlib
    return importlib.import_module(pkg.__name__ + ".twophase")


def _say(*a):
    print("[twophase]", *a, flush=True)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300)


def test_upstream_matches_oracle_bitwise(pkg, tp, tpo):
    """Same P and Vn in: the GPU sub-steps reproduce the oracle's arithmetic
    (no contraction, same order) bit for bit."""
    _say("upstream")
    nb, aol, vol, K, Q, S, n, _ = _random_box(pkg, ns=(12, 10, 8))
    dn = np.array([n // 3], dtype=np.int64)
    P, Vn = tpo.tpfa(K, np.zeros(n), dn, Q, aol, nb)
    for T in (0.3, 2.0):
        So, nts = tpo.upstream(S, FLUID, Q, T, P, Vn, nb, vol)
        Sg = tp.upstream(S, FLUID, Q, T, P, Vn, nb, vol, dn, aol)
        assert tp.upstream.last_substeps == nts > 0
        assert np.array_equal(Sg, So)


def test_tpfa_matches_oracle(pkg, tp, tpo):
    _say("tpfa")
    nb, aol, vol, K, Q, S, n, _ = _random_box(pkg, ns=(20, 16, 10))
    dn = np.array([1, n // 3, n - 5], dtype=np.int64)
    dh = np.zeros(n)
    dh[dn - 1] = [1.0, 0.5, -0.25]
    P, Vn = tp.tpfa(K, dh, dn, Q, aol, nb)
    Po, Vo = tpo.tpfa(K, dh, dn, Q, aol, nb)
    assert rel(P, Po) < 1e-10 and rel(Vn, Vo) < 1e-9
    assert np.array_equal(P[dn - 1], dh[dn - 1])


def test_quarter_five_spot(pkg, tp, tpo):
    """The reference test's 16x16 case (no Dirichlet cell): against the textbook
    computation per step and the reference's total_norm gate."""
    _say("quarter five-spot")
    N = 16
    nb, aol, vol, q, n = quarter_five_spot(pkg, N)
    Ps, Ss = tp.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), np.zeros(n), [], q, vol,
                              aol, FLUID, 0.7 / 25, nb, 5, everyStep=True)
    Pt, St = tpo.textbook_quarter_five_spot(N, 5)
    for k in range(5):
        pt, st = _to_grid_order(Pt[k], N), _to_grid_order(St[k], N)
        assert rel(Ps[k], pt) < 1e-9
        assert np.abs(Ss[k] - st).max() < 1e-9
    P, S = tp.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), np.zeros(n), [], q, vol, aol,
                            FLUID, 0.7 / 25, nb, 5)
    assert np.array_equal(P, Ps[-1]) and np.array_equal(S, Ss[-1])
    total = np.linalg.norm(P - _to_grid_order(Pt[-1], N)) + \
        np.linalg.norm(S - _to_grid_order(St[-1], N))
    assert total < 1e-5


def test_buckley_leverett(pkg, tp, tpo):
    _say("Buckley-Leverett")
    coords, nb, aol, vol, dn, Q, n = buckley_leverett_column(pkg)
    dh = np.zeros(n)
    P, S = tp.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), dh, dn, Q, vol, aol, FLUID,
                            0.7 / 25, nb, 25)
    Po, So = tpo.solvetwophase(np.zeros(n), np.zeros(n), np.ones(n), dh, dn, Q, vol, aol, FLUID,
                               0.7 / 25, nb, 25)
    assert rel(P, Po) < 1e-10 and np.abs(S - So).max() < 1e-10
    S_BL = tpo.buckley_leverett(coords[0], Q[0] * 0.7)
    assert abs(np.mean(S_BL - S)) < 1e-2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_five_spot_3d_vs_oracle(pkg, tp, tpo, precision):
    """A heterogeneous 3-D five-spot with a clamped fluid (swc, sor > 0) over
    several steps; the mixed pressure solve within its tolerance."""
    _say("3-D five-spot", precision)
    rng = np.random.default_rng(8)
    ns = (24, 20, 6)
    coords, nb, aol, vol = pkg.grid.regulargrid3d((0, 0, 0), (23.0, 19.0, 5.0), ns)
    n = coords.shape[1]
    K = np.exp(0.7 * rng.normal(size=n))
    Q = np.zeros(n)
    Q[[0, 5, 2000]] = [2.0, 1.0, 1.5]
    dn = np.array([n - i for i in range(6)], dtype=np.int64)
    dh = np.zeros(n)
    fluid = (1.0, 5.0, 0.1, 0.15)
    S0 = np.full(n, 0.1)
    kw = dict(precision=precision, reltol=1e-12)
    st = {}
    P, S = tp.solvetwophase(np.zeros(n), S0, K, dh, dn, Q, vol, aol, fluid, 3.0, nb, 4,
                            stats=st, **kw)
    Po, So = tpo.solvetwophase(np.zeros(n), S0, K, dh, dn, Q, vol, aol, fluid, 3.0, nb, 4)
    assert st["substeps"] > 4
    assert rel(P, Po) < (1e-10 if precision == "fp64" else 1e-8)
    assert np.abs(S - So).max() < (1e-9 if precision == "fp64" else 1e-6)
    assert S.min() >= 0.1 and S.max() <= 0.85 and S.max() > 0.5


def test_nt_zero_and_errors(pkg, tp):
    _say("edge cases")
    nb, aol, vol, q, n = quarter_five_spot(pkg, 8)
    h0 = np.linspace(0, 1, n)
    S0 = np.full(n, 0.2)
    P, S = tp.solvetwophase(h0, S0, np.ones(n), np.zeros(n), [], q, vol, aol, FLUID, 0.1, nb, 0)
    assert np.array_equal(P, h0) and np.array_equal(S, S0)
    # no flow at all: no sub-step, S unchanged
    S1 = tp.upstream(S0, FLUID, np.zeros(n), 0.5, np.zeros(n), np.zeros(nb.shape[0]), nb, vol)
    assert tp.upstream.last_substeps == 0 and np.array_equal(S1, S0)
    # a cell of zero pore volume with inflow: the CFL step is zero
    P, Vn = tp.tpfa(np.ones(n), np.zeros(n), [], q, aol, nb)
    v0 = vol.copy()
    v0[3] = 0.0
    with pytest.raises(pkg._lib.DpfError):
        tp.upstream(S0, FLUID, q, 0.5, P, Vn, nb, v0)


def test_rrule_solvepressure(pkg, tp, oracle):
    """src/twophase.jl:59-80: the transmissivity system with edge
    permeabilities (tx = 2 aol); against the oracle's groundwater rrule at
    2 K_e, whose K-tangent is halved by the chain rule, and, without a
    Dirichlet cell, against central differences of the mean-zero pressure."""
    _say("rrule_solvepressure")
    nb, aol, vol, K, Q, S, n, rng = _random_box(pkg, ns=(10, 8, 6))
    Ke = 0.5 * (K[nb[:, 0] - 1] + K[nb[:, 1] - 1])
    dn = np.array([2, n - 3], dtype=np.int64)
    dh = np.zeros(n)
    dh[dn - 1] = [1.0, -1.0]
    h, pb = tp.rrule_solvepressure(Ke, nb, aol, dn, dh, Q)
    ho, pbo = oracle.steadystate_rrule(2.0 * Ke, nb, aol, dn, dh, Q, reltol=1e-14)
    assert rel(h, ho) < 1e-10
    delta = rng.normal(size=n)
    g, go = pb(delta), pbo(delta)
    assert rel(g[1], 2.0 * go[1]) < 1e-8 and rel(g[5], go[5]) < 1e-8 and rel(g[6], go[6]) < 1e-8
    # singular system (no Dirichlet cell, balanced sources): mean-zero pressure
    Qb = Q - Q.mean()
    h0, pb0 = tp.rrule_solvepressure(Ke, nb, aol, [], np.zeros(n), Qb)
    assert abs(h0.mean()) < 1e-12 * np.abs(h0).max()
    d = rng.normal(size=n)
    gK = pb0(d)[1]
    for e in np.argsort(-np.abs(gK))[:3]:
        eps = 1e-6 * Ke[e]
        Kp, Km = Ke.copy(), Ke.copy()
        Kp[e] += eps
        Km[e] -= eps
        fd = (d @ tp.solvepressure(Kp, nb, aol, [], np.zeros(n), Qb) -
              d @ tp.solvepressure(Km, nb, aol, [], np.zeros(n), Qb)) / (2 * eps)
        assert abs(fd - gK[e]) < 1e-5 * abs(gK[e])


@pytest.mark.parametrize("pinned", [False, True])
def test_solvetwophase_gradient_vs_oracle(pkg, tp, tpo, pinned):
    """Reverse mode of the IMPES loop against the oracle's (itself pinned by
    central differences), J = c_P . P + c_S . S after three steps."""
    _say("twophase gradient", pinned)
    nb, aol, vol, K, Q, S, n, rng = _random_box(pkg, ns=(12, 10, 6))
    if pinned:
        dn = np.array([], dtype=np.int64)
        Q = Q - Q.mean()
    else:
        dn = np.arange(6, n + 1, 6, dtype=np.int64)
    dh = np.zeros(n)
    fluid = (1.0, 3.0, 0.05, 0.1)
    S0 = np.full(n, 0.05)
    cP, cS = rng.normal(size=n), rng.normal(size=n)
    args = (np.zeros(n), S0, K, dh, dn, Q, vol, aol, fluid, 0.8, nb, 3)
    (P, Sg), pb = tp.rrule_solvetwophase(*args)
    dh0, dS0, dK = pb(cP, cS)
    Po, So, dKo, dS0o = tpo.solvetwophase_grad(*args, cP, cS)
    assert rel(P, Po) < 1e-10 and np.abs(Sg - So).max() < 1e-10
    assert rel(dK, dKo) < 1e-8 and rel(dS0, dS0o) < 1e-8
    assert not dh0.any()


def test_findgrad_quarter_five_spot(pkg, tp):
    """test/twophase.jl:192-205: the gradient of P at the centre cell after
    two steps, against one-sided differences with delta 1e-8 and rtol 0.1 at
    the three largest entries, as the reference checks it."""
    _say("FindGrad")
    N = 16
    nb, aol, vol, q, n = quarter_five_spot(pkg, N)
    crit = int(np.ceil(N * N / 2)) - 1
    K = np.ones(n)
    args = [np.zeros(n), np.zeros(n), K, np.zeros(n), [], q, vol, aol, FLUID, 0.7 / 25, nb, 2]
    (P, S), pb = tp.rrule_solvetwophase(*args)
    seed = np.zeros(n)
    seed[crit] = 1.0
    dK = pb(seed, None)[2]
    f0 = P[crit]
    for i in np.argsort(-np.abs(dK))[:3]:
        Kp = K.copy()
        Kp[i] += 1e-8
        a = list(args)
        a[2] = Kp
        fd = (tp.solvetwophase(*a)[0][crit] - f0) / 1e-8
        assert abs(fd - dK[i]) <= 0.1 * abs(dK[i])
