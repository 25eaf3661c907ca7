"""Pins the steady Richards oracle (oracle/richards_oracle.py): the van
Genuchten closed forms, finite differences of the residual and of the solve,
and the saturated limit, where Richards' equation is the groundwater equation
for h = psi + z."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def ro():
    from oracle import richards_oracle
    return richards_oracle


def richards_box(pkg, ns=(6, 5, 3), seed=0, psiD=0.0, Q=1e-6, alpha=0.375, N=1.25):
    """The set-up of examples/richards_steadystate_3d/inputdeck.jl:10-61 on a
    small grid: injection at the first cell, psi fixed at the last, per-edge
    van Genuchten parameters, geometric-mean K from a random logK."""
    rng = np.random.default_rng(seed)
    maxs = (50.0 * (ns[0] - 1) / 29, 50.0 * (ns[1] - 1) / 29, 1.0 * (ns[2] - 1) / 2)
    coords, nb, aol, vol = pkg.grid.regulargrid3d((0, 0, 0), maxs, ns)
    n = coords.shape[1]
    logK = -9.0 + rng.normal(size=n)
    Ks = np.exp(0.5 * (logK[nb[:, 0] - 1] + logK[nb[:, 1] - 1]))
    Qs = np.zeros(n)
    Qs[0] = Q
    dn = np.array([n], dtype=np.int64)
    dpsi = np.zeros(n)
    dpsi[n - 1] = psiD
    alphas = np.full(nb.shape[0], alpha)
    Ns = np.full(nb.shape[0], N)
    return coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns


def test_kr_closed_forms(ro):
    psi = np.array([0.0, 2.0, -0.5, -3.0])
    assert np.array_equal(ro.kr(psi[:2], 0.4, 1.7), [1.0, 1.0])
    # N = 2: m = 1/2, kr = (1 - s / sqrt(1 + s^2))^2 / (1 + s^2)^(1/4)
    s = 0.4 * np.abs(psi[2:])
    closed = (1 - s / np.sqrt(1 + s * s)) ** 2 / (1 + s * s) ** 0.25
    assert np.allclose(ro.kr(psi[2:], 0.4, 2.0), closed, rtol=1e-14)


def test_kr_partials_fd(ro):
    rng = np.random.default_rng(1)
    psi = -rng.random(20) * 5 - 0.01
    al = 0.1 + rng.random(20)
    N = 1.1 + rng.random(20)
    k, dp, da, dN = ro.kr_partials(psi, al, N)
    h = 1e-6
    for arr, d in ((0, dp), (1, da), (2, dN)):
        args_p = [psi.copy(), al.copy(), N.copy()]
        args_m = [psi.copy(), al.copy(), N.copy()]
        args_p[arr] += h
        args_m[arr] -= h
        fd = (ro.kr(*args_p) - ro.kr(*args_m)) / (2 * h)
        assert np.allclose(d, fd, rtol=1e-6, atol=1e-12)
    assert not any(x.any() for x in ro.kr_partials(np.array([0.0, 1.0]), 0.4, 1.5)[1:])


def test_jacobian_fd(pkg, ro):
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, seed=2)
    dpsi[n - 1] = -0.3
    rng = np.random.default_rng(3)
    psi = -2.0 + 3.0 * rng.random(n - 1)  # both signs
    one = np.ones(n)
    args = (Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs, one, 2.0 * one)
    J = ro.richards_jacobian(psi, *args).toarray()
    h = 1e-6
    for j in rng.choice(n - 1, 8, replace=False):
        e = np.zeros(n - 1)
        e[j] = h
        fd = (ro.richards_residuals(psi + e, *args) - ro.richards_residuals(psi - e, *args)) / (2 * h)
        assert np.allclose(J[:, j], fd, rtol=1e-6, atol=1e-9 * np.abs(J).max())


def test_saturated_limit_is_groundwater(pkg, ro, oracle):
    """psi >= 0 everywhere makes kr = 1: the solution is h - z with h the
    groundwater heads for the same K_e, sources and Dirichlet heads psi_D + z_D."""
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, psiD=50.0, Q=1e-9)
    psi = ro.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    assert psi.min() > 0
    z = coords[2]
    h = oracle.groundwater_steadystate(Ks, nb, aol, dn, dpsi + z, Qs, reltol=1e-14)
    assert np.allclose(psi, h - z, rtol=0, atol=1e-9 * np.abs(h).max())


def test_example_problem_solves(pkg, ro):
    """The example's 30 x 30 x 3 grid (alpha 0.375, N 1.25, 1e-6 m^3/s at the
    first cell, psi = 0 at the last): Newton converges (the injection keeps
    the domain saturated); with extraction instead, unsaturated heads develop
    and Newton still converges."""
    for Q in (1e-6, -3e-6):
        coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(30, 30, 3), Q=Q)
        psi, info = ro.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas,
                                            Ns, Qs, return_info=True)
        one = np.ones(n)
        r = ro.richards_residuals(psi[:-1], Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs, one,
                                  one)
        assert np.abs(r).max() <= 1e-13 and info["iterations"] < 40
        assert (psi.min() < -1e-3) == (Q < 0)


def test_rrule_fd(pkg, ro):
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, seed=4)
    dpsi[n - 1] = -0.2
    rng = np.random.default_rng(5)
    alphas = alphas * (0.8 + 0.4 * rng.random(alphas.size))
    Ns = Ns * (0.9 + 0.2 * rng.random(Ns.size))
    w = rng.normal(size=n)
    base = dict(Ks=Ks, dirichletpsis=dpsi, alphas=alphas, Ns=Ns, Qs=Qs)

    def J(**kw):
        a = dict(base, **kw)
        return w @ ro.richards_steadystate(np.zeros(n), a["Ks"], nb, aol, dn, a["dirichletpsis"],
                                           coords, a["alphas"], a["Ns"], a["Qs"])
    psi, pb = ro.richards_rrule(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    g = pb(w)
    for name in ("Ks", "alphas", "Ns", "Qs", "dirichletpsis"):
        v = base[name]
        idx = np.argsort(-np.abs(g[name]))[:3]
        for i in idx:
            eps = 1e-6 * max(abs(v[i]), 1e-3 if name != "Ks" else 0.0) if name != "Qs" else 1e-9
            vp, vm = v.copy(), v.copy()
            vp[i] += eps
            vm[i] -= eps
            fd = (J(**{name: vp}) - J(**{name: vm})) / (2 * eps)
            assert abs(fd - g[name][i]) <= 1e-5 * abs(g[name][i]) + 1e-12, (name, i, fd, g[name][i])
