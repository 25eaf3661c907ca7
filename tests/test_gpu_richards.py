"""Steady Richards on the GPU (dpf_richards_residuals / _steady / _pullback)
against the oracle, which finite differences and the saturated groundwater
limit pin."""
import numpy as np
import pytest

from test_oracle_richards import richards_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ro():
    from oracle import richards_oracle
    return richards_oracle


@pytest.fixture(scope="module")
def api(pkg):
    import importlib
    return importlib.import_module(pkg.__name__ + ".api")


@pytest.fixture(scope="module")
def rich(pkg):
    import importlib
    return importlib.import_module(pkg.__name__ + ".richards")


def _say(*a):
    print("[richards]", *a, flush=True)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300)


def test_residuals_vs_oracle(pkg, rich, ro):
    _say("residuals")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(12, 10, 6), seed=7)
    rng = np.random.default_rng(8)
    dn = np.array([3, n // 2, n], dtype=np.int64)
    dpsi[dn - 1] = [-0.4, 0.3, 0.0]
    alphas = alphas * (0.5 + rng.random(alphas.size))
    Ns = 1.1 + rng.random(Ns.size)
    psi = -2.0 + 3.0 * rng.random(n - dn.size)
    Ss, V = 0.5 + rng.random(n), 0.5 + rng.random(n)
    r = rich.richards_residuals(psi, Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs, Ss, V)
    ro_r = ro.richards_residuals(psi, Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs, Ss, V)
    assert rel(r, ro_r) < 1e-13


@pytest.mark.parametrize("Q", [1e-6, -3e-6])
def test_example_problem_vs_oracle(pkg, rich, ro, Q):
    """examples/richards_steadystate_3d/inputdeck.jl's 30 x 30 x 3 problem, with
    the example's injection and with extraction (unsaturated heads)."""
    _say("example", Q)
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(30, 30, 3), Q=Q)
    st = {}
    psi = rich.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs,
                                    stats=st, reltol=1e-12)
    pso = ro.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    assert rel(psi, pso) < 1e-9
    assert st["converged"] == 1 and 0 < st["outer"] < 40 and st["relres"] <= 1e-12
    assert (psi.min() < -1e-3) == (Q < 0)


def test_unsaturated_box_and_gradients(pkg, rich, ro):
    """Strongly unsaturated (Dirichlet psi = -0.2, varied van Genuchten
    parameters): heads and every rrule tangent against the oracle's."""
    _say("gradients")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(12, 10, 6), seed=4)
    dpsi[n - 1] = -0.2
    rng = np.random.default_rng(5)
    alphas = alphas * (0.8 + 0.4 * rng.random(alphas.size))
    Ns = Ns * (0.9 + 0.2 * rng.random(Ns.size))
    psi, pb = rich.rrule_richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas,
                                              Ns, Qs, reltol=1e-12)
    pso, pbo = ro.richards_rrule(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    assert rel(psi, pso) < 1e-9 and psi.min() < -0.1
    w = rng.normal(size=n)
    g, go = pb(w), pbo(w)
    assert rel(g[2], go["Ks"]) < 1e-7
    assert rel(g[6], go["dirichletpsis"]) < 1e-7
    assert rel(g[8], go["alphas"]) < 1e-7
    assert rel(g[9], go["Ns"]) < 1e-7
    assert rel(g[10], go["Qs"]) < 1e-7
    assert not g[1].any()


def test_saturated_limit_matches_groundwater(pkg, rich, api):
    """kr = 1 everywhere: psi = h - z with h from the groundwater solver."""
    _say("saturated")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(12, 10, 6),
                                                                    psiD=50.0, Q=1e-9)
    psi = rich.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs,
                                    reltol=1e-12)
    z = coords[2]
    h = api.groundwater_steadystate(Ks, nb, aol, dn, dpsi + z, Qs, reltol=1e-13)
    assert psi.min() > 0 and np.abs(psi - (h - z)).max() < 1e-8 * np.abs(h).max()


def test_errors(pkg, rich):
    _say("errors")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(4, 3, 3))
    with pytest.raises(pkg._lib.DpfError):
        rich.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    with pytest.raises(ValueError):
        rich.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords[:2], alphas, Ns, Qs)


def test_parallel_edges(pkg, rich, ro):
    """Edges listed twice (reversed) with their aol split in half give the
    same heads; both copies add into one slot of the Newton matrix and of its
    transpose."""
    _say("parallel edges")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(10, 8, 5), seed=9)
    dpsi[n - 1] = -0.2
    rng = np.random.default_rng(10)
    ne = nb.shape[0]
    dup = rng.choice(ne, ne // 3, replace=False)
    nb2 = np.vstack([nb, nb[dup][:, ::-1]])
    aol2 = np.concatenate([aol[:ne], aol[dup] * 0.5])
    aol2[dup] *= 0.5
    ext = lambda a: np.concatenate([a, a[dup]])  # noqa: E731
    args2 = (ext(Ks), nb2, aol2, dn, dpsi, coords, ext(alphas), ext(Ns), Qs)
    psi1 = rich.richards_steadystate(np.zeros(n), Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs,
                                     reltol=1e-12)
    psi2, pb = rich.rrule_richards_steadystate(np.zeros(n), *args2, reltol=1e-12)
    pso, pbo = ro.richards_rrule(np.zeros(n), *args2)
    assert rel(psi2, pso) < 1e-9 and rel(psi2, psi1) < 1e-9 and psi2.min() < -0.1
    w = rng.normal(size=n)
    g, go = pb(w), pbo(w)
    assert rel(g[2], go["Ks"]) < 1e-7
    assert rel(g[10], go["Qs"]) < 1e-7


def test_fp32_preconditioner(pkg, rich, ro):
    """precision mixed_pc: the V-cycle preconditioner in FP32; heads and
    tangents agree with the oracle as closely as the FP64 path's."""
    _say("fp32 preconditioner")
    coords, nb, aol, n, Ks, Qs, dn, dpsi, alphas, Ns = richards_box(pkg, ns=(12, 10, 6), seed=4)
    dpsi[n - 1] = -0.2
    args = (Ks, nb, aol, dn, dpsi, coords, alphas, Ns, Qs)
    st = {}
    psi, pb = rich.rrule_richards_steadystate(np.zeros(n), *args, stats=st, reltol=1e-12,
                                              precision="mixed_pc")
    pso, pbo = ro.richards_rrule(np.zeros(n), *args)
    assert st["converged"] == 1 and rel(psi, pso) < 1e-9
    w = np.random.default_rng(6).normal(size=n)
    g, go = pb(w), pbo(w)
    for k, key in ((2, "Ks"), (6, "dirichletpsis"), (8, "alphas"), (9, "Ns"), (10, "Qs")):
        assert rel(g[k], go[key]) < 1e-7, key
