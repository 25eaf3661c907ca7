"""Advection-dispersion on the GPU (dpf_transport_*) against the oracle, the
reference test's physics gates and its gradient check."""
import numpy as np
import pytest

from test_oracle_transport import column

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def to():
    from oracle import transport_oracle
    return transport_oracle


@pytest.fixture(scope="module")
def tr(pkg):
    import importlib
    return importlib.import_module(pkg.__name__ + ".transport")


def _say(*a):
    print("[transport]", *a, flush=True)


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - b) / max(np.linalg.norm(b), 1e-300)


def _box2d(pkg, seed=2):
    rng = np.random.default_rng(seed)
    coords, nb, aol, vol = pkg.grid.regulargrid2d((0, 0), (4.0, 3.0), (30, 24), 1.0)
    coords = np.vstack([coords, np.zeros((1, coords.shape[1]))])
    n, ne = coords.shape[1], nb.shape[0]
    vx, vy, vz = 0.5 + rng.random(ne), 0.3 * rng.normal(size=ne), np.zeros(ne)
    Ds = 0.02 + 0.02 * rng.random(ne)
    dn = np.array([1, n], dtype=np.int64)
    ud = np.zeros(n)
    ud[n - 1] = 0.2
    Q = np.zeros(n)
    Q[n // 2] = 0.3
    u0 = np.exp(-((coords[0] - 1.5) ** 2 + (coords[1] - 1.5) ** 2))
    u0[[0, n - 1]] = ud[[0, n - 1]]
    return coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q, rng


@pytest.mark.parametrize("up", [True, False])
def test_residuals_vs_oracle(pkg, tr, to, up):
    _say("residuals", up)
    coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q, rng = _box2d(pkg)
    u = rng.normal(size=n)
    r = tr.transport_residuals(u, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, up)
    ro_ = to.transport_residuals(u, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, up)
    assert rel(r, ro_) < 1e-13


def test_steps_and_gradient_vs_oracle(pkg, tr, to):
    _say("steps + gradient")
    coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q, rng = _box2d(pkg)
    u, pb = tr.rrule_transport_steps(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0,
                                     0.7, 40, reltol=1e-13)
    g = rng.normal(size=n)
    uo, go = to.transport_adjoint(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 0.7,
                                  40, g)
    assert rel(u, uo) < 1e-11
    gr = pb(g)
    for k in ("vxs", "vys", "Ds", "Qs"):
        assert rel(gr[k], go[k]) < 1e-9, k
    allu = tr.transport_steps(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 0.7, 40,
                              return_all=True, reltol=1e-13)
    assert allu.shape == (n, 41) and np.array_equal(allu[:, -1], u)


def test_reference_column(pkg, tr):
    """test/transport.jl:12-70 with 1000 fixed steps: the physics gates and the
    one-sided gradient check at the three largest entries."""
    _say("reference column")
    coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q = column(pkg)
    args = [u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 1.0, 1000]
    u1, pb = tr.rrule_transport_steps(*args, reltol=1e-12)
    x = coords[0]
    assert np.isclose((u1 * vol).sum(), (u0 * vol).sum(), rtol=1e-2)
    assert np.isclose((u1 * x * vol).sum(), 2.0, rtol=1e-2)
    assert np.isclose(((x - 2) ** 2 * u1 * vol).sum(), 0.09, rtol=1e-1)
    node = int(np.argmin((coords[0] - 1.5) ** 2 + (coords[1] - 1.0) ** 2))
    seed = np.zeros(n)
    seed[node] = 1.0
    grad = pb(seed)["vxs"]
    for i in np.argsort(-np.abs(grad))[:3]:
        vp = vx.copy()
        vp[i] += 1e-8
        a = list(args)
        a[1] = vp
        fd = (tr.transport_steps(*a, reltol=1e-12)[node] - u1[node]) / 1e-8
        assert np.isclose(grad[i], fd, atol=1e-4, rtol=1e-3)


def test_parallel_edges_and_unit_step(pkg, tr, to):
    """Edges listed twice with their aol split in half give the same system
    (both copies add into one matrix slot); a step of exactly 1 is refused
    when there are Dirichlet cells (their backward-Euler row is singular)."""
    _say("parallel edges")
    coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q, rng = _box2d(pkg, seed=5)
    ne = nb.shape[0]
    dup = rng.choice(ne, ne // 3, replace=False)
    nb2 = np.vstack([nb, nb[dup][:, ::-1]])
    aol2 = np.concatenate([aol[:ne], aol[dup] * 0.5])
    aol2[dup] *= 0.5
    ext = lambda a: np.concatenate([a, a[dup]])  # noqa: E731
    args2 = (ext(vx), ext(vy), ext(vz), ext(Ds), nb2, aol2, dn, ud, Q, vol, coords, 0.0, 0.7, 20)
    u2 = tr.transport_steps(u0, *args2, reltol=1e-13)
    uo2 = to.transport_steps(u0, *args2)
    u1 = tr.transport_steps(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 0.7, 20,
                            reltol=1e-13)
    assert rel(u2, uo2) < 1e-11 and rel(u2, u1) < 1e-11
    g = rng.normal(size=n)
    _, pb = tr.rrule_transport_steps(u0, *args2, reltol=1e-13)
    _, go = to.transport_adjoint(u0, *args2, g)
    gr = pb(g)
    for k in ("vxs", "vys", "Ds", "Qs"):
        assert rel(gr[k], go[k]) < 1e-9, k
    with pytest.raises(Exception, match="singular"):
        tr.transport_steps(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 3.0, 3)
