"""Pins the advection-dispersion oracle (oracle/transport_oracle.py): the
residual's affinity and a hand-computed case, the physics gates of the
reference's test/transport.jl, and finite differences of the adjoint."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def to():
    from oracle import transport_oracle
    return transport_oracle


def column(pkg, ns=4000, L=4.0, c=1.0, sigma=0.3, D=1e-9):
    """test/transport.jl:12-29: a 1-D column as a 3-D mesh, a Gaussian pulse
    at x = 1, velocity c on every edge, Dirichlet u = 0 at both ends."""
    coords, nb, aol, vol = pkg.grid.regulargrid1d(0.0, L, ns, 1.0, 1.0)
    coords = np.vstack([coords, np.zeros((2, coords.shape[1]))])
    n = coords.shape[1]
    u0 = np.exp(-((coords[0] - 1.0) ** 2) / (2 * sigma ** 2)) / (sigma * np.sqrt(2 * np.pi))
    ne = nb.shape[0]
    vx, vy, vz = np.full(ne, c), np.zeros(ne), np.zeros(ne)
    Ds = np.full(ne, D)
    dn = np.array([1, n], dtype=np.int64)
    return coords, nb, aol[:ne], vol, n, u0, vx, vy, vz, Ds, dn, np.zeros(n), np.zeros(n)


def test_residual_affine_and_hand_case(pkg, to):
    rng = np.random.default_rng(0)
    coords, nb, aol, vol = pkg.grid.regulargrid3d((0, 0, 0), (3.0, 2.0, 1.0), (4, 3, 2))
    n, ne = coords.shape[1], nb.shape[0]
    v = [rng.normal(size=ne) for _ in range(3)]
    Ds = rng.random(ne)
    dn = np.array([2, n], dtype=np.int64)
    ud = rng.normal(size=n)
    Q = rng.normal(size=n)
    for up in (True, False):
        args = (*v, Ds, nb, aol, dn, ud, Q, vol, coords, up)
        J = to.transport_jacobian(nb, aol, dn, vol, coords, *v, Ds, up)
        u = rng.normal(size=n)
        assert np.allclose(to.transport_residuals(u, *args),
                           J @ u + to.transport_residuals(np.zeros(n), *args), rtol=1e-13,
                           atol=1e-13)
    # two cells 1 m apart, flow from cell 1 to 2 at 2 m/s, aol 3, volumes 1 and 2
    c2 = np.array([[0.0, 1.0], [0.0, 0.0], [0.0, 0.0]])
    r = to.transport_residuals(np.array([1.0, 0.25]), [2.0], [0.0], [0.0], [0.5], [[1, 2]], [3.0],
                               [], np.zeros(2), np.zeros(2), [1.0, 2.0], c2)
    # cell 2 is downstream: -(v.(x2 - x1))(u2 - u1) aol / V2 + D (u1 - u2) aol / V2
    assert np.allclose(r, [0.5 * (0.25 - 1.0) * 3.0 / 1.0,
                           -2.0 * (0.25 - 1.0) * 3.0 / 2.0 + 0.5 * (1.0 - 0.25) * 3.0 / 2.0],
                       rtol=1e-15)


def test_reference_physics_gates(pkg, to):
    """test/transport.jl:37-41 with fixed backward-Euler steps."""
    coords, nb, aol, vol, n, u0, vx, vy, vz, Ds, dn, ud, Q = column(pkg)
    u1 = to.transport_steps(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 1.0, 1000)
    x = coords[0]
    assert np.isclose((u1 * vol).sum(), (u0 * vol).sum(), rtol=1e-2)
    assert np.isclose((u0 * x * vol).sum(), 1.0, rtol=1e-2)
    assert np.isclose((u1 * x * vol).sum(), 2.0, rtol=1e-2)
    assert np.isclose(((x - 1) ** 2 * u0 * vol).sum(), 0.09, rtol=1e-2)
    assert np.isclose(((x - 2) ** 2 * u1 * vol).sum(), 0.09, rtol=1e-1)


def test_adjoint_fd(pkg, to):
    rng = np.random.default_rng(2)
    coords, nb, aol, vol = pkg.grid.regulargrid2d((0, 0), (4.0, 3.0), (9, 7), 1.0)
    coords = np.vstack([coords, np.zeros((1, coords.shape[1]))])
    n, ne = coords.shape[1], nb.shape[0]
    vx, vy, vz = 0.5 + rng.random(ne), rng.normal(size=ne) * 0.3, np.zeros(ne)
    Ds = 0.05 + 0.05 * rng.random(ne)
    dn = np.array([1, n], dtype=np.int64)
    ud = np.zeros(n)
    ud[n - 1] = 0.2
    Q = np.zeros(n)
    Q[n // 2] = 0.3
    u0 = np.exp(-((coords[0] - 1.5) ** 2 + (coords[1] - 1.5) ** 2))
    u0[[0, n - 1]] = ud[[0, n - 1]]
    g = rng.normal(size=n)
    base = dict(vxs=vx, vys=vy, vzs=vz, Ds=Ds, Qs=Q)

    def J(**kw):
        a = dict(base, **kw)
        return g @ to.transport_steps(u0, a["vxs"], a["vys"], a["vzs"], a["Ds"], nb, aol, dn, ud,
                                      a["Qs"], vol, coords, 0.0, 0.7, 40)
    _, gr = to.transport_adjoint(u0, vx, vy, vz, Ds, nb, aol, dn, ud, Q, vol, coords, 0.0, 0.7,
                                 40, g)
    for key in ("vxs", "vys", "Ds", "Qs"):
        for i in np.argsort(-np.abs(gr[key]))[:3]:
            eps = 1e-6
            vp, vm = base[key].copy(), base[key].copy()
            vp[i] += eps
            vm[i] -= eps
            fd = (J(**{key: vp}) - J(**{key: vm})) / (2 * eps)
            assert abs(fd - gr[key][i]) <= 1e-6 * abs(gr[key][i]) + 1e-12, (key, i)
