"""GPU parity of the fixed-step backward-Euler transient solve and its
checkpointed adjoint against the FP64 oracle (oracle/gw_oracle.c)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _problem(pkg, ns=(20, 18, 8), sigma=1.5):
    w = pkg.workloads.steady_2m(sigma=sigma, ns=ns, maxs=(ns[0] * 0.5, ns[1] * 0.5, ns[2] * 0.5))
    Ss = np.full(w.n, 1e-5)
    return w, Ss


@pytest.mark.parametrize("precision,ck", [("fp64", 7), ("fp64", 0), ("mixed", 5),
                                          ("mixed_pc", 4)])
def test_transient_forward_adjoint_vs_oracle(pkg, oracle, precision, ck):
    from dpfehm_jl_amd import api
    print("transient", precision, ck, flush=True)
    w, Ss = _problem(pkg)
    K = w.Ks()
    nf = w.n - w.dirichletnodes.size
    N, dt = 24, 3600.0
    qs = pkg.workloads.transient_schedule(N, dt, ramp_steps=10)
    u0 = np.zeros(nf)
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss,
            w.volumes)
    ref = oracle.transient(*args, u0, N, dt, qs, reltol=1e-14)
    st = {}
    uN, pb = api.rrule_groundwater_transient(*args, u0, N, dt, qs, checkpoint_every=ck,
                                             reltol=1e-12, precision=precision, stats=st)
    print(st, flush=True)
    tol = 1e-10 if precision == "fp64" else 1e-8
    assert rel(uN, ref[-1]) < tol
    rng = np.random.default_rng(3)
    g = np.zeros(nf)
    g[rng.choice(nf, 30, replace=False)] = rng.normal(size=30)
    G = np.zeros_like(ref)
    G[-1] = g
    dK_ref, dQ_ref = oracle.transient_adjoint(K, w.neighbors, w.areasoverlengths, w.dirichletnodes,
                                              w.dirichleths, Ss, w.volumes, N, dt, qs, ref, G,
                                              reltol=1e-14)
    dK, _, dQ, _ = pb(g)
    assert rel(dK, dK_ref) < 1e-7
    assert rel(dQ, dQ_ref) < 1e-7


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_pc"])
def test_stiff_adjoint_long_horizon(pkg, oracle, precision):
    """dt = 3600 s on a 4 m deep box with K/Ss ~ 12 m^2/s: every step damps
    the slowest mode by ~1e-4, so the adjoint falls below 1e-160 within 40
    steps of the reverse sweep and squared FP64 norms of it underflow.  The
    states reach the steady state after a few steps, so with a constant Q the
    gradient over 100 steps equals the gradient over the last 20 (earlier
    steps add terms ~1e-80 smaller), which the oracle gives."""
    from dpfehm_jl_amd import api
    print("stiff adjoint", precision, flush=True)
    w, Ss = _problem(pkg)
    K = w.Ks()
    nf = w.n - w.dirichletnodes.size
    dt = 3600.0
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss,
            w.volumes)
    rng = np.random.default_rng(4)
    g = np.zeros(nf)
    g[rng.choice(nf, 30, replace=False)] = rng.normal(size=30)
    out = {}
    for N in (20, 100):
        uN, pb = api.rrule_groundwater_transient(*args, np.zeros(nf), N, dt, np.ones(N + 1),
                                                 checkpoint_every=25, reltol=1e-12,
                                                 precision=precision)
        dK_, _, dQ_, _ = pb(g)
        out[N] = (uN, dK_, dQ_)
    ref = oracle.transient(*args, np.zeros(nf), 20, dt, np.ones(21), reltol=1e-14)
    G = np.zeros_like(ref)
    G[-1] = g
    dK_ref, dQ_ref = oracle.transient_adjoint(K, w.neighbors, w.areasoverlengths, w.dirichletnodes,
                                              w.dirichleths, Ss, w.volumes, 20, dt, np.ones(21),
                                              ref, G, reltol=1e-14)
    assert rel(out[20][0], ref[-1]) < 1e-9
    assert rel(out[100][0], ref[-1]) < 1e-9
    for N in (20, 100):
        assert rel(out[N][1], dK_ref) < 1e-7
        assert rel(out[N][2], dQ_ref) < 1e-7


def test_transient_states_and_steady_limit(pkg):
    from dpfehm_jl_amd import api
    w, Ss = _problem(pkg, ns=(16, 14, 6))
    K = w.Ks()
    nf = w.n - w.dirichletnodes.size
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss,
            w.volumes)
    states = api.groundwater_transient(*args, np.zeros(nf), 5, 1e9, return_states=True,
                                       reltol=1e-13)
    assert states.shape == (6, nf) and np.all(states[0] == 0)
    h = api.groundwater_steadystate(K, w.neighbors, w.areasoverlengths, w.dirichletnodes,
                                    w.dirichleths, w.Qs, reltol=1e-13)
    isfree = np.ones(w.n, dtype=bool)
    isfree[w.dirichletnodes - 1] = False
    assert rel(states[-1], h[isfree]) < 1e-8


@pytest.mark.parametrize("precision,ck", [("fp64", 3), ("fp64", 0), ("mixed", 4),
                                          ("mixed_pc", 1)])
def test_transient_step_sequence_every_tangent_vs_oracle(pkg, oracle, precision, ck):
    """A caller-given, non-uniform step sequence and an objective of several
    states (u_0 included): all four tangents of the reverse mode (Ks,
    dirichleths, Qs, u0) against the oracle's."""
    from dpfehm_jl_amd import api
    print("transient dts", precision, ck, flush=True)
    w, Ss = _problem(pkg, ns=(18, 16, 8))
    rng = np.random.default_rng(11)
    w.dirichleths[w.dirichletnodes - 1] = rng.normal(size=w.dirichletnodes.size)
    K = w.Ks()
    nf = w.n - w.dirichletnodes.size
    N = 13
    dts = 3600.0 * rng.uniform(0.05, 2.0, N)
    qs = np.minimum(1.0, np.concatenate([[0.0], np.cumsum(dts)]) / 20000.0)
    u0 = 0.01 * rng.normal(size=nf)
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss,
            w.volumes)
    ref = oracle.transient(*args, u0, N, dts, qs, reltol=1e-14)
    st = {}
    uN, pb = api.rrule_groundwater_transient(*args, u0, N, dts, qs, checkpoint_every=ck,
                                             reltol=1e-12, precision=precision, stats=st)
    states = api.groundwater_transient(*args, u0, N, dts, qs, return_states=True,
                                       reltol=1e-12, precision=precision)
    tol = 1e-10 if precision == "fp64" else 1e-8
    assert rel(uN, ref[-1]) < tol
    assert rel(states, ref) < tol and np.array_equal(states[-1], uN)
    G = np.zeros_like(ref)
    for k in (0, 4, 9, 11):
        G[k, rng.choice(nf, 25, replace=False)] = rng.normal(size=25)
    dK_r, ddh_r, dQ_r, du0_r = oracle.transient_adjoint(
        K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, Ss, w.volumes, N,
        dts, qs, ref, G, reltol=1e-14, full=True)
    dK, ddh, dQ, du0 = pb(G)
    print(rel(dK, dK_r), rel(ddh, ddh_r), rel(dQ, dQ_r), rel(du0, du0_r), flush=True)
    assert rel(dK, dK_r) < 1e-7
    assert rel(ddh, ddh_r) < 1e-7
    assert rel(dQ, dQ_r) < 1e-7
    assert rel(du0, du0_r) < 1e-7
    isfree = np.ones(w.n, dtype=bool)
    isfree[w.dirichletnodes - 1] = False
    assert np.all(ddh[isfree] == 0.0) and np.all(dQ[~isfree] == 0.0)
    # a seed at u_0 alone: du0 is the seed, every other tangent zero
    G0 = np.zeros_like(ref)
    G0[0] = rng.normal(size=nf)
    dK0, ddh0, dQ0, du00 = pb(G0)
    assert np.array_equal(du00, G0[0])
    assert not dK0.any() and not ddh0.any() and not dQ0.any()
