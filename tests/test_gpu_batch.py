"""Batched ensembles: uneven meshes packed into one block-diagonal topology,
solved in one set of launches with per-member CG scalars; every member is
checked against the FP64 oracle solved alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _members(pkg):
    W = pkg.workloads
    return [
        W.box_problem("a", (12, 10, 6), (0, 0, 0), (6.0, 5.0, 3.0), lam=4.0, sigma=1.5,
                      field_seed=100, dirichlet="top", sources=(), n_injectors=3, n_obs=20,
                      obs_seed=3),
        W.box_problem("b", (16, 8, 8), (0, 0, 0), (8.0, 4.0, 4.0), lam=6.0, sigma=1.5,
                      field_seed=101, dirichlet="last", sources=((1, 1e-4),), n_obs=20,
                      obs_seed=3),
        W.box_problem("c", (10, 10, 10), (0, 0, 0), (5.0, 5.0, 5.0), lam=2.0, sigma=2.5,
                      field_seed=102, dirichlet="top", sources=(), n_injectors=2, n_obs=20,
                      obs_seed=3),
    ]


@pytest.mark.parametrize("precision,tol_h", [("fp64", 1e-10), ("mixed", 1e-8),
                                           ("mixed_pc", 1e-8)])
def test_batch_matches_members_solved_alone(pkg, oracle, precision, tol_h):
    from dpfehm_jl_amd import api
    print("batch", precision, flush=True)
    ms = _members(pkg)
    topo = api.BatchTopology([(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes) for w in ms])
    assert topo.nmembers == 3
    logK = np.concatenate([w.logK for w in ms])
    Ke = api.hm_forward(topo, logK)
    dh = np.concatenate([w.dirichleths for w in ms])
    Q = np.concatenate([w.Qs for w in ms])
    st = {}
    h, pb = api.rrule_batch_steadystate(topo, Ke, dh, Q, reltol=1e-12, precision=precision,
                                        stats=st)
    print(st, flush=True)
    relres = topo.member_relres()
    assert np.all(relres <= 1e-12), relres
    delta = np.zeros(topo.n)
    for m, w in enumerate(ms):
        delta[topo.cell0[m] + w.obs - 1] = 1.0 + m
    dK, ddh, dQ = pb(delta)
    dlogK = api.hm_pullback(topo, logK, Ke, dK)
    for m, w in enumerate(ms):
        c0, c1 = topo.cell0[m], topo.cell0[m + 1]
        e0, e1 = topo.edge0[m], topo.edge0[m + 1]
        K_m = oracle.hm_forward(w.logK, w.neighbors)
        assert rel(Ke[e0:e1], K_m) < 1e-14
        h_ref, pb_ref = oracle.steadystate_rrule(K_m, w.neighbors, w.areasoverlengths,
                                                 w.dirichletnodes, w.dirichleths, w.Qs,
                                                 reltol=1e-14)
        assert rel(h[c0:c1], h_ref) < tol_h, m
        _, dK_ref, _, _, _, ddh_ref, dQ_ref = pb_ref(delta[c0:c1])
        assert rel(dK[e0:e1], dK_ref) < 1e-7, m
        assert rel(ddh[c0:c1], ddh_ref) < 1e-7, m
        assert rel(dQ[c0:c1], dQ_ref) < 1e-7, m
        assert rel(dlogK[c0:c1], oracle.hm_pullback(w.logK, w.neighbors, dK_ref)) < 1e-7, m


def test_batch_member_with_zero_rhs(pkg):
    """A member with no sources and h = 0 on its boundary converges at once to
    h = 0 while the others iterate."""
    from dpfehm_jl_amd import api
    ms = _members(pkg)
    ms[1].Qs[:] = 0.0
    topo = api.BatchTopology([(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes) for w in ms])
    h = api.batch_steadystate(topo, np.concatenate([w.Ks() for w in ms]),
                              np.concatenate([w.dirichleths for w in ms]),
                              np.concatenate([w.Qs for w in ms]), reltol=1e-11)
    assert np.all(h[topo.cell0[1]:topo.cell0[2]] == 0.0)
    assert np.abs(h[:topo.cell0[1]]).max() > 0 and np.abs(h[topo.cell0[2]:]).max() > 0
