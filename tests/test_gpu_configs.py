"""The benchmark configurations at their full sizes, on the GPU.

* C1   100k cells (single Dirichlet cell): dJ/dlogK of J = sum h_obs^2 against
       the oracle, both precisions.
* C3   2M cells at sigma 1, 2 and 3: FP64 and mixed relres, an independent
       groundwater_residuals check, mixed vs FP64 heads and logK gradients.
* C5a  the 96-member packed ensemble (23.9 M cells) in mixed precision with
       each PCG iteration replayed as a hipGraph: every member's relres, and
       one member of each grid shape against the oracle solved alone.
* C5b  2M cells x 200 backward-Euler steps of 3600 s with the injection ramp
       and checkpoints every 20 steps: checkpointed against every-state
       adjoint, mixed against FP64 final heads, central differences of the
       logK gradient.

Tolerances (from the configurations): heads 1e-10 (FP64) / 1e-8 (mixed)
relative L2, gradients 1e-6 relative."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

skip_big = pytest.mark.skipif(bool(os.environ.get("DPF_SKIP_2M")), reason="DPF_SKIP_2M set")


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def api(pkg):
    from dpfehm_jl_amd import api as A
    return A


@pytest.mark.parametrize("precision,tol_h", [("fp64", 1e-10), ("mixed", 1e-8),
                                           ("mixed_pc", 1e-8)])
def test_c1_dlogK_100k(pkg, api, oracle, precision, tol_h):
    """C1's output is dJ/dlogK over all 100k cells, J = sum over the 50
    observation cells of h^2."""
    _say("c1", precision)
    w = pkg.workloads.steady_100k()
    args = (w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs)
    K_ref = oracle.hm_forward(w.logK, w.neighbors)
    h_ref, pb_ref = oracle.steadystate_rrule(K_ref, *args, reltol=1e-14)
    delta = np.zeros(w.n)
    delta[w.obs - 1] = 2.0 * h_ref[w.obs - 1]
    g_ref = oracle.hm_pullback(w.logK, w.neighbors, pb_ref(delta)[1])
    K = api.logKs2Ks_harmonic(w.logK, w.neighbors, w.dirichletnodes, w.areasoverlengths)
    st = {}
    h, pb = api.rrule_groundwater_steadystate(K, *args, reltol=1e-10, precision=precision,
                                              stats=st)
    assert st["converged"] == 1 and st["relres"] <= 1e-10
    assert rel(h, h_ref) < tol_h
    d = np.zeros(w.n)
    d[w.obs - 1] = 2.0 * h[w.obs - 1]
    g = api.logKs2Ks_harmonic_pullback(w.logK, K, pb(d)[1], w.neighbors, w.dirichletnodes,
                                       w.areasoverlengths)
    _say("c1", precision, st["iterations"], rel(h, h_ref), rel(g, g_ref))
    assert rel(g, g_ref) < 1e-6


@skip_big
@pytest.mark.parametrize("sigma", [1.0, 2.0, 3.0])
def test_c3_2m_sigma_sweep(pkg, api, sigma):
    """C3 (and C4's objective): the 2M-cell problem with logK of standard
    deviation sigma, FP64 and mixed precision, each to relres 1e-10.  The
    default hierarchy (smoother degrees 1 / 1 / 2, prolongator damping 1.5 /
    2) converges in 18-20 PCG iterations a solve here; the bound (21 FP64, 22
    mixed, summed over refinement steps) fails on a weaker hierarchy, e.g. the
    damping 4/3 with degree 2 on the second level (21-23 / 23)."""
    _say("c3 sigma", sigma)
    w = pkg.workloads.steady_2m(sigma=sigma)
    args = (w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs)
    K = api.logKs2Ks_harmonic(w.logK, w.neighbors, w.dirichletnodes, w.areasoverlengths)
    K_obs = api.logKs2Ks_harmonic(pkg.workloads.observation_logK(w), w.neighbors,
                                  w.dirichletnodes, w.areasoverlengths)
    h_obs = api.groundwater_steadystate(K_obs, *args, reltol=1e-12)[w.obs - 1]
    h_obs = h_obs + pkg.workloads.observation_noise(w.obs.size)
    isfree = np.ones(w.n, dtype=bool)
    isfree[w.dirichletnodes - 1] = False
    nf = int(isfree.sum())
    one = np.ones(w.n)
    b = -api.groundwater_residuals(np.zeros(nf), K, *args, one, one)
    res = {}
    for precision in ("fp64", "mixed"):
        st = {}
        h, pb = api.rrule_groundwater_steadystate(K, *args, reltol=1e-10, precision=precision,
                                                  stats=st)
        assert st["converged"] == 1 and st["relres"] <= 1e-10, st
        r = api.groundwater_residuals(h[isfree], K, *args, one, one)
        indep = np.linalg.norm(r) / np.linalg.norm(b)
        assert indep <= 1.5e-10, indep
        delta = np.zeros(w.n)
        delta[w.obs - 1] = 2.0 * (h[w.obs - 1] - h_obs)
        st2 = {}
        dK = pb(delta, stats=st2)[1]
        assert st2["converged"] == 1 and st2["relres"] <= 1e-10, st2
        g = api.logKs2Ks_harmonic_pullback(w.logK, K, dK, w.neighbors, w.dirichletnodes,
                                           w.areasoverlengths)
        pb.state.close()
        res[precision] = (h, g)
        _say("c3", sigma, precision, st["iterations"], st2["iterations"], st["relres"], indep)
        cap = 21 if precision == "fp64" else 22
        assert st["iterations"] <= cap and st2["iterations"] <= cap, (st, st2)
    eh = rel(res["mixed"][0], res["fp64"][0])
    eg = rel(res["mixed"][1], res["fp64"][1])
    _say("c3", sigma, "mixed vs fp64", eh, eg)
    assert eh <= 1e-8
    assert eg <= 1e-6


@skip_big
def test_c5a_ensemble_96_members(pkg, api, oracle):
    """C5a: 96 realizations (32 x 64^3, 32 x 96x96x24, 32 x 128x128x16; 23.9 M
    cells) packed into one batch topology, forward + adjoint in mixed
    precision with hipGraph-replayed iterations and the autotuner (lanes per
    row of the coarse levels, smoother degrees) on its first solve."""
    _say("c5a build")
    ms = pkg.workloads.ensemble_members()
    topo = api.BatchTopology([(m["n"], m["nb"], m["aol"], m["dn"]) for m in ms])
    assert topo.nmembers == 96 and topo.n == 23855104
    logK = np.concatenate([m["logK"] for m in ms])
    Ke = api.hm_forward(topo, logK)
    dh = np.zeros(topo.n)
    Q = np.concatenate([m["Q"] for m in ms])
    st = {}
    h, pb = api.rrule_batch_steadystate(topo, Ke, dh, Q, reltol=1e-10, precision="mixed",
                                        graph=True, autotune=True, stats=st)
    _say("c5a forward", st, "lanes", topo.lanes()[0], "degrees", topo.tuned_degrees()["mixed"])
    assert topo.lanes()[1], "the autotuner timed no lanes-per-row layout"
    relres = topo.member_relres()
    assert np.all(relres <= 1e-10), relres.max()
    delta = np.zeros(topo.n)
    for m, mm in enumerate(ms):
        c0 = topo.cell0[m]
        delta[c0 + mm["obs"] - 1] = 2.0 * h[c0 + mm["obs"] - 1]
    st2 = {}
    dK, ddh, dQ = pb(delta, stats=st2)
    _say("c5a adjoint", st2)
    relres = topo.member_relres()
    assert np.all(relres <= 1e-10), relres.max()
    dlogK = api.hm_pullback(topo, logK, Ke, dK)
    pb.state.close()
    for m in (0, 32, 64):
        mm = ms[m]
        c0, c1 = topo.cell0[m], topo.cell0[m + 1]
        K_m = oracle.hm_forward(mm["logK"], mm["nb"])
        h_ref, pb_ref = oracle.steadystate_rrule(K_m, mm["nb"], mm["aol"], mm["dn"],
                                                 np.zeros(mm["n"]), mm["Q"], reltol=1e-14)
        eh = rel(h[c0:c1], h_ref)
        g_ref = oracle.hm_pullback(mm["logK"], mm["nb"], pb_ref(delta[c0:c1])[1])
        eg = rel(dlogK[c0:c1], g_ref)
        _say("c5a member", m, mm["ns"], eh, eg)
        assert eh <= 1e-8
        assert eg <= 1e-6


@skip_big
def test_c5b_transient_2m_200_steps(pkg, api):
    """C5b: 200 backward-Euler steps of 3600 s on the 2M-cell grid, Ss = 1e-5
    1/m, the 4 injectors ramped over the first 50 steps, J = sum over the 200
    observation cells of u_N^2, checkpoints every 20 steps."""
    _say("c5b")
    w = pkg.workloads.steady_2m()
    N, dt = 200, 3600.0
    qs = pkg.workloads.transient_schedule(N, dt, 50)
    Ss = np.full(w.n, 1e-5)
    isfree = np.ones(w.n, dtype=bool)
    isfree[w.dirichletnodes - 1] = False
    nf = int(isfree.sum())
    obs_f = (np.cumsum(isfree) - 1)[w.obs - 1]
    u0 = np.zeros(nf)

    def Kof(logK):
        return api.logKs2Ks_harmonic(logK, w.neighbors, w.dirichletnodes, w.areasoverlengths)

    def run(logK, precision, ck, reltol):
        st = {}
        uN, pb = api.rrule_groundwater_transient(
            Kof(logK), w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs,
            Ss, w.volumes, u0, N, dt, qs, checkpoint_every=ck, reltol=reltol,
            precision=precision, stats=st)
        return uN, pb, st

    K = Kof(w.logK)
    uN, pb, st = run(w.logK, "fp64", 20, 1e-12)
    _say("c5b fp64 forward", st)
    g = np.zeros(nf)
    g[obs_f] = 2.0 * uN[obs_f]
    out20 = pb(g)
    dK20 = out20[0]
    pb.state.close()
    uN1, pb1, _ = run(w.logK, "fp64", 1, 1e-12)
    assert np.array_equal(uN1, uN)
    out1 = pb1(g)
    pb1.state.close()
    # dKs, ddirichleths, dQs, du0
    e_ck = max(rel(a, b) for a, b in zip(out20, out1))
    _say("c5b checkpointed vs every state", e_ck)
    assert e_ck <= 1e-12
    uNm, pbm, stm = run(w.logK, "mixed", 20, 1e-12)
    _say("c5b mixed forward", stm)
    assert rel(uNm, uN) <= 1e-8
    gm = np.zeros(nf)
    gm[obs_f] = 2.0 * uNm[obs_f]
    dKm = pbm(gm)[0]
    pbm.state.close()
    dlogK = api.logKs2Ks_harmonic_pullback(w.logK, K, dK20, w.neighbors, w.dirichletnodes,
                                           w.areasoverlengths)
    dlogKm = api.logKs2Ks_harmonic_pullback(w.logK, K, dKm, w.neighbors, w.dirichletnodes,
                                            w.areasoverlengths)
    assert rel(dlogKm, dlogK) <= 1e-6

    def J(logK):
        u = api.groundwater_transient(Kof(logK), w.neighbors, w.areasoverlengths,
                                      w.dirichletnodes, w.dirichleths, w.Qs, Ss, w.volumes, u0,
                                      N, dt, qs, reltol=1e-13)
        return np.sum(u[obs_f] ** 2)

    eps, gmax = 1e-6, np.abs(dlogK).max()
    for c in np.argsort(-np.abs(dlogK))[:3]:
        lp, lm = w.logK.copy(), w.logK.copy()
        lp[c] += eps
        lm[c] -= eps
        fd = (J(lp) - J(lm)) / (2 * eps)
        _say("c5b fd", c + 1, dlogK[c], fd)
        assert abs(dlogK[c] - fd) <= 1e-5 * gmax + 1e-5 * abs(fd), (c, dlogK[c], fd)
