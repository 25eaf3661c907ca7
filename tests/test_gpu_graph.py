"""DPF_FLAG_GRAPH: each PCG iteration replayed as one captured hipGraph.  The
graph holds the same kernels in the same order as the launched form, so every
result must be bitwise equal to the eager one; the replays are interleaved
with eager calls on the same topology, across several systems, in the steady,
batched and transient paths."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[graph]", *a, flush=True)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_pc"])
def test_steady_graph_bitwise(pkg, precision):
    from dpfehm_jl_amd import api
    _say("steady", precision)
    w = pkg.workloads.steady_2m(ns=(30, 28, 12), maxs=(15.0, 14.0, 6.0))
    args = (w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs)
    rng = np.random.default_rng(1)
    delta = np.zeros(w.n)
    delta[w.obs - 1] = rng.normal(size=w.obs.size)
    fields = [w.logK, 1.3 * w.logK + 2.0, w.logK + 0.5 * rng.normal(size=w.n)]
    for rep in range(2):
        for k, lk in enumerate(fields):
            K = api.logKs2Ks_harmonic(lk, w.neighbors, w.dirichletnodes, w.areasoverlengths)
            outs = []
            for graph in (False, True):
                st = {}
                h, pb = api.rrule_groundwater_steadystate(K, *args, precision=precision,
                                                          graph=graph, stats=st)
                outs.append((h, pb(delta)[1], st["iterations"]))
            (h0, g0, i0), (h1, g1, i1) = outs
            assert i0 == i1 and np.array_equal(h0, h1) and np.array_equal(g0, g1), (rep, k)


def test_batch_graph_bitwise(pkg):
    from dpfehm_jl_amd import api
    from test_gpu_batch import _members
    _say("batch")
    ms = _members(pkg)
    topo = api.BatchTopology([(w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes) for w in ms])
    logK = np.concatenate([w.logK for w in ms])
    Ke = api.hm_forward(topo, logK)
    dh = np.concatenate([w.dirichleths for w in ms])
    Q = np.concatenate([w.Qs for w in ms])
    delta = np.zeros(topo.n)
    delta[::17] = 1.0
    for precision in ("mixed", "fp64"):
        outs = []
        for graph in (False, True, False, True):
            h, pb = api.rrule_batch_steadystate(topo, Ke, dh, Q, precision=precision, graph=graph)
            outs.append((h, pb(delta)[0]))
        for h, dK in outs[1:]:
            assert np.array_equal(h, outs[0][0]) and np.array_equal(dK, outs[0][1]), precision


def test_transient_graph_bitwise(pkg):
    from dpfehm_jl_amd import api
    from test_gpu_transient import _problem
    _say("transient")
    w, Ss = _problem(pkg)
    K = w.Ks()
    nf = w.n - w.dirichletnodes.size
    N, dt = 12, 3600.0
    qs = pkg.workloads.transient_schedule(N, dt, ramp_steps=5)
    args = (K, w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss,
            w.volumes, np.zeros(nf), N, dt, qs)
    g = np.random.default_rng(4).normal(size=nf)
    outs = []
    for graph in (False, True):
        uN, pb = api.rrule_groundwater_transient(*args, checkpoint_every=4, precision="mixed",
                                                 graph=graph)
        outs.append((uN, *pb(g)))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)
