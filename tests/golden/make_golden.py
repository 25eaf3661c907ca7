"""Writes tests/golden/golden_small.npz.

The reference (DPFEHM.jl, Julia) cannot run in this image and ships no data
files, so these vectors are produced by an independent numpy/scipy
restatement of the closed form of src/groundwater.jl:1-21,75-109 (edge
transmissibilities T_e = K_e aol_e, b = Q + sum of T_e dh over Dirichlet
neighbours, A = sum_e T_e (e_a - e_b)(e_a - e_b)' restricted to free cells,
direct sparse solve, adjoint lambda = A^-1 delta, dK_e = -aol_e (h~a - h~b)
(lambda~a - lambda~b)) and of the harmonic-mean map of src/richards.jl:50-52.
It shares no code with oracle/gw_oracle.c; tests/test_golden.py checks the
oracle and the GPU library against these vectors.

Run:  python tests/golden/make_golden.py
"""
import importlib.util
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _pkg():
    d = os.path.join(ROOT, "dpfehm.jl_amd")
    spec = importlib.util.spec_from_file_location("dpfehm_jl_amd", os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    m = importlib.util.module_from_spec(spec)
    sys.modules["dpfehm_jl_amd"] = m
    spec.loader.exec_module(m)
    return m


def closed_form(nb1, aol, dn1, dh, Q, logK, delta_fn):
    n = Q.size
    a, b = nb1[:, 0] - 1, nb1[:, 1] - 1
    k = np.exp(logK)
    K = 2.0 / (1.0 / k[a] + 1.0 / k[b])
    T = K * aol
    free = np.ones(n, dtype=bool)
    free[dn1 - 1] = False
    idx = -np.ones(n, dtype=np.int64)
    idx[free] = np.arange(free.sum())
    nf = int(free.sum())
    rows, cols, vals = [], [], []
    rhs = Q[free].copy()
    for x, y in ((a, b), (b, a)):
        m = free[x]
        rows.append(idx[x[m]]); cols.append(idx[x[m]]); vals.append(T[m])
        ff = m & free[y]
        rows.append(idx[x[ff]]); cols.append(idx[y[ff]]); vals.append(-T[ff])
        fd = m & ~free[y]
        np.add.at(rhs, idx[x[fd]], T[fd] * dh[y[fd]])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nf, nf)).tocsc()
    hf = spla.spsolve(A, rhs)
    h = dh.copy()
    h[free] = hf
    delta = delta_fn(h)
    lam = spla.spsolve(A.T.tocsc(), delta[free])
    lt = np.zeros(n)
    lt[free] = lam
    dK = -aol * (h[a] - h[b]) * (lt[a] - lt[b])
    ddh = np.where(free, 0.0, delta)
    fd_e = free[a] & ~free[b]
    np.add.at(ddh, b[fd_e], T[fd_e] * lt[a[fd_e]])
    fd_e = free[b] & ~free[a]
    np.add.at(ddh, a[fd_e], T[fd_e] * lt[b[fd_e]])
    dQ = np.where(free, lt, 0.0)
    dlogK = np.zeros(n)
    np.add.at(dlogK, a, dK * K * k[b] / (k[a] + k[b]))
    np.add.at(dlogK, b, dK * K * k[a] / (k[a] + k[b]))
    return dict(Ke=K, h=h, dK=dK, ddh=ddh, dQ=dQ, dlogK=dlogK, delta=delta)


def main():
    P = _pkg()
    out = {}
    cases = {
        # the 2M-cell construction (top-face Dirichlet, 4 injectors) at 6x5x4
        "tiny": P.workloads.steady_2m(sigma=1.0, ns=(6, 5, 4), maxs=(2.5, 2.0, 1.5)),
        # the 100k construction (one Dirichlet cell, Q at cell 1) at 20x20x5
        "c1small": P.workloads.box_problem("c1small", (20, 20, 5), (0, 0, 0), (10.0, 10.0, 2.0),
                                           lam=10.0, sigma=1.0, field_seed=0, dirichlet="last",
                                           sources=((1, 1e-4),), n_obs=20, obs_seed=1),
        # sigma = 3 contrast on the 2M construction at 16x14x8
        "sigma3": P.workloads.steady_2m(sigma=3.0, ns=(16, 14, 8), maxs=(8.0, 7.0, 3.5)),
    }
    for name, w in cases.items():
        rng = np.random.Generator(np.random.Philox(42))
        h_obs = rng.normal(size=w.obs.size) * 1e-3

        def delta_fn(h, w=w, h_obs=h_obs):
            d = np.zeros(w.n)
            d[w.obs - 1] = 2.0 * (h[w.obs - 1] - h_obs)
            return d

        r = closed_form(w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths,
                        w.Qs, w.logK, delta_fn)
        out[f"{name}/neighbors"] = w.neighbors
        out[f"{name}/aol"] = w.areasoverlengths
        out[f"{name}/dirichletnodes"] = w.dirichletnodes
        out[f"{name}/dirichleths"] = w.dirichleths
        out[f"{name}/Qs"] = w.Qs
        out[f"{name}/logK"] = w.logK
        out[f"{name}/obs"] = w.obs
        out[f"{name}/h_obs"] = h_obs
        for k, v in r.items():
            out[f"{name}/{k}"] = v
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)
    print("wrote", os.path.join(HERE, "golden_small.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
