"""Jacobian export (dpf_jacobian_pattern / dpf_jacobian_values): the
generated groundwater_h, groundwater_Ks, groundwater_dirichleths and
groundwater_Qs of src/groundwater.jl:1 as CSR over the free cells.

* Small mesh with a parallel edge and a self-edge: groundwater_h equals the
  oracle's CSR (same pattern, values to 1e-14); the Ks, dirichleths and Qs
  Jacobians equal the dense matrices built column by column from the oracle's
  residual, which is linear in Ks and Qs and affine in dirichleths.
* 2M cells: J_h h = r(h) - r(0) and J_Ks Ks = r(Ks) - r(0 Ks) with the GPU
  residual (size-independent linearity properties)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api(pkg):
    from dpfehm_jl_amd import api as A
    return A


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _mesh():
    nx, ny, nz = 9, 8, 5
    idx = np.arange(nx * ny * nz).reshape(nz, ny, nx) + 1
    nb = np.concatenate([np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], 1),
                         np.stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()], 1),
                         np.stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()], 1)])
    top = idx[-1].ravel()
    # a parallel edge between two free cells, one to a Dirichlet cell, a self-edge
    extra = np.array([[idx[1, 2, 3], idx[1, 2, 4]], [idx[-2, 3, 3], idx[-1, 3, 3]],
                      [idx[2, 2, 2], idx[2, 2, 2]]])
    nb = np.concatenate([nb, extra]).astype(np.int64)
    rng = np.random.default_rng(7)
    n, ne = idx.size, nb.shape[0]
    return dict(n=n, nb=nb, aol=rng.uniform(0.5, 2.0, ne), dn=top,
                K=np.exp(rng.normal(-9.0, 1.0, ne)), dh=rng.uniform(-1.0, 1.0, n),
                Q=rng.normal(0.0, 1e-4, n), Ss=rng.uniform(1e-6, 1e-4, n),
                V=rng.uniform(0.5, 2.0, n))


def test_jacobians_match_oracle(pkg, api, oracle):
    print("jacobians small", flush=True)
    m = _mesh()
    n, nb, aol, dn = m["n"], m["nb"], m["aol"], m["dn"]
    K, dh, Q, Ss, V = m["K"], m["dh"], m["Q"], m["Ss"], m["V"]
    nf = n - dn.size
    h = np.random.default_rng(8).normal(size=nf)
    args = (nb, aol, dn, dh, Q, Ss, V)

    def r(h_=h, K_=K, dh_=dh, Q_=Q):
        return oracle.groundwater_residuals(h_, K_, nb, aol, dn, dh_, Q_, Ss, V)

    Jh = api.groundwater_h(h, K, *args)
    Jh_ref = oracle.groundwater_h(h, K, *args)
    assert Jh.shape == (nf, nf)
    assert np.array_equal(Jh.indptr, Jh_ref.indptr)
    assert np.array_equal(Jh.indices, Jh_ref.indices)
    assert rel(Jh.data, Jh_ref.data) < 1e-14
    # Ks: r is linear in Ks
    JK = api.groundwater_Ks(h, K, *args)
    ne = nb.shape[0]
    assert JK.shape == (nf, ne)
    r0 = r(K_=np.zeros(ne))
    dense = np.empty((nf, ne))
    for e in range(ne):
        k = np.zeros(ne)
        k[e] = 1.0
        dense[:, e] = r(K_=k) - r0
    assert rel(JK.toarray(), dense) < 1e-12
    # the self-edge adds no entry; the parallel edges each have their own column
    assert JK[:, ne - 1].nnz == 0
    assert JK[:, ne - 3].nnz == 2 and JK[:, ne - 2].nnz == 1
    # dirichleths: affine
    JD = api.groundwater_dirichleths(h, K, *args)
    assert JD.shape == (nf, n)
    r1 = r()
    dense = np.empty((nf, n))
    for i in range(n):
        d = dh.copy()
        d[i] += 1.0
        dense[:, i] = r(dh_=d) - r1
    assert rel(JD.toarray(), dense) < 1e-9
    assert JD.nnz == np.count_nonzero(dense)
    # Qs: linear
    JQ = api.groundwater_Qs(h, K, *args)
    dense = np.empty((nf, n))
    for i in range(n):
        q = Q.copy()
        q[i] += 1.0
        dense[:, i] = r(Q_=q) - r1
    assert rel(JQ.toarray(), dense) < 1e-9
    assert JQ.nnz == nf


def test_jacobian_linearity_2m(pkg, api):
    """J_h h = r(h) - r(0) and J_Ks Ks = r(Ks) - r(0 Ks) on the 2M-cell grid."""
    print("jacobians 2m", flush=True)
    w = pkg.workloads.steady_2m()
    nf = w.n - w.dirichletnodes.size
    Ss, V = np.full(w.n, 1e-5), w.volumes
    K = w.Ks()
    h = np.random.default_rng(9).normal(size=nf)
    args = (w.neighbors, w.areasoverlengths, w.dirichletnodes, w.dirichleths, w.Qs, Ss, V)

    def r(h_, K_):
        return api.groundwater_residuals(h_, K_, *args)

    Jh = api.groundwater_h(h, K, *args)
    assert Jh.shape == (nf, nf) and Jh.nnz == api.get_topology(
        w.n, w.neighbors, w.areasoverlengths, w.dirichletnodes).nnz
    assert rel(Jh @ h, r(h, K) - r(np.zeros(nf), K)) < 1e-12
    JK = api.groundwater_Ks(h, K, *args)
    assert rel(JK @ K, r(h, K) - r(h, np.zeros_like(K))) < 1e-12
