/*
 * gw_oracle.c -- FP64 CPU restatement of DPFEHM's single-phase groundwater path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is built into
 * oracle/_build/libgw_oracle.so and loaded by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  Nothing
 * in the product (dpfehm.jl_amd/) links, loads or calls it.
 *
 * Indices are 0-based throughout (the reference is 1-based; the Python
 * wrapper oracle/gw_oracle.py converts).  Scalar, single-threaded C.
 *
 * What each function restates (reference file:line) and what pins it is
 * listed in the header of oracle/gw_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* getfreenodes: src/utilities.jl:10-24 (intended O(n) form; the effective
 * src/twophase.jl:15-20 form yields identical arrays).  n2f[i] = free index
 * or -1; f2n[j] = cell index.  Returns the number of free cells. */
int64_t orc_freenodes(int64_t n, int64_t nd, const int64_t *dnodes,
                      uint8_t *isfree, int64_t *n2f, int64_t *f2n) {
  for (int64_t i = 0; i < n; i++) isfree[i] = 1;
  for (int64_t k = 0; k < nd; k++) isfree[dnodes[k]] = 0;
  int64_t j = 0;
  for (int64_t i = 0; i < n; i++) {
    if (isfree[i]) {
      n2f[i] = j;
      f2n[j] = i;
      j++;
    } else {
      n2f[i] = -1;
    }
  }
  return j;
}

/* groundwater_residuals: src/groundwater.jl:1-21.  Same term order as the
 * macro body: the source term of every free cell first, then every edge in
 * list order, each of its two orientations in turn. */
void orc_residuals(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                   const double *aol, const uint8_t *isfree, const int64_t *n2f,
                   const double *h, const double *dh, const double *Q,
                   const double *Ss, const double *V, int64_t nf, double *r) {
  for (int64_t j = 0; j < nf; j++) r[j] = 0.0;
  for (int64_t i = 0; i < n; i++)
    if (isfree[i]) r[n2f[i]] += -Q[i] / (Ss[i] * V[i]);
  for (int64_t e = 0; e < ne; e++) {
    int64_t a = nb[2 * e], b = nb[2 * e + 1];
    for (int o = 0; o < 2; o++) {
      int64_t n1 = o ? b : a, n2 = o ? a : b;
      if (!isfree[n1]) continue;
      int64_t j1 = n2f[n1];
      double s1 = Ss[n1] * V[n1];
      if (isfree[n2])
        r[j1] += K[e] * (h[j1] - h[n2f[n2]]) * aol[e] / s1;
      else
        r[j1] += K[e] * (h[j1] - dh[n2]) * aol[e] / s1;
    }
  }
}

/* Jacobian dr/dh (the generated groundwater_h of src/groundwater.jl:1) as a
 * CSR matrix with sorted columns and duplicates summed, i.e. what
 * SparseArrays.sparse(I, J, V) makes of the COO triplets.  Two-call protocol:
 * with col == NULL only rowptr is filled and nnz returned. */
static int cmp_i64(const void *x, const void *y) {
  int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return (a > b) - (a < b);
}

int64_t orc_jacobian_h(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                       const double *aol, const uint8_t *isfree, const int64_t *n2f,
                       const double *Ss, const double *V, int64_t nf,
                       int64_t *rowptr, int64_t *col, double *val) {
  /* COO triplets in generation order */
  int64_t cap = 4 * ne + 1, m = 0;
  int64_t *ci = (int64_t *)malloc(sizeof(int64_t) * cap);
  int64_t *cj = (int64_t *)malloc(sizeof(int64_t) * cap);
  double *cv = (double *)malloc(sizeof(double) * cap);
  for (int64_t e = 0; e < ne; e++) {
    int64_t a = nb[2 * e], b = nb[2 * e + 1];
    for (int o = 0; o < 2; o++) {
      int64_t n1 = o ? b : a, n2 = o ? a : b;
      if (!isfree[n1]) continue;
      int64_t j1 = n2f[n1];
      double t = K[e] * aol[e] / (Ss[n1] * V[n1]);
      ci[m] = j1; cj[m] = j1; cv[m] = t; m++;
      if (isfree[n2]) { ci[m] = j1; cj[m] = n2f[n2]; cv[m] = -t; m++; }
    }
  }
  /* bucket by row, preserving generation order inside a row */
  int64_t *cnt = (int64_t *)calloc((size_t)nf + 1, sizeof(int64_t));
  for (int64_t k = 0; k < m; k++) cnt[ci[k] + 1]++;
  for (int64_t j = 0; j < nf; j++) cnt[j + 1] += cnt[j];
  int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (nf + 1));
  memcpy(pos, cnt, sizeof(int64_t) * (nf + 1));
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (m + 1));
  for (int64_t k = 0; k < m; k++) perm[pos[ci[k]]++] = k;
  /* per row: sort (col, order) keys, merge duplicates */
  int64_t nnz = 0;
  int64_t *keys = (int64_t *)malloc(sizeof(int64_t) * 2 * (m + 1));
  rowptr[0] = 0;
  for (int64_t j = 0; j < nf; j++) {
    int64_t s = cnt[j], t = cnt[j + 1], len = t - s;
    for (int64_t q = 0; q < len; q++) {
      int64_t k = perm[s + q];
      keys[2 * q] = cj[k] * (m + 1) + q; /* col major key, stable by order */
      keys[2 * q + 1] = k;
    }
    qsort(keys, (size_t)len, 2 * sizeof(int64_t), cmp_i64);
    int64_t last = -1;
    for (int64_t q = 0; q < len; q++) {
      int64_t k = keys[2 * q + 1];
      if (cj[k] != last) {
        if (col) { col[nnz] = cj[k]; val[nnz] = cv[k]; }
        nnz++;
        last = cj[k];
      } else if (col) {
        val[nnz - 1] += cv[k];
      }
    }
    rowptr[j + 1] = nnz;
  }
  free(ci); free(cj); free(cv); free(cnt); free(pos); free(perm); free(keys);
  return nnz;
}

/* ---- linear solver: CG preconditioned with incomplete Cholesky IC(0) ----
 * A stand-in for AlgebraicMultigrid._solve (src/groundwater.jl:45-49): only
 * the solution matters for parity, so any solver iterated to a tight
 * residual serves.  IC(0) exists for the M-matrices this path produces. */
typedef struct {
  int64_t n;
  const int64_t *rp, *ci;
  const double *v;
  int64_t *lrp, *lci; /* lower triangle incl. diagonal (last in row) */
  double *lv;
} orc_ic0;

static int ic0_build(orc_ic0 *f, int64_t n, const int64_t *rp, const int64_t *ci,
                     const double *v) {
  f->n = n; f->rp = rp; f->ci = ci; f->v = v;
  f->lrp = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  int64_t c = 0;
  f->lrp[0] = 0;
  for (int64_t i = 0; i < n; i++) {
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) if (ci[k] <= i) c++;
    f->lrp[i + 1] = c;
  }
  f->lci = (int64_t *)malloc(sizeof(int64_t) * (c + 1));
  f->lv = (double *)malloc(sizeof(double) * (c + 1));
  double *w = (double *)calloc((size_t)n, sizeof(double));
  int64_t *mark = (int64_t *)malloc(sizeof(int64_t) * n);
  for (int64_t i = 0; i < n; i++) mark[i] = -1;
  int ok = 1;
  for (int64_t i = 0; i < n; i++) {
    int64_t p = f->lrp[i];
    for (int64_t k = rp[i]; k < rp[i + 1]; k++)
      if (ci[k] <= i) { f->lci[p] = ci[k]; f->lv[p] = v[k]; p++; }
    /* row i of L: columns sorted ascending, diagonal last */
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1]; q++) {
      w[f->lci[q]] = f->lv[q];
      mark[f->lci[q]] = i;
    }
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1] - 1; q++) {
      int64_t kcol = f->lci[q];
      double s = w[kcol];
      /* subtract sum_{m<k} L_im L_km over row k of L */
      for (int64_t r = f->lrp[kcol]; r < f->lrp[kcol + 1] - 1; r++) {
        int64_t mcol = f->lci[r];
        if (mark[mcol] == i) s -= w[mcol] * f->lv[r];
      }
      s /= f->lv[f->lrp[kcol + 1] - 1];
      w[kcol] = s;
    }
    double d = w[i];
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1] - 1; q++) d -= w[f->lci[q]] * w[f->lci[q]];
    if (!(d > 0.0)) { ok = 0; d = 1e-300; }
    w[i] = sqrt(d);
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1]; q++) f->lv[q] = w[f->lci[q]];
  }
  free(w); free(mark);
  return ok;
}

static void ic0_apply(const orc_ic0 *f, const double *r, double *z) {
  int64_t n = f->n;
  for (int64_t i = 0; i < n; i++) { /* L y = r */
    double s = r[i];
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1] - 1; q++) s -= f->lv[q] * z[f->lci[q]];
    z[i] = s / f->lv[f->lrp[i + 1] - 1];
  }
  for (int64_t i = n - 1; i >= 0; i--) { /* L^T z = y, column sweep */
    z[i] /= f->lv[f->lrp[i + 1] - 1];
    double zi = z[i];
    for (int64_t q = f->lrp[i]; q < f->lrp[i + 1] - 1; q++) z[f->lci[q]] -= f->lv[q] * zi;
  }
}

static void ic0_free(orc_ic0 *f) { free(f->lrp); free(f->lci); free(f->lv); }

static void csr_mv(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
                   const double *x, double *y) {
  for (int64_t i = 0; i < n; i++) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

static double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Solves A x = b (x holds the initial guess).  Stops when
 * ||b - A x|| <= reltol * ||b||.  info[0] = iterations, info[1] = final
 * relative residual (recomputed explicitly).  Returns 0 on convergence,
 * 1 if maxiter was reached, 2 if IC(0) broke down. */
int orc_pcg(int64_t n, const int64_t *rp, const int64_t *ci, const double *v,
            const double *b, double *x, double reltol, int64_t maxiter, double *info) {
  orc_ic0 f;
  int icok = ic0_build(&f, n, rp, ci, v);
  double *r = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *p = (double *)malloc(sizeof(double) * n), *q = (double *)malloc(sizeof(double) * n);
  csr_mv(n, rp, ci, v, x, q);
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
  double bn = sqrt(dot(n, b, b));
  if (bn == 0.0) bn = 1.0;
  ic0_apply(&f, r, z);
  memcpy(p, z, sizeof(double) * n);
  double rz = dot(n, r, z);
  int64_t it = 0;
  int status = 1;
  for (; it < maxiter; it++) {
    if (sqrt(dot(n, r, r)) <= reltol * bn) { status = 0; break; }
    csr_mv(n, rp, ci, v, p, q);
    double alpha = rz / dot(n, p, q);
    for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    ic0_apply(&f, r, z);
    double rz2 = dot(n, r, z);
    double beta = rz2 / rz;
    rz = rz2;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  if (status && sqrt(dot(n, r, r)) <= reltol * bn) status = 0;
  csr_mv(n, rp, ci, v, x, q);
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
  info[0] = (double)it;
  info[1] = sqrt(dot(n, r, r)) / bn;
  free(r); free(z); free(p); free(q);
  ic0_free(&f);
  return icok ? status : 2;
}

/* ---- harmonic-mean edge conductivity and its pullback ----
 * K_e = hm(exp(logK_a), exp(logK_b)), hm(x,y) = 2/(1/x+1/y)
 * (src/richards.jl:50-52).  Pullback: dlogK_c += dK_e * dK_e/dlogK_c with
 * dK_e/dlogK_a = K_e * k_b / (k_a + k_b), accumulated over edges in list
 * order (the scatter-add Zygote performs for the getindex/broadcast). */
void orc_hm_forward(int64_t ne, const int64_t *nb, const double *logK, double *Ke) {
  for (int64_t e = 0; e < ne; e++) {
    double ka = exp(logK[nb[2 * e]]), kb = exp(logK[nb[2 * e + 1]]);
    Ke[e] = 2.0 / (1.0 / ka + 1.0 / kb);
  }
}

void orc_hm_pullback(int64_t n, int64_t ne, const int64_t *nb, const double *logK,
                     const double *dKe, double *dlogK) {
  for (int64_t i = 0; i < n; i++) dlogK[i] = 0.0;
  for (int64_t e = 0; e < ne; e++) {
    int64_t a = nb[2 * e], b = nb[2 * e + 1];
    double ka = exp(logK[a]), kb = exp(logK[b]);
    double Ke = 2.0 / (1.0 / ka + 1.0 / kb);
    dlogK[a] += dKe[e] * Ke * kb / (ka + kb);
    dlogK[b] += dKe[e] * Ke * ka / (ka + kb);
  }
}

/* ---- geometric-mean edge conductivity and its pullback ----
 * K_e = exp(0.5 (logK_a + logK_b)), the logK -> K_e map of
 * examples/gw_steadystate_3d/inputdeck.jl:51.  Pullback:
 * dlogK_c += dK_e * K_e / 2 for both ends, over edges in list order. */
void orc_geo_forward(int64_t ne, const int64_t *nb, const double *logK, double *Ke) {
  for (int64_t e = 0; e < ne; e++) Ke[e] = exp(0.5 * (logK[nb[2 * e]] + logK[nb[2 * e + 1]]));
}

void orc_geo_pullback(int64_t n, int64_t ne, const int64_t *nb, const double *logK,
                      const double *dKe, double *dlogK) {
  for (int64_t i = 0; i < n; i++) dlogK[i] = 0.0;
  for (int64_t e = 0; e < ne; e++) {
    double g = dKe[e] * 0.5 * exp(0.5 * (logK[nb[2 * e]] + logK[nb[2 * e + 1]]));
    dlogK[nb[2 * e]] += g;
    dlogK[nb[2 * e + 1]] += g;
  }
}

/* ---- steady state: src/groundwater.jl:75-83 ----
 * Assemble at hfree = 0 with Ss = V = 1: b = -r(0), A = dr/dh; solve;
 * h = hfree on free cells, dirichleths on Dirichlet cells.
 * hfree_out (nf) keeps the free solution for the pullback.  info as orc_pcg. */
int orc_steady(int64_t n, int64_t ne, const int64_t *nb, const double *K,
               const double *aol, int64_t nd, const int64_t *dnodes,
               const double *dh, const double *Q, double reltol, int64_t maxiter,
               double *h_out, double *hfree_out, double *info) {
  uint8_t *isfree = (uint8_t *)malloc((size_t)n);
  int64_t *n2f = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *f2n = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t nf = orc_freenodes(n, nd, dnodes, isfree, n2f, f2n);
  double *ones = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) ones[i] = 1.0;
  double *zero = (double *)calloc((size_t)nf + 1, sizeof(double));
  double *b = (double *)malloc(sizeof(double) * (nf + 1));
  orc_residuals(n, ne, nb, K, aol, isfree, n2f, zero, dh, Q, ones, ones, nf, b);
  for (int64_t j = 0; j < nf; j++) b[j] = -b[j];
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (nf + 1));
  int64_t nnz = orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, rp, NULL, NULL);
  int64_t *ci = (int64_t *)malloc(sizeof(int64_t) * (nnz + 1));
  double *v = (double *)malloc(sizeof(double) * (nnz + 1));
  orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, rp, ci, v);
  for (int64_t j = 0; j < nf; j++) hfree_out[j] = 0.0;
  int st = orc_pcg(nf, rp, ci, v, b, hfree_out, reltol, maxiter, info);
  for (int64_t i = 0; i < n; i++) h_out[i] = isfree[i] ? hfree_out[n2f[i]] : dh[i];
  free(isfree); free(n2f); free(f2n); free(ones); free(zero); free(b);
  free(rp); free(ci); free(v);
  return st;
}

/* ---- rrule pullback: src/groundwater.jl:93-107 ----
 * lambda = A'^{-1} delta[free]; the three tangents are the transposed
 * Jacobian products the reference forms, written out per edge orientation:
 *   dKs  = -(dr/dKs)' lambda      dr_j1/dK_e = (h~_1 - h~_2) aol_e / s_1
 *   ddh  = -(dr/ddh)' lambda + delta .* !isfree,   dr_j1/ddh_2 = -K_e aol_e / s_1
 *   dQs  = -(dr/dQs)' lambda      dr_j/dQ_i = -1 / s_i
 * with s = Ss*V = 1 as in the steady solve. */
int orc_steady_pullback(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                        const double *aol, int64_t nd, const int64_t *dnodes,
                        const double *dh, const double *hfree, const double *delta,
                        double reltol, int64_t maxiter, double *dK, double *ddh,
                        double *dQ, double *lambda_out, double *info) {
  uint8_t *isfree = (uint8_t *)malloc((size_t)n);
  int64_t *n2f = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *f2n = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t nf = orc_freenodes(n, nd, dnodes, isfree, n2f, f2n);
  double *ones = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) ones[i] = 1.0;
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (nf + 1));
  int64_t nnz = orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, rp, NULL, NULL);
  int64_t *ci = (int64_t *)malloc(sizeof(int64_t) * (nnz + 1));
  double *v = (double *)malloc(sizeof(double) * (nnz + 1));
  orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, rp, ci, v);
  /* A' as CSR: transpose explicitly (A is symmetric here, but the reference
   * solves with SparseMatrixCSC(A') and so does this restatement). */
  int64_t *trp = (int64_t *)calloc((size_t)nf + 1, sizeof(int64_t));
  int64_t *tci = (int64_t *)malloc(sizeof(int64_t) * (nnz + 1));
  double *tv = (double *)malloc(sizeof(double) * (nnz + 1));
  for (int64_t k = 0; k < nnz; k++) trp[ci[k] + 1]++;
  for (int64_t j = 0; j < nf; j++) trp[j + 1] += trp[j];
  int64_t *tp = (int64_t *)malloc(sizeof(int64_t) * (nf + 1));
  memcpy(tp, trp, sizeof(int64_t) * (nf + 1));
  for (int64_t j = 0; j < nf; j++)
    for (int64_t k = rp[j]; k < rp[j + 1]; k++) { tci[tp[ci[k]]] = j; tv[tp[ci[k]]++] = v[k]; }
  double *rhs = (double *)malloc(sizeof(double) * (nf + 1));
  for (int64_t j = 0; j < nf; j++) rhs[j] = delta[f2n[j]];
  for (int64_t j = 0; j < nf; j++) lambda_out[j] = 0.0;
  int st = orc_pcg(nf, trp, tci, tv, rhs, lambda_out, reltol, maxiter, info);
  const double *lam = lambda_out;
  for (int64_t e = 0; e < ne; e++) dK[e] = 0.0;
  for (int64_t i = 0; i < n; i++) {
    ddh[i] = isfree[i] ? 0.0 : delta[i];
    dQ[i] = isfree[i] ? lam[n2f[i]] : 0.0; /* -(-1/s) lambda */
  }
  for (int64_t e = 0; e < ne; e++) {
    int64_t a = nb[2 * e], b = nb[2 * e + 1];
    for (int o = 0; o < 2; o++) {
      int64_t n1 = o ? b : a, n2 = o ? a : b;
      if (!isfree[n1]) continue;
      int64_t j1 = n2f[n1];
      double h2 = isfree[n2] ? hfree[n2f[n2]] : dh[n2];
      dK[e] -= (hfree[j1] - h2) * aol[e] * lam[j1];
      if (!isfree[n2]) ddh[n2] += K[e] * aol[e] * lam[j1];
    }
  }
  free(isfree); free(n2f); free(f2n); free(ones); free(rp); free(ci); free(v);
  free(trp); free(tci); free(tv); free(tp); free(rhs);
  return st;
}

/* ---- transient, backward Euler with a given step sequence ----
 * The reference integrates du/dt = f(u) = -groundwater_residuals(u, ..., Ss, V)
 * (examples/gw_transient_3d/inputdeck.jl:78-85) with DifferentiableBackwardEuler.
 * For this linear f, one backward-Euler step u_n = u_{n-1} + dt_n f(u_n) is
 *     (S + dt_n A) u_n = S u_{n-1} + dt_n b(t_n),   S = diag(Ss V) on free cells,
 * with A = dr/dh and b = -r(0) assembled at Ss = V = 1.  Q(t_n) = qscale[n] Q;
 * dts[n-1] is the length of step n.
 * states (nsteps+1) x nf row-major receives u_0 .. u_N. */
static void assemble_sys(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                         const double *aol, const uint8_t *isfree, const int64_t *n2f,
                         int64_t nf, int64_t **rp, int64_t **ci, double **v) {
  double *ones = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) ones[i] = 1.0;
  *rp = (int64_t *)malloc(sizeof(int64_t) * (nf + 1));
  int64_t nnz = orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, *rp, NULL, NULL);
  *ci = (int64_t *)malloc(sizeof(int64_t) * (nnz + 1));
  *v = (double *)malloc(sizeof(double) * (nnz + 1));
  orc_jacobian_h(n, ne, nb, K, aol, isfree, n2f, ones, ones, nf, *rp, *ci, *v);
  free(ones);
}

int orc_transient(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                  const double *aol, int64_t nd, const int64_t *dnodes, const double *dh,
                  const double *Q, const double *Ss, const double *V, const double *u0,
                  int64_t nsteps, const double *dts, const double *qscale, double reltol,
                  int64_t maxiter, double *states, double *info) {
  uint8_t *isfree = (uint8_t *)malloc((size_t)n);
  int64_t *n2f = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *f2n = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t nf = orc_freenodes(n, nd, dnodes, isfree, n2f, f2n);
  int64_t *rp, *ci;
  double *v;
  assemble_sys(n, ne, nb, K, aol, isfree, n2f, nf, &rp, &ci, &v);
  double *s = (double *)malloc(sizeof(double) * (nf + 1));
  for (int64_t j = 0; j < nf; j++) s[j] = Ss[f2n[j]] * V[f2n[j]];
  double *mv = (double *)malloc(sizeof(double) * (rp[nf] + 1));
  /* b split as Q part and Dirichlet part: b(t) = qscale * Q + bD */
  double *ones = (double *)malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) ones[i] = 1.0;
  double *zero = (double *)calloc((size_t)nf + 1, sizeof(double));
  double *zq = (double *)calloc((size_t)n, sizeof(double));
  double *bD = (double *)malloc(sizeof(double) * (nf + 1));
  orc_residuals(n, ne, nb, K, aol, isfree, n2f, zero, dh, zq, ones, ones, nf, bD);
  for (int64_t j = 0; j < nf; j++) bD[j] = -bD[j];
  double *rhs = (double *)malloc(sizeof(double) * (nf + 1));
  memcpy(states, u0, sizeof(double) * nf);
  int st = 0;
  double itsum = 0.0, worst = 0.0, inf[2];
  for (int64_t t = 1; t <= nsteps; t++) {
    const double *up = states + (t - 1) * nf;
    double *u = states + t * nf;
    double dt = dts[t - 1];
    /* M_t = S + dt_t A */
    for (int64_t j = 0; j < nf; j++)
      for (int64_t k = rp[j]; k < rp[j + 1]; k++) mv[k] = dt * v[k] + (ci[k] == j ? s[j] : 0.0);
    for (int64_t j = 0; j < nf; j++)
      rhs[j] = s[j] * up[j] + dt * (qscale[t] * Q[f2n[j]] + bD[j]);
    memcpy(u, up, sizeof(double) * nf);
    int r = orc_pcg(nf, rp, ci, mv, rhs, u, reltol, maxiter, inf);
    if (r) st = r;
    itsum += inf[0];
    if (inf[1] > worst) worst = inf[1];
  }
  info[0] = itsum;
  info[1] = worst;
  free(isfree); free(n2f); free(f2n); free(rp); free(ci); free(v); free(s); free(mv);
  free(ones); free(zero); free(zq); free(bD); free(rhs);
  return st;
}

/* Discrete adjoint of orc_transient for J = J(u_0..u_N) with dJdu
 * ((nsteps+1) x nf, row n = dJ/du_n), M_n = S + dt_n A (symmetric):
 *   M_N mu_N = dJ/du_N,   M_n mu_n = dJ/du_n + S mu_{n+1}   (n = N-1 .. 1)
 *   dJ/du_0  = dJ/du_0 + S mu_1
 *   dJ/dK_e  = -sum_n dt_n aol_e (u~n_a - u~n_b)(mu~n_a - mu~n_b)
 *   dJ/dQ_i  = sum_n dt_n qscale[n] mu_n[j(i)] on free cells, 0 on Dirichlet cells
 *   dJ/ddh_o = sum_n dt_n sum over edges e = (o, free i) of K_e aol_e mu_n[j(i)]
 *              on Dirichlet cells (the tangent of dt_n b_D, b_D_j = sum K_e aol_e dh_o),
 *              0 on free cells
 * where ~ puts dirichleths (states) or 0 (multipliers) on Dirichlet cells.
 * ddh and du0 may be NULL. */
int orc_transient_adjoint(int64_t n, int64_t ne, const int64_t *nb, const double *K,
                          const double *aol, int64_t nd, const int64_t *dnodes,
                          const double *dh, const double *Ss, const double *V,
                          int64_t nsteps, const double *dts, const double *qscale,
                          const double *states, const double *dJdu, double reltol,
                          int64_t maxiter, double *dK, double *dQ, double *ddh, double *du0,
                          double *info) {
  uint8_t *isfree = (uint8_t *)malloc((size_t)n);
  int64_t *n2f = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t *f2n = (int64_t *)malloc(sizeof(int64_t) * n);
  int64_t nf = orc_freenodes(n, nd, dnodes, isfree, n2f, f2n);
  int64_t *rp, *ci;
  double *v;
  assemble_sys(n, ne, nb, K, aol, isfree, n2f, nf, &rp, &ci, &v);
  double *s = (double *)malloc(sizeof(double) * (nf + 1));
  for (int64_t j = 0; j < nf; j++) s[j] = Ss[f2n[j]] * V[f2n[j]];
  double *mv = (double *)malloc(sizeof(double) * (rp[nf] + 1));
  double *mu = (double *)calloc((size_t)nf + 1, sizeof(double));
  double *rhs = (double *)malloc(sizeof(double) * (nf + 1));
  for (int64_t e = 0; e < ne; e++) dK[e] = 0.0;
  for (int64_t i = 0; i < n; i++) dQ[i] = 0.0;
  if (ddh)
    for (int64_t i = 0; i < n; i++) ddh[i] = 0.0;
  int st = 0;
  double itsum = 0.0, worst = 0.0, inf[2];
  for (int64_t t = nsteps; t >= 1; t--) {
    double dt = dts[t - 1];
    for (int64_t j = 0; j < nf; j++)
      for (int64_t k = rp[j]; k < rp[j + 1]; k++) mv[k] = dt * v[k] + (ci[k] == j ? s[j] : 0.0);
    for (int64_t j = 0; j < nf; j++) rhs[j] = dJdu[t * nf + j] + s[j] * mu[j];
    int r = orc_pcg(nf, rp, ci, mv, rhs, mu, reltol, maxiter, inf); /* M symmetric */
    if (r) st = r;
    itsum += inf[0];
    if (inf[1] > worst) worst = inf[1];
    const double *u = states + t * nf;
    for (int64_t e = 0; e < ne; e++) {
      int64_t a = nb[2 * e], b = nb[2 * e + 1];
      double ua = isfree[a] ? u[n2f[a]] : dh[a], ub = isfree[b] ? u[n2f[b]] : dh[b];
      double ma = isfree[a] ? mu[n2f[a]] : 0.0, mb = isfree[b] ? mu[n2f[b]] : 0.0;
      dK[e] -= dt * aol[e] * (ua - ub) * (ma - mb);
      if (ddh && a != b) {
        if (!isfree[a] && isfree[b]) ddh[a] += dt * K[e] * aol[e] * mb;
        if (!isfree[b] && isfree[a]) ddh[b] += dt * K[e] * aol[e] * ma;
      }
    }
    for (int64_t j = 0; j < nf; j++) dQ[f2n[j]] += dt * qscale[t] * mu[j];
  }
  if (du0)
    for (int64_t j = 0; j < nf; j++) du0[j] = dJdu[j] + (nsteps > 0 ? s[j] * mu[j] : 0.0);
  info[0] = itsum;
  info[1] = worst;
  free(isfree); free(n2f); free(f2n); free(rp); free(ci); free(v); free(s); free(mv);
  free(mu); free(rhs);
  return st;
}
