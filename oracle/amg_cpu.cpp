// amg_cpu.cpp -- FP64 multi-core CPU restatement of the steady groundwater
// forward + adjoint evaluation with smoothed-aggregation AMG-PCG.
//
// TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the
// GPU (and tests/ check).  Nothing in the product (dpfehm.jl_amd/) links,
// loads or calls it.  It is a port, not the reference: the reference is pure
// Julia (src/groundwater.jl:75-109 with AlgebraicMultigrid.jl, absent here),
// and Julia is not installed.
//
// One evaluation of the inverse-problem objective, all on the host:
//   K_e = hm(exp(logK_a), exp(logK_b))                      (src/richards.jl:50-52)
//   A = d r/d h and b = -r(0) at Ss = V = 1                 (src/groundwater.jl:1-21,75-79)
//   h = A^-1 b by CG preconditioned with one SA-AMG V-cycle (the role of
//       amg_solver, src/groundwater.jl:45-49), scattered to every cell
//   J = sum_obs (h - h_obs)^2, delta = dJ/dh
//   lambda = A^-1 delta_free (A symmetric), dK_e = -aol_e (h~_a - h~_b)(lambda~_a - lambda~_b)
//                                                         (src/groundwater.jl:93-101)
//   dJ/dlogK_c = sum_{e at c} dK_e K_e k_other / (k_c + k_other)
// Same method as the GPU library: aggregation of the matrix graph (every
// coupling strong, three-phase), P = (I - omega D_l1^-1 A) T (omega 1.5 on
// the finest level, 2 below), R = P', Galerkin
// coarse operators, l1-Jacobi Chebyshev smoothing (degree 1 on the finest
// and the second level, 2 below, interval [1/15, 1]), dense Cholesky of the
// coarsest level.
// The conductivity-independent part (aggregates, every sparsity pattern) is
// built once per mesh; each evaluation redoes all numeric work.  OpenMP over
// rows throughout.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using I = int64_t;

struct Csr {
  I n = 0, m = 0;
  std::vector<I> rp;
  std::vector<int32_t> ci;
  std::vector<double> v;
};

struct Level {
  Csr A;                       // level operator (pattern; values per evaluation)
  std::vector<double> invd;    // 1 / (a_ii + sum_j |a_ij|)
  std::vector<int32_t> agg;    // row -> coarse row
  Csr P, R;                    // prolongator and R = P'
  std::vector<int32_t> pslot;  // per A entry: slot of agg(col) in P's row
  std::vector<int32_t> pself;  // per row: slot of agg(row) in P's row
  std::vector<I> rsrc;         // per R entry: the P entry it mirrors
  Csr AP;                      // A P (pattern; values per evaluation)
  std::vector<double> x, b, r, d, t;  // V-cycle work vectors
};

struct Handle {
  I n = 0, ne = 0, nf = 0;
  std::vector<int32_t> ea, eb;  // edges, 0-based
  std::vector<double> aol;
  std::vector<int32_t> n2f, f2n;
  std::vector<I> inc_ptr;       // cell -> incident edges (self-edges left out)
  std::vector<int32_t> inc_edge, inc_other, inc_slot;  // slot in the free row, -1: other is Dirichlet
  std::vector<int32_t> diag;    // slot of the diagonal in each fine row
  std::vector<Level> lev;
  int nc_dense = 0;
  std::vector<double> chol;     // coarsest level, dense Cholesky factor (row-major, lower)
  int threads = 1;
};

// Symbolic C = A B (sorted columns) with a thread-local marker, row-parallel.
void spgemm_pattern(const Csr& A, const Csr& B, Csr& C) {
  C.n = A.n;
  C.m = B.m;
  C.rp.assign(A.n + 1, 0);
  std::vector<std::vector<int32_t>> rows(A.n);
#pragma omp parallel
  {
    std::vector<I> mark(B.m, -1);
#pragma omp for schedule(dynamic, 1024)
    for (I i = 0; i < A.n; i++) {
      std::vector<int32_t>& row = rows[i];
      for (I p = A.rp[i]; p < A.rp[i + 1]; p++) {
        I k = A.ci[p];
        for (I q = B.rp[k]; q < B.rp[k + 1]; q++) {
          int32_t c = B.ci[q];
          if (mark[c] != i) {
            mark[c] = i;
            row.push_back(c);
          }
        }
      }
      std::sort(row.begin(), row.end());
    }
  }
  for (I i = 0; i < A.n; i++) C.rp[i + 1] = C.rp[i] + (I)rows[i].size();
  C.ci.resize(C.rp[A.n]);
#pragma omp parallel for schedule(static)
  for (I i = 0; i < A.n; i++) std::copy(rows[i].begin(), rows[i].end(), C.ci.begin() + C.rp[i]);
  C.v.assign(C.ci.size(), 0.0);
}

// Numeric C = A B on C's pattern.
void spgemm_values(const Csr& A, const Csr& B, Csr& C) {
#pragma omp parallel
  {
    std::vector<int32_t> pos(B.m, 0);
#pragma omp for schedule(dynamic, 1024)
    for (I i = 0; i < A.n; i++) {
      for (I p = C.rp[i]; p < C.rp[i + 1]; p++) {
        pos[C.ci[p]] = (int32_t)(p - C.rp[i]);
        C.v[p] = 0.0;
      }
      double* out = C.v.data() + C.rp[i];
      for (I p = A.rp[i]; p < A.rp[i + 1]; p++) {
        double a = A.v[p];
        I k = A.ci[p];
        for (I q = B.rp[k]; q < B.rp[k + 1]; q++) out[pos[B.ci[q]]] += a * B.v[q];
      }
    }
  }
}

// Three-phase aggregation of the graph of A (every coupling strong).
I aggregate(const Csr& A, std::vector<int32_t>& agg) {
  I n = A.n, nc = 0;
  agg.assign(n, -1);
  for (I i = 0; i < n; i++) {
    bool ok = true;
    for (I p = A.rp[i]; p < A.rp[i + 1] && ok; p++)
      if (A.ci[p] != i && agg[A.ci[p]] != -1) ok = false;
    if (!ok || agg[i] != -1) continue;
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++) agg[A.ci[p]] = (int32_t)nc;
    agg[i] = (int32_t)nc++;
  }
  std::vector<int32_t> agg1 = agg;
  for (I i = 0; i < n; i++) {
    if (agg[i] != -1) continue;
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++)
      if (agg1[A.ci[p]] != -1) {
        agg[i] = agg1[A.ci[p]];
        break;
      }
  }
  for (I i = 0; i < n; i++) {
    if (agg[i] != -1) continue;
    agg[i] = (int32_t)nc;
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++)
      if (agg[A.ci[p]] == -1) agg[A.ci[p]] = (int32_t)nc;
    nc++;
  }
  return nc;
}

// Symbolic transfer operators of level L and the next level's pattern.
void build_transfer(Level& L, Level& next) {
  const Csr& A = L.A;
  I n = A.n;
  I nc = aggregate(A, L.agg);
  Csr& P = L.P;
  P.n = n;
  P.m = nc;
  P.rp.assign(n + 1, 0);
  std::vector<std::vector<int32_t>> rows(n);
#pragma omp parallel for schedule(dynamic, 2048)
  for (I i = 0; i < n; i++) {
    std::vector<int32_t>& row = rows[i];
    row.push_back(L.agg[i]);
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++) row.push_back(L.agg[A.ci[p]]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
  }
  for (I i = 0; i < n; i++) P.rp[i + 1] = P.rp[i] + (I)rows[i].size();
  P.ci.resize(P.rp[n]);
  P.v.assign(P.ci.size(), 0.0);
  L.pslot.resize(A.ci.size());
  L.pself.resize(n);
#pragma omp parallel for schedule(dynamic, 2048)
  for (I i = 0; i < n; i++) {
    std::copy(rows[i].begin(), rows[i].end(), P.ci.begin() + P.rp[i]);
    auto b = rows[i].begin(), e = rows[i].end();
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++)
      L.pslot[p] = (int32_t)(std::lower_bound(b, e, L.agg[A.ci[p]]) - b);
    L.pself[i] = (int32_t)(std::lower_bound(b, e, L.agg[i]) - b);
  }
  // R = P'
  Csr& R = L.R;
  R.n = nc;
  R.m = n;
  R.rp.assign(nc + 1, 0);
  for (I q = 0; q < (I)P.ci.size(); q++) R.rp[P.ci[q] + 1]++;
  for (I k = 0; k < nc; k++) R.rp[k + 1] += R.rp[k];
  R.ci.resize(P.ci.size());
  R.v.assign(P.ci.size(), 0.0);
  L.rsrc.resize(P.ci.size());
  std::vector<I> fill(R.rp.begin(), R.rp.end() - 1);
  for (I i = 0; i < n; i++)
    for (I q = P.rp[i]; q < P.rp[i + 1]; q++) {
      I dst = fill[P.ci[q]]++;
      R.ci[dst] = (int32_t)i;
      L.rsrc[dst] = q;
    }
  spgemm_pattern(A, P, L.AP);
  next = Level();
  spgemm_pattern(R, L.AP, next.A);
}

struct Cheb {
  double c0 = 0;
  std::vector<double> c1, c2;
  void make(int degree, double ratio) {
    double lmax = 1.0, lmin = lmax / ratio;
    double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    double sigma = theta / delta, rho = 1.0 / sigma;
    c0 = 1.0 / theta;
    c1.assign(degree, 0.0);
    c2.assign(degree, 0.0);
    c2[0] = c0;
    for (int k = 1; k < degree; k++) {
      double rn = 1.0 / (2.0 * sigma - rho);
      c1[k] = rn * rho;
      c2[k] = 2.0 * rn / delta;
      rho = rn;
    }
  }
};

void spmv(const Csr& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (I i = 0; i < A.n; i++) {
    double s = 0.0;
    for (I p = A.rp[i]; p < A.rp[i + 1]; p++) s += A.v[p] * x[A.ci[p]];
    y[i] = s;
  }
}

double dot(I n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (I i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// Chebyshev sweeps k = k0 .. m-1 on A x = b (l1-Jacobi), d carried.
void cheb_sweeps(Level& L, const Cheb& ch, int k0, int m, const double* b, double* x) {
  I n = L.A.n;
  for (int k = k0; k < m; k++) {
    double c1 = k == 0 ? 0.0 : ch.c1[k];
    double c2 = k == 0 ? ch.c0 : ch.c2[k];
#pragma omp parallel for schedule(static)
    for (I i = 0; i < n; i++) {
      double ax = 0.0;
      for (I p = L.A.rp[i]; p < L.A.rp[i + 1]; p++) ax += L.A.v[p] * x[L.A.ci[p]];
      double dn = c2 * L.invd[i] * (b[i] - ax);
      if (c1 != 0.0) dn += c1 * L.d[i];
      L.d[i] = dn;
      L.t[i] = x[i] + dn;
    }
    std::memcpy(x, L.t.data(), n * sizeof(double));
  }
}

void vcycle(Handle& H, const std::vector<Cheb>& chs, size_t l, const double* b, double* x) {
  Level& L = H.lev[l];
  I n = L.A.n;
  if (l + 1 == H.lev.size()) {
    // dense Cholesky solve of the coarsest level
    int nc = H.nc_dense;
    const double* C = H.chol.data();
    for (int i = 0; i < nc; i++) {
      double s = b[i];
      for (int k = 0; k < i; k++) s -= C[(I)i * nc + k] * x[k];
      x[i] = s / C[(I)i * nc + i];
    }
    for (int i = nc - 1; i >= 0; i--) {
      double s = x[i];
      for (int k = i + 1; k < nc; k++) s -= C[(I)k * nc + i] * x[k];
      x[i] = s / C[(I)i * nc + i];
    }
    return;
  }
  const Cheb& ch = chs[l];
  int m = (int)ch.c1.size();
  // pre-smoothing from the zero guess
#pragma omp parallel for schedule(static)
  for (I i = 0; i < n; i++) {
    L.d[i] = ch.c0 * L.invd[i] * b[i];
    x[i] = L.d[i];
  }
  cheb_sweeps(L, ch, 1, m, b, x);
  spmv(L.A, x, L.r.data());
#pragma omp parallel for schedule(static)
  for (I i = 0; i < n; i++) L.r[i] = b[i] - L.r[i];
  Level& N = H.lev[l + 1];
  spmv(L.R, L.r.data(), N.b.data());
  vcycle(H, chs, l + 1, N.b.data(), N.x.data());
#pragma omp parallel for schedule(static)
  for (I i = 0; i < n; i++) {
    double s = 0.0;
    for (I p = L.P.rp[i]; p < L.P.rp[i + 1]; p++) s += L.P.v[p] * N.x[L.P.ci[p]];
    x[i] += s;
  }
  cheb_sweeps(L, ch, 0, m, b, x);
}

// Numeric refresh of the hierarchy for the fine values in lev[0].A.
int refresh(Handle& H) {
  for (size_t l = 0; l < H.lev.size(); l++) {
    Level& L = H.lev[l];
    I n = L.A.n;
    L.invd.resize(n);
#pragma omp parallel for schedule(static)
    for (I i = 0; i < n; i++) {
      double s = 0.0;
      for (I p = L.A.rp[i]; p < L.A.rp[i + 1]; p++) s += std::fabs(L.A.v[p]);
      L.invd[i] = 1.0 / s;  // a_ii > 0: |a_ii| + sum |a_ij|
    }
    if (l + 1 == H.lev.size()) break;
#pragma omp parallel for schedule(static)
    for (I i = 0; i < n; i++) {
      double* pr = L.P.v.data() + L.P.rp[i];
      for (I q = L.P.rp[i]; q < L.P.rp[i + 1]; q++) pr[q - L.P.rp[i]] = 0.0;
      // omega = 4 / (3 rho) with the estimates rho = 0.89 (finest level) and
      // 0.67 (Galerkin levels) of rho(D_l1^-1 A), as the GPU library uses
      double f = (l == 0 ? 1.5 : 2.0) * L.invd[i];
      for (I p = L.A.rp[i]; p < L.A.rp[i + 1]; p++) pr[L.pslot[p]] -= f * L.A.v[p];
      pr[L.pself[i]] += 1.0;
    }
#pragma omp parallel for schedule(static)
    for (I q = 0; q < (I)L.R.v.size(); q++) L.R.v[q] = L.P.v[L.rsrc[q]];
    spgemm_values(L.A, L.P, L.AP);
    spgemm_values(L.R, L.AP, H.lev[l + 1].A);
  }
  // dense Cholesky of the coarsest operator
  Level& C = H.lev.back();
  int nc = (int)C.A.n;
  H.nc_dense = nc;
  H.chol.assign((I)nc * nc, 0.0);
  for (I i = 0; i < nc; i++)
    for (I p = C.A.rp[i]; p < C.A.rp[i + 1]; p++) H.chol[i * nc + C.A.ci[p]] = C.A.v[p];
  double* W = H.chol.data();
  for (int j = 0; j < nc; j++) {
    double s = W[(I)j * nc + j];
    for (int k = 0; k < j; k++) s -= W[(I)j * nc + k] * W[(I)j * nc + k];
    if (!(s > 0.0)) return 1;
    double d = std::sqrt(s);
    W[(I)j * nc + j] = d;
    for (int i = j + 1; i < nc; i++) {
      double t = W[(I)i * nc + j];
      for (int k = 0; k < j; k++) t -= W[(I)i * nc + k] * W[(I)j * nc + k];
      W[(I)i * nc + j] = t / d;
    }
  }
  return 0;
}

// PCG on the fine level to ||b - A x|| <= reltol ||b||; returns iterations or -1.
int pcg(Handle& H, const std::vector<Cheb>& chs, const double* b, double* x, double reltol,
        int maxiter) {
  Level& L = H.lev[0];
  I n = L.A.n;
  std::vector<double> r(b, b + n), z(n), p(n), q(n);
  std::fill(x, x + n, 0.0);
  double bb = dot(n, b, b), stop = reltol * reltol * bb;
  if (bb == 0.0) return 0;
  vcycle(H, chs, 0, r.data(), z.data());
  double rz = dot(n, r.data(), z.data());
  std::copy(z.begin(), z.end(), p.begin());
  for (int it = 1; it <= maxiter; it++) {
    spmv(L.A, p.data(), q.data());
    double alpha = rz / dot(n, p.data(), q.data());
    double rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
    for (I i = 0; i < n; i++) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rr += r[i] * r[i];
    }
    if (rr <= stop) return it;
    vcycle(H, chs, 0, r.data(), z.data());
    double rz2 = dot(n, r.data(), z.data()), beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (I i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
  }
  return -1;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

extern "C" {

// Mesh set-up (conductivity-independent).  nb: 2 ne int64 pairs, 1-based;
// dnodes: 1-based.  nthreads <= 0: OpenMP's default.  Returns NULL on an
// invalid mesh.
void* acpu_create(int64_t n, int64_t ne, const int64_t* nb, const double* aol, int64_t nd,
                  const int64_t* dnodes, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  Handle* H = new Handle();
  H->threads = omp_get_max_threads();
  H->n = n;
  H->ne = ne;
  H->ea.resize(ne);
  H->eb.resize(ne);
  H->aol.assign(aol, aol + ne);
  for (I e = 0; e < ne; e++) {
    H->ea[e] = (int32_t)(nb[2 * e] - 1);
    H->eb[e] = (int32_t)(nb[2 * e + 1] - 1);
  }
  std::vector<uint8_t> isfree(n, 1);
  for (I k = 0; k < nd; k++) isfree[dnodes[k] - 1] = 0;
  H->n2f.assign(n, -1);
  for (I i = 0; i < n; i++)
    if (isfree[i]) {
      H->n2f[i] = (int32_t)H->f2n.size();
      H->f2n.push_back((int32_t)i);
    }
  I nf = H->nf = (I)H->f2n.size();
  H->inc_ptr.assign(n + 1, 0);
  for (I e = 0; e < ne; e++)
    if (H->ea[e] != H->eb[e]) {
      H->inc_ptr[H->ea[e] + 1]++;
      H->inc_ptr[H->eb[e] + 1]++;
    }
  for (I i = 0; i < n; i++) H->inc_ptr[i + 1] += H->inc_ptr[i];
  H->inc_edge.resize(H->inc_ptr[n]);
  H->inc_other.resize(H->inc_ptr[n]);
  H->inc_slot.assign(H->inc_ptr[n], -1);
  std::vector<I> pos(H->inc_ptr.begin(), H->inc_ptr.end() - 1);
  for (I e = 0; e < ne; e++) {
    int32_t a = H->ea[e], b = H->eb[e];
    if (a == b) continue;
    H->inc_edge[pos[a]] = (int32_t)e;
    H->inc_other[pos[a]++] = b;
    H->inc_edge[pos[b]] = (int32_t)e;
    H->inc_other[pos[b]++] = a;
  }
  H->lev.resize(1);
  Csr& A = H->lev[0].A;
  A.n = A.m = nf;
  A.rp.assign(nf + 1, 0);
  std::vector<std::vector<int32_t>> rows(nf);
#pragma omp parallel for schedule(dynamic, 2048)
  for (I j = 0; j < nf; j++) {
    I i = H->f2n[j];
    std::vector<int32_t>& row = rows[j];
    row.push_back((int32_t)j);
    for (I p = H->inc_ptr[i]; p < H->inc_ptr[i + 1]; p++)
      if (H->n2f[H->inc_other[p]] >= 0) row.push_back(H->n2f[H->inc_other[p]]);
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
  }
  for (I j = 0; j < nf; j++) {
    if (H->inc_ptr[H->f2n[j] + 1] == H->inc_ptr[H->f2n[j]]) {
      delete H;
      return nullptr;  // a free cell without edges: singular
    }
    A.rp[j + 1] = A.rp[j] + (I)rows[j].size();
  }
  A.ci.resize(A.rp[nf]);
  A.v.assign(A.ci.size(), 0.0);
  H->diag.resize(nf);
#pragma omp parallel for schedule(dynamic, 2048)
  for (I j = 0; j < nf; j++) {
    std::copy(rows[j].begin(), rows[j].end(), A.ci.begin() + A.rp[j]);
    auto b = rows[j].begin(), e = rows[j].end();
    H->diag[j] = (int32_t)(std::lower_bound(b, e, (int32_t)j) - b);
    I i = H->f2n[j];
    for (I p = H->inc_ptr[i]; p < H->inc_ptr[i + 1]; p++) {
      int32_t o = H->n2f[H->inc_other[p]];
      if (o >= 0) H->inc_slot[p] = (int32_t)(std::lower_bound(b, e, o) - b);
    }
  }
  // hierarchy patterns: coarsen to <= 128 rows or until a level shrinks by < 15 %
  while (H->lev.size() < 25 && H->lev.back().A.n > 128) {
    Level next;
    build_transfer(H->lev.back(), next);
    if ((double)next.A.n > 0.85 * (double)H->lev.back().A.n) {
      H->lev.back().agg.clear();
      break;
    }
    H->lev.push_back(std::move(next));
  }
  if (H->lev.back().A.n > 4096) {  // the dense coarsest solve would not fit
    delete H;
    return nullptr;
  }
  for (Level& L : H->lev) {
    I m = L.A.n;
    L.x.assign(m, 0.0);
    L.b.assign(m, 0.0);
    L.r.assign(m, 0.0);
    L.d.assign(m, 0.0);
    L.t.assign(m, 0.0);
  }
  return H;
}

void acpu_destroy(void* h) { delete (Handle*)h; }

// info: [0] levels, [1] fine nnz, [2] operator complexity, [3] threads
void acpu_info(void* h, double* info) {
  Handle* H = (Handle*)h;
  double nnz = 0;
  for (Level& L : H->lev) nnz += (double)L.A.ci.size();
  info[0] = (double)H->lev.size();
  info[1] = (double)H->lev[0].A.ci.size();
  info[2] = nnz / (double)H->lev[0].A.ci.size();
  info[3] = H->threads;
}

// One forward + adjoint evaluation (module header).  obs: nobs cells,
// 1-based; hobs: their observed heads.  Outputs J, dlogK (n) and h (n, may be
// NULL).  info: [0] forward PCG iterations, [1] adjoint, [2] ms in the
// numeric set-up (map, assembly, hierarchy), [3] ms in the forward solve, [4]
// ms in the adjoint solve and gradient.  Returns 0, or 1 on a solver failure.
int acpu_eval(void* h, const double* logK, const double* dh, const double* Q,
              const int64_t* obs, const double* hobs, int64_t nobs, double reltol,
              double* J, double* dlogK, double* hout, double* info) {
  Handle* H = (Handle*)h;
  I n = H->n, ne = H->ne, nf = H->nf;
  double t0 = now_ms();
  std::vector<double> k(n), Ke(ne);
#pragma omp parallel for schedule(static)
  for (I i = 0; i < n; i++) k[i] = std::exp(logK[i]);
#pragma omp parallel for schedule(static)
  for (I e = 0; e < ne; e++) Ke[e] = 2.0 / (1.0 / k[H->ea[e]] + 1.0 / k[H->eb[e]]);
  // assembly of A and b
  Csr& A = H->lev[0].A;
  std::vector<double> b(nf);
#pragma omp parallel for schedule(static)
  for (I j = 0; j < nf; j++) {
    I i = H->f2n[j];
    double* row = A.v.data() + A.rp[j];
    for (I p = A.rp[j]; p < A.rp[j + 1]; p++) row[p - A.rp[j]] = 0.0;
    double dsum = 0.0, bj = Q[i];
    for (I p = H->inc_ptr[i]; p < H->inc_ptr[i + 1]; p++) {
      int32_t e = H->inc_edge[p];
      double t = Ke[e] * H->aol[e];
      dsum += t;
      if (H->inc_slot[p] >= 0)
        row[H->inc_slot[p]] -= t;
      else
        bj += t * dh[H->inc_other[p]];
    }
    row[H->diag[j]] = dsum;
    b[j] = bj;
  }
  if (refresh(*H)) return 1;
  std::vector<Cheb> chs(H->lev.size());
  for (size_t l = 0; l < chs.size(); l++) chs[l].make(l <= 1 ? 1 : 2, 15.0);
  double t1 = now_ms();
  std::vector<double> hf(nf), hfull(n);
  int it1 = pcg(*H, chs, b.data(), hf.data(), reltol, 5000);
  if (it1 < 0) return 1;
#pragma omp parallel for schedule(static)
  for (I i = 0; i < n; i++) hfull[i] = H->n2f[i] >= 0 ? hf[H->n2f[i]] : dh[i];
  double t2 = now_ms();
  // objective and seed
  std::vector<double> delta(nf, 0.0);
  double jv = 0.0;
  for (I q = 0; q < nobs; q++) {
    I c = obs[q] - 1;
    double r = hfull[c] - hobs[q];
    jv += r * r;
    if (H->n2f[c] >= 0) delta[H->n2f[c]] += 2.0 * r;
  }
  std::vector<double> lam(nf);
  int it2 = pcg(*H, chs, delta.data(), lam.data(), reltol, 5000);
  if (it2 < 0) return 1;
  std::vector<double> dK(ne);
#pragma omp parallel for schedule(static)
  for (I e = 0; e < ne; e++) {
    int32_t a = H->ea[e], bb = H->eb[e];
    int32_t fa = H->n2f[a], fb = H->n2f[bb];
    double ha = hfull[a], hb = hfull[bb];
    double la = fa >= 0 ? lam[fa] : 0.0, lb = fb >= 0 ? lam[fb] : 0.0;
    dK[e] = -H->aol[e] * (ha - hb) * (la - lb);
  }
#pragma omp parallel for schedule(static)
  for (I c = 0; c < n; c++) {
    double s = 0.0;
    for (I p = H->inc_ptr[c]; p < H->inc_ptr[c + 1]; p++) {
      int32_t e = H->inc_edge[p];
      double ko = k[H->inc_other[p]];
      s += dK[e] * Ke[e] * ko / (k[c] + ko);
    }
    dlogK[c] = s;
  }
  double t3 = now_ms();
  *J = jv;
  if (hout) std::memcpy(hout, hfull.data(), n * sizeof(double));
  info[0] = it1;
  info[1] = it2;
  info[2] = t1 - t0;
  info[3] = t2 - t1;
  info[4] = t3 - t2;
  return 0;
}

}  // extern "C"
