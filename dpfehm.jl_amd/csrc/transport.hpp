// Advection-dispersion (src/transport.jl:2-33): the residual r(u) = J u + r0,
// affine in u, applied matrix-free over the cell incidence, and the pieces of
// backward Euler and its adjoint.
#pragma once

#include "kernels.hpp"

namespace dpf {

// Per-edge and per-cell inputs of one transport problem on the device.
struct TrIn {
  const double *vx, *vy, *vz, *D;  // per edge
  const double *ud, *Q, *V;        // per cell
  const double *cx, *cy, *cz;      // cell coordinates
  int upwind;
};

// The coefficients of the edge term at cell c towards its neighbour o
// (direction (c, o) of src/transport.jl:15-31): w = v . (x_c - x_o); the
// advection term -(w)(u_c - u_o) aol / V_c is active when (w > 0) == upwind;
// diffusion D (u_o - u_c) aol / V_c.  Returns the coefficient of u_c (adv
// part) and the diffusion coefficient.
__device__ __forceinline__ void tr_edge(const TrIn& in, const double* __restrict__ aol, int32_t e,
                                        int32_t c, int32_t o, double& adv, double& dif,
                                        double& w) {
  w = in.vx[e] * (in.cx[c] - in.cx[o]) + in.vy[e] * (in.cy[c] - in.cy[o]) +
      in.vz[e] * (in.cz[c] - in.cz[o]);
  bool act = (w > 0.0) == (in.upwind != 0);
  double s = aol[e] / in.V[c];
  adv = act ? -w * s : 0.0;
  dif = in.D[e] * s;
}

// y_c = alpha x_c + beta (J x)_c + gamma r0_c for every cell (Dirichlet cells
// d: J x = x, r0 = -ud).  r(u) = alpha 0, beta 1, gamma 1 at x = u; backward
// Euler's matrix I - dt J is alpha 1, beta -dt, gamma 0; its right-hand side
// u_k + dt r0 is alpha 1, beta 0, gamma dt.
__global__ void k_tr_rows(int64_t n, const int32_t* __restrict__ n2f,
                          const int64_t* __restrict__ inc_ptr,
                          const int32_t* __restrict__ inc_edge,
                          const int32_t* __restrict__ inc_other, const double* __restrict__ aol,
                          TrIn in, const double* __restrict__ x, double alpha, double beta,
                          double gamma, double* __restrict__ y) {
  int64_t c = logical_row();
  if (c >= n) return;
  double xc = x ? x[c] : 0.0;
  if (n2f[c] < 0) {
    y[c] = alpha * xc + beta * xc - gamma * in.ud[c];
    return;
  }
  double jx = 0.0, r0 = in.Q[c] / in.V[c];
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
    int32_t e = inc_edge[p], o = inc_other[p];
    double adv, dif, w;
    tr_edge(in, aol, e, (int32_t)c, o, adv, dif, w);
    bool of = n2f[o] >= 0;
    double xo = of && x ? x[o] : 0.0;
    double uo = of ? 0.0 : in.ud[o];
    // adv (u_c - u_o) + dif (u_o - u_c), split into the part linear in x and
    // the constant part (Dirichlet neighbours)
    jx += adv * (xc - xo) + dif * (xo - xc);
    r0 += -adv * uo + dif * uo;
  }
  y[c] = alpha * xc + beta * jx + gamma * r0;
}

// Backward Euler's matrix M = I - dt J restricted to the free cells, on the
// fine pattern (row j = free cell i = f2n[j]), built once per solve since J is
// constant: M_jj = 1 - dt sum (adv - dif), M_jk = -dt (dif - adv) for a free
// neighbour k (parallel edges add into one slot).  MT (nullable) receives the
// transpose (entry (k, j) at row k's slot for the same edge; only thread j
// writes column j, so the adds do not race), dg the diagonal, c0 (nullable)
// the constant part of the free rows of r: Q_i / V_i + sum over Dirichlet
// neighbours o of (dif - adv) ud_o.  A Dirichlet row of J is the identity and
// no free row depends on a Dirichlet cell's state, so the free block is the
// whole coupled system.  M and MT are zeroed by the caller.
__global__ void k_tr_matrix(int64_t nf, const int32_t* __restrict__ f2n,
                            const int32_t* __restrict__ n2f, const int64_t* __restrict__ inc_ptr,
                            const int32_t* __restrict__ inc_edge,
                            const int32_t* __restrict__ inc_other,
                            const int32_t* __restrict__ mirror,
                            const uint16_t* __restrict__ inc_slot,
                            const uint16_t* __restrict__ diag_slot, SellD<double> A,
                            const double* __restrict__ aol, TrIn in, double dt,
                            double* __restrict__ M, double* __restrict__ MT,
                            double* __restrict__ dg, double* __restrict__ c0) {
  int64_t j = logical_row();
  if (j >= nf) return;
  int32_t i = f2n[j];
  int64_t base;
  int w;
  sell_row(A, j, base, w);
  double jd = 0.0, r0 = in.Q[i] / in.V[i];
  for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
    int32_t o = inc_other[p];
    double adv, dif, wv;
    tr_edge(in, aol, inc_edge[p], i, o, adv, dif, wv);
    jd += adv - dif;
    int32_t fo = n2f[o];
    if (fo >= 0) {
      double v = -dt * (dif - adv);
      if (M) M[base + (int64_t)inc_slot[p] * 64] += v;
      if (MT) {
        // row fo differs from lane to lane: its slice pointer is a vector load
        int64_t bo = A.sptr[fo >> 6] + (fo & 63);
        MT[bo + (int64_t)inc_slot[mirror[p]] * 64] += v;
      }
    } else {
      r0 += (dif - adv) * in.ud[o];
    }
  }
  double d = 1.0 - dt * jd;
  int64_t dpos = base + (int64_t)diag_slot[j] * 64;
  if (M) M[dpos] = d;
  if (MT) MT[dpos] = d;
  dg[j] = d;
  if (c0) c0[j] = r0;
}

// One backward-Euler step's set-up over all cells: a free cell c = f2n[j]
// gets rhs_j = u_c + dt c0_j and the initial guess x_j = u_c; a Dirichlet
// cell, whose row (1 - dt) u' = u - dt ud is decoupled, is solved in place
// into unext.
__global__ void k_tr_step_in(int64_t n, const int32_t* __restrict__ n2f,
                             const double* __restrict__ ud, const double* __restrict__ c0,
                             double dt, const double* __restrict__ u, double* __restrict__ unext,
                             double* __restrict__ rhs, double* __restrict__ x) {
  int64_t c = logical_row();
  if (c >= n) return;
  int32_t f = n2f[c];
  double uc = u[c];
  if (f < 0) {
    unext[c] = (uc - dt * ud[c]) / (1.0 - dt);
    return;
  }
  rhs[f] = uc + dt * c0[f];
  x[f] = uc;
}

// unext[f2n[j]] = x[j] (free cells).
__global__ void k_tr_step_out(int64_t nf, const int32_t* __restrict__ f2n,
                              const double* __restrict__ x, double* __restrict__ unext) {
  int64_t j = logical_row();
  if (j < nf) unext[f2n[j]] = x[j];
}

// x[j] = src[f2n[j]] (free cells).
__global__ void k_tr_gather(int64_t nf, const int32_t* __restrict__ f2n,
                            const double* __restrict__ src, double* __restrict__ x) {
  int64_t j = logical_row();
  if (j < nf) x[j] = src[f2n[j]];
}

// Adjoint accumulation of one backward-Euler step with adjoint mu (per free
// cell) and state u (dt dr/dp^T mu): per edge the velocity components and D,
// over both directions of the edge whose first cell is free.
__global__ void k_tr_grad_edges(int64_t ne, const int2* __restrict__ edges,
                                const int32_t* __restrict__ n2f, const double* __restrict__ aol,
                                TrIn in, const double* __restrict__ u,
                                const double* __restrict__ mu, double dt,
                                double* __restrict__ dvx, double* __restrict__ dvy,
                                double* __restrict__ dvz, double* __restrict__ dD) {
  int64_t e = logical_row();
  if (e >= ne) return;
  int2 ab = edges[e];
  if (ab.x == ab.y) return;
  double gx = 0.0, gy = 0.0, gz = 0.0, gd = 0.0;
  for (int dir = 0; dir < 2; dir++) {
    int32_t c = dir ? ab.y : ab.x, o = dir ? ab.x : ab.y;
    int32_t fc = n2f[c];
    if (fc < 0) continue;
    double adv, dif, w;
    tr_edge(in, aol, (int32_t)e, c, o, adv, dif, w);
    double uo = n2f[o] >= 0 ? u[o] : in.ud[o];
    double s = aol[e] / in.V[c] * mu[fc] * dt;
    if ((w > 0.0) == (in.upwind != 0)) {
      double g = -(u[c] - uo) * s;
      gx += g * (in.cx[c] - in.cx[o]);
      gy += g * (in.cy[c] - in.cy[o]);
      gz += g * (in.cz[c] - in.cz[o]);
    }
    gd += (uo - u[c]) * s;
  }
  if (dvx) dvx[e] += gx;
  if (dvy) dvy[e] += gy;
  if (dvz) dvz[e] += gz;
  if (dD) dD[e] += gd;
}

// dQ_c += dt mu_j / V_c at free cells c = f2n[j]
__global__ void k_tr_grad_q(int64_t nf, const int32_t* __restrict__ f2n, TrIn in,
                            const double* __restrict__ mu, double dt, double* __restrict__ dQ) {
  int64_t j = logical_row();
  if (j >= nf) return;
  int32_t c = f2n[j];
  dQ[c] += dt * mu[j] / in.V[c];
}

}  // namespace dpf
