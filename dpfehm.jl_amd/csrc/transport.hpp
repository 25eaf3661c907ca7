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

// y = (I - dt J)^T x: y_j = x_j - dt [J_jj x_j + sum over free neighbours i
// of J_ij x_i] (J_ij from i's edge term; a Dirichlet column has only J_jj = 1).
__global__ void k_tr_rows_t(int64_t n, const int32_t* __restrict__ n2f,
                            const int64_t* __restrict__ inc_ptr,
                            const int32_t* __restrict__ inc_edge,
                            const int32_t* __restrict__ inc_other,
                            const double* __restrict__ aol, TrIn in,
                            const double* __restrict__ x, double dt, double* __restrict__ y) {
  int64_t j = logical_row();
  if (j >= n) return;
  double xj = x[j];
  bool jf = n2f[j] >= 0;
  double s = 0.0;
  if (!jf) {
    s = xj;  // Dirichlet row of J: identity, and no free row couples to j
  } else {
    for (int64_t p = inc_ptr[j]; p < inc_ptr[j + 1]; p++) {
      int32_t e = inc_edge[p], i = inc_other[p];
      double adv, dif, w;
      tr_edge(in, aol, e, (int32_t)j, i, adv, dif, w);
      s += (adv - dif) * xj;  // J_jj contributions of j's own row
      if (n2f[i] >= 0) {
        double adv2, dif2, w2;
        tr_edge(in, aol, e, i, (int32_t)j, adv2, dif2, w2);
        s += (-adv2 + dif2) * x[i];  // J_ij, i's row
      }
    }
  }
  y[j] = xj - dt * s;
}

// diag of I - dt J (the Jacobi preconditioner of the backward-Euler solves).
__global__ void k_tr_diag(int64_t n, const int32_t* __restrict__ n2f,
                          const int64_t* __restrict__ inc_ptr,
                          const int32_t* __restrict__ inc_edge,
                          const int32_t* __restrict__ inc_other, const double* __restrict__ aol,
                          TrIn in, double dt, double* __restrict__ diag) {
  int64_t c = logical_row();
  if (c >= n) return;
  double jcc = 0.0;
  if (n2f[c] < 0) {
    jcc = 1.0;
  } else {
    for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
      double adv, dif, w;
      tr_edge(in, aol, inc_edge[p], (int32_t)c, inc_other[p], adv, dif, w);
      jcc += adv - dif;
    }
  }
  diag[c] = 1.0 - dt * jcc;
}

// z = x / d
__global__ void k_tr_jacobi(int64_t n, const double* __restrict__ x,
                            const double* __restrict__ d, double* __restrict__ z) {
  int64_t i = logical_row();
  if (i < n) z[i] = x[i] / d[i];
}

// Adjoint accumulation of one backward-Euler step with adjoint mu and state u
// (dt dr/dp^T mu): per edge the velocity components and D, over both
// directions of the edge whose first cell is free.
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
    if (n2f[c] < 0) continue;
    double adv, dif, w;
    tr_edge(in, aol, (int32_t)e, c, o, adv, dif, w);
    double uo = n2f[o] >= 0 ? u[o] : in.ud[o];
    double s = aol[e] / in.V[c] * mu[c] * dt;
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

// dQ_c += dt mu_c / V_c at free cells
__global__ void k_tr_grad_q(int64_t n, const int32_t* __restrict__ n2f, TrIn in,
                            const double* __restrict__ mu, double dt, double* __restrict__ dQ) {
  int64_t c = logical_row();
  if (c < n && n2f[c] >= 0) dQ[c] += dt * mu[c] / in.V[c];
}

}  // namespace dpf
