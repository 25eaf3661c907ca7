// Two-phase flow, IMPES (src/twophase.jl): total-mobility pressure solve on the
// groundwater operator, Darcy fluxes, and the explicit upstream saturation
// transport with its CFL sub-steps.
//
// Every kernel here turns floating-point contraction off, so that each
// product and sum is rounded on its own, in the order of the reference's
// element-wise expressions; the results then agree bit for bit with a plain
// IEEE restatement of those expressions.
#pragma once

#include "kernels.hpp"

namespace dpf {

// Fluid of src/twophase.jl:201-206 and test/twophase.jl:11-16; span =
// (1 - swc) - sor, the denominator of the effective saturation.
struct Fluid {
  double vw, vo, swc, sor, span;
};

// Water and oil mobilities of saturation s (relativeperm, src/twophase.jl:201-206).
__device__ __forceinline__ void tp_mobility(double s, const Fluid& f, double& mw, double& mo) {
#pragma clang fp contract(off)
  double S = (s - f.swc) / f.span;
  mw = S * S / f.vw;
  double o = 1.0 - S;
  mo = o * o / f.vo;
}

// Water fractional flow mw / (mw + mo) (src/twophase.jl:229-230).
__device__ __forceinline__ double tp_fracflow(double s, const Fluid& f) {
#pragma clang fp contract(off)
  double mw, mo;
  tp_mobility(s, f, mw, mo);
  return mw / (mw + mo);
}

// Km_c = (Mw + Mo)(S_c) K_c (src/twophase.jl:247-249).
__global__ void k_tp_kmob(int64_t n, const double* __restrict__ S, const double* __restrict__ K,
                          Fluid f, double* __restrict__ Km) {
#pragma clang fp contract(off)
  int64_t c = logical_row();
  if (c >= n) return;
  double mw, mo;
  tp_mobility(S[c], f, mw, mo);
  Km[c] = (mw + mo) * K[c];
}

// Edge factor of the pressure operator: K_a + K_b.  The reference's
// transmissibility (2 aol_e)(0.5 (K_a + K_b)) (src/twophase.jl:26,190) equals
// aol_e (K_a + K_b) exactly (scalings by powers of two), so the groundwater
// operator with K_e = K_a + K_b is the reference's transmissivity operator.
__global__ void k_tp_ksum(int64_t ne, const int2* __restrict__ edges,
                          const double* __restrict__ K, double* __restrict__ Ksum) {
  int64_t e = logical_row();
  if (e >= ne) return;
  int2 ab = edges[e];
  Ksum[e] = K[ab.x] + K[ab.y];
}

// Darcy flux Vn_e = (P_a - P_b)(aol_e Ksum_e), which is the reference's
// P_n (tx K_e) (src/twophase.jl:194-196) exactly.
__global__ void k_tp_flux(int64_t ne, const int2* __restrict__ edges,
                          const double* __restrict__ aol, const double* __restrict__ Ksum,
                          const double* __restrict__ P, double* __restrict__ Vn) {
#pragma clang fp contract(off)
  int64_t e = logical_row();
  if (e >= ne) return;
  int2 ab = edges[e];
  Vn[e] = (P[ab.x] - P[ab.y]) * (aol[e] * Ksum[e]);
}

// NaN-propagating minimum, as Julia's min / minimum.
__device__ __forceinline__ double tp_min(double a, double b) {
  return (a != a || b != b) ? a + b : (a < b ? a : b);
}

// Per cell c the pore-volume-to-inflow ratio pv_c / (Vi_c + max(Q_c, 0)),
// Vi_c summing the fluxes that enter c, in edge order (src/twophase.jl:213-220);
// one minimum per block into part[block].
__global__ void __launch_bounds__(BLK) k_tp_cfl(int64_t n, const int2* __restrict__ edges,
                                                const int64_t* __restrict__ inc_ptr,
                                                const int32_t* __restrict__ inc_edge,
                                                const double* __restrict__ Vn,
                                                const double* __restrict__ Q,
                                                const double* __restrict__ pv,
                                                double* __restrict__ part) {
#pragma clang fp contract(off)
  int64_t c = logical_row();
  double v = INFINITY;
  if (c < n) {
    double vi = 0.0;
    for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
      int32_t e = inc_edge[p];
      double f = Vn[e];
      if (f < 0.0) {
        if (edges[e].x == c) vi = vi - f;
      } else if (edges[e].y == c) {
        vi = vi + f;
      }
    }
    double q = Q[c];
    double fi = (q != q) ? q : (q > 0.0 ? q : 0.0);
    v = pv[c] / (vi + fi);
  }
  // block minimum (the order does not matter for a minimum)
  for (int o = 32; o >= 1; o >>= 1) v = tp_min(v, __shfl_xor(v, o));
  __shared__ double sh[BLK / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = sh[0];
    for (int k = 1; k < BLK / 64; k++) m = tp_min(m, sh[k]);
    part[logical_row() / BLK] = m;
  }
}

// Fluxes and pressures are fixed during one upstream call, so the upwind
// choice of every (cell, edge) pair is too: w[p] = |Vn_e| where the
// neighbour k is upwind of c (P_k - P_c >= 0), -|Vn_e| where c is.  The
// reference's term (src/twophase.jl:96-107) is then f_k w[p] or f_c w[p]
// with the sign of w[p] choosing: f (v > 0 ? v : -v) = f |v| and
// -f v = f (-v) hold exactly in IEEE arithmetic, so the product is the
// reference's bit for bit (a zero flux gives a signed zero, which leaves
// the sum unchanged).  Entry p of the incidence of cell c.
__global__ void k_tp_weights(int64_t n, const int64_t* __restrict__ inc_ptr,
                             const int32_t* __restrict__ inc_edge,
                             const int32_t* __restrict__ inc_other,
                             const double* __restrict__ P, const double* __restrict__ Vn,
                             double* __restrict__ w) {
#pragma clang fp contract(off)
  int64_t c = logical_row();
  if (c >= n) return;
  double pc = P[c];
  int64_t p0 = inc_ptr[c], p1 = inc_ptr[c + 1];
  constexpr int U = 8;
  for (int64_t b = p0; b < p1; b += U) {
    int32_t e[U], k[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool ok = b + u < p1;
      e[u] = ok ? inc_edge[b + u] : 0;
      k[u] = ok ? inc_other[b + u] : 0;
    }
    double v[U], pk[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool ok = b + u < p1;
      v[u] = ok ? Vn[e[u]] : 0.0;
      pk[u] = ok ? P[k[u]] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (b + u < p1) {
        double a = fabs(v[u]);
        w[b + u] = pk[u] - pc >= 0.0 ? a : -a;
      }
  }
}

// f_c = mw / (mw + mo) at S_c.
__global__ void k_tp_frac(int64_t n, const double* __restrict__ S, Fluid f,
                          double* __restrict__ fw) {
  int64_t c = logical_row();
  if (c < n) fw[c] = tp_fracflow(S[c], f);
}

// One CFL sub-step of the upstream transport (src/twophase.jl:227-233): the
// residual of saturationcalc (src/twophase.jl:85-112, every cell an unknown)
// R_c = min(Q_c, 0) f_c + sum over the edges at c, in edge order, of the
// upwind term, then S_c <- clamp(S_c + (R_c dtx_c + max(Q_c, 0) dtx_c), swc,
// 1 - sor) with dtx_c = tau / pv_c, tau = T / Nts.  fw holds f at S (every
// cell) and receives f at the new S, so no sub-step evaluates the fractional
// flow of a cell more than once; the loads of a cell's incidence go out eight
// at a time.
__global__ void k_tp_upstream(int64_t n, const int64_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_other,
                              const double* __restrict__ w, const double* __restrict__ S,
                              const double* __restrict__ fw, double* __restrict__ Snew,
                              double* __restrict__ fwnew, const double* __restrict__ Q,
                              const double* __restrict__ pv, double tau, Fluid f, double lo,
                              double hi) {
#pragma clang fp contract(off)
  int64_t c = logical_row();
  if (c >= n) return;
  double sc = S[c];
  double fc = fw[c];
  double q = Q[c];
  double pvc = pv[c];  // loaded with the cell's other entries, not after the sum
  double qneg = (q != q) ? q : (q < 0.0 ? q : 0.0);
  double qpos = (q != q) ? q : (q > 0.0 ? q : 0.0);
  double R = qneg * fc;
  int64_t p0 = inc_ptr[c], p1 = inc_ptr[c + 1];
  constexpr int U = 8;
  for (int64_t b = p0; b < p1; b += U) {
    double wv[U], fo[U];
    int32_t k[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool ok = b + u < p1;
      wv[u] = ok ? w[b + u] : 0.0;
      k[u] = ok ? inc_other[b + u] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; u++) fo[u] = b + u < p1 ? fw[k[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++)
      if (b + u < p1) R = R + (signbit(wv[u]) ? fc : fo[u]) * wv[u];
  }
  double dtx = tau / pvc;
  double s = sc + (R * dtx + qpos * dtx);
  s = s > hi ? hi : (s < lo ? lo : s);
  Snew[c] = s;
  fwnew[c] = tp_fracflow(s, f);
}

// ------------------------------------------------------------------ reverse mode
//
// Adjoint of the IMPES loop with sub-step counts, upwind choices and clamp
// activity held fixed (piecewise constant in K).  Gradients are sums of
// products whose order is fixed by the incidence lists, so they are
// reproducible; contraction is allowed here.

// dMt/ds and df/ds at saturation s (Mt = Mw + Mo, f = Mw / Mt).
__device__ __forceinline__ void tp_slopes(double s, const Fluid& f, double& mt, double& dmt,
                                          double& dfds) {
  double S = (s - f.swc) / f.span;
  double mw = S * S / f.vw, mo = (1.0 - S) * (1.0 - S) / f.vo;
  double dmw = 2.0 * S / f.vw / f.span, dmo = -2.0 * (1.0 - S) / f.vo / f.span;
  mt = mw + mo;
  dmt = dmw + dmo;
  dfds = (dmw * mo - mw * dmo) / (mt * mt);
}

// Reverse sub-steps, per cell c.  Sub-step k maps (S_k, f_k) to S_{k+1}.
//
// Part 2 of sub-step k (do_b): df_c from the cell's own terms (mass term,
// and the edges where c is downwind) and from the neighbours' terms that
// take f_c (c upwind of them: the mirror entry's weight is non-negative),
// then dS_k = du_k + f'(S_k) df_c.  Inputs S_k (Sb), du_k and dR_k (dRb, read
// at the neighbours too).
//
// Part 1 of sub-step k - 1 (do_a): the forward update recomputed from the
// taped S_{k-1}, f_{k-1} (Sa, fa) to find the clamp's activity; du = dS_k
// where lo <= u <= hi, dR = du dtx; dabs[p] += d(term)/d|Vn| dR_c over the
// cell's incidence (f of the upwind cell, or -f_c).  It needs dS_k at c
// only, so the two parts run in one pass: dS_k stays in a register.  With
// do_b false, dS_k is read from dS; with do_a false, it is written there.
//
// du is per cell (read by part 2, then rewritten by part 1 at the same
// cell); dR is double-buffered (dRb read at the neighbours, dRa written).
// The arithmetic of each part, and its order, are those of the separate
// passes.
__global__ void k_tp_adj(int64_t n, const int64_t* __restrict__ inc_ptr,
                         const int32_t* __restrict__ inc_other,
                         const int32_t* __restrict__ mirror, const double* __restrict__ w,
                         const double* __restrict__ Q, const double* __restrict__ pv, Fluid f,
                         double tau, double lo, double hi, int do_b, int do_a,
                         const double* __restrict__ Sb, const double* __restrict__ dRb,
                         const double* __restrict__ Sa, const double* __restrict__ fa,
                         double* __restrict__ du, double* __restrict__ dRa,
                         double* __restrict__ dabs, double* __restrict__ dS) {
  int64_t c = logical_row();
  if (c >= n) return;
  double q = Q[c];
  double qneg = (q != q) ? q : (q < 0.0 ? q : 0.0);
  double qpos = (q != q) ? q : (q > 0.0 ? q : 0.0);
  int64_t p0 = inc_ptr[c], p1 = inc_ptr[c + 1];
  const bool one = p1 - p0 <= 8;  // the whole incidence in one batch (3-D grids)
  constexpr int U = 8;
  double wv[U];
  int32_t o[U];
  auto load = [&](int64_t b) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool ok = b + u < p1;
      wv[u] = ok ? w[b + u] : 0.0;
      o[u] = ok ? inc_other[b + u] : 0;
    }
  };
  if (one) load(p0);
  // part 1's own-cell entries, in flight during part 2
  double fc = do_a ? fa[c] : 0.0, sc = do_a ? Sa[c] : 0.0, pvc = do_a ? pv[c] : 1.0;
  double dsn;
  if (do_b) {
    double rk = dRb[c];
    double df = qneg * rk;
    for (int64_t b = p0; b < p1; b += U) {
      if (!one) load(b);
      double wm[U], ro[U];
      int32_t m[U];
#pragma unroll
      for (int u = 0; u < U; u++) m[u] = b + u < p1 ? mirror[b + u] : 0;
#pragma unroll
      for (int u = 0; u < U; u++) {
        bool ok = b + u < p1;
        wm[u] = ok ? w[m[u]] : -1.0;
        ro[u] = ok ? dRb[o[u]] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (signbit(wv[u])) df += wv[u] * rk;
        if (!signbit(wm[u])) df += wm[u] * ro[u];
      }
    }
    double mt, dmt, dfds;
    tp_slopes(Sb[c], f, mt, dmt, dfds);
    dsn = du[c] + dfds * df;
    if (!do_a) {
      dS[c] = dsn;
      return;
    }
  } else {
    dsn = dS[c];
  }
  double fo[U];
  auto gather_f = [&](int64_t b) {
#pragma unroll
    for (int u = 0; u < U; u++) fo[u] = b + u < p1 ? fa[o[u]] : 0.0;
  };
  double R = qneg * fc;
  for (int64_t b = p0; b < p1; b += U) {
    if (!one) load(b);
    gather_f(b);
    {
#pragma clang fp contract(off)
#pragma unroll
      for (int u = 0; u < U; u++)
        if (b + u < p1) R = R + (signbit(wv[u]) ? fc : fo[u]) * wv[u];
    }
  }
  double dtx = tau / pvc;
  double uu;
  {
#pragma clang fp contract(off)
    uu = sc + (R * dtx + qpos * dtx);
  }
  double d = (uu >= lo && uu <= hi) ? dsn : 0.0;
  double r = d * dtx;
  du[c] = d;
  dRa[c] = r;
  for (int64_t b = p0; b < p1; b += U) {
    if (!one) {
      load(b);
      gather_f(b);
    }
    double old[U];
#pragma unroll
    for (int u2 = 0; u2 < U; u2++) old[u2] = b + u2 < p1 ? dabs[b + u2] : 0.0;
#pragma unroll
    for (int u2 = 0; u2 < U; u2++)
      if (b + u2 < p1) dabs[b + u2] = old[u2] + (signbit(wv[u2]) ? -fc : fo[u2]) * r;
  }
}

// Per edge: dVn_e = sign(Vn_e) (dabs at a + dabs at b); dKsum_e += dVn_e
// (P_a - P_b) aol_e.  epos[e] = the edge's entries in a's and b's incidence.
__global__ void k_tp_flux_adj(int64_t ne, const int2* __restrict__ edges,
                              const int2* __restrict__ epos, const double* __restrict__ aol,
                              const double* __restrict__ P, const double* __restrict__ Vn,
                              const double* __restrict__ dabs, double* __restrict__ dVn,
                              double* __restrict__ dKsum) {
  int64_t e = logical_row();
  if (e >= ne) return;
  int2 ab = edges[e], pp = epos[e];
  double v = Vn[e];
  double sg = v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0);
  double g = ab.x == ab.y ? 0.0 : sg * (dabs[pp.x] + dabs[pp.y]);
  dVn[e] = g;
  dKsum[e] = g * (P[ab.x] - P[ab.y]) * aol[e];
}

// Per cell: dP_c = seed_c + sum over its edges of +-dVn_e aol_e Ksum_e (+ at
// the edge's first cell, - at its second).
__global__ void k_tp_dP(int64_t n, const int2* __restrict__ edges,
                        const int64_t* __restrict__ inc_ptr,
                        const int32_t* __restrict__ inc_edge, const double* __restrict__ aol,
                        const double* __restrict__ Ksum, const double* __restrict__ dVn,
                        const double* __restrict__ seed, double* __restrict__ dP) {
  int64_t c = logical_row();
  if (c >= n) return;
  double s = seed ? seed[c] : 0.0;
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
    int32_t e = inc_edge[p];
    double g = dVn[e] * aol[e] * Ksum[e];
    s += edges[e].x == c ? g : -g;
  }
  dP[c] = s;
}

// Per cell: dKm_c = sum of dKsum over the cell's edges (Ksum_e = Km_a + Km_b);
// through Km = Mt(S) K: dK_c += Mt dKm_c, dS_c += K_c Mt'(S_c) dKm_c.
__global__ void k_tp_kmob_adj(int64_t n, const int64_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_edge,
                              const double* __restrict__ dKsum, const double* __restrict__ S,
                              const double* __restrict__ K, Fluid f, double* __restrict__ dK,
                              double* __restrict__ dS) {
  int64_t c = logical_row();
  if (c >= n) return;
  double g = 0.0;
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) g += dKsum[inc_edge[p]];
  double mt, dmt, dfds;
  tp_slopes(S[c], f, mt, dmt, dfds);
  dK[c] += mt * g;
  dS[c] += K[c] * dmt * g;
}

// Per-block sums of x (block_sum order) into part[block].
__global__ void __launch_bounds__(BLK) k_tp_blocksum(int64_t n, const double* __restrict__ x,
                                                     double* __restrict__ part) {
  __shared__ double sh[4];
  int64_t i = logical_row();
  double s = block_sum(i < n ? x[i] : 0.0, sh);
  if (threadIdx.x == 0) part[i / BLK] = s;
}

// *mean = (sum of the nb block sums, in block order) / n (one thread).
__global__ void k_tp_mean(int64_t n, int64_t nb, const double* __restrict__ part,
                          double* __restrict__ mean) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int64_t b = 0; b < nb; b++) s += part[b];
  *mean = s / (double)n;
}

// x -= *mean
__global__ void k_tp_sub(int64_t n, double* __restrict__ x, const double* __restrict__ mean) {
  int64_t i = logical_row();
  if (i < n) x[i] -= *mean;
}

}  // namespace dpf
