// Steady Richards equation (src/richards.jl:97-205): van Genuchten relative
// permeability with its partials, residual, Newton matrix on the fine pattern,
// and the tangents of the rrule.
#pragma once

#include "kernels.hpp"

namespace dpf {

// kr(psi, alpha, N) of src/richards.jl:18-27 and its partials in psi, alpha
// and N (what ForwardDiff takes of it, src/richards.jl:36-43), written out by
// the chain rule through s = |alpha psi|, D = 1 + s^N, A = s^(N-1),
// B = D^-m, num = 1 - A B, kr = num^2 D^(-m/2), m = (N - 1) / N.  All partials
// are 0 for psi >= 0 (kr = 1).
struct Kr {
  double k, dpsi, dalpha, dN;
};

__device__ __forceinline__ Kr vg_kr(double psi, double alpha, double N) {
  Kr r{1.0, 0.0, 0.0, 0.0};
  if (!(psi < 0.0)) return r;
  double m = (N - 1.0) / N;
  double ap = alpha * psi;
  double s = fabs(ap);
  double L = log(s);
  double sN = exp(N * L);
  double A = exp((N - 1.0) * L);
  double D = 1.0 + sN;
  double LD = log1p(sN);
  double B = exp(-m * LD);
  double E = exp(-0.5 * m * LD);
  double num = 1.0 - A * B;
  r.k = num * num * E;
  // d/ds
  double dD = N * A;
  double dA = (N - 1.0) * A / s;
  double dB = -m * B * dD / D;
  double dnum = -(dA * B + A * dB);
  double dE = -0.5 * m * E * dD / D;
  double dk_ds = 2.0 * num * dnum * E + num * num * dE;
  double sg = ap > 0.0 ? 1.0 : -1.0;
  r.dpsi = dk_ds * sg * alpha;
  r.dalpha = dk_ds * sg * psi;
  // d/dN at fixed s (dm/dN = 1/N^2)
  double dm = 1.0 / (N * N);
  double DN = sN * L;
  double AN = A * L;
  double BN = B * (-dm * LD - m * DN / D);
  double numN = -(AN * B + A * BN);
  double EN = E * (-0.5 * dm * LD - 0.5 * m * DN / D);
  r.dN = 2.0 * num * numN * E + num * num * EN;
  return r;
}

// Per edge, the pieces every kernel below uses: kr at both ends (psi of a
// Dirichlet end from dirichletpsis), hm = 2 / (1/ka + 1/kb) and its partials
// (src/richards.jl:50-52), and the head difference (psi + z)_b - (psi + z)_a.
struct RichEdge {
  Kr ka, kb;
  double H, Ha, Hb, dlt;
};

__device__ __forceinline__ RichEdge rich_edge(int64_t e, const int2* __restrict__ edges,
                                              const int32_t* __restrict__ n2f,
                                              const double* __restrict__ psi,
                                              const double* __restrict__ dpsi,
                                              const double* __restrict__ z,
                                              const double* __restrict__ alpha,
                                              const double* __restrict__ Nv) {
  RichEdge g;
  int2 ab = edges[e];
  int32_t fa = n2f[ab.x], fb = n2f[ab.y];
  double pa = fa >= 0 ? psi[fa] : dpsi[ab.x];
  double pb = fb >= 0 ? psi[fb] : dpsi[ab.y];
  g.ka = vg_kr(pa, alpha[e], Nv[e]);
  g.kb = vg_kr(pb, alpha[e], Nv[e]);
  double ka = g.ka.k, kb = g.kb.k;
  g.H = 2.0 / (1.0 / ka + 1.0 / kb);
  double s = (ka + kb) * (ka + kb);
  g.Ha = 2.0 * kb * kb / s;
  g.Hb = 2.0 * ka * ka / s;
  g.dlt = pb + z[ab.y] - pa - z[ab.x];
  return g;
}

// Picard edge factor K_e hm(kr_a, kr_b): the groundwater operator with it is
// the symmetric part of the Newton matrix, whose AMG V-cycle preconditions
// the Newton solves.
__global__ void k_rich_picard(int64_t ne, const int2* __restrict__ edges,
                              const int32_t* __restrict__ n2f, const double* __restrict__ psi,
                              const double* __restrict__ dpsi, const double* __restrict__ z,
                              const double* __restrict__ alpha, const double* __restrict__ Nv,
                              const double* __restrict__ K, double* __restrict__ Kp) {
  int64_t e = logical_row();
  if (e >= ne) return;
  RichEdge g = rich_edge(e, edges, n2f, psi, dpsi, z, alpha, Nv);
  Kp[e] = K[e] * g.H;
}

// Per edge: the flux F = hm K dlt aol and its partials in psi_a, psi_b
// (flux terms of src/richards.jl:106-114; the equation at a adds F / sv_a, the
// one at b subtracts F / sv_b).
__global__ void k_rich_edge_flux(int64_t ne, const int2* __restrict__ edges,
                                 const int32_t* __restrict__ n2f, const double* __restrict__ psi,
                                 const double* __restrict__ dpsi, const double* __restrict__ z,
                                 const double* __restrict__ alpha, const double* __restrict__ Nv,
                                 const double* __restrict__ K, const double* __restrict__ aol,
                                 double* __restrict__ F, double* __restrict__ dFa,
                                 double* __restrict__ dFb) {
  int64_t e = logical_row();
  if (e >= ne) return;
  RichEdge g = rich_edge(e, edges, n2f, psi, dpsi, z, alpha, Nv);
  double ka = K[e] * aol[e];
  F[e] = g.H * K[e] * g.dlt * aol[e];
  if (dFa) dFa[e] = (g.Ha * g.ka.dpsi * g.dlt - g.H) * ka;
  if (dFb) dFb[e] = (g.Hb * g.kb.dpsi * g.dlt + g.H) * ka;
}

// r_j = Q_i / sv_i + sum over the edges at i (edge order) of +-F_e / sv_i
// for free cell i = f2n[j] (src/richards.jl:99-117); partials of |r| (max)
// and r.r per block into part (2 per block).
__global__ void __launch_bounds__(BLK) k_rich_residual(
    int64_t nf, const int32_t* __restrict__ f2n, const int2* __restrict__ edges,
    const int64_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_edge,
    const double* __restrict__ F, const double* __restrict__ Q, const double* __restrict__ sv,
    double* __restrict__ r, double* __restrict__ part) {
  int64_t j = logical_row();
  double a = 0.0, q2 = 0.0;
  if (j < nf) {
    int32_t i = f2n[j];
    double s = sv ? sv[i] : 1.0;
    double v = Q[i] / s;
    for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
      int32_t e = inc_edge[p];
      double f = F[e] / s;
      v += edges[e].x == i ? f : -f;
    }
    r[j] = v;
    a = fabs(v);
    q2 = v * v;
    if (v != v) a = INFINITY;
  }
  __shared__ double sh[4], sm[4];
  double tot = block_sum(q2, sh);
  double mx = wave_max(a);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t lb = logical_row() / BLK;
    part[2 * lb] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
    part[2 * lb + 1] = tot;
  }
}

// The Newton matrix M = -dr/dpsi on the fine pattern (row j = free cell i,
// if M is given), and, if MT is given, its transpose there (entry (k, j) at row k's slot for
// the same edge).  The arrays are zeroed by the caller; parallel edges add
// into one slot (only thread j writes row j of M and column j of MT).  M is the Picard (groundwater)
// operator plus the kr-slope terms: M_jj = sum (H - Hi kr_i' dlt_io) K aol / sv_i,
// M_jk = -(Ho kr_o' dlt_io + H) K aol / sv_i, dlt_io = (psi + z)_o - (psi + z)_i.
__global__ void k_rich_newton_matrix(
    int64_t nf, const int32_t* __restrict__ f2n, const int32_t* __restrict__ n2f,
    const int2* __restrict__ edges, const int64_t* __restrict__ inc_ptr,
    const int32_t* __restrict__ inc_edge, const int32_t* __restrict__ inc_other,
    const int32_t* __restrict__ mirror, const uint16_t* __restrict__ inc_slot,
    const uint16_t* __restrict__ diag_slot, SellD<double> A, const double* __restrict__ dFa,
    const double* __restrict__ dFb, const double* __restrict__ sv, double* __restrict__ M,
    double* __restrict__ MT) {
  int64_t j = logical_row();
  if (j >= nf) return;
  int32_t i = f2n[j];
  double s = sv ? sv[i] : 1.0;
  int64_t base;
  int w;
  sell_row(A, j, base, w);
  double diag = 0.0;
  for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
    int32_t e = inc_edge[p];
    bool first = edges[e].x == i;
    // the equation at i holds +F (i = a) or -F (i = b); M = -dr/dpsi
    double self = first ? -dFa[e] : dFb[e];
    double oth = first ? -dFb[e] : dFa[e];
    diag += self / s;
    int32_t fo = n2f[inc_other[p]];
    if (fo >= 0) {
      double v = oth / s;
      if (M) M[base + (int64_t)inc_slot[p] * 64] += v;
      if (MT) {
        // row fo differs from lane to lane: its slice pointer is a vector load
        int64_t bo = A.sptr[fo >> 6] + (fo & 63);
        MT[bo + (int64_t)inc_slot[mirror[p]] * 64] += v;
      }
    }
  }
  int64_t dpos = base + (int64_t)diag_slot[j] * 64;
  if (M) M[dpos] = diag;
  if (MT) MT[dpos] = diag;
}

// Tangents of the rrule (src/richards.jl:185-203) per edge from the adjoint
// lambda (free cells): g = lambda_a [a free] - lambda_b [b free];
// dK = -g H dlt aol, dalpha = -g (Ha kra_alpha + Hb krb_alpha) K dlt aol,
// dN likewise; ga, gb = -g dF/dpsi at a Dirichlet end (0 at a free end).
__global__ void k_rich_tangents(int64_t ne, const int2* __restrict__ edges,
                                const int32_t* __restrict__ n2f, const double* __restrict__ psi,
                                const double* __restrict__ dpsi, const double* __restrict__ z,
                                const double* __restrict__ alpha, const double* __restrict__ Nv,
                                const double* __restrict__ K, const double* __restrict__ aol,
                                const double* __restrict__ lam, double* __restrict__ dK,
                                double* __restrict__ dal, double* __restrict__ dN,
                                double* __restrict__ ga, double* __restrict__ gb) {
  int64_t e = logical_row();
  if (e >= ne) return;
  int2 ab = edges[e];
  RichEdge g = rich_edge(e, edges, n2f, psi, dpsi, z, alpha, Nv);
  int32_t fa = n2f[ab.x], fb = n2f[ab.y];
  double gg = (ab.x == ab.y) ? 0.0 : (fa >= 0 ? lam[fa] : 0.0) - (fb >= 0 ? lam[fb] : 0.0);
  double ka = K[e] * aol[e];
  if (dK) dK[e] = -gg * g.H * g.dlt * aol[e];
  if (dal) dal[e] = -gg * (g.Ha * g.ka.dalpha + g.Hb * g.kb.dalpha) * ka * g.dlt;
  if (dN) dN[e] = -gg * (g.Ha * g.ka.dN + g.Hb * g.kb.dN) * ka * g.dlt;
  double dfa = (g.Ha * g.ka.dpsi * g.dlt - g.H) * ka;
  double dfb = (g.Hb * g.kb.dpsi * g.dlt + g.H) * ka;
  ga[e] = fa >= 0 ? 0.0 : -gg * dfa;
  gb[e] = fb >= 0 ? 0.0 : -gg * dfb;
}

// Per cell: ddirichletpsis_c = delta_c + sum of ga / gb over its edges
// (Dirichlet cells; 0 at free cells); dQ_c = -lambda (free) or 0.
__global__ void k_rich_cell_tangents(int64_t n, const int32_t* __restrict__ n2f,
                                     const int2* __restrict__ edges,
                                     const int64_t* __restrict__ inc_ptr,
                                     const int32_t* __restrict__ inc_edge,
                                     const double* __restrict__ ga, const double* __restrict__ gb,
                                     const double* __restrict__ lam,
                                     const double* __restrict__ delta, double* __restrict__ ddp,
                                     double* __restrict__ dQ) {
  int64_t c = logical_row();
  if (c >= n) return;
  int32_t f = n2f[c];
  if (dQ) dQ[c] = f >= 0 ? -lam[f] : 0.0;
  if (!ddp) return;
  if (f >= 0) {
    ddp[c] = 0.0;
    return;
  }
  double s = delta[c];
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
    int32_t e = inc_edge[p];
    s += edges[e].x == c ? ga[e] : gb[e];
  }
  ddp[c] = s;
}

// ---------------------------------------------------------------- BiCGStab pieces

// y = M x on the fine pattern (any value array of the fine layout).
__global__ void k_rich_spmv(SellD<double> A, const double* __restrict__ x,
                            double* __restrict__ y) {
  int64_t i = logical_row();
  if (i >= A.nrows) return;
  int64_t base;
  int w;
  sell_row(A, i, base, w);
  y[i] = sell_dot(A, i, base, w, x);
}

// The V-cycle's zero-guess first step on b: d = x = c0 invd b.
__global__ void k_rich_pre(int64_t n, const double* __restrict__ b,
                           const double* __restrict__ invd, double c0, double* __restrict__ d,
                           double* __restrict__ x) {
  int64_t i = logical_row();
  if (i >= n) return;
  double v = c0 * invd[i] * b[i];
  if (d) d[i] = v;
  x[i] = v;
}

// Fixed grid of the reduction kernels below: a thread's grid-stride
// sequence, and so every sum, is the same on every run.
constexpr int DOT_GRID = 1024;

// The FP32 V-cycle's input and zero-guess first step from an FP64 vector:
// b32 = fl32(b), d = x = c0 invd b32.
__global__ void k_rich_pre32(int64_t n, const double* __restrict__ b,
                             const float* __restrict__ invd, float c0, float* __restrict__ b32,
                             float* __restrict__ d, float* __restrict__ x) {
  int64_t i = logical_row();
  if (i >= n) return;
  float bv = (float)b[i];
  b32[i] = bv;
  float v = c0 * invd[i] * bv;
  d[i] = v;
  x[i] = v;
}

// z = (double) x
__global__ void k_widen(int64_t n, const float* __restrict__ x, double* __restrict__ z) {
  int64_t i = logical_row();
  if (i < n) z[i] = (double)x[i];
}

// Up to three dot products in one pass: part[3 * block + k] = block sum of
// x_k . y_k (null x_k: 0), grid-stride over a grid of at most DOT_GRID blocks.
__global__ void __launch_bounds__(BLK) k_dots3(int64_t n, const double* __restrict__ x0,
                                               const double* __restrict__ y0,
                                               const double* __restrict__ x1,
                                               const double* __restrict__ y1,
                                               const double* __restrict__ x2,
                                               const double* __restrict__ y2,
                                               double* __restrict__ part) {
  __shared__ double sh[4];
  double a = 0.0, b = 0.0, c = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLK;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += stride) {
    a += x0[i] * y0[i];
    if (x1) b += x1[i] * y1[i];
    if (x2) c += x2[i] * y2[i];
  }
  a = block_sum(a, sh);
  if (threadIdx.x == 0) part[3 * blockIdx.x] = a;
  b = block_sum(b, sh);
  if (threadIdx.x == 0) part[3 * blockIdx.x + 1] = b;
  c = block_sum(c, sh);
  if (threadIdx.x == 0) part[3 * blockIdx.x + 2] = c;
}

// BiCGStab's closing update of one iteration in one pass:
// x += a ph + w sh; r = sv - w tv; and the partial sums of r.r and rh.r
// (part as k_dots3's, third sum 0).
__global__ void __launch_bounds__(BLK) k_bicg_xr(int64_t n, double a,
                                                 const double* __restrict__ ph, double w,
                                                 const double* __restrict__ shv,
                                                 const double* __restrict__ sv,
                                                 const double* __restrict__ tv,
                                                 const double* __restrict__ rh,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* __restrict__ part) {
  __shared__ double sh[4];
  double rr = 0.0, hr = 0.0;
  const int64_t stride = (int64_t)gridDim.x * BLK;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < n; i += stride) {
    x[i] = x[i] + a * ph[i] + w * shv[i];
    double ri = sv[i] - w * tv[i];
    r[i] = ri;
    rr += ri * ri;
    hr += rh[i] * ri;
  }
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) part[3 * blockIdx.x] = rr;
  hr = block_sum(hr, sh);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 1] = hr;
    part[3 * blockIdx.x + 2] = 0.0;
  }
}

// out[k] = sum over the nb blocks of part[3 b + k], k < 3, in a fixed order
// (one block: strided per-thread sums, then the block tree); out is pinned
// host memory, stored to at system scope.
__global__ void __launch_bounds__(BLK) k_dots3_fin(int64_t nb, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  __shared__ double sh[4];
  for (int k = 0; k < 3; k++) {
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += BLK) s += part[3 * b + k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) mput(&out[k], s);
  }
}

// BiCGStab's vector updates with the Jacobi preconditioner z = x / d fused
// in: p = r + beta (p - omega v), ph = p / d (a = beta, b = -beta omega,
// x = r, y = v); or s = r - alpha v, sh = s / d (pass p = s, beta 0: a = 0,
// b = -alpha).  out = x + a out + b y; z = out / d.
__global__ void k_bicg_jacobi(int64_t n, const double* __restrict__ x, double a, double b,
                             const double* __restrict__ y, const double* __restrict__ d,
                             double* __restrict__ out, double* __restrict__ z) {
  int64_t i = logical_row();
  if (i >= n) return;
  double v = a != 0.0 ? x[i] + a * out[i] + b * y[i] : x[i] + b * y[i];
  out[i] = v;
  z[i] = v / d[i];
}

// out = a x + b y
__global__ void k_lin2(int64_t n, double a, const double* __restrict__ x, double b,
                       const double* __restrict__ y, double* __restrict__ out) {
  int64_t i = logical_row();
  if (i < n) out[i] = a * x[i] + b * y[i];
}

// out = a * b elementwise
__global__ void k_rich_mul(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                           double* __restrict__ out) {
  int64_t i = logical_row();
  if (i < n) out[i] = a[i] * b[i];
}

// y = a x + b y (+ c z when z)
__global__ void k_rich_axpby(int64_t n, double a, const double* __restrict__ x, double b,
                             double* __restrict__ y, double c, const double* __restrict__ z) {
  int64_t i = logical_row();
  if (i >= n) return;
  double v = a * x[i] + b * y[i];
  if (z) v += c * z[i];
  y[i] = v;
}

}  // namespace dpf
