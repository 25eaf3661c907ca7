// Device kernels of the groundwater solve and adjoint (gfx950, wave64).
//
// All kernels are memory-bound row or edge loops; none uses MFMA.
//  * Matrices are SELL-64: one matrix row per lane, the 64 rows of a slice
//    share a wavefront, entries column-major inside the slice, so every
//    value/column load of a wavefront is one contiguous 256/512-byte burst.
//  * Row-parallel grids remap the hardware block index so that each of the
//    8 XCDs (own L2 each) processes one contiguous band of rows: the x-vector
//    gathers of a band then stay inside that XCD's L2.
//  * Nothing scatters: assembly, transfer operators and gradients are row
//    or cell gathers through index maps built on the host, so results are
//    bitwise reproducible.  Reductions are per-block shuffles writing one
//    partial per block, summed in fixed order by a one-workgroup finish
//    kernel (no atomics of any kind).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dpf {

constexpr int BLK = 256;
// coarsest levels up to this many rows are inverted in LDS (nc^2 doubles <= 64 KiB)
constexpr int DENSE_LDS_MAX = 90;
constexpr uint16_t NO_SLOT_D = 0xFFFF;

// Host-visible mirror of the solve's progress, in pinned coherent host
// memory: the finish kernel writes it with system-scope stores, so the host
// follows convergence without synchronising the stream or copying.
struct HostMirror {
  int seq;        // solve sequence number of the fields below
  int iters;      // largest iteration count over the members
  int done;       // 0 while any member runs; else 3 if any broke down, 2 if any hit the cap, 1
  int outer_seq;  // sequence number of the refinement residuals
  int active;     // members whose refinement residual is above their target
  int cut;        // set by the assembly when an edge's conductance K_e aol_e is 0
  double rr;
  double rr_stop;  // a single system's stop threshold on rr (FIN_START), -1 for a batch
};

__device__ __forceinline__ void mput(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void mput(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void mpublish(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Device-side scalars of one Krylov solve, one record per member of a batch
// (a single system is a batch of one).  All FP64, also in mixed mode.
struct SolveScalars {
  double rr;       // current ||r||^2
  double rr_stop;  // converged when rr <= rr_stop
  double rr_freeze;  // stop updating the member when rr <= rr_freeze (<= rr_stop)
  double rz, rz_old;
  double pq, alpha;
  double outer_s;  // refinement: ||b - A x|| in FP64
  double bnorm;    // refinement: ||b||
  double target;   // refinement: stop when outer_s <= target
  double bscale;   // refinement: power of two that brings max|b| into [1/2, 1)
  double xb;       // warm start: x0 . b
  double cwarm;    // warm start: x0 is replaced by cwarm x0
  int iters, maxiter;
  int done;  // 0 running, 1 converged, 2 iteration cap, 3 breakdown
  int first;
  int conv;  // rr has reached rr_stop (the member may still be iterating)
  int pad;
};

// A SELL-64 operator on the device.  soff (fine operator only, else null):
// per slice 9 ints, [0] = width of an offset slice (0: explicit columns),
// [1..8] the offsets o_k, so that slot k of row r has column r + o_k.
// col16 / cbase (compact columns, else null): the column of slot e of slice s
// is cbase[s] + col16[e] (every column of a slice within 65535 of its
// smallest; padding slots point at that smallest column), in place of col.
template <typename T>
struct SellD {
  int64_t nrows;
  const int64_t* sptr;
  const int32_t* col;
  T* val;
  const int32_t* soff = nullptr;
  const uint16_t* col16 = nullptr;
  const int32_t* cbase = nullptr;
  // bf16 values (the upper 16 bits of the FP32 value), read in place of val
  // when set (the FP32 V-cycle's transfer and coarse split operators)
  const uint16_t* v16 = nullptr;
};

// bf16 <-> FP32: a bf16 value is the upper half of an FP32 one; rounding to
// nearest, ties to even (finite values)
__device__ __forceinline__ float bf16_to_f32(uint16_t u) {
  return __uint_as_float((uint32_t)u << 16);
}
__device__ __forceinline__ uint16_t f32_to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// The offsets of row r's slice if it is an offset slice, else null.  Every
// lane of a wave is in the same slice, so the lookup is a scalar load.
template <typename T>
__device__ __forceinline__ const int32_t* slice_offsets(const SellD<T>& A, int64_t r) {
  if (!A.soff) return nullptr;
  int sl = __builtin_amdgcn_readfirstlane((int)(r >> 6));
  const int32_t* q = A.soff + 9 * (int64_t)sl;
  return q[0] ? q + 1 : nullptr;
}

__device__ __forceinline__ int64_t logical_row() {
  // Hardware dispatches block b to XCD b % 8.  Give XCD x the contiguous
  // logical blocks [start_x, start_x + count_x).
  int b = blockIdx.x, nb = gridDim.x;
  int q = nb >> 3, r = nb & 7;
  int x = b & 7, p = b >> 3;
  int lb = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + p;
  return (int64_t)lb * BLK + threadIdx.x;
}

// Cross-lane moves of the data-parallel primitives (DPP): a register of
// another lane of the same row of 16 (or, for row_bcast, of the previous
// rows), with 0 in the rows outside row_mask.  32-bit moves; a double moves
// as two halves.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_move(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROWMASK, 0xF, false);
  int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWMASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over each group of 2^lg consecutive lanes (lg <= 6), in a fixed
// pattern: lane pairs, quads, half-rows and rows (every lane of a group of
// up to 16 then holds the sum), then rows combined by row broadcasts, which
// leave the sum of a group of 32 or 64 in the group's last lane.  The result
// is read from the group's LAST lane.  DPP moves instead of LDS permutes: a
// six-step wave sum is six dependent ALU steps instead of six LDS round trips.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int lg) {
  if (lg >= 1) v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
  if (lg >= 2) v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
  if (lg >= 3) v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  if (lg >= 4) v += dpp_move<0x140, 0xF>(v);  // row_mirror
  if (lg >= 5) v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1, 3
  if (lg >= 6) v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Sum over the wavefront, valid in lane 63.
__device__ __forceinline__ double wave_sum(double v) { return group_sum(v, 6); }

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Sum over the block; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 63) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return s;
}

// Batched solves pack the members' fine rows contiguously, each member with at
// least BLK rows, so a logical block of the fine grid holds rows of at most
// two members.  bm[lb] = {first member of block lb, first row of the next
// member (or INT32_MAX)}; bm == nullptr for a single system.
__device__ __forceinline__ int member_of(const int2* __restrict__ bm, int64_t row, bool& second) {
  if (!bm) {
    second = false;
    return 0;
  }
  int2 e = bm[row / BLK];
  second = row >= e.y;
  return second ? e.x + 1 : e.x;
}

// Block partials of a grid-wide sum, two per logical block: slot 0 for the
// block's first member, slot 1 for its second (single systems use slot 0).
// The sums over blocks are finished by k_finish in the next launch on the
// stream; the kernel boundary makes the partials visible without any
// device-scope fence or atomic (on a multi-XCD part a release fence per block
// writes back the XCD's L2 and costs far more than the kernel's own traffic).
// two = false (a single system, no member table): slot 1 is 0 and only one
// block reduction runs.
__device__ __forceinline__ void block_partial2(double v, bool second,
                                               double* __restrict__ partial, bool two) {
  __shared__ double sh[4];
  double a = block_sum(second ? 0.0 : v, sh);
  double b = two ? block_sum(second ? v : 0.0, sh) : 0.0;
  if (threadIdx.x == 0) {
    int64_t lb = logical_row() / BLK;
    partial[2 * lb] = a;
    partial[2 * lb + 1] = b;
  }
}

// Block maxima of non-negative values, laid out like block_partial2.
__device__ __forceinline__ void block_partial2_max(double v, bool second,
                                                   double* __restrict__ partial) {
  __shared__ double sh[2][4];
  double a = wave_max(second ? 0.0 : v);
  double b = wave_max(second ? v : 0.0);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = a;
    sh[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t lb = logical_row() / BLK;
    partial[2 * lb] = fmax(fmax(sh[0][0], sh[0][1]), fmax(sh[0][2], sh[0][3]));
    partial[2 * lb + 1] = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
  }
}

enum FinishMode {
  FIN_START = 0,
  FIN_PQ = 1,
  FIN_XR = 2,
  FIN_RZ = 3,
  FIN_OUTER = 4,
  FIN_AMAX = 5,
  FIN_WXB = 6,
  FIN_WXR = 7
};

// Per member: the scalar step of the Krylov solve after the block partials.
//   FIN_START  rr, stop threshold (relative, or the member's absolute
//              refinement target when abs_target), counters
//   FIN_PQ     alpha = rz / p.q          FIN_XR  rr, iteration count, convergence
//   FIN_RZ     rz_old = rz, rz = r.z     FIN_OUTER  FP64 residual norm (and, when
//              abs_target, the member's ||b|| and target: the first residual)
//   FIN_AMAX   tot = max|b| over the member: bscale = 2^-e with max|b| = f 2^e,
//              f in [1/2, 1) (1 for a zero or non-finite b), so that the
//              refinement works on b bscale, whose squared norms neither
//              underflow nor overflow in FP64; scaling by a power of two is
//              exact, so the iterates are those of the unscaled system times
//              bscale wherever the unscaled ones are representable
//   FIN_WXB    xb = x0 . b           FIN_WXR  tot = x0 . r0: cwarm = x0.b / x0.A x0,
//              the multiple of the initial guess closest to the solution in
//              the A-norm (x0.A x0 = x0.b - x0.r0; 1 when x0 = 0)
// deep: a batch's relative (inner) solve keeps iterating a converged member
// while other members still run (its updates cost nothing extra: the batched
// kernels sweep every row anyway) until its residual is 100 times below the
// target, so that the next refinement round starts deeper for it.
__device__ void finish_member(int mode, double tot, SolveScalars& m, double reltol,
                              double abstol, int maxiter, int abs_target, int deep) {
  switch (mode) {
    case FIN_START: {
      m.rr = tot;
      double stop;
      if (abs_target) {
        stop = m.target * m.target;
      } else {
        // the inner solve of a refinement step, whose right-hand side is the
        // FP64 residual over its norm outer_s: relative tolerance reltol,
        // unless the step can meet the refinement target by itself (target /
        // outer_s at most ten times below reltol), in which case it aims at
        // half the target so that no further step is needed for want of a
        // few digits
        double rel = reltol;
        double rho = m.outer_s > 0.0 ? m.target / m.outer_s : 0.0;
        if (rho >= 0.1 * reltol) rel = 0.5 * rho;
        stop = rel * rel * tot;
        if (stop < abstol * abstol) stop = abstol * abstol;
      }
      m.rr_stop = stop;
      m.rr_freeze = deep && !abs_target ? 1e-4 * stop : stop;
      m.iters = 0;
      m.maxiter = maxiter;
      m.first = 1;
      m.rz = 0.0;
      m.rz_old = 0.0;
      m.conv = tot <= stop;
      m.done = (tot <= m.rr_freeze) ? 1 : (tot == tot ? 0 : 3);
      break;
    }
    case FIN_PQ:
      if (m.done) break;
      m.pq = tot;
      m.first = 0;
      if (!(tot > 0.0))
        m.done = m.conv ? 1 : 3;
      else
        m.alpha = m.rz / tot;
      break;
    case FIN_XR: {
      if (m.done) break;
      m.rr = tot;
      int it = m.iters + 1;
      m.iters = it;
      if (tot <= m.rr_stop) m.conv = 1;
      if (tot <= m.rr_freeze)
        m.done = 1;
      else if (!(tot == tot))
        m.done = 3;
      else if (it >= m.maxiter)
        m.done = m.conv ? 1 : 2;
      break;
    }
    case FIN_RZ:
      if (m.done) break;
      m.rz_old = m.rz;
      m.rz = tot;
      if (!(tot > 0.0)) m.done = m.conv ? 1 : 3;
      break;
    case FIN_WXB:
      m.xb = tot;
      break;
    case FIN_WXR: {
      double xax = m.xb - tot;
      m.cwarm = (xax > 0.0 && m.xb == m.xb) ? m.xb / xax : 1.0;
      break;
    }
    case FIN_AMAX: {
      int e = (tot > 0.0 && tot <= 1.79769313486231570e308) ? ilogb(tot) + 1 : 0;
      e = e > 1000 ? 1000 : (e < -1000 ? -1000 : e);
      m.bscale = ldexp(1.0, -e);
      break;
    }
    default: {
      double sn = sqrt(tot);
      m.outer_s = sn;
      if (abs_target) {
        m.bnorm = sn;
        double tg = reltol * sn, ta = abstol * m.bscale;
        m.target = tg > ta ? tg : ta;
      }
    }
  }
}

// The end of a Krylov step's finish (every mode but FIN_OUTER): the early-exit
// flag and the host mirror from the folded member states (flags: bit 0 a
// member still running, bit 1 one capped, bit 2 one broken down).
__device__ __forceinline__ void publish_solve(int mode, int flags, int maxit, double rr,
                                              double rr_stop, int seq, int* gdone,
                                              HostMirror* hm) {
  int running = flags & 1, capped = flags & 2, broke = flags & 4;
  int dn = running ? 0 : (broke ? 3 : (capped ? 2 : 1));
  *gdone = dn;
  // the host follows the iteration count (FIN_XR) and the end of the solve;
  // the other steps of a running iteration leave the mirror as it is
  if (mode == FIN_START || mode == FIN_XR || dn != 0) {
    mput(&hm->rr, rr);
    mput(&hm->iters, maxit);
    mput(&hm->done, dn);
  }
  if (mode == FIN_START) {
    mput(&hm->rr_stop, rr_stop);
    mpublish(&hm->seq, seq);
  }
}

// A single system's member flags, as k_finish folds them.
__device__ __forceinline__ int member_flags(const SolveScalars& q) {
  return (q.done == 0 && !q.conv ? 1 : 0) | (q.done == 2 ? 2 : 0) | (q.done == 3 ? 4 : 0);
}

// Sums the block partials of the previous kernel per member, in a fixed
// order, applies finish_member, then folds the members' states into the
// global early-exit flag *gdone and the host mirror.  One workgroup; with
// several members each wave takes members m = wave, wave + 16, ... and folds
// them in registers, and thread 0 combines the 16 wave results.  The solve
// stops once no member is below its target and still unconverged.
__global__ void __launch_bounds__(1024) k_finish(int mode, int nm, int nblk,
                                                 const double* __restrict__ partial,
                                                 const int2* __restrict__ bm,
                                                 const int* __restrict__ mblk,
                                                 SolveScalars* sc, int* gdone, double reltol,
                                                 double abstol, int maxiter, int abs_target,
                                                 int seq, HostMirror* hm,
                                                 double* __restrict__ hm_outer) {
  // the early-exit flag is tested once the partials' loads are in flight
  const int gd = (mode != FIN_START && mode < FIN_OUTER) ? *gdone : 0;
  constexpr int NW = 16;  // waves of the 1024-thread workgroup
  __shared__ double sh[NW], sh_rr[NW];
  __shared__ int sh_flags[NW], sh_maxit[NW], sh_active[NW];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool mx = mode == FIN_AMAX;
  // this wave's fold of its members: running (below target, unconverged),
  // capped, broken down; largest iteration count and rr; members above their
  // refinement target (FIN_OUTER)
  int flags = 0, maxit = 0, active = 0;
  double rr = 0.0;
  auto fold = [&](const SolveScalars& q) {
    flags |= member_flags(q);
    maxit = q.iters > maxit ? q.iters : maxit;
    rr = q.rr > rr ? q.rr : rr;
    active += q.outer_s > q.target ? 1 : 0;
  };
  if (nm == 1) {
    // the member's record is fetched while the partials load
    SolveScalars q;
    if (threadIdx.x == 0) q = sc[0];
    // eight independent loads per thread in flight at a time
    double s = 0.0;
    for (int b0 = threadIdx.x; b0 < nblk; b0 += 8 * 1024) {
      double v[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        int b = b0 + k * 1024;
        v[k] = b < nblk ? partial[2 * b] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) s = mx ? fmax(s, v[k]) : s + v[k];
    }
    if (gd) return;  // uniform: before the barrier
    s = mx ? wave_max(s) : wave_sum(s);
    if (lane == 63) sh[wave] = s;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot = 0.0;
    for (int k = 0; k < NW; k++) tot = mx ? fmax(tot, sh[k]) : tot + sh[k];
    finish_member(mode, tot, q, reltol, abstol, maxiter, abs_target, 0);
    sc[0] = q;
    if (mode > FIN_OUTER) return;
    fold(q);
    if (mode == FIN_OUTER) mput(&hm_outer[0], q.outer_s);
  } else {
    if (gd) return;
    for (int m = wave; m < nm; m += NW) {
      // every block of the member's range but the first starts with the
      // member (slot 0); the first may start with the previous member
      int b0 = mblk[2 * m], b1 = mblk[2 * m + 1];
      double s = 0.0;
      for (int bb = b0 + 1 + lane; bb <= b1; bb += 8 * 64) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
          int b = bb + k * 64;
          v[k] = b <= b1 ? partial[2 * b] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; k++) s = mx ? fmax(s, v[k]) : s + v[k];
      }
      if (lane == 0) {
        double v = partial[2 * b0 + (bm[b0].x == m ? 0 : 1)];
        s = mx ? fmax(s, v) : s + v;
      }
      s = mx ? wave_max(s) : wave_sum(s);
      if (lane == 63) {
        SolveScalars q = sc[m];
        finish_member(mode, s, q, reltol, abstol, maxiter, abs_target, 1);
        sc[m] = q;
        fold(q);
        if (mode == FIN_OUTER) mput(&hm_outer[m], q.outer_s);
      }
    }
    if (mode > FIN_OUTER) return;
    if (lane == 63) {
      sh_flags[wave] = flags;
      sh_maxit[wave] = maxit;
      sh_active[wave] = active;
      sh_rr[wave] = rr;
    }
    // the barrier also completes every wave's mirror stores before thread 0
    // publishes below
    __syncthreads();
    if (threadIdx.x != 0) return;
    flags = 0;
    maxit = 0;
    active = 0;
    rr = 0.0;
    for (int k = 0; k < NW; k++) {
      flags |= sh_flags[k];
      maxit = sh_maxit[k] > maxit ? sh_maxit[k] : maxit;
      active += sh_active[k];
      rr = sh_rr[k] > rr ? sh_rr[k] : rr;
    }
  }
  if (mode == FIN_OUTER) {
    mput(&hm->active, active);
    mpublish(&hm->outer_seq, seq);
    return;
  }
  publish_solve(mode, flags, maxit, rr, nm == 1 ? sc[0].rr_stop : -1.0, seq, gdone, hm);
}

// FIN_XR of a single system, done by block 0 of the next kernel in place of a
// k_finish launch of its own: the step's results (rr, the iteration count,
// convergence) are needed only by the host and by the early exit of later
// kernels (which only saves work), not by that kernel's other blocks.
// mode < 0: none.
struct FoldFin {
  const double* partial;
  SolveScalars* sc;
  int* gdone;
  HostMirror* hm;
  double reltol, abstol;
  int nblk, mode, maxiter, abs_target, seq;
};

// The fold, called by every thread of block 0 (BLK threads) of a kernel that
// has not taken the early exit: the partials are summed exactly as k_finish's
// 1024 threads sum them (each thread here stands for four of them, the
// 16 wave sums combined in the same order), so the results are bitwise
// those of the k_finish launch it replaces.
__device__ void fold_finish(const FoldFin& f) {
  constexpr int VT = 1024, NW = VT / 64, J = VT / BLK;
  static_assert(VT % BLK == 0 && J % 2 == 0, "four virtual threads a thread");
  __shared__ double sh[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // two virtual threads at a time: sixteen loads in flight
#pragma unroll
  for (int j = 0; j < J; j += 2) {
    const int vt0 = threadIdx.x + j * BLK;  // virtual wave: vt >> 6 = wave + (BLK / 64) j
    double s0 = 0.0, s1 = 0.0;
    for (int b0 = vt0; b0 < f.nblk; b0 += 8 * VT) {
      double v0[8], v1[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        int b = b0 + k * VT;
        v0[k] = b < f.nblk ? f.partial[2 * b] : 0.0;
        v1[k] = b + BLK < f.nblk ? f.partial[2 * (b + BLK)] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        s0 += v0[k];
        s1 += v1[k];
      }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane == 63) {
      sh[wave + (BLK / 64) * j] = s0;
      sh[wave + (BLK / 64) * (j + 1)] = s1;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tot = 0.0;
  for (int k = 0; k < NW; k++) tot += sh[k];
  SolveScalars q = f.sc[0];
  finish_member(f.mode, tot, q, f.reltol, f.abstol, f.maxiter, f.abs_target, 0);
  f.sc[0] = q;
  // (k_finish's fold of one member: largest count and rr from 0)
  publish_solve(f.mode, member_flags(q), q.iters > 0 ? q.iters : 0, q.rr > 0.0 ? q.rr : 0.0,
                q.rr_stop, f.seq, f.gdone, f.hm);
}

// Start and width of row r's SELL entries.  Every caller runs the 64 rows of
// one slice in one wavefront, so the slice index is wave-uniform and its
// pointer loads are scalar loads.
template <typename T>
__device__ __forceinline__ void sell_row(const SellD<T>& A, int64_t r, int64_t& base, int& w) {
  int s = __builtin_amdgcn_readfirstlane((int)(r >> 6));
  int64_t p0 = A.sptr[s];
  base = p0 + (r & 63);
  w = (int)((A.sptr[s + 1] - p0) >> 6);
}

// The scalar part of a SELL row -- its slice's start and width and, for an
// offset slice, the slice's column offsets -- from loads that issue together,
// so that a kernel can put them in flight before it waits for anything else
// (the early-exit flag in particular).  Every lane of a wave is in the same
// slice; r must be a valid row.
struct SellRow {
  int64_t base;
  int w;
  int offs;  // nonzero: an offset slice, slot k of row r has column r + o[k]
  int32_t o[8];
  int32_t cb;  // compact columns: the slice's base column
};

template <typename T>
__device__ __forceinline__ SellRow sell_row_at(const SellD<T>& A, int64_t r) {
  SellRow R;
  int s = __builtin_amdgcn_readfirstlane((int)(r >> 6));
  int64_t p0 = A.sptr[s];
  R.base = p0 + (r & 63);
  R.w = (int)((A.sptr[s + 1] - p0) >> 6);
  R.offs = 0;
  R.cb = A.cbase ? A.cbase[s] : 0;
#pragma unroll
  for (int u = 0; u < 8; u++) R.o[u] = 0;
  if (A.soff) {
    // the record always holds 8 offsets (zero past the width): no
    // condition on its first word before the others load
    const int32_t* q = A.soff + 9 * (int64_t)s;
    R.offs = q[0];
#pragma unroll
    for (int u = 0; u < 8; u++) R.o[u] = q[1 + u];
  }
  return R;
}

// ------------------------------------------------------------------ edge maps

// k_c = exp(logK_c), once per cell (the maps below gather it per edge end).
__global__ void k_exp(int64_t n, const double* __restrict__ logK, double* __restrict__ k) {
  int64_t c = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (c < n) k[c] = exp(logK[c]);
}

// K_e = hm(k_a, k_b), hm(x, y) = 2 / (1/x + 1/y), k = exp(logK) per cell.
__global__ void k_hm_forward(int64_t ne, const int2* __restrict__ edges,
                             const double* __restrict__ k, double* __restrict__ Ke) {
  int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (e >= ne) return;
  int2 ab = edges[e];
  double ka = k[ab.x], kb = k[ab.y];
  Ke[e] = 2.0 / (1.0 / ka + 1.0 / kb);
}

// K_e = exp((logK_a + logK_b) / 2), the geometric mean of exp(logK).
__global__ void k_geo_forward(int64_t ne, const int2* __restrict__ edges,
                              const double* __restrict__ logK, double* __restrict__ Ke) {
  int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (e >= ne) return;
  int2 ab = edges[e];
  Ke[e] = exp(0.5 * (logK[ab.x] + logK[ab.y]));
}

// dlogK_c = sum_{e incident to c} dK_e K_e / 2, in edge-list order per cell.
__global__ void k_geo_pullback(int64_t n, const int64_t* __restrict__ inc_ptr,
                               const int32_t* __restrict__ inc_edge,
                               const double* __restrict__ Ke, const double* __restrict__ dKe,
                               double* __restrict__ dlogK) {
  int64_t c = logical_row();
  if (c >= n) return;
  double s = 0.0;
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
    int32_t e = inc_edge[p];
    s += dKe[e] * 0.5 * Ke[e];
  }
  dlogK[c] = s;
}

// dlogK_c = sum_{e incident to c} dK_e K_e k_other / (k_c + k_other), in
// edge-list order per cell (k = exp(logK) per cell).
__global__ void k_hm_pullback(int64_t n, const int64_t* __restrict__ inc_ptr,
                              const int32_t* __restrict__ inc_edge,
                              const int32_t* __restrict__ inc_other,
                              const double* __restrict__ k, const double* __restrict__ Ke,
                              const double* __restrict__ dKe, double* __restrict__ dlogK) {
  int64_t c = logical_row();
  if (c >= n) return;
  double kc = k[c];
  double s = 0.0;
  for (int64_t p = inc_ptr[c]; p < inc_ptr[c + 1]; p++) {
    int32_t e = inc_edge[p];
    double ko = k[inc_other[p]];
    s += dKe[e] * Ke[e] * ko / (kc + ko);
  }
  dlogK[c] = s;
}

// ------------------------------------------------------------------ assembly

// Row j of  M = tscale * dr/dh (+ diag(svec) if given), and optionally
// b_j = Q_i + sum over Dirichlet neighbours of K_e aol_e dh_o, for free cell
// i = f2n[j].  Values are written in FP64 and, if val32 is given, also in FP32
// and, if a16 is given, in the bf16 form of k_rowdiag (diagonal slot 0, the
// FP32 diagonal plus the off-diagonals' rounding error in dg32).
// invd = 1 / (m_jj + sum_k |m_jk|) (l1-Jacobi diagonal).
__global__ void __launch_bounds__(BLK) k_assemble(int64_t nf, const int32_t* __restrict__ f2n,
                           const int64_t* __restrict__ inc_ptr,
                           const int32_t* __restrict__ inc_edge,
                           const int32_t* __restrict__ inc_other,
                           const uint16_t* __restrict__ inc_slot,
                           const uint16_t* __restrict__ diag_slot,
                           const double* __restrict__ Ke, const double* __restrict__ aol,
                           const double* __restrict__ dh, const double* __restrict__ Q,
                           double tscale, const double* __restrict__ svec, SellD<double> A,
                           float* __restrict__ val32, double* __restrict__ invd64,
                           float* __restrict__ invd32, double* __restrict__ b,
                           uint16_t* __restrict__ a16, float* __restrict__ dg32,
                           int* __restrict__ cut) {
  // rows up to AW slots accumulate in LDS ([slot][lane]) and are written once;
  // wider rows (unstructured meshes) accumulate in place
  constexpr int AW = 8;
  __shared__ double acc[AW][BLK];
  int64_t j = logical_row();
  if (j >= nf) return;
  int64_t base;
  int w;
  sell_row(A, j, base, w);
  const int tid = threadIdx.x;
  const bool lds = w <= AW;
  auto slot = [&](int k) -> double& { return lds ? acc[k][tid] : A.val[base + (int64_t)k * 64]; };
  for (int k = 0; k < w; k++) slot(k) = 0.0;
  int32_t i = f2n[j];
  double dsum = 0.0, bj = Q ? Q[i] : 0.0;
  int64_t p1 = inc_ptr[i + 1];
  constexpr int U = 8;
  for (int64_t p0 = inc_ptr[i]; p0 < p1; p0 += U) {
    int32_t e[U];
    uint16_t sl[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      bool ok = p0 + u < p1;
      e[u] = ok ? inc_edge[p0 + u] : 0;
      sl[u] = ok ? inc_slot[p0 + u] : NO_SLOT_D;
    }
    double t[U];
#pragma unroll
    for (int u = 0; u < U; u++) t[u] = p0 + u < p1 ? Ke[e[u]] * aol[e[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (p0 + u >= p1) break;
      // a zero conductance may cut a component off every Dirichlet cell:
      // flagged for the steady solve's host check
      if (t[u] == 0.0 && cut) mput(cut, 1);
      dsum += t[u];
      if (sl[u] != NO_SLOT_D)
        slot(sl[u]) -= tscale * t[u];
      else
        bj += t[u] * dh[inc_other[p0 + u]];
    }
  }
  int ds = diag_slot[j];
  double dv = tscale * dsum + (svec ? svec[j] : 0.0);
  slot(ds) = dv;
  double l1 = dv;
  for (int k = 0; k < w; k++)
    if (k != ds) l1 += fabs(slot(k));
  double inv = 1.0 / l1;
  if (invd64) invd64[j] = inv;
  if (invd32) invd32[j] = (float)inv;
  float err = 0.0f;  // bf16 copy (with val32): rounding error of the off-diagonals
  for (int k = 0; k < w; k++) {
    double v = slot(k);
    int64_t e = base + (int64_t)k * 64;
    if (lds) A.val[e] = v;
    if (val32) val32[e] = (float)v;
    if (a16) {
      uint16_t bv = k == ds ? (uint16_t)0 : f32_to_bf16((float)v);
      a16[e] = bv;
      if (k != ds) err += (float)v - bf16_to_f32(bv);
    }
  }
  if (a16) dg32[j] = (float)dv + err;
  if (b) b[j] = bj;
}

// groundwater_residuals: r_j = -Q_i/s_i + sum_e K_e (h_i - h~_o) aol_e / s_i,
// term order of src/groundwater.jl:4-20 (source first, then edges in order).
__global__ void k_residuals(int64_t nf, const int32_t* __restrict__ f2n,
                            const int64_t* __restrict__ inc_ptr,
                            const int32_t* __restrict__ inc_edge,
                            const int32_t* __restrict__ inc_other,
                            const int32_t* __restrict__ n2f, const double* __restrict__ Ke,
                            const double* __restrict__ aol, const double* __restrict__ h,
                            const double* __restrict__ dh, const double* __restrict__ Q,
                            const double* __restrict__ Ss, const double* __restrict__ V,
                            double* __restrict__ r) {
  int64_t j = logical_row();
  if (j >= nf) return;
  int32_t i = f2n[j];
  double s = Ss[i] * V[i];
  double acc = -Q[i] / s;
  double hi = h[j];
  for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
    int32_t e = inc_edge[p], o = inc_other[p];
    int32_t fo = n2f[o];
    double ho = fo >= 0 ? h[fo] : dh[o];
    acc += Ke[e] * (hi - ho) * aol[e] / s;
  }
  r[j] = acc;
}

// Jacobians of groundwater_residuals over the free rows (the matrices the
// generated groundwater_h / _Ks / _dirichleths / _Qs of src/groundwater.jl:1
// return), values in the CSR order of their patterns: columns ascending,
// contributions to one entry summed in incidence (edge) order.  pos[p]: the
// value index of incidence p of the row's cell, -1 where the incidence adds
// no entry to this Jacobian; dpos[j]: row j's diagonal (kind 0).  s = Ss V:
//   kind 0 (h)           diagonal += K aol / s;  [free other] -= K aol / s
//   kind 1 (Ks)          [edge] = (h_i - h~_o) aol / s
//   kind 2 (dirichleths) [Dirichlet other] -= K aol / s
//   kind 3 (Qs)          [cell] = -1 / s
__global__ void k_jacobian(int kind, int64_t nf, const int32_t* __restrict__ f2n,
                           const int64_t* __restrict__ inc_ptr,
                           const int32_t* __restrict__ inc_edge,
                           const int32_t* __restrict__ inc_other,
                           const int32_t* __restrict__ n2f, const int64_t* __restrict__ rowptr,
                           const int32_t* __restrict__ pos, const int32_t* __restrict__ dpos,
                           const double* __restrict__ Ke, const double* __restrict__ aol,
                           const double* __restrict__ h, const double* __restrict__ dh,
                           const double* __restrict__ Ss, const double* __restrict__ V,
                           double* __restrict__ vals) {
  int64_t j = logical_row();
  if (j >= nf) return;
  int32_t i = f2n[j];
  double s = Ss[i] * V[i];
  if (kind == 3) {
    vals[j] = -1.0 / s;
    return;
  }
  if (kind != 1)
    for (int64_t q = rowptr[j]; q < rowptr[j + 1]; q++) vals[q] = 0.0;
  const double hi = kind == 1 ? h[j] : 0.0;
  double diag = 0.0;
  for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
    int32_t e = inc_edge[p], q = pos[p];
    if (kind == 1) {
      int32_t o = inc_other[p], fo = n2f[o];
      double ho = fo >= 0 ? h[fo] : dh[o];
      vals[q] = (hi - ho) * aol[e] / s;
    } else {
      double t = Ke[e] * aol[e] / s;
      if (kind == 0) diag += t;
      if (q >= 0) vals[q] -= t;
    }
  }
  if (kind == 0) vals[dpos[j]] += diag;
}

// ------------------------------------------------------------------ AMG numeric setup

// P(i,:) = e_agg(i) - omega invd_i sum_k a_ik e_agg(k) (invd: the l1 diagonal,
// whose spectral radius bound is 1; omega = 4 / (3 rho) with the caller's
// estimate rho of rho(D_l1^-1 A)).
// p16 (may be null): P's values also in bf16.
template <typename T>
__device__ __forceinline__ void put_val(T* val, uint16_t* v16, int64_t e, T v) {
  val[e] = v;
  if (v16) v16[e] = f32_to_bf16((float)v);
}

template <typename T>
__global__ void __launch_bounds__(BLK) k_psmooth(SellD<T> A, const T* __restrict__ invd,
                                                 const uint16_t* __restrict__ p_slot,
                                                 const uint16_t* __restrict__ p_self, SellD<T> P,
                                                 uint16_t* __restrict__ p16, double omega) {
  // the lane's P row accumulates in LDS ([slot][lane]: no bank conflicts)
  // instead of through read-modify-writes of global memory
  constexpr int PW = 16;
  __shared__ T acc[PW][BLK];
  int64_t i = logical_row();
  if (i >= A.nrows) return;
  int64_t pb, ab;
  int pw, aw;
  sell_row(P, i, pb, pw);
  sell_row(A, i, ab, aw);
  T f = (T)omega * invd[i];
  int tid = threadIdx.x;
  if (pw <= PW) {
    for (int k = 0; k < pw; k++) acc[k][tid] = (T)0;
    constexpr int U = 8;
    for (int k0 = 0; k0 < aw; k0 += U) {
      uint16_t sl[U];
      T v[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        int64_t e = ab + (int64_t)(k0 + u) * 64;
        bool ok = k0 + u < aw;
        sl[u] = ok ? p_slot[e] : NO_SLOT_D;
        v[u] = ok ? A.val[e] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (sl[u] != NO_SLOT_D) acc[sl[u]][tid] -= f * v[u];
    }
    acc[p_self[i]][tid] += (T)1;
    for (int k = 0; k < pw; k++) put_val(P.val, p16, pb + (int64_t)k * 64, acc[k][tid]);
  } else {
    for (int k = 0; k < pw; k++) P.val[pb + (int64_t)k * 64] = (T)0;
    for (int k = 0; k < aw; k++) {
      int64_t e = ab + (int64_t)k * 64;
      uint16_t sl = p_slot[e];
      if (sl != NO_SLOT_D) P.val[pb + (int64_t)sl * 64] -= f * A.val[e];
    }
    P.val[pb + (int64_t)p_self[i] * 64] += (T)1;
    if (p16)
      for (int k = 0; k < pw; k++)
        p16[pb + (int64_t)k * 64] = f32_to_bf16((float)P.val[pb + (int64_t)k * 64]);
  }
}

// R = P': every R entry copies the P entry it mirrors (also: a row-split
// copy of an operator).  out16 (may be null): the bf16 copy too, gathered
// from src16 where given (a bf16 operator whose diagonal lives apart), else
// rounded from the value.
template <typename T>
__global__ void k_rgather(int64_t nent, const int32_t* __restrict__ r_src,
                          const T* __restrict__ pval, T* __restrict__ rval,
                          const uint16_t* __restrict__ src16, uint16_t* __restrict__ out16) {
  int64_t e = logical_row();
  if (e >= nent) return;
  int32_t s = r_src[e];
  T v = s >= 0 ? pval[s] : (T)0;
  rval[e] = v;
  if (out16) out16[e] = src16 ? (s >= 0 ? src16[s] : (uint16_t)0) : f32_to_bf16((float)v);
}

// Galerkin products grouped by destination: list k holds the contributions
// of destination dest[k] (k itself when dest is null):
//   out[dest[k]] = sum_{q in [off[k], off[k+1])} X[s1[q]] * Y[s2[q]].
// Each list is spread over G = 2^lg consecutive lanes (lane g takes q = off[k]
// + g, g + G, ...) and combined by wave shuffles in a fixed pattern, so the
// result is reproducible.  Used for A*P (X = A, Y = P) and for the next
// level's operator R*(A*P) (X = P, since R(I,i) = P(i,I); Y = A*P); empty
// lists (SELL padding) give 0.
template <typename T>
__global__ void k_prod_gather(int64_t nlists, int lg, const int64_t* __restrict__ off,
                              const int32_t* __restrict__ dest, const int32_t* __restrict__ s1,
                              const int32_t* __restrict__ s2, const T* __restrict__ X,
                              const T* __restrict__ Y, T* __restrict__ out) {
  // each XCD takes a contiguous range of lists, so that the rows of X and Y
  // the lists gather from stay in that XCD's L2
  int64_t sub = logical_row();
  int64_t k = sub >> lg;
  int G = 1 << lg;
  T acc = (T)0;
  if (k < nlists) {
    // this lane's products q = off[k] + g, + G, ... in order, loads in groups of 8
    constexpr int U = 8;
    int64_t q1 = off[k + 1];
    for (int64_t q0 = off[k] + (sub & (G - 1)); q0 < q1; q0 += (int64_t)U * G) {
      int32_t a[U], b[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        int64_t q = q0 + (int64_t)u * G;
        a[u] = q < q1 ? s1[q] : 0;
        b[u] = q < q1 ? s2[q] : 0;
      }
      T xa[U], yb[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        bool ok = q0 + (int64_t)u * G < q1;
        xa[u] = ok ? X[a[u]] : (T)0;
        yb[u] = ok ? Y[b[u]] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (q0 + (int64_t)u * G < q1) acc += xa[u] * yb[u];
    }
  }
  acc = group_sum(acc, lg);
  if ((sub & (G - 1)) == G - 1 && k < nlists) out[dest ? dest[k] : k] = acc;
}

// A*P by rows (a GalerkinPlan's row-wise form): fine row i is spread over
// G = 2^lg lanes; lane g walks the row's A slots s = g, g + G, ... -- every
// slot of an offset slice (G = 1), or those below arowlen[i] (null: the
// slice width) -- and, for slot s with column j, the entries t < rowlen(j) of
// P's row j, adding a_s p_jt into the AP entry its product code names (255:
// the row holds no entry in that slot, the product is dropped).  The lanes'
// partial rows accumulate in LDS ([entry][lane]: no bank conflicts whatever
// the entries), are summed over the G lanes in a fixed pattern, and the
// group's last lane writes the row's ap_w or fewer entries in CSR order
// (ap_rp).  Dynamic LDS: ap_w * BLK * sizeof(T).
template <typename T>
__global__ void __launch_bounds__(BLK) k_ap_rows(SellD<T> A, const int32_t* __restrict__ arowlen,
                                                 int lg, int ap_w, SellD<T> P,
                                                 const int32_t* __restrict__ prowlen,
                                                 const int64_t* __restrict__ apc_ptr,
                                                 const uint8_t* __restrict__ apc,
                                                 const int64_t* __restrict__ ap_rp,
                                                 T* __restrict__ apv) {
  extern __shared__ unsigned char ap_lds[];
  T* acc = (T*)ap_lds;
  const int64_t sub = logical_row();
  const int64_t i = sub >> lg;
  const int G = 1 << lg, g = (int)(sub & (G - 1));
  const bool valid = i < A.nrows;
  const int tid = threadIdx.x;
  for (int k = 0; k < ap_w; k++) acc[k * BLK + tid] = (T)0;
  if (valid) {
    const SellRow R = sell_row_at(A, i);
    const int64_t cq = apc_ptr[sub >> 6] + (sub & 63);
    int64_t q = 0;
    // one A slot: the P row's entries, their codes and values in groups of 8
    auto slot = [&](T a, int32_t j) {
      const int64_t pb = P.sptr[j >> 6] + (j & 63);
      const int pw = prowlen[j];
      for (int t0 = 0; t0 < pw; t0 += 8) {
        T pv[8];
        uint8_t cd[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const bool ok = t0 + u < pw;
          pv[u] = ok ? P.val[pb + (int64_t)(t0 + u) * 64] : (T)0;
          cd[u] = ok ? apc[cq + (q + t0 + u) * 64] : (uint8_t)255;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
          if (cd[u] != 255) acc[cd[u] * BLK + tid] += a * pv[u];
      }
      q += pw;
    };
    if (R.offs) {
      // an offset slice (G = 1): slot s has column i + o[s] (unrolled: the
      // offsets stay in registers)
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (s < R.w) slot(A.val[R.base + (int64_t)s * 64], (int32_t)i + R.o[s]);
    } else {
      const int wa = arowlen ? arowlen[i] : R.w;
      for (int s = g; s < wa; s += G) {
        const int64_t e = R.base + (int64_t)s * 64;
        const int32_t j = A.col16 ? R.cb + (int32_t)A.col16[e] : A.col[e];
        slot(A.val[e], j);
      }
    }
  }
  // every lane of a group takes part in the sums (groups lie within a wave)
  const int64_t rb = valid ? ap_rp[i] : 0;
  const int m = valid ? (int)(ap_rp[i + 1] - rb) : 0;
  for (int k = 0; k < ap_w; k++) {
    T v = group_sum(acc[k * BLK + tid], lg);
    if (g == G - 1 && k < m) apv[rb + k] = v;
  }
}

// The next level's entries on and above the diagonal by rows (a
// GalerkinPlan's row-wise form), over R's row-split copy Rs: sub-row (I, g)
// holds the entries t of R's row I with t = g mod 2^lg (at most 8), each with
// fine row i_t and value P(i_t, I) = pval[rs_src].  The sub-row's products
// AP(i_t, k) P(i_t, I), coded t | k << 3 | d << 8, come sorted by
// destination d and then t, so each destination's products are one run summed
// in a register (ascending fine row) and stored once into the lane's partial
// row in LDS; the partial rows are summed over the lanes in a fixed pattern
// and the group's last lane writes the upper entries.  Dynamic LDS:
// rap_w * BLK * sizeof(T).
template <typename T>
__global__ void __launch_bounds__(BLK) k_rap_rows(
    SellD<T> Rs, const int32_t* __restrict__ rsrowlen, int lg, int rap_w, int64_t nc,
    const int32_t* __restrict__ rs_src, const int64_t* __restrict__ rcd_ptr,
    const uint16_t* __restrict__ rcd, const uint16_t* __restrict__ rnp,
    const T* __restrict__ pval, const int64_t* __restrict__ ap_rp, const T* __restrict__ apv,
    const int64_t* __restrict__ csptr, const int32_t* __restrict__ crowlen,
    const uint16_t* __restrict__ cdiag, T* __restrict__ cval) {
  extern __shared__ unsigned char rap_lds[];
  T* acc = (T*)rap_lds;
  const int64_t sub = logical_row();
  const int64_t I = sub >> lg;
  const int G = 1 << lg, g = (int)(sub & (G - 1));
  const bool valid = I < nc;
  const int tid = threadIdx.x;
  for (int d = 0; d < rap_w; d++) acc[d * BLK + tid] = (T)0;
  if (valid) {
    const SellRow R = sell_row_at(Rs, sub);
    const int rw = rsrowlen[sub];
    const int np = rnp[sub];
    const int64_t cq = rcd_ptr[sub >> 6] + (sub & 63);
    // the sub-row's entries: fine row, P value, AP row start
    int32_t src[8], fi[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int64_t e = R.base + (int64_t)t * 64;
      const bool ok = t < rw;
      fi[t] = ok ? (Rs.col16 ? R.cb + (int32_t)Rs.col16[e] : Rs.col[e]) : 0;
      src[t] = ok ? rs_src[e] : 0;
    }
    T pv[8];
    int64_t ab[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const bool ok = t < rw;
      pv[t] = ok ? pval[src[t]] : (T)0;
      ab[t] = ok ? ap_rp[fi[t]] : 0;
    }
    T run = (T)0;
    int dcur = -1;
    for (int q0 = 0; q0 < np; q0 += 8) {
      uint16_t cd[8];
#pragma unroll
      for (int u = 0; u < 8; u++) cd[u] = q0 + u < np ? rcd[cq + (int64_t)(q0 + u) * 64] : (uint16_t)0;
      T av[8], pu[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = cd[u] & 7;
        int64_t a = ab[0];
        T p = pv[0];
#pragma unroll
        for (int v = 1; v < 8; v++)
          if (t == v) {
            a = ab[v];
            p = pv[v];
          }
        pu[u] = p;
        av[u] = q0 + u < np ? apv[a + ((cd[u] >> 3) & 31)] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (q0 + u >= np) break;
        const int d = cd[u] >> 8;
        if (d != dcur) {
          if (dcur >= 0) acc[dcur * BLK + tid] = run;
          run = (T)0;
          dcur = d;
        }
        run += pu[u] * av[u];
      }
    }
    if (dcur >= 0) acc[dcur * BLK + tid] = run;
  }
  const int ds = valid ? cdiag[I] : 0;
  const int up = valid ? crowlen[I] - ds : 0;
  const int64_t cb = valid ? csptr[I >> 6] + (I & 63) : 0;
  for (int d = 0; d < rap_w; d++) {
    T v = group_sum(acc[d * BLK + tid], lg);
    if (g == G - 1 && d < up) cval[cb + (int64_t)(ds + d) * 64] = v;
  }
}

// val[dst[m]] = val[src[m]]: the entries below the diagonal of a symmetric
// Galerkin operator from their mirrors above it (disjoint sets).
template <typename T>
__global__ void k_mirror_copy(int64_t nm, const int32_t* __restrict__ dst,
                              const int32_t* __restrict__ src, T* __restrict__ val) {
  int64_t m = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (m < nm) val[dst[m]] = val[src[m]];
}

// invd_i = 1 / (a_ii + sum_{j != i} |a_ij|) for a level operator.  With a16
// (FP32 only), also the operator's bf16 copy: off-diagonal values rounded to
// bf16, the diagonal slot 0, and dg_i = a_ii + sum_j (a_ij - bf16(a_ij)) over
// the off-diagonals, so that the copy's rows (with dg as their diagonal) sum
// exactly as the operator's do: the operator's near-null space, and with it
// positive definiteness, survive the rounding.
template <typename T>
__global__ void k_rowdiag(SellD<T> A, const uint16_t* __restrict__ diag_slot,
                          T* __restrict__ invd, uint16_t* __restrict__ a16, T* __restrict__ dg) {
  int64_t i = logical_row();
  if (i >= A.nrows) return;
  int64_t base;
  int w;
  sell_row(A, i, base, w);
  int ds = diag_slot[i];
  T dv = A.val[base + (int64_t)ds * 64];
  T l1 = dv;
  float err = 0.0f;
  constexpr int U = 8;
  for (int k0 = 0; k0 < w; k0 += U) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = k0 + u < w ? A.val[base + (int64_t)(k0 + u) * 64] : (T)0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (k0 + u >= w) continue;
      int64_t e = base + (int64_t)(k0 + u) * 64;
      if (k0 + u != ds) {
        l1 += fabs(v[u]);
        if (a16) {
          uint16_t b = f32_to_bf16((float)v[u]);
          a16[e] = b;
          err += (float)v[u] - bf16_to_f32(b);
        }
      } else if (a16) {
        a16[e] = 0;
      }
    }
  }
  invd[i] = (T)1 / l1;
  if (a16) dg[i] = (T)((float)dv + err);
}

// In-place Gauss-Jordan inverse of the (SPD) coarsest operator, in FP64, by
// one workgroup.  W is nc*nc, row-major.  status[0] = 1 on a non-positive pivot.
template <typename T>
__global__ void __launch_bounds__(1024) k_dense_inverse(SellD<T> Ac, const int32_t* __restrict__ rowlen,
                                                        double* __restrict__ W, int* status) {
  int nc = (int)Ac.nrows;
  int tid = threadIdx.x, nt = blockDim.x;
  // up to DENSE_LDS_MAX rows the elimination runs in LDS and W is written once
  extern __shared__ double wl[];
  double* Wk = nc <= DENSE_LDS_MAX && nc * nc <= 8 * nt ? wl : W;
  for (int64_t t = tid; t < (int64_t)nc * nc; t += nt) Wk[t] = 0.0;
  __syncthreads();
  // entry k of row I for every (I, k) with k < rowlen[I] <= nc, spread over
  // the whole workgroup (a row's columns are distinct: plain stores)
  for (int64_t e = tid; e < (int64_t)nc * nc; e += nt) {
    int I = (int)(e % nc), k = (int)(e / nc);
    if (k >= rowlen[I]) continue;
    int64_t p = Ac.sptr[I >> 6] + (I & 63) + (int64_t)k * 64;
    Wk[(int64_t)I * nc + Ac.col[p]] = (double)Ac.val[p];
  }
  __syncthreads();
  __shared__ int bad;
  if (tid == 0) bad = 0;
  __syncthreads();
  if (Wk != W) {
    // Small systems: each thread keeps at most 8 fixed positions of W in
    // registers; the owners of step k's pivot row and column publish them in
    // LDS (double-buffered by step parity, so one barrier per step suffices),
    // and every thread updates its own positions.  Same arithmetic, in the
    // same order, as the in-memory elimination below.
    __shared__ double prow[2][DENSE_LDS_MAX], pcol[2][DENSE_LDS_MAX];
    int pi[8], pj[8];
    double v[8];
    const int nn = nc * nc;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int t = tid + q * nt;
      pi[q] = t < nn ? t / nc : -1;  // -1: no position
      pj[q] = t < nn ? t - pi[q] * nc : -1;
      v[q] = t < nn ? Wk[t] : 0.0;
    }
    const int npos = (nn + nt - 1) / nt;  // positions per thread, uniform
    for (int k = 0; k < nc; k++) {
      const int b = k & 1;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if (q >= npos) break;
        if (pi[q] == k) prow[b][pj[q]] = v[q];
        if (pj[q] == k) pcol[b][pi[q]] = v[q];
      }
      __syncthreads();
      double piv = prow[b][k];
      if (!(piv > 0.0)) {
        if (tid == 0) bad = 1;
        piv = 1.0;
      }
      double ip = 1.0 / piv;
      // branch-free: the elimination update, the pivot row / column scaling
      // and the pivot itself, selected per position
#pragma unroll
      for (int q = 0; q < 8; q++) {
        if (q >= npos) break;
        int i = pi[q], j = pj[q];
        bool ok = i >= 0, ri = i == k, cj = j == k;
        double a = pcol[b][ok ? i : 0], c = prow[b][ok ? j : 0];
        double upd = v[q] - a * c * ip;
        double scl = v[q] * (ri ? ip : -ip);
        double nv = (!ri && !cj) ? upd : ((ri && cj) ? ip : scl);
        v[q] = ok ? nv : v[q];
      }
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (pi[q] >= 0) W[(int64_t)pi[q] * nc + pj[q]] = v[q];
    __syncthreads();
    if (tid == 0) status[0] = bad;
    return;
  }
  for (int k = 0; k < nc; k++) {
    double piv = Wk[(int64_t)k * nc + k];
    if (!(piv > 0.0)) {
      if (tid == 0) bad = 1;
      piv = 1.0;
    }
    double ip = 1.0 / piv;
    for (int64_t t = tid; t < (int64_t)nc * nc; t += nt) {
      int i = (int)(t / nc), j = (int)(t % nc);
      if (i != k && j != k) Wk[t] -= Wk[(int64_t)i * nc + k] * Wk[(int64_t)k * nc + j] * ip;
    }
    __syncthreads();
    for (int t = tid; t < nc; t += nt) {
      if (t != k) {
        Wk[(int64_t)k * nc + t] *= ip;
        Wk[(int64_t)t * nc + k] *= -ip;
      }
    }
    __syncthreads();
    if (tid == 0) Wk[(int64_t)k * nc + k] = ip;
    __syncthreads();
  }
  if (tid == 0) status[0] = bad;
}

// The same inverse for a coarsest operator of at most 64 rows, by one
// 256-thread workgroup: a 16 x 16 grid of threads, each holding a 4 x 4 tile
// of the zero-padded 64 x 64 matrix in registers (the padding never mixes
// with the matrix: its pivot-row and pivot-column entries are 0).  Per pivot
// step the owners of row and column k publish them in LDS (double-buffered
// by step parity: one barrier a step) and every thread updates its 16
// positions with 4 + 4 pivot values, the same arithmetic per position as
// k_dense_inverse.
template <typename T>
__global__ void __launch_bounds__(256) k_dense_inverse64(SellD<T> Ac,
                                                         const int32_t* __restrict__ rowlen,
                                                         double* __restrict__ W, int* status) {
  constexpr int N = 64;
  const int nc = (int)Ac.nrows;  // <= N
  const int tid = threadIdx.x;
  __shared__ double Ad[N][N + 1];  // (+1: the column reads below do not conflict)
  __shared__ double prow[2][N], pcol[2][N];
  __shared__ int bad;
  for (int t = tid; t < N * (N + 1); t += 256) (&Ad[0][0])[t] = 0.0;
  if (tid == 0) bad = 0;
  __syncthreads();
  {
    // row I = lane, entries k = wave, wave + 4, ...: eight independent loads
    // a lane in flight (one slice: nc <= 64)
    const int I = tid & 63;
    const int len = I < nc ? rowlen[I] : 0;
    const int64_t p0 = Ac.sptr[0] + I;
    for (int k0 = tid >> 6; k0 < len; k0 += 32) {
      int32_t c[8];
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int k = k0 + 4 * u;
        const bool ok = k < len;
        c[u] = ok ? Ac.col[p0 + (int64_t)k * 64] : 0;
        v[u] = ok ? (double)Ac.val[p0 + (int64_t)k * 64] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (k0 + 4 * u < len) Ad[I][c[u]] = v[u];
    }
  }
  __syncthreads();
  const int ty = tid >> 4, tx = tid & 15;
  double v[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) v[a][b] = Ad[4 * ty + a][4 * tx + b];
  for (int k = 0; k < nc; k++) {
    const int bb = k & 1, kt = k >> 2, kq = k & 3;
    // publish row k and column k (static register indices: the step's
    // position inside the tile is selected, not indexed)
#pragma unroll
    for (int a = 0; a < 4; a++) {
      if (ty == kt && a == kq)
#pragma unroll
        for (int b = 0; b < 4; b++) prow[bb][4 * tx + b] = v[a][b];
      if (tx == kt && a == kq)
#pragma unroll
        for (int b = 0; b < 4; b++) pcol[bb][4 * ty + b] = v[b][a];
    }
    __syncthreads();
    double piv = prow[bb][k];
    if (!(piv > 0.0)) {
      if (tid == 0) bad = 1;
      piv = 1.0;
    }
    const double ip = 1.0 / piv;
    double pc[4], pr[4];
#pragma unroll
    for (int a = 0; a < 4; a++) {
      pc[a] = pcol[bb][4 * ty + a];
      pr[a] = prow[bb][4 * tx + a];
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const bool ri = 4 * ty + a == k, cj = 4 * tx + b == k;
        const double upd = v[a][b] - pc[a] * pr[b] * ip;
        const double scl = v[a][b] * (ri ? ip : -ip);
        v[a][b] = (!ri && !cj) ? upd : ((ri && cj) ? ip : scl);
      }
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int i = 4 * ty + a, j = 4 * tx + b;
      if (i < nc && j < nc) W[(int64_t)i * nc + j] = v[a][b];
    }
  __syncthreads();
  if (tid == 0) status[0] = bad;
}

// ------------------------------------------------------------------ V-cycle

// Groups of N consecutive entries of one SELL row (N known at compile time):
// all values and columns load, then all gathers, then the sums in ascending
// entry order.  A whole group sits in one basic block, so its loads issue
// back to back.  (Entry loads guarded by per-entry conditions make the
// compiler wait for each load at the join of its branch, which serializes
// the group.)
template <int N, bool C16, bool B16, typename T, typename TX>
__device__ __forceinline__ T dot_group(const SellD<T>& A, int64_t e0, int32_t cb,
                                       const TX* __restrict__ x, T s) {
  T v[N];
  int32_t c[N];
  TX xv[N];
#pragma unroll
  for (int u = 0; u < N; u++) {
    const int64_t e = e0 + (int64_t)u * 64;
    if constexpr (B16)
      v[u] = (T)bf16_to_f32(A.v16[e]);
    else
      v[u] = A.val[e];
    if constexpr (C16)
      c[u] = cb + (int32_t)A.col16[e];  // compact: the slice's base + 16-bit offset
    else
      c[u] = A.col[e];
  }
#pragma unroll
  for (int u = 0; u < N; u++) xv[u] = x[c[u]];
#pragma unroll
  for (int u = 0; u < N; u++) s += v[u] * (T)xv[u];
  return s;
}

// The row sum over explicit columns: groups of 8, then one group of the
// remaining 1..7 entries.
template <bool C16, bool B16, typename T, typename TX>
__device__ __forceinline__ T explicit_dot(const SellD<T>& A, int64_t base, int w, int32_t cb,
                                          const TX* __restrict__ x) {
  T s = (T)0;
  int k0 = 0;
  for (; k0 + 8 <= w; k0 += 8)
    s = dot_group<8, C16, B16>(A, base + (int64_t)k0 * 64, cb, x, s);
  const int64_t e0 = base + (int64_t)k0 * 64;
  switch (w - k0) {
    case 1: s = dot_group<1, C16, B16>(A, e0, cb, x, s); break;
    case 2: s = dot_group<2, C16, B16>(A, e0, cb, x, s); break;
    case 3: s = dot_group<3, C16, B16>(A, e0, cb, x, s); break;
    case 4: s = dot_group<4, C16, B16>(A, e0, cb, x, s); break;
    case 5: s = dot_group<5, C16, B16>(A, e0, cb, x, s); break;
    case 6: s = dot_group<6, C16, B16>(A, e0, cb, x, s); break;
    case 7: s = dot_group<7, C16, B16>(A, e0, cb, x, s); break;
    default: break;
  }
  return s;
}

// The same over an offset slice's N slots: columns r + o[u], no column loads.
template <int N, bool B16, typename T, typename TX>
__device__ __forceinline__ T dot_group_off(const SellD<T>& A, int64_t e0, int32_t r,
                                           const int32_t* o, const TX* __restrict__ x, T s) {
  T v[N];
  TX xv[N];
#pragma unroll
  for (int u = 0; u < N; u++) {
    const int64_t e = e0 + (int64_t)u * 64;
    if constexpr (B16)
      v[u] = (T)bf16_to_f32(A.v16[e]);
    else
      v[u] = A.val[e];
    xv[u] = x[r + o[u]];
  }
#pragma unroll
  for (int u = 0; u < N; u++) s += v[u] * (T)xv[u];
  return s;
}

template <bool B16, typename T, typename TX>
__device__ __forceinline__ T offset_dot(const SellD<T>& A, int64_t base, int w, int32_t r,
                                        const int32_t* o, const TX* __restrict__ x) {
  const T s = (T)0;
  switch (w) {
    case 1: return dot_group_off<1, B16>(A, base, r, o, x, s);
    case 2: return dot_group_off<2, B16>(A, base, r, o, x, s);
    case 3: return dot_group_off<3, B16>(A, base, r, o, x, s);
    case 4: return dot_group_off<4, B16>(A, base, r, o, x, s);
    case 5: return dot_group_off<5, B16>(A, base, r, o, x, s);
    case 6: return dot_group_off<6, B16>(A, base, r, o, x, s);
    case 7: return dot_group_off<7, B16>(A, base, r, o, x, s);
    default: return dot_group_off<8, B16>(A, base, r, o, x, s);
  }
}

// The p-forming row sum sum_u A[e_u] (z[c_u] + beta p[c_u]) over a group.
template <int N, bool OFF, typename T, typename TZ>
__device__ __forceinline__ T dot_group_zp(const SellD<T>& A, int64_t e0, int32_t r,
                                          const int32_t* o, const TZ* __restrict__ z,
                                          const T* __restrict__ p, T beta, T s) {
  T v[N], zv[N], pv[N];
  int32_t c[N];
#pragma unroll
  for (int u = 0; u < N; u++) {
    v[u] = A.val[e0 + (int64_t)u * 64];
    c[u] = OFF ? r + o[u] : A.col[e0 + (int64_t)u * 64];
  }
#pragma unroll
  for (int u = 0; u < N; u++) {
    zv[u] = (T)z[c[u]];
    pv[u] = p[c[u]];
  }
#pragma unroll
  for (int u = 0; u < N; u++) s += v[u] * (zv[u] + beta * pv[u]);
  return s;
}

// sum_k A.val[e_k] x[A.col[e_k]] over one SELL row (e_k = base + 64 k, k < w),
// added in ascending k, in groups of 8 entries and one final group of the
// remaining 1..7 (the width is the same for every lane of a slice, so the
// group sizes are uniform): a lane keeps up to 8 independent gather chains in
// flight, where one chain at a time would leave the sweeps latency-bound.
// Columns come from col, or compact from cbase + col16; values from val, or
// from v16 (bf16).
// In an offset slice (fine operator) the columns come from the slice record
// already in registers, so the row's value and x loads all issue at once.
template <typename T, typename TX>
__device__ __forceinline__ T sell_dot(const SellD<T>& A, const SellRow& R, int64_t r,
                                      const TX* __restrict__ x) {
  const int64_t base = R.base;
  const int w = R.w;
  if (R.offs)
    return A.v16 ? offset_dot<true>(A, base, w, (int32_t)r, R.o, x)
                 : offset_dot<false>(A, base, w, (int32_t)r, R.o, x);
  if (A.col16)
    return A.v16 ? explicit_dot<true, true>(A, base, w, R.cb, x)
                 : explicit_dot<true, false>(A, base, w, R.cb, x);
  return A.v16 ? explicit_dot<false, true>(A, base, w, R.cb, x)
               : explicit_dot<false, false>(A, base, w, R.cb, x);
}

// The same row sum for a caller that has only the row's start and width.
template <typename T, typename TX>
__device__ __forceinline__ T sell_dot(const SellD<T>& A, int64_t r, int64_t base, int w,
                                      const TX* __restrict__ x) {
  SellRow R = sell_row_at(A, r);
  R.base = base;
  R.w = w;
  return sell_dot(A, R, r, x);
}

// sum_k A.val[e_k] (z[c_k] + beta p[c_k]), same grouping and order as sell_dot
template <typename T, typename TZ>
__device__ __forceinline__ T sell_dot_zp(const SellD<T>& A, const SellRow& R, int64_t r,
                                         const TZ* __restrict__ z, const T* __restrict__ p,
                                         T beta) {
  const int64_t base = R.base;
  const int w = R.w;
  const int32_t rr = (int32_t)r;
  T s = (T)0;
  if (R.offs) {
    switch (w) {
      case 1: return dot_group_zp<1, true>(A, base, rr, R.o, z, p, beta, s);
      case 2: return dot_group_zp<2, true>(A, base, rr, R.o, z, p, beta, s);
      case 3: return dot_group_zp<3, true>(A, base, rr, R.o, z, p, beta, s);
      case 4: return dot_group_zp<4, true>(A, base, rr, R.o, z, p, beta, s);
      case 5: return dot_group_zp<5, true>(A, base, rr, R.o, z, p, beta, s);
      case 6: return dot_group_zp<6, true>(A, base, rr, R.o, z, p, beta, s);
      case 7: return dot_group_zp<7, true>(A, base, rr, R.o, z, p, beta, s);
      default: return dot_group_zp<8, true>(A, base, rr, R.o, z, p, beta, s);
    }
  }
  int k0 = 0;
  for (; k0 + 8 <= w; k0 += 8)
    s = dot_group_zp<8, false>(A, base + (int64_t)k0 * 64, rr, R.o, z, p, beta, s);
  const int64_t e0 = base + (int64_t)k0 * 64;
  switch (w - k0) {
    case 1: s = dot_group_zp<1, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 2: s = dot_group_zp<2, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 3: s = dot_group_zp<3, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 4: s = dot_group_zp<4, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 5: s = dot_group_zp<5, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 6: s = dot_group_zp<6, false>(A, e0, rr, R.o, z, p, beta, s); break;
    case 7: s = dot_group_zp<7, false>(A, e0, rr, R.o, z, p, beta, s); break;
    default: break;
  }
  return s;
}

// The scalar part of row sub of a row-split operator (clamped into the
// matrix for the lanes past its end, which add nothing).
template <typename T>
__device__ __forceinline__ SellRow split_row_at(const SellD<T>& A, int64_t sub) {
  return sell_row_at(A, sub < A.nrows ? sub : A.nrows - 1);
}

// (A x)_r with the rows of A split over G = 2^lg consecutive lanes: the
// group's last lane returns the full row sum (fixed reduction pattern).  The
// caller's lanes must all reach this call (no early exit before it).
template <typename T>
__device__ __forceinline__ T split_row_dot(const SellD<T>& A, const SellRow& R, int64_t sub,
                                           int lg, const T* __restrict__ x) {
  T ax = sub < A.nrows ? sell_dot(A, R, sub, x) : (T)0;
  return group_sum(ax, lg);
}

// The V-cycle and CG kernels below test the early-exit flag (*done) only
// once their first loads are in flight: the flag's load then overlaps with
// the row's slice pointers and own entries instead of preceding them.  The
// test comes before every store and every block reduction, and is uniform.

// One Chebyshev step of the l1-Jacobi-preconditioned level operator:
//   d <- c1 d + c2 invd (b - A xin);  xout = xin + d,
// A given with its rows split over 2^lg lanes (lg = 0 on the finest level).
// d is read from din (only when c1 != 0: not on the first sweep after a
// coarse correction) and written to dout (null on a level's last pre- and
// post-smoothing sweeps, whose d no later step reads).
// With fuse_dot, also writes the block partials of b . xout (the fine-level
// sweep that completes z = M^-1 r, where b is the CG residual).
// FINE marks the finest level's instantiation (a separate symbol in profiles).
// dg (null: none): the diagonal of a bf16 operator, kept in FP32 outside its
// values (whose diagonal slots hold 0); the row's own lane adds dg_r x_r.
template <typename T, bool FINE>
__global__ void k_cheb(SellD<T> A, int lg, int64_t n, const T* __restrict__ xin,
                       T* __restrict__ xout, const T* __restrict__ b, const T* din, T* dout,
                       const T* __restrict__ invd, const T* __restrict__ dg, T c1, T c2,
                       const int* __restrict__ done, int fuse_dot,
                       const int2* __restrict__ bm, double* __restrict__ partial) {
  const int dn = *done;
  int64_t sub = logical_row();
  int64_t r = sub >> lg;
  const SellRow R = split_row_at(A, sub);
  // the row's own vector entries are loaded before the row sum, so that they
  // travel with the operator's loads instead of after them
  const bool own = (sub & ((1 << lg) - 1)) == (1 << lg) - 1 && r < n;
  T br = own ? b[r] : (T)0, ir = own ? invd[r] : (T)0, xr = own ? xin[r] : (T)0;
  T dr = own && c1 != (T)0 ? din[r] : (T)0;
  T gr = own && dg ? dg[r] : (T)0;
  if (dn || R.w < 0) return;
  T ax = split_row_dot(A, R, sub, lg, xin);
  if (dg) ax += gr * xr;
  double local = 0.0;
  if (own) {
    T dnew = c2 * ir * (br - ax);
    if (c1 != (T)0) dnew = c1 * dr + dnew;
    if (dout) dout[r] = dnew;
    T xo = xr + dnew;
    xout[r] = xo;
    local = (double)br * (double)xo;
  }
  if (fuse_dot) {  // finest level only (lg == 0): row == sub
    bool second;
    member_of(bm, sub, second);
    block_partial2(local, second, partial, bm != nullptr);
  }
}

// r = b - A x  (rows of A split over 2^lg lanes).  FOLD: the finest level's
// instantiation, whose block 0 also does the CG step's FIN_XR (ff.mode >= 0).
template <typename T, bool FOLD>
__global__ void k_residual(SellD<T> A, int lg, int64_t n, const T* __restrict__ x,
                           const T* __restrict__ b, const T* __restrict__ dg,
                           T* __restrict__ r, const int* __restrict__ done, FoldFin ff) {
  const int dn = *done;
  int64_t sub = logical_row();
  int64_t i = sub >> lg;
  const SellRow R = split_row_at(A, sub);
  const bool own = (sub & ((1 << lg) - 1)) == (1 << lg) - 1 && i < n;
  T bi = own ? b[i] : (T)0;  // loaded with the operator, not after it
  T gx = own && dg ? dg[i] * x[i] : (T)0;
  if (dn || R.w < 0) return;  // (R.w >= 0 always: every thread of block 0 reaches the fold)
  T ax = split_row_dot(A, R, sub, lg, x);
  if (own) r[i] = bi - (ax + gx);
  if constexpr (FOLD) {
    if (ff.mode >= 0 && blockIdx.x == 0) fold_finish(ff);
  }
}

// bc = R rf with every row of R split over G = 2^lg consecutive lanes
// (sub-row I*G + g); the lane partials are combined by wave shuffles in a
// fixed pattern.  If init, also the zero-guess first Chebyshev step on the
// coarse level: dc = xc = c0 invdc bc.
template <typename T>
__global__ void k_restrict(SellD<T> Rs, int lg, int64_t nc, const T* __restrict__ rf,
                           T* __restrict__ bc, const T* __restrict__ invdc,
                           T* __restrict__ dc, T* __restrict__ xc, T c0, int init,
                           const int* __restrict__ done) {
  const int dn = *done;
  int64_t sr = logical_row();
  int64_t I = sr >> lg;
  const SellRow R = split_row_at(Rs, sr);
  const bool own = (sr & ((1 << lg) - 1)) == (1 << lg) - 1 && I < nc;
  T iv = own && init ? invdc[I] : (T)0;  // loaded with R, not after the row sum
  if (dn || R.w < 0) return;
  T s = split_row_dot(Rs, R, sr, lg, rf);
  if (own) {
    bc[I] = s;
    if (init) {
      T dd = c0 * iv * s;
      dc[I] = dd;
      xc[I] = dd;
    }
  }
}

// Dense coarsest solve folded into the prolongation that follows it: every
// workgroup loads W (nc <= COARSE_FOLD_MAX rows, at most 32 KiB) and bc into
// LDS with independent loads, forms xc = W bc there (each row in the order
// k_coarse_solve uses), then xf += P xc for its rows.  One launch instead of
// two; results identical to k_coarse_solve + k_prolong.
constexpr int COARSE_FOLD_MAX = 64;
static_assert(COARSE_FOLD_MAX * COARSE_FOLD_MAX % BLK == 0 && COARSE_FOLD_MAX <= BLK,
              "the fold kernel's W staging assumes these");
template <typename T>
__global__ void __launch_bounds__(BLK)
    k_prolong_coarse(SellD<T> P, int nc, const double* __restrict__ W,
                     const T* __restrict__ bc, T* __restrict__ xf,
                     const int* __restrict__ done) {
  const int dn = *done;
  __shared__ double ws[COARSE_FOLD_MAX * COARSE_FOLD_MAX];
  __shared__ double bs[COARSE_FOLD_MAX], xs[COARSE_FOLD_MAX];
  const int64_t i = logical_row();
  const SellRow R = sell_row_at(P, i < P.nrows ? i : P.nrows - 1);
  const T xi = i < P.nrows ? xf[i] : (T)0;  // in flight during the coarse solve
  // all of a thread's loads of W in flight at once, then the LDS stores
  constexpr int PER = COARSE_FOLD_MAX * COARSE_FOLD_MAX / BLK;
  const int nn = nc * nc;
  double wv[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int t = threadIdx.x + k * BLK;
    wv[k] = t < nn ? W[t] : 0.0;
  }
  double bv = threadIdx.x < nc ? (double)bc[threadIdx.x] : 0.0;
  if (dn || R.w < 0) return;  // uniform: before the block's barriers
  if (threadIdx.x < nc) bs[threadIdx.x] = bv;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int t = threadIdx.x + k * BLK;
    if (t < nn) ws[t] = wv[k];
  }
  __syncthreads();
  if (threadIdx.x < nc) {
    int ic = threadIdx.x;
    double a = 0.0;
#pragma unroll 8
    for (int j = 0; j < nc; j++) a += ws[j * nc + ic] * bs[j];
    xs[ic] = (double)(T)a;
  }
  __syncthreads();
  if (i >= P.nrows) return;
  xf[i] = xi + sell_dot(P, R, i, (const double*)xs);
}

// xf += P xc
template <typename T>
__global__ void k_prolong(SellD<T> P, const T* __restrict__ xc, T* __restrict__ xf,
                          const int* __restrict__ done) {
  const int dn = *done;
  int64_t i = logical_row();
  if (i >= P.nrows) return;
  const SellRow R = sell_row_at(P, i);
  T xi = xf[i];  // loaded with P, not after the row sum
  if (dn || R.w < 0) return;
  xf[i] = xi + sell_dot(P, R, i, xc);
}

// Coarsest level: x = W b with the FP64 inverse W (symmetric, so column j
// of W is row j and the loads are coalesced).
template <typename T>
__global__ void k_coarse_solve(int nc, const double* __restrict__ W, const T* __restrict__ b,
                               T* __restrict__ x, const int* __restrict__ done) {
  const int dn = *done;
  extern __shared__ double bs[];
  for (int j = threadIdx.x; j < nc; j += blockDim.x) bs[j] = (double)b[j];
  if (dn) return;  // uniform: before the barrier
  __syncthreads();
  for (int i = threadIdx.x; i < nc; i += blockDim.x) {
    // sixteen loads of W in flight at a time; the sum stays in j order
    double s = 0.0;
    int j = 0;
    for (; j + 16 <= nc; j += 16) {
      double v[16];
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = W[(int64_t)(j + k) * nc + i];
#pragma unroll
      for (int k = 0; k < 16; k++) s += v[k] * bs[j + k];
    }
    for (; j < nc; j++) s += W[(int64_t)j * nc + i] * bs[j];
    x[i] = (T)s;
  }
}

// rz = b . z over the grid (used when the V-cycle ends without a fine-level
// smoothing sweep to fuse it into).
template <typename T>
__global__ void k_rz_dot(int64_t n, const T* __restrict__ b, const T* __restrict__ z,
                         const int* __restrict__ done, const int2* __restrict__ bm,
                         double* __restrict__ partial) {
  const int dn = *done;
  int64_t i = logical_row();
  bool second;
  member_of(bm, i, second);
  double local = i < n ? (double)b[i] * (double)z[i] : 0.0;
  if (dn) return;
  block_partial2(local, second, partial, bm != nullptr);
}

// ------------------------------------------------------------------ CG
// Every vector kernel reads the scalars of its row's member; a member whose
// solve has finished is frozen (its x and r no longer change) while the
// others iterate.

// x = 0, r = b, rr = b.b (partials); first smoothing step of the V-cycle on r
// (d0 = z0 = c0 invd r).
// r = b, x = 0, partials of b.b, and the V-cycle's zero-guess first step
// d0 = z0 = c0 D^-1 r.  The Krylov vectors are in TK and the V-cycle in TP;
// when they differ, rp receives r rounded to TP, the V-cycle's input (rp is
// null when TK == TP: the V-cycle reads r itself).  With b64 (b null), the
// right-hand side is the refinement residual b64 / outer_s of the members
// above their refinement target, rounded to TK, and 0 for the others (whose
// solve then ends at once) -- the FP32 inner solve's right-hand side.
template <typename TK, typename TP>
__global__ void k_cg_start(int64_t n, const TK* __restrict__ b, const double* __restrict__ b64,
                           const SolveScalars* __restrict__ sc, TK* __restrict__ r,
                           TK* __restrict__ x, TP* __restrict__ rp, const TP* __restrict__ invd,
                           TP* __restrict__ d0, TP* __restrict__ z0, TP c0,
                           const int2* __restrict__ bm, double* __restrict__ partial) {
  int64_t i = logical_row();
  bool second;
  const int mi = member_of(bm, i, second);
  double local = 0.0;
  if (i < n) {
    TK bv;
    if (b64) {
      const SolveScalars& m = sc[mi];
      const double sv = m.outer_s;
      bv = (sv > m.target && sv > 0.0) ? (TK)(b64[i] / sv) : (TK)0;
    } else {
      bv = b[i];
    }
    r[i] = bv;
    x[i] = (TK)0;
    TP rv = (TP)bv;
    if (rp) rp[i] = rv;
    TP dd = c0 * invd[i] * rv;
    if (d0) d0[i] = dd;  // null when no sweep reads it (finest degree 1)
    z0[i] = dd;
    local = (double)bv * (double)bv;
  }
  block_partial2(local, second, partial, bm != nullptr);
}

// p_new = z + beta p_old,  q = A p_new (formed from z and p_old at the
// neighbours, so p is never read while it is rewritten), partials of p.q.
// Neighbours belong to the same member (the operator is block diagonal).
// (z is the V-cycle's output, in its precision TZ; p, q and A are in T)
template <typename T, typename TZ>
__global__ void k_cg_pq(SellD<T> A, const TZ* __restrict__ z, const T* __restrict__ pold,
                        T* __restrict__ pnew, T* __restrict__ q, const int* __restrict__ done,
                        const SolveScalars* __restrict__ sc, const int2* __restrict__ bm,
                        double* __restrict__ partial) {
  const int dn = *done;
  int64_t i = logical_row();
  bool second;
  int mi = member_of(bm, i, second);
  const bool in = i < A.nrows;
  const SellRow R = sell_row_at(A, in ? i : A.nrows - 1);
  const SolveScalars& m = sc[mi];
  const int first = m.first;
  const double rz = m.rz, rz_old = m.rz_old;
  T zi = in ? (T)z[i] : (T)0;  // loaded with A, not after it
  T poi = in && !first ? pold[i] : (T)0;
  if (dn || R.w < 0) return;
  double local = 0.0;
  if (in) {
    T beta = first ? (T)0 : (T)(rz / rz_old);
    T s, pn;
    if (first) {
      s = sell_dot(A, R, i, z);
      pn = zi;
    } else {
      s = sell_dot_zp(A, R, i, z, pold, beta);
      pn = zi + beta * poi;
    }
    pnew[i] = pn;
    q[i] = s;
    local = (double)pn * (double)s;
  }
  block_partial2(local, second, partial, bm != nullptr);
}

// x += alpha p, r -= alpha q (members still running), partials of r.r; first
// smoothing step of the next V-cycle on r.
template <typename TK, typename TP>
__global__ void k_cg_xr(int64_t n, TK* __restrict__ x, TK* __restrict__ r,
                        const TK* __restrict__ p, const TK* __restrict__ q,
                        TP* __restrict__ rp, const TP* __restrict__ invd, TP* __restrict__ d0,
                        TP* __restrict__ z0, TP c0, const int* __restrict__ done,
                        const SolveScalars* __restrict__ sc, const int2* __restrict__ bm,
                        double* __restrict__ partial) {
  const int dn = *done;
  int64_t i = logical_row();
  bool second;
  int mi = member_of(bm, i, second);
  const bool in = i < n;
  const int mdone = sc[mi].done;
  const double alpha64 = sc[mi].alpha;
  TK rv = in ? r[i] : (TK)0, xv = in ? x[i] : (TK)0, pv = in ? p[i] : (TK)0;
  TK qv = in ? q[i] : (TK)0;
  TP iv = in ? invd[i] : (TP)0;
  if (dn) return;
  double local = 0.0;
  if (in) {
    if (!mdone) {
      TK alpha = (TK)alpha64;
      x[i] = xv + alpha * pv;
      rv = rv - alpha * qv;
      r[i] = rv;
    }
    TP rpv = (TP)rv;
    if (rp) rp[i] = rpv;
    TP dd = c0 * iv * rpv;
    if (d0) d0[i] = dd;  // null when no sweep reads it (finest degree 1)
    z0[i] = dd;
    local = (double)rv * (double)rv;
  }
  block_partial2(local, second, partial, bm != nullptr);
}

// ------------------------------------------------------------------ FP64 outer residual / refinement

// Partials of max|b| per member (FIN_AMAX).
__global__ void k_amax(int64_t n, const double* __restrict__ b, const int2* __restrict__ bm,
                       double* __restrict__ partial) {
  int64_t i = logical_row();
  bool second;
  member_of(bm, i, second);
  // NaN propagates as +inf, which the finish maps to scale 1
  double v = i < n ? (b[i] == b[i] ? fabs(b[i]) : INFINITY) : 0.0;
  block_partial2_max(v, second, partial);
}

// y = x bscale (inverse = 0) or x / bscale (inverse = 1), per member
__global__ void k_scale_copy(int64_t n, const double* __restrict__ x, double* __restrict__ y,
                             const SolveScalars* __restrict__ sc, const int2* __restrict__ bm,
                             int inverse) {
  int64_t i = logical_row();
  if (i >= n) return;
  bool second;
  double f = sc[member_of(bm, i, second)].bscale;
  y[i] = inverse ? x[i] / f : x[i] * f;
}

// Partials of x0 . (b bscale) and x0 . r0 per member (warm start projection).
__global__ void k_warm_dots(int64_t n, const double* __restrict__ x, const double* __restrict__ b,
                            const double* __restrict__ r, const SolveScalars* __restrict__ sc,
                            const int2* __restrict__ bm, double* __restrict__ pxb,
                            double* __restrict__ pxr) {
  int64_t i = logical_row();
  bool second;
  int mi = member_of(bm, i, second);
  double xb = 0.0, xr = 0.0;
  if (i < n) {
    double xi = x[i];
    xb = xi * (b[i] * sc[mi].bscale);
    xr = xi * r[i];
  }
  block_partial2(xb, second, pxb, bm != nullptr);
  block_partial2(xr, second, pxr, bm != nullptr);
}

// x = c x0, r = b bscale - c A x0 = (1 - c) b bscale + c r0, partials of ||r||^2
__global__ void k_warm_apply(int64_t n, double* __restrict__ x, const double* __restrict__ b,
                             double* __restrict__ r, const SolveScalars* __restrict__ sc,
                             const int2* __restrict__ bm, double* __restrict__ partial) {
  int64_t i = logical_row();
  bool second;
  int mi = member_of(bm, i, second);
  double local = 0.0;
  if (i < n) {
    const SolveScalars& m = sc[mi];
    double c = m.cwarm;
    x[i] *= c;
    double rv = (1.0 - c) * (b[i] * m.bscale) + c * r[i];
    r[i] = rv;
    local = rv * rv;
  }
  block_partial2(local, second, partial, bm != nullptr);
}

// r = b bscale - A x in FP64 (bscale = 1 when sc is null; x = null stands
// for x = 0), partials of ||r||^2 per member.
__global__ void k_outer_residual(SellD<double> A, const double* __restrict__ x,
                                 const double* __restrict__ b, double* __restrict__ r,
                                 const SolveScalars* __restrict__ sc,
                                 const int2* __restrict__ bm, double* __restrict__ partial) {
  int64_t i = logical_row();
  bool second;
  int mi = member_of(bm, i, second);
  double local = 0.0;
  if (i < A.nrows) {
    double bi = sc ? b[i] * sc[mi].bscale : b[i];  // loaded with A, not after it
    double ax = 0.0;
    if (x) {
      int64_t base;
      int w;
      sell_row(A, i, base, w);
      ax = sell_dot(A, i, base, w, x);
    }
    double rv = bi - ax;
    r[i] = rv;
    local = rv * rv;
  }
  block_partial2(local, second, partial, bm != nullptr);
}

// x64 += s * x32 for members above their refinement target
__global__ void k_refine(int64_t n, double* __restrict__ x, const float* __restrict__ d,
                         const SolveScalars* __restrict__ sc, const int2* __restrict__ bm) {
  int64_t i = logical_row();
  if (i >= n) return;
  bool second;
  const SolveScalars& m = sc[member_of(bm, i, second)];
  if (m.outer_s > m.target) x[i] += m.outer_s * (double)d[i];
}

// y += a x
__global__ void k_axpy_f64(int64_t n, double a, const double* __restrict__ x,
                           double* __restrict__ y) {
  int64_t i = logical_row();
  if (i < n) y[i] += a * x[i];
}

__global__ void k_zero_f64(int64_t n, double* __restrict__ x) {
  int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// ------------------------------------------------------------------ scatter / gather / adjoint

// h_i = hfree_{j(i)} on free cells, dirichleths_i on Dirichlet cells
__global__ void k_scatter_h(int64_t n, const int32_t* __restrict__ n2f,
                            const double* __restrict__ hfree, const double* __restrict__ dh,
                            double* __restrict__ h) {
  int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int32_t j = n2f[i];
  h[i] = j >= 0 ? hfree[j] : dh[i];
}

// out_j = v_{f2n[j]}
__global__ void k_gather_free(int64_t nf, const int32_t* __restrict__ f2n,
                              const double* __restrict__ v, double* __restrict__ out) {
  int64_t j = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (j >= nf) return;
  out[j] = v[f2n[j]];
}

// dK_e = -c aol_e (h~_a - h~_b)(lambda~_a - lambda~_b), lambda~ = 0 on Dirichlet
// cells; added to dK when accumulate is set (the transient sum over steps).
__global__ void k_edge_grad(int64_t ne, const int2* __restrict__ edges,
                            const int32_t* __restrict__ n2f, const double* __restrict__ aol,
                            const double* __restrict__ hfree, const double* __restrict__ dh,
                            const double* __restrict__ lam, double c, int accumulate,
                            double* __restrict__ dK) {
  int64_t e = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (e >= ne) return;
  int2 ab = edges[e];
  int32_t fa = n2f[ab.x], fb = n2f[ab.y];
  double ha = fa >= 0 ? hfree[fa] : dh[ab.x];
  double hb = fb >= 0 ? hfree[fb] : dh[ab.y];
  double la = fa >= 0 ? lam[fa] : 0.0;
  double lb = fb >= 0 ? lam[fb] : 0.0;
  double g = -c * aol[e] * (ha - hb) * (la - lb);
  dK[e] = accumulate ? dK[e] + g : g;
}

// ------------------------------------------------------------------ backward Euler

// s_j = Ss_i V_i for free cell i = f2n[j]
__global__ void k_storage(int64_t nf, const int32_t* __restrict__ f2n,
                          const double* __restrict__ Ss, const double* __restrict__ V,
                          double* __restrict__ svec) {
  int64_t j = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (j >= nf) return;
  int32_t i = f2n[j];
  svec[j] = Ss[i] * V[i];
}

// rhs_j = s_j u_j + dt (q Q_i + bD_j): the right-hand side of (S + dt A) u_n = S u_{n-1} + dt b(t_n)
__global__ void k_be_rhs(int64_t nf, const int32_t* __restrict__ f2n,
                         const double* __restrict__ svec, const double* __restrict__ u,
                         const double* __restrict__ Q, double q, double dt,
                         const double* __restrict__ bD, double* __restrict__ rhs) {
  int64_t j = logical_row();
  if (j >= nf) return;
  rhs[j] = svec[j] * u[j] + dt * (q * Q[f2n[j]] + bD[j]);
}

// adjoint right-hand side: rhs_j = s_j mu_j (if mu) + g_j (if g)
__global__ void k_be_adj_rhs(int64_t nf, const double* __restrict__ svec,
                             const double* __restrict__ mu, const double* __restrict__ g,
                             double* __restrict__ rhs) {
  int64_t j = logical_row();
  if (j >= nf) return;
  double v = mu ? svec[j] * mu[j] : 0.0;
  rhs[j] = g ? v + g[j] : v;
}

// The reverse sweep's multiplier sums over steps (free cells):
//   macc (+)= cm mu,  qacc (+)= cq mu   (either may be null)
__global__ void k_adj_acc(int64_t nf, const double* __restrict__ mu, double cm, double cq,
                          int accumulate, double* __restrict__ macc, double* __restrict__ qacc) {
  int64_t j = logical_row();
  if (j >= nf) return;
  double m = mu[j];
  if (macc) macc[j] = accumulate ? macc[j] + cm * m : cm * m;
  if (qacc) qacc[j] = accumulate ? qacc[j] + cq * m : cq * m;
}

// dQ_i = lambda_{j(i)} (free), 0 (Dirichlet);
// ddh_i = 0 (free), delta_i + sum_{e: other free} K_e aol_e lambda_other (Dirichlet;
// delta = null stands for 0)
__global__ void k_ddh_dq(int64_t n, const int32_t* __restrict__ n2f,
                         const int64_t* __restrict__ inc_ptr, const int32_t* __restrict__ inc_edge,
                         const int32_t* __restrict__ inc_other, const double* __restrict__ Ke,
                         const double* __restrict__ aol, const double* __restrict__ lam,
                         const double* __restrict__ delta, double* __restrict__ ddh,
                         double* __restrict__ dQ) {
  int64_t i = logical_row();
  if (i >= n) return;
  int32_t j = n2f[i];
  if (j >= 0) {
    if (dQ) dQ[i] = lam[j];
    if (ddh) ddh[i] = 0.0;
    return;
  }
  if (dQ) dQ[i] = 0.0;
  if (ddh) {
    double s = delta ? delta[i] : 0.0;
    for (int64_t p = inc_ptr[i]; p < inc_ptr[i + 1]; p++) {
      int32_t fo = n2f[inc_other[p]];
      if (fo >= 0) {
        int32_t e = inc_edge[p];
        s += Ke[e] * aol[e] * lam[fo];
      }
    }
    ddh[i] = s;
  }
}

}  // namespace dpf
