// libdpfgpu: C-ABI driver of the groundwater solve and adjoint on one MI355X.
//
// Per topology (dpf_topology_create): host symbolic setup (dpf_host.h), then
// every index array is uploaded once.  Per solve: assembly of A and b from
// the edge conductivities, numeric refresh of the smoothed-aggregation
// hierarchy (prolongator smoothing, R = P', Galerkin products, dense inverse
// of the coarsest operator), then CG preconditioned by one V-cycle with
// l1-Jacobi Chebyshev smoothing.  In mixed precision the hierarchy and the
// inner CG run in FP32 inside FP64 iterative refinement; the refinement
// residual, the update, every CG scalar and the coarsest solve stay FP64.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/dpfgpu.h"
#include "dpf_host.h"
#include "kernels.hpp"
#include "twophase.hpp"
#include "richards.hpp"
#include "transport.hpp"

namespace dpf {

static thread_local std::string g_err;

struct Err {
  int code;
  std::string msg;
};

#define HIPTRY(x)                                                                       \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      throw Err{DPF_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)};              \
  } while (0)

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + BLK - 1) / BLK); }

static double now_wall_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// ------------------------------------------------------------------ device buffers

// Device buffers released by forward states and per-call temporaries, kept
// for reuse: hipFree synchronises the whole device and hipMalloc of large
// buffers costs hundreds of microseconds, both per evaluation otherwise.
// Shared (reference-counted) by a topology and its states, so that a state
// may outlive its topology's handle.  Release is stream-ordered: a buffer
// given back by a stream-ordered owner carries an event recorded on the
// stream of its last use, and whoever takes it next waits for that event on
// its own stream (a GPU-side wait), or on the host when it names no stream.
// An owner that names no stream gives the buffer back with no work pending
// (its call synchronised its stream).
struct BufPool {
  struct Entry {
    void* p;
    hipEvent_t ev;  // work that may still use p (null: none)
  };
  std::mutex mu;
  std::multimap<size_t, Entry> free;
  std::vector<hipEvent_t> spare;  // events no entry holds
  size_t cached = 0, limit;
  explicit BufPool(size_t lim) : limit(lim) {}
  BufPool(const BufPool&) = delete;
  BufPool& operator=(const BufPool&) = delete;
  ~BufPool() {
    for (auto& kv : free) {
      if (kv.second.ev) {
        (void)hipEventSynchronize(kv.second.ev);
        (void)hipEventDestroy(kv.second.ev);
      }
      (void)hipFree(kv.second.p);
    }
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
  }
  // a cached buffer of at least b bytes (and at most 2b), or null; with
  // ordered, stream s waits for the work that may still use it
  void* take(size_t b, size_t& got, bool ordered, hipStream_t s) {
    Entry e;
    {
      std::lock_guard<std::mutex> g(mu);
      auto it = free.lower_bound(b);
      if (it == free.end() || it->first > 2 * b) return nullptr;
      e = it->second;
      got = it->first;
      cached -= got;
      free.erase(it);
    }
    if (e.ev) {
      hipError_t r = ordered ? hipStreamWaitEvent(s, e.ev, 0) : hipEventSynchronize(e.ev);
      if (r != hipSuccess) {
        (void)hipEventSynchronize(e.ev);
        (void)hipEventDestroy(e.ev);
        give(e.p, got, false, nullptr);
        throw Err{DPF_EHIP, std::string("pooled buffer wait: ") + hipGetErrorString(r)};
      }
      std::lock_guard<std::mutex> g(mu);
      spare.push_back(e.ev);
    }
    return e.p;
  }
  // with ordered, the buffer's last use was queued on stream s
  void give(void* p, size_t b, bool ordered, hipStream_t s) {
    hipEvent_t ev = nullptr;
    if (ordered) {
      {
        std::lock_guard<std::mutex> g(mu);
        if (!spare.empty()) {
          ev = spare.back();
          spare.pop_back();
        }
      }
      if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) ev = nullptr;
      if (!ev || hipEventRecord(ev, s) != hipSuccess) {
        // no event to order by: wait for the stream here instead
        if (ev) (void)hipEventDestroy(ev);
        ev = nullptr;
        (void)hipStreamSynchronize(s);
      }
    }
    {
      std::lock_guard<std::mutex> g(mu);
      if (cached + b <= limit) {
        free.emplace(b, Entry{p, ev});
        cached += b;
        return;
      }
    }
    if (ev) {
      (void)hipEventSynchronize(ev);
      (void)hipEventDestroy(ev);
    }
    (void)hipFree(p);
  }
};

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  BufPool* pool = nullptr;  // where the memory goes back to (null: hipFree)
  // a pooled buffer used on one stream: taken after the work that still used
  // it (waited for on that stream), given back ordered after its own
  bool ordered = false;
  hipStream_t stream = nullptr;
  DBuf() = default;
  explicit DBuf(BufPool* bp) : pool(bp) {}
  DBuf(BufPool* bp, hipStream_t s) : pool(bp), ordered(true), stream(s) {}
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (!p) return;
    if (pool)
      pool->give(p, bytes, ordered, stream);
    else
      (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (p && bytes >= b) return;
    release();
    if (b == 0) b = 8;
    if (pool) {
      size_t got = 0;
      if ((p = pool->take(b, got, ordered, stream))) {
        bytes = got;
        return;
      }
    }
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Err{DPF_ENOMEM, "hipMalloc(" + std::to_string(b) + " bytes) failed: " +
                                hipGetErrorString(e)};
    }
    bytes = b;
  }
  // alloc(b) whose take from the pool waits on stream s (GPU side) for work
  // that may still use the pooled buffer; the buffer's own release mode is
  // kept
  void alloc_on(size_t b, hipStream_t s) {
    const bool o0 = ordered;
    const hipStream_t s0 = stream;
    ordered = true;
    stream = s;
    try {
      alloc(b);
    } catch (...) {
      ordered = o0;
      stream = s0;
      throw;
    }
    ordered = o0;
    stream = s0;
  }
  template <typename T>
  T* as() const {
    return (T*)p;
  }
};

template <typename T>
static T* upload(DBuf& d, const std::vector<T>& v) {
  d.alloc(v.size() * sizeof(T));
  if (!v.empty()) HIPTRY(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d.as<T>();
}

struct SellDev {
  int64_t nrows = 0, ncols = 0, entries = 0;
  DBuf sptr, col, rowlen, soff, col16, cbase;
  bool offsets = false, compact = false;
  void up(const SellHost& h) {
    nrows = h.nrows;
    ncols = h.ncols;
    entries = h.entries();
    compact = false;
    upload(sptr, h.sptr);
    upload(col, h.col);
    upload(rowlen, h.rowlen);
  }
  // Compact columns (16-bit offsets from each slice's smallest column, the
  // padding slots pointing at that column) when every slice's columns span
  // at most 65535; 32-bit columns otherwise.  For operators whose columns
  // only the row sums read (split level operators, prolongators,
  // restrictions).
  void up_compact(const SellHost& h) {
    const int64_t ns = h.nslices();
    std::vector<int32_t> base(ns, 0);
    std::vector<uint16_t> c16(h.entries(), 0);
    int bad = 0;
    for (int64_t sl = 0; sl < ns; sl++) {
      int64_t r0 = sl * SELL_C, r1 = std::min<int64_t>(h.nrows, r0 + SELL_C);
      int32_t lo = INT32_MAX, hi = INT32_MIN;
      for (int64_t r = r0; r < r1; r++)
        for (int k = 0; k < h.rowlen[r]; k++) {
          int32_t c = h.at(r, k);
          lo = std::min(lo, c);
          hi = std::max(hi, c);
        }
      if (lo > hi) lo = hi = 0;  // a slice without entries
      if ((int64_t)hi - lo > 65535) bad = 1;
      base[sl] = lo;
      for (int64_t r = r0; r < r1; r++)
        for (int k = 0; k < h.rowlen[r]; k++) c16[h.idx(r, k)] = (uint16_t)(h.at(r, k) - lo);
    }
    if (bad) {
      up(h);
      return;
    }
    nrows = h.nrows;
    ncols = h.ncols;
    entries = h.entries();
    upload(sptr, h.sptr);
    upload(rowlen, h.rowlen);
    upload(col16, c16);
    upload(cbase, base);
    compact = true;
  }
  // the fine operator in its offset-slice layout (rowlen is not uploaded:
  // a row's entries are not the first slots of an offset slice)
  void up_fine(const SellHost& h, const FineLayout& F) {
    nrows = h.nrows;
    ncols = h.ncols;
    entries = F.sptr.back();
    upload(sptr, F.sptr);
    upload(col, F.col);
    upload(soff, F.soff);
    offsets = true;
  }
  template <typename T>
  SellD<T> view(T* val) const {
    return SellD<T>{nrows, sptr.as<int64_t>(), col.as<int32_t>(), val,
                    offsets ? soff.as<int32_t>() : nullptr,
                    compact ? col16.as<uint16_t>() : nullptr,
                    compact ? cbase.as<int32_t>() : nullptr};
  }
};

struct LevelDev {
  SellDev A;
  DBuf diag_slot;
  bool transfer = false;
  SellDev P, Rs, As;
  int rs_lg = 0, as_lg = 0;
  DBuf as_src;
  int64_t nc = 0;  // rows of the next level
  DBuf p_slot, p_self, rs_src, ap_off, ap_a, ap_p, rap_off, rap_dest, rap_p, rap_ap;
  DBuf rap_mdst, rap_msrc;
  int64_t n_rap = 0, n_mirror = 0;  // Galerkin lists (upper entries), mirrored entries
  int ap_lg = 0, rap_lg = 0;
  int64_t ap_nnz = 0;
  // row-wise Galerkin products (GalerkinPlan): A*P and/or R*(A*P) by rows
  bool ap_rows = false, rap_rows = false;
  int apr_lg = 0, ap_w = 0, rap_w = 0;  // lanes per fine row of the row-wise A*P
  int64_t ap_prods = 0, rap_prods = 0;
  DBuf ap_rp, apc_ptr, apc, rcd_ptr, rcd, rnp;
};

// Numeric values and work vectors of a hierarchy in precision T.
template <typename T>
struct LevelNum {
  DBuf val, invd, pval, rval, apval, asval;
  DBuf x, xs, b, r, d;
  // FP32 hierarchy: bf16 copies of the transfer operators and of the coarse
  // split operator, and that operator's compensated FP32 diagonal
  DBuf p16, r16, a16, as16, dg;
};

template <typename T>
struct Hier {
  bool allocated = false;
  std::vector<LevelNum<T>> lev;
  DBuf W;  // dense FP64 inverse of the coarsest operator
  DBuf status;
  // CG vectors on the fine level (x, r, p ping-pong, q); z is the V-cycle's
  // fine-level output.
  DBuf cx, cr, cp0, cp1, cq;
  uint64_t version = 0;  // numeric refresh this hierarchy holds
};

// Chebyshev coefficients for the interval [lmax/ratio, lmax] of the
// l1-Jacobi-preconditioned operator (lmax = 1 bounds its spectrum).
struct Cheb {
  double c0;               // zero-guess first step: d = c0 invd r
  std::vector<double> c1;  // step k >= 1: d = c1[k] d + c2[k] invd r
  std::vector<double> c2;
  double ratio = 15.0;     // interval [lmax / ratio, lmax]
  void make(int degree, double lmax, double ratio_) {
    ratio = ratio_;
    double lmin = lmax / ratio;
    double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    double sigma = theta / delta;
    c0 = 1.0 / theta;
    c1.assign(degree, 0.0);
    c2.assign(degree, 0.0);
    c2[0] = c0;
    double rho = 1.0 / sigma;
    for (int k = 1; k < degree; k++) {
      double rn = 1.0 / (2.0 * sigma - rho);
      c1[k] = rn * rho;
      c2[k] = 2.0 * rn / delta;
      rho = rn;
    }
  }
};

// Kernel classes of the timing table (dpf_kernel_table).
enum KCat {
  KC_CHEB, KC_RESIDUAL, KC_RESTRICT, KC_PROLONG, KC_PROLONG_COARSE, KC_COARSE_SOLVE,
  KC_RZ_DOT, KC_CG_START, KC_CG_PQ, KC_CG_XR, KC_FINISH, KC_OUTER_RESIDUAL, KC_REFINE_VEC,
  KC_ASSEMBLE, KC_PSMOOTH, KC_RGATHER, KC_PROD_AP, KC_PROD_RAP, KC_MIRROR, KC_ROWDIAG,
  KC_DENSE_INVERSE, KC_KMAP, KC_GRADIENT, KC_SCATTER, KC_AP_ROWS, KC_RAP_ROWS, KC_NCAT
};
static const char* const kcat_name[KC_NCAT] = {
  "cheb_sweep", "residual", "restrict", "prolong", "prolong_coarse_solve", "coarse_solve",
  "rz_dot", "cg_start", "cg_pq", "cg_xr", "finish", "outer_residual", "refine_vector",
  "assemble", "psmooth", "rgather", "galerkin_ap", "galerkin_rap", "mirror_copy", "rowdiag",
  "dense_inverse", "k_map", "gradient", "scatter_gather", "galerkin_ap_rows",
  "galerkin_rap_rows"};

struct KRec {
  int cat, level, w, iter, phase;
  double bytes;
  hipEvent_t a, b;
};
struct KAcc {
  int64_t launches = 0;
  double ms = 0, bytes = 0;
};

}  // namespace dpf

using namespace dpf;

struct dpf_topology {
  TopologyHost H;
  std::mutex mu;
  hipStream_t own_stream = nullptr;
  int device = 0;
  std::shared_ptr<BufPool> pool;       // reused state and temporary buffers
  hipEvent_t ev_setup[2] = {nullptr, nullptr};  // GPU time of a call's numeric setup
  // ordering of DPF_DEVICE calls with a NULL stream against the NULL stream
  hipEvent_t ev_null[2] = {nullptr, nullptr};
  // the end of the last call's work (on whatever stream it ran): every call
  // starts after it, so the calls on a topology run in order on the GPU
  hipEvent_t ev_call = nullptr;
  bool call_seen = false;
  // topology arrays on the device
  DBuf edges, aol, n2f, f2n, inc_ptr, inc_edge, inc_other, inc_slot;
  std::vector<std::unique_ptr<LevelDev>> lev;
  int64_t max_grid = 0;
  // fine FP64 operator (values of levels[0] pattern), its l1 diagonal, b
  DBuf A64, invd64, b64;
  // inputs staged from the host / kept for the pullback
  DBuf Ke, dh, Q, stage_n, stage_n2, stage_e, stage_f, stage_f2, stage_f3;
  // refinement vectors
  DBuf x64, r64;
  Hier<double> h64;
  Hier<float> h32;
  // reductions and solve scalars
  DBuf partial, partial2, scal, gdone, zero_flag;
  // members of a batch (one for a single system): fine-row ranges and, per
  // logical block of the fine grid, the block's first member and the row
  // where the next member starts
  int nm = 1;
  std::vector<int64_t> m_cell0, m_row0, m_edge0;  // nm + 1 offsets each
  DBuf bm, mblk;
  double* hm_outer_host = nullptr;  // per-member refinement residual norms (pinned)
  std::vector<double> m_relres;     // per-member relative residual of the last solve
  double* hm_outer_dev = nullptr;
  SolveScalars* scal_host = nullptr;
  HostMirror* hm_host = nullptr;  // pinned coherent host memory the finish kernel writes
  HostMirror* hm_dev = nullptr;   // the same memory as seen by the device
  double* dots_host = nullptr;    // pinned coherent: the three sums of a BiCGStab reduction
  double* dots_dev = nullptr;
  // DPF_FLAG_GRAPH: captured PCG iterations, keyed by precision, smoother,
  // buffers and the p ping-pong parity
  bool use_graphs = false;
  std::map<std::string, hipGraphExec_t> iter_graphs;
  int seq = 0;                    // solve / residual sequence numbers
  uint64_t version = 0;        // bumped at every numeric refresh
  uint64_t assembled_for = 0;  // version whose K the fine operator holds
  // fine operator layout: offset slices, and nonzeros whose column is stored
  int64_t fine_offset_slices = 0, fine_explicit_nnz = 0;
  // two-phase reverse mode: per incidence entry the same edge's entry in the
  // other cell's list, per edge its two entries (built on first use)
  DBuf tp_mirror, tp_epos;
  // Jacobian export (built on first use): per kind (h, Ks, dirichleths, Qs)
  // the CSR pattern on the host and its row pointers on the device, per
  // incidence the value index (h, Ks, dirichleths), the diagonals' indices
  bool jac_maps = false, jac_dev = false;
  std::vector<int64_t> jac_rp[4];
  std::vector<int32_t> jac_cols[4];
  std::vector<int32_t> jac_pos[3], jac_dpos;
  DBuf jac_rp_d[4], jac_pos_d[3], jac_dpos_d;
  bool tp_maps = false;
  // Chebyshev interval [1/ratio, 1] of D_l1^-1 A: default, and tuned per precision
  double cheb_ratio = 15.0;
  double ratio_tuned[3] = {15.0, 15.0, 15.0};
  // Chebyshev degrees of the finest level, the second level and every deeper
  // level per precision (index = precision: FP64, mixed refinement, FP64
  // Krylov with FP32 V-cycle): the fixed default 1/1/2 unless autotuned
  // (DPF_FLAG_AUTOTUNE) or pinned, and the tuner's record.  The second level
  // holds about half the finest level's nonzeros and its sweeps are
  // bandwidth-bound, so it gets one sweep a V-cycle; the deeper levels are
  // small enough that a sweep costs about a launch, and their stronger
  // smoothing is what keeps the V-cycle's convergence.
  int deg_fine[3] = {1, 1, 1}, deg_coarse[3] = {1, 1, 1}, deg_deep[3] = {2, 2, 2};
  // Damping of the prolongator smoothing P = (I - omega D_l1^-1 A) T on the
  // finest level and on the Galerkin levels: omega = 4 / (3 rho) with a fixed
  // estimate rho of rho(D_l1^-1 A) instead of its bound 1 (rho = 0.89, omega
  // 1.5; rho = 0.67, omega 2, where the l1 diagonal overestimates the
  // spectrum more).  No eigenvalue estimate runs at set-up.  On spatially
  // correlated conductivity fields these take fewer PCG iterations than 4/3
  // (2M cells: 18-19 against 20-22); on fields rough from cell to cell, 4/3
  // can take fewer, and dpf_topology_set_damping sets other values.
  double omega_fine = 1.5, omega_coarse = 2.0;
  bool tuned[3] = {false, false, false};
  // lanes per row of the coarse levels' split copies: the autotuner's sweep
  // times per level and lane exponent 0..6 (0 = not timed), and whether the
  // layout was tuned or pinned (either fixes it for the topology's life)
  std::vector<std::array<double, 7>> lane_us;
  bool lanes_fixed = false;
  double tune_ms[3][7] = {};  // autotuner trials: 5 degree pairs, then 2 intervals
  int tune_iters[3][7] = {};
  // optional event timing (opts->flags): every finest-level smoothing sweep
  // (sweep timing) or every launch (kernel timing), per kernel class, level
  // and value size
  bool timing = false, ktiming = false;
  // the FP32 V-cycle reads the bf16 copies (off with DPF_FLAG_FP32_STORAGE)
  bool vbf16 = true;
  std::vector<hipEvent_t> evpool;
  size_t evused = 0;
  double sweep_ms = 0;
  int sweep_n = 0;
  int kiter = -1, kphase = 0;  // Krylov iteration and phase of the launches being queued
  std::vector<KRec> kpend;
  std::map<std::tuple<int, int, int>, KAcc> ktable;
  ~dpf_topology() {
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
  }
  int levels() const { return (int)H.levels.size(); }
};

struct dpf_state {
  dpf_topology* topo;
  std::shared_ptr<BufPool> pool;  // declared before the buffers: outlives them
  uint64_t version;  // topo->version of the forward solve
  DBuf Ke, dh, hfree, lam, delta, Q;
  int precision;
  // transient runs
  int kind = 0;  // 0 steady, 1 backward Euler
  int nsteps = 0, ck = 0;
  double dt = 0;            // two-phase and transport runs: the (uniform) step length
  std::vector<double> dts;  // backward Euler: step lengths (nsteps)
  double cur_dt = 0;        // the step length the hierarchy was last prepared for
  std::vector<double> qs;   // Q scale per time level (nsteps + 1)
  DBuf svec;               // S = Ss V on free cells
  DBuf ckpt;               // checkpointed states u_0, u_ck, u_2ck, ...
  // two-phase runs (kind 2): per IMPES step the saturation it started from,
  // the edge factor Ksum, pressure and fluxes (step-major), and its sub-step
  // count; the run's K, Q, pore volumes, fluid, dt and pressure constant
  DBuf tp_S, tp_Ksum, tp_P, tp_Vn, tp_K, tp_pv;
  // steady Richards (kind 3): elevation and per-edge van Genuchten parameters
  // (hfree holds the free heads, Ke the edge K, dh the Dirichlet heads)
  DBuf r_z, r_al, r_N;
  // transport (kind 4): every state u_0 .. u_N and the run's inputs
  DBuf tr_u, tr_v, tr_D, tr_ud, tr_Q, tr_V, tr_c;
  int tr_upwind = 1;
  std::vector<int64_t> tp_nts;
  double tp_fluid[4] = {0, 0, 0, 0};
  int tp_zero_mean = 0;
  explicit dpf_state(dpf_topology* t);
  std::vector<DBuf*> buffers();
  // the buffers go back to the pool ordered after the work queued on s (not
  // ordered: with no work pending)
  void release_on(hipStream_t s, bool ordered = true);
};

dpf_state::dpf_state(dpf_topology* t) : topo(t), pool(t->pool), version(0), precision(0) {
  for (DBuf* b : buffers()) b->pool = pool.get();
}

std::vector<DBuf*> dpf_state::buffers() {
  return {&Ke, &dh, &hfree, &lam, &delta, &Q, &svec, &ckpt, &tp_S, &tp_Ksum, &tp_P,
          &tp_Vn, &tp_K, &tp_pv, &r_z, &r_al, &r_N, &tr_u, &tr_v, &tr_D, &tr_ud,
          &tr_Q, &tr_V, &tr_c};
}

void dpf_state::release_on(hipStream_t s, bool ordered) {
  for (DBuf* b : buffers()) {
    b->ordered = ordered;
    b->stream = s;
  }
}

namespace {

void set_err(const std::string& m) { g_err = m; }

// The stream a call queues its work on.  DPF_DEVICE with a stream: that
// stream.  Otherwise the topology's own (non-blocking) stream; in DPF_DEVICE
// mode with a NULL stream the call is then ordered as if it ran on the legacy
// NULL stream: its work starts after everything the NULL stream holds when
// the call begins (the caller's device inputs may still be being written
// there), and work the caller queues on the NULL stream after the call
// returns starts after the call's work -- two event waits, no host
// synchronisation.  Whatever the stream, the call's work starts after the
// previous call's on the same topology (an event wait on the GPU), since
// calls share the topology's buffers and a call in DPF_DEVICE mode may
// return before its work is done -- except on a stream under graph capture,
// whose caller orders the captured work itself.  Construct it with the
// topology's lock held.
struct CallStream {
  dpf_topology* t;
  hipStream_t s;
  bool order_back = false;
  bool chain = true;
  CallStream(dpf_topology* t_, const dpf_opts* o) : t(t_) {
    if (o && o->memory == DPF_DEVICE && o->stream) {
      s = (hipStream_t)o->stream;
    } else {
      s = t->own_stream;
      if (o && o->memory == DPF_DEVICE) {
        HIPTRY(hipEventRecord(t->ev_null[0], nullptr));
        HIPTRY(hipStreamWaitEvent(s, t->ev_null[0], 0));
        order_back = true;
      }
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cap) == hipSuccess && cap != hipStreamCaptureStatusNone)
      chain = false;
    if (chain && t->call_seen) HIPTRY(hipStreamWaitEvent(s, t->ev_call, 0));
  }
  ~CallStream() {
    if (chain && hipEventRecord(t->ev_call, s) == hipSuccess) t->call_seen = true;
    if (!order_back) return;
    if (hipEventRecord(t->ev_null[1], s) == hipSuccess)
      (void)hipStreamWaitEvent(nullptr, t->ev_null[1], 0);
  }
  CallStream(const CallStream&) = delete;
  CallStream& operator=(const CallStream&) = delete;
};

// Event pair around one timed launch.
hipEvent_t timing_event(dpf_topology* t) {
  if (t->evused == t->evpool.size()) {
    hipEvent_t e;
    HIPTRY(hipEventCreate(&e));
    t->evpool.push_back(e);
  }
  return t->evpool[t->evused++];
}

// A launch of kernel class `cat` on AMG level `level` (-1: none) with values
// of w bytes, moving `bytes` algorithmic bytes.  With kernel timing on
// (DPF_FLAG_KERNEL_TIMING; or sweep timing, for finest-level sweeps), the
// launch carries a start and a stop event that the runtime stamps at the
// dispatch's own begin and end (the timestamps a kernel trace reports), so
// the timing adds no packets between kernels; harvest_timing adds them up.
template <typename F, typename... Args>
void klaunch(dpf_topology* t, int cat, int level, int w, double bytes, F kernel, dim3 grid,
             dim3 block, size_t shmem, hipStream_t s, Args... args) {
  if (t->ktiming || (t->timing && cat == KC_CHEB && level == 0)) {
    hipEvent_t a = timing_event(t);
    hipEvent_t b = timing_event(t);
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, args...);
    t->kpend.push_back(KRec{cat, level, w, t->kiter, t->kphase, bytes, a, b});
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
  }
}

// Adds the recorded durations (finest-level sweeps to the sweep total, every
// launch to the kernel table).  its >= 0 (end of a Krylov solve with its
// iterations): launches the early-exit flag turned into no-ops are left out,
// i.e. every launch of a later iteration and the V-cycle of iteration its
// (launches outside the Krylov loop are tagged iteration -1).
void harvest_timing(dpf_topology* t, int its = -1) {
  if (!t->kpend.empty()) HIPTRY(hipEventSynchronize(t->kpend.back().b));
  for (const KRec& r : t->kpend) {
    if (its >= 0 && r.iter >= 0 && (r.iter > its || (r.iter == its && r.phase == 1))) continue;
    float ms = 0;
    HIPTRY(hipEventElapsedTime(&ms, r.a, r.b));
    if (t->timing && r.cat == KC_CHEB && r.level == 0) {
      t->sweep_ms += ms;
      t->sweep_n++;
    }
    if (t->ktiming) {
      KAcc& e = t->ktable[std::make_tuple(r.cat, r.level, r.w)];
      e.launches++;
      e.ms += ms;
      e.bytes += r.bytes;
    }
  }
  t->kpend.clear();
  t->evused = 0;
}

// Sizes for the algorithmic byte counts of level l: rows, nonzeros of the
// level operator and the bytes of the columns the smoothing sweeps read (the
// fine operator's offset slices store none; compact columns take 2 bytes),
// nonzeros of the prolongator, the bytes of one P or R column, and rows of
// the next level.
struct LvlSize {
  double rows, nnz, cols, pnnz, pcol, nc;
};
LvlSize lvl_size(const dpf_topology* t, int l) {
  const LevelHost& H = t->H.levels[l];
  const LevelDev& D = *t->lev[l];
  LvlSize z;
  z.rows = (double)H.A.nrows;
  z.nnz = (double)H.A.nnz;
  if (l == 0)
    z.cols = 4.0 * (double)t->fine_explicit_nnz;
  else
    z.cols = (D.as_lg > 0 && D.As.compact ? 2.0 : 4.0) * (double)H.A.nnz;
  z.pnnz = H.agg.empty() ? 0.0 : (double)H.P.nnz;
  z.pcol = D.transfer && D.P.compact ? 2.0 : 4.0;
  z.nc = H.agg.empty() ? 0.0 : (double)H.R.nrows;
  return z;
}

// Kernel timing of one C-ABI call (DPF_FLAG_KERNEL_TIMING): on from its
// start, harvested into the topology's table once its stream has drained.
void ktiming_begin(dpf_topology* t, const dpf_opts* o) {
  t->kpend.clear();
  t->evused = 0;
  t->kiter = -1;
  t->kphase = 0;
  t->ktiming = (o->flags & DPF_FLAG_KERNEL_TIMING) != 0;
}
void ktiming_end(dpf_topology* t) {
  if (t->ktiming) harvest_timing(t);
  t->ktiming = false;
}

template <typename T>
void ensure_hier(dpf_topology* t, Hier<T>& h) {
  if (h.allocated) return;
  int L = t->levels();
  h.lev.clear();
  h.lev = std::vector<LevelNum<T>>(L);
  for (int l = 0; l < L; l++) {
    LevelDev& d = *t->lev[l];
    LevelNum<T>& n = h.lev[l];
    size_t nr = (size_t)d.A.nrows;
    if (!(l == 0 && sizeof(T) == sizeof(double))) n.val.alloc(d.A.entries * sizeof(T));
    // coarse operators: the Galerkin refresh never writes SELL padding
    if (l > 0) HIPTRY(hipMemsetAsync(n.val.p, 0, d.A.entries * sizeof(T), nullptr));
    if (!(l == 0 && sizeof(T) == sizeof(double))) n.invd.alloc(nr * sizeof(T));
    if (d.transfer) {
      n.pval.alloc(d.P.entries * sizeof(T));
      n.rval.alloc(d.Rs.entries * sizeof(T));
      n.apval.alloc(d.ap_nnz * sizeof(T));
    }
    if (d.as_lg > 0) n.asval.alloc(d.As.entries * sizeof(T));
    if (std::is_same<T, float>::value) {
      // the bf16 copies the FP32 V-cycle reads: transfer operators and the
      // coarse levels' split operators
      if (d.transfer) {
        n.p16.alloc(d.P.entries * sizeof(uint16_t));
        n.r16.alloc(d.Rs.entries * sizeof(uint16_t));
      }
      if (l > 0 && d.as_lg > 0) {
        n.a16.alloc(d.A.entries * sizeof(uint16_t));
        n.dg.alloc(nr * sizeof(T));
      }
      if (l > 0 && d.as_lg > 0) n.as16.alloc(d.As.entries * sizeof(uint16_t));
    }
    n.x.alloc(nr * sizeof(T));
    n.xs.alloc(nr * sizeof(T));
    n.d.alloc(nr * sizeof(T));
    n.r.alloc(nr * sizeof(T));
    if (l > 0) n.b.alloc(nr * sizeof(T));
  }
  // the zeroing above (null stream) completes before any stream uses the levels
  HIPTRY(hipStreamSynchronize(nullptr));
  int64_t nc = t->lev[L - 1]->A.nrows;
  if (t->H.coarse_dense) h.W.alloc((size_t)nc * nc * sizeof(double));
  h.status.alloc(sizeof(int));
  size_t nf = (size_t)t->H.nf;
  h.cx.alloc(nf * sizeof(T));
  h.cr.alloc(nf * sizeof(T));
  h.cp0.alloc(nf * sizeof(T));
  h.cp1.alloc(nf * sizeof(T));
  h.cq.alloc(nf * sizeof(T));
  h.allocated = true;
}

template <typename T>
T* level_val(dpf_topology* t, Hier<T>& h, int l) {
  if (l == 0 && sizeof(T) == sizeof(double)) return (T*)t->A64.p;
  return h.lev[l].val.template as<T>();
}
template <typename T>
T* level_invd(dpf_topology* t, Hier<T>& h, int l) {
  if (l == 0 && sizeof(T) == sizeof(double)) return (T*)t->invd64.p;
  return h.lev[l].invd.template as<T>();
}

// Fine assembly from t->Ke (+ Q, dh): A64 (+ FP32 copy into h32 level 0), b64.
void assemble(dpf_topology* t, hipStream_t s, bool want32, const double* Q, bool want_b,
              double tscale = 1.0, const double* svec = nullptr) {
  LevelDev& L0 = *t->lev[0];
  float* v32 = nullptr;
  float* i32 = nullptr;
  // the finest level has no bf16 copy (a16, dg stay null)
  uint16_t* a16 = nullptr;
  float* dg = nullptr;
  if (want32) {
    ensure_hier(t, t->h32);
    v32 = t->h32.lev[0].val.as<float>();
    i32 = t->h32.lev[0].invd.as<float>();
  }
  // incidence (edge, other, slot) and the K, aol gathers per entry; per row
  // its maps, Q and diagonal slot; the values (FP64, and FP32 if wanted), the
  // l1 diagonals and b written
  const LvlSize z0 = lvl_size(t, 0);
  const double ninc = (double)t->H.inc_ptr.back();
  const double ab = ninc * 26.0 + z0.rows * 22.0 + z0.nnz * (want32 ? 12.0 : 8.0) +
                    z0.rows * ((want32 ? 12.0 : 8.0) + (want_b ? 8.0 : 0.0));
  klaunch(t, KC_ASSEMBLE, 0, 8, ab, k_assemble, dim3(nblocks(t->H.nf)), dim3(BLK), 0, s,
          (int64_t)t->H.nf,
                     t->f2n.as<int32_t>(), t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                     t->inc_other.as<int32_t>(), t->inc_slot.as<uint16_t>(),
                     L0.diag_slot.as<uint16_t>(), t->Ke.as<double>(), t->aol.as<double>(),
                     t->dh.as<double>(), Q, tscale, svec,
                     L0.A.view<double>(t->A64.as<double>()), v32, t->invd64.as<double>(), i32,
                     want_b ? t->b64.as<double>() : (double*)nullptr, a16, dg,
                     &t->hm_dev->cut);
  HIPTRY(hipGetLastError());
}

// Numeric refresh of levels 1.. and the coarsest inverse, from level 0 values.
template <typename T>
void refresh(dpf_topology* t, Hier<T>& h, hipStream_t s) {
  ensure_hier(t, h);
  int L = t->levels();
  for (int l = 0; l + 1 < L; l++) {
    LevelDev& d = *t->lev[l];
    LevelDev& dn = *t->lev[l + 1];
    SellD<T> A = d.A.view<T>(level_val(t, h, l));
    SellD<T> P = d.P.view<T>(h.lev[l].pval.template as<T>());
    T* apv = h.lev[l].apval.template as<T>();
    T* cval = level_val(t, h, l + 1);
    const double w = sizeof(T);
    const double b16w = std::is_same<T, float>::value ? 1.0 : 0.0;  // bf16 copies written
    const LevelHost& LH = t->H.levels[l];
    const LvlSize z = lvl_size(t, l), zn = lvl_size(t, l + 1);
    const double nap = (double)LH.ap_off.back(), nrap = (double)LH.rap_off.back();
    klaunch(t, KC_PSMOOTH, l, (int)w,
            z.nnz * (w + 2.0) + z.rows * (w + 2.0) + z.pnnz * (w + 2.0 * b16w),
            k_psmooth<T>, dim3(nblocks(d.A.nrows)), dim3(BLK), 0, s, A,
            (const T*)level_invd(t, h, l), d.p_slot.as<uint16_t>(), d.p_self.as<uint16_t>(), P,
            h.lev[l].p16.template as<uint16_t>(), l == 0 ? t->omega_fine : t->omega_coarse);
    klaunch(t, KC_RGATHER, l, (int)w, z.pnnz * (4.0 + 2.0 * w + 2.0 * b16w), k_rgather<T>,
            dim3(nblocks(d.Rs.entries)), dim3(BLK), 0, s, d.Rs.entries, d.rs_src.as<int32_t>(),
            (const T*)P.val, h.lev[l].rval.template as<T>(), (const uint16_t*)nullptr,
            h.lev[l].r16.template as<uint16_t>());
    if (d.ap_rows) {
      // one fine row per lane: the row's A entries and, per entry, the P row
      // it selects (value, start, length), a code byte and a P value per
      // product, the row-wise AP entries written
      klaunch(t, KC_AP_ROWS, l, (int)w,
              z.nnz * (w + 12.0) + (double)d.ap_prods * (1.0 + w) + (double)d.ap_nnz * w,
              k_ap_rows<T>, dim3(nblocks(d.A.nrows << d.apr_lg)), dim3(BLK),
              (size_t)d.ap_w * BLK * sizeof(T), s, A,
              d.A.offsets ? (const int32_t*)nullptr : d.A.rowlen.as<int32_t>(), d.apr_lg,
              d.ap_w, P, d.P.rowlen.as<int32_t>(), d.apc_ptr.as<int64_t>(),
              d.apc.as<uint8_t>(), d.ap_rp.as<int64_t>(), apv);
    } else {
      klaunch(t, KC_PROD_AP, l, (int)w, nap * (8.0 + 2.0 * w) + (double)d.ap_nnz * (8.0 + w),
              k_prod_gather<T>, dim3(nblocks(d.ap_nnz << d.ap_lg)), dim3(BLK), 0, s, d.ap_nnz,
              d.ap_lg, d.ap_off.as<int64_t>(), (const int32_t*)nullptr, d.ap_a.as<int32_t>(),
              d.ap_p.as<int32_t>(), (const T*)A.val, (const T*)P.val, apv);
    }
    if (d.rap_rows) {
      // coarse rows over Rs's lanes: Rs's entries (column, P-value map, P
      // value, AP row start), a product code and an AP value per product, the
      // upper entries written
      klaunch(t, KC_RAP_ROWS, l, (int)w,
              (double)d.Rs.entries * (14.0 + w) + (double)d.rap_prods * (2.0 + w) +
                  (double)d.n_rap * w,
              k_rap_rows<T>, dim3(nblocks(d.Rs.nrows)), dim3(BLK),
              (size_t)d.rap_w * BLK * sizeof(T), s, d.Rs.view<T>(nullptr),
              d.Rs.rowlen.as<int32_t>(), d.rs_lg, d.rap_w, dn.A.nrows, d.rs_src.as<int32_t>(),
              d.rcd_ptr.as<int64_t>(), d.rcd.as<uint16_t>(), d.rnp.as<uint16_t>(),
              (const T*)P.val, d.ap_rp.as<int64_t>(), (const T*)apv, dn.A.sptr.as<int64_t>(),
              dn.A.rowlen.as<int32_t>(), dn.diag_slot.as<uint16_t>(), cval);
    } else {
      klaunch(t, KC_PROD_RAP, l, (int)w, nrap * (8.0 + 2.0 * w) + (double)d.n_rap * (12.0 + w),
              k_prod_gather<T>, dim3(nblocks(d.n_rap << d.rap_lg)), dim3(BLK), 0, s, d.n_rap,
              d.rap_lg, d.rap_off.as<int64_t>(), d.rap_dest.as<int32_t>(), d.rap_p.as<int32_t>(),
              d.rap_ap.as<int32_t>(), (const T*)P.val, (const T*)apv, cval);
    }
    if (d.n_mirror > 0)
      klaunch(t, KC_MIRROR, l + 1, (int)w, (double)d.n_mirror * (8.0 + 2.0 * w),
              k_mirror_copy<T>, dim3(nblocks(d.n_mirror)), dim3(BLK), 0, s, d.n_mirror,
              d.rap_mdst.as<int32_t>(), d.rap_msrc.as<int32_t>(), cval);
    klaunch(t, KC_ROWDIAG, l + 1, (int)w,
            zn.nnz * (w + 2.0 * b16w) + zn.rows * (2.0 + w + w * b16w), k_rowdiag<T>,
            dim3(nblocks(dn.A.nrows)), dim3(BLK), 0, s, dn.A.view<T>(cval),
            dn.diag_slot.as<uint16_t>(), level_invd(t, h, l + 1),
            h.lev[l + 1].a16.template as<uint16_t>(), h.lev[l + 1].dg.template as<T>());
    if (dn.as_lg > 0)
      klaunch(t, KC_RGATHER, l + 1, (int)w, zn.nnz * (4.0 + 2.0 * w + 4.0 * b16w), k_rgather<T>,
              dim3(nblocks(dn.As.entries)), dim3(BLK), 0, s, dn.As.entries,
              dn.as_src.as<int32_t>(), (const T*)cval, h.lev[l + 1].asval.template as<T>(),
              (const uint16_t*)h.lev[l + 1].a16.template as<uint16_t>(),
              h.lev[l + 1].as16.template as<uint16_t>());
  }
  if (t->H.coarse_dense) {
    LevelDev& dc = *t->lev[L - 1];
    int ncd = (int)dc.A.nrows;
    if (ncd <= 64) {
      klaunch(t, KC_DENSE_INVERSE, L - 1, (int)sizeof(T), 2.0 * ncd * ncd * 8.0,
              k_dense_inverse64<T>, dim3(1), dim3(256), 0, s,
              dc.A.view<T>(level_val(t, h, L - 1)), dc.A.rowlen.as<int32_t>(),
              h.W.template as<double>(), h.status.template as<int>());
    } else {
      size_t lds = ncd <= DENSE_LDS_MAX ? (size_t)ncd * ncd * sizeof(double) : 0;
      klaunch(t, KC_DENSE_INVERSE, L - 1, (int)sizeof(T), 2.0 * ncd * ncd * 8.0,
              k_dense_inverse<T>, dim3(1), dim3(1024), lds, s,
              dc.A.view<T>(level_val(t, h, L - 1)), dc.A.rowlen.as<int32_t>(),
              h.W.template as<double>(), h.status.template as<int>());
    }
  }
  HIPTRY(hipGetLastError());
}

template <typename T>
auto cheb_kernel(int level) -> decltype(&k_cheb<T, true>) {
  return level == 0 ? &k_cheb<T, true> : &k_cheb<T, false>;
}

// The residual kernel of level l (the finest level's can fold a FIN_XR).
template <typename T>
auto residual_kernel(int level) -> decltype(&k_residual<T, true>) {
  return level == 0 ? &k_residual<T, true> : &k_residual<T, false>;
}

// No finish folded into a kernel.
FoldFin no_fold() {
  FoldFin f{};
  f.mode = -1;
  return f;
}

// The member table of a batch (null for a single system).
const int2* bm_ptr(dpf_topology* t) { return t->nm > 1 ? t->bm.as<int2>() : nullptr; }

// Member maps of the fine grid (logical blocks of BLK rows).  A single system
// is one member spanning every block.
void setup_members(dpf_topology* t) {
  int64_t nf = t->H.nf;
  int64_t nblk = (nf + BLK - 1) / BLK;
  if (t->m_row0.empty()) t->m_row0 = {0, nf};
  int nm = t->nm;
  std::vector<int> mblk(2 * nm);
  std::vector<int2> bm(nblk);
  for (int m = 0; m < nm; m++) {
    mblk[2 * m] = (int)(t->m_row0[m] / BLK);
    mblk[2 * m + 1] = (int)((t->m_row0[m + 1] - 1) / BLK);
  }
  int m = 0;
  for (int64_t b = 0; b < nblk; b++) {
    int64_t r0 = b * BLK;
    while (m + 1 < nm && t->m_row0[m + 1] <= r0) m++;
    int64_t nxt = m + 1 < nm ? t->m_row0[m + 1] : INT32_MAX;
    bm[b] = make_int2(m, (int)std::min<int64_t>(nxt, INT32_MAX));
  }
  upload(t->mblk, mblk);
  upload(t->bm, bm);
}


int mirror_load(const int* p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
void mirror_store(int* p, int v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }
double mirror_load(const double* p) {
  double v;
  __atomic_load(p, &v, __ATOMIC_ACQUIRE);
  return v;
}

// Waits (spinning) until the mirror's field at `seqp` equals `want`.  Fails
// if the stream drains or errs without it, or after a generous time limit.
void wait_published(dpf_topology* t, hipStream_t s, const int* seqp, int want) {
  double t0 = now_wall_ms();
  unsigned spins = 0;
  while (mirror_load(seqp) != want) {
    if ((++spins & 1023) == 0) {
      hipError_t q = hipStreamQuery(s);
      if (q != hipSuccess && q != hipErrorNotReady)
        throw Err{DPF_EHIP, std::string("stream error while waiting: ") + hipGetErrorString(q)};
      if (q == hipSuccess && mirror_load(seqp) != want)
        throw Err{DPF_EHIP, "stream drained without publishing the solve state"};
      if (now_wall_ms() - t0 > 300000.0) throw Err{DPF_EHIP, "timed out waiting for the GPU"};
    }
    __builtin_ia32_pause();
  }
}

// Finishes the grid sum the previous kernel (nparts blocks) left in t->partial.
void finish(dpf_topology* t, hipStream_t s, int mode, unsigned nparts, double reltol = 0,
            double abstol = 0, int maxiter = 0, int abs_target = 0, int seq = 0,
            const double* partial = nullptr) {
  klaunch(t, KC_FINISH, -1, 8, 16.0 * (double)nparts + 256.0, k_finish, dim3(1), dim3(1024), 0,
          s, mode, t->nm, (int)nparts,
          partial ? partial : (const double*)t->partial.as<double>(), bm_ptr(t),
          (const int*)t->mblk.as<int>(), t->scal.as<SolveScalars>(), t->gdone.as<int>(), reltol,
          abstol, maxiter, abs_target, seq, t->hm_dev, t->hm_outer_dev);
}

// Algorithmic bytes of a Chebyshev sweep on level l (values of w bytes):
// the operator, reads of x, b and the l1 diagonal, the write of x', and d
// read (c1 != 0) and written (when a later sweep reads it).
// (With bf16 values, 2 bytes per value and the compensated diagonal read.)
double cheb_bytes(const dpf_topology* t, int l, double w, bool din, bool dout, bool b16) {
  LvlSize z = lvl_size(t, l);
  return z.nnz * (b16 ? 2.0 : w) + z.cols +
         z.rows * w * (4.0 + (din ? 1.0 : 0.0) + (dout ? 1.0 : 0.0) + (b16 ? 1.0 : 0.0));
}

// One V-cycle on levels l..: input b at level l (for l = 0 the CG residual),
// level x/d already hold the zero-guess first Chebyshev step.  Returns the
// buffer holding x_l on exit through *xout.
template <typename T>
void vcycle(dpf_topology* t, Hier<T>& h, hipStream_t s, int l, const T* bl, T* xl, T* xsl,
            T** xout, const std::vector<Cheb>& chs, int* done, bool fuse_dot_at_fine,
            const FoldFin* fold = nullptr) {
  const Cheb& ch = chs[l];
  int L = t->levels();
  LevelDev& d = *t->lev[l];
  LevelNum<T>& n = h.lev[l];
  SellD<T> A = d.A.view<T>(level_val(t, h, l));
  const T* invd = level_invd(t, h, l);
  T* dd = n.d.template as<T>();
  double* part = t->partial.as<double>();
  unsigned g = nblocks(d.A.nrows);
  int m = (int)ch.c1.size();
  const int w = (int)sizeof(T);
  const LvlSize z = lvl_size(t, l);
  // the smoothing sweeps and the residual use the row-split copy on coarse
  // levels; the FP32 V-cycle reads it, and the transfer operators, in bf16
  int lg = d.as_lg;
  SellD<T> Asm = lg > 0 ? d.As.view<T>(n.asval.template as<T>()) : A;
  constexpr bool is32 = std::is_same<T, float>::value;
  const bool b16 = is32 && t->vbf16;
  // (not the finest level: its FP32 values, which the CG product reads in the
  // same iteration, stay in the Infinity Cache between the kernels that read
  // them, where a bf16 copy would add its own traffic)
  const bool a16 = b16 && l > 0 && lg > 0;
  const T* dg = nullptr;
  if (a16) {
    Asm.v16 = n.as16.template as<uint16_t>();
    dg = n.dg.template as<T>();
  }
  const double wv = a16 ? 2.0 : (double)w, wt = b16 ? 2.0 : (double)w;
  unsigned gs = nblocks(Asm.nrows);
  int64_t nr = d.A.nrows;
  T* cur = xl;
  T* oth = xsl;
  // a folded finish rides on the finest level's residual, its first kernel
  // when the finest smoother has degree 1
  if (fold && (l != 0 || l == L - 1 || m != 1))
    throw Err{DPF_EHIP, "a finish can be folded only into the finest level's first kernel"};
  const FoldFin nofold = no_fold();
  if (l == L - 1) {
    if (t->H.coarse_dense) {
      int nc = (int)d.A.nrows;
      klaunch(t, KC_COARSE_SOLVE, l, w, (double)nc * nc * 8.0 + 2.0 * nc * w,
              k_coarse_solve<T>, dim3(1), dim3(256), nc * sizeof(double), s, nc,
              (const double*)h.W.template as<double>(), bl, cur, (const int*)done);
      // a one-level hierarchy: the CG scalar r.z still has to be formed
      if (fuse_dot_at_fine) {
        klaunch(t, KC_RZ_DOT, l, w, 2.0 * z.rows * w + 16.0 * g, k_rz_dot<T>, dim3(g),
                dim3(BLK), 0, s, d.A.nrows, bl, (const T*)cur, (const int*)done, bm_ptr(t),
                part);
        finish(t, s, FIN_RZ, g);
      }
    } else {
      // no dense solve: a longer smoothing sweep (degree 2m) stands in
      Cheb c2;
      c2.make(2 * m, 1.0, ch.ratio);
      for (int k = 1; k < 2 * m; k++) {
        int fuse = (fuse_dot_at_fine && k == 2 * m - 1) ? 1 : 0;
        bool wd = k != 2 * m - 1;
        klaunch(t, KC_CHEB, l, w, cheb_bytes(t, l, w, true, wd, a16), cheb_kernel<T>(l),
                dim3(gs), dim3(BLK), 0, s, Asm, lg, nr, (const T*)cur, oth, bl, (const T*)dd,
                wd ? dd : (T*)nullptr, invd, dg, (T)c2.c1[k], (T)c2.c2[k], (const int*)done,
                fuse, bm_ptr(t), part);
        if (fuse) finish(t, s, FIN_RZ, g);
        std::swap(cur, oth);
      }
    }
    *xout = cur;
    return;
  }
  // pre-smoothing (step 0 done by the producer of b)
  for (int k = 1; k < m; k++) {
    // the last pre-smoothing sweep's d is dead: the post-smoothing restarts at c1 = 0
    bool wd = k != m - 1;
    klaunch(t, KC_CHEB, l, w, cheb_bytes(t, l, w, true, wd, a16), cheb_kernel<T>(l), dim3(gs),
            dim3(BLK), 0, s, Asm, lg, nr, (const T*)cur, oth, bl, (const T*)dd,
            wd ? dd : (T*)nullptr, invd, dg, (T)ch.c1[k], (T)ch.c2[k], (const int*)done, 0,
            bm_ptr(t), part);
    std::swap(cur, oth);
  }
  T* rl = n.r.template as<T>();
  klaunch(t, KC_RESIDUAL, l, w, z.nnz * wv + z.cols + (a16 ? 4.0 : 3.0) * z.rows * w,
          residual_kernel<T>(l), dim3(gs), dim3(BLK), 0, s, Asm, lg, nr, (const T*)cur, bl, dg,
          rl, (const int*)done, fold ? *fold : nofold);
  LevelDev& dn = *t->lev[l + 1];
  LevelNum<T>& nn = h.lev[l + 1];
  bool coarse_next = (l + 1 == L - 1) && t->H.coarse_dense;
  SellD<T> Rs = d.Rs.view<T>(n.rval.template as<T>());
  if (b16) Rs.v16 = n.r16.template as<uint16_t>();
  klaunch(t, KC_RESTRICT, l, w,
          z.pnnz * (wt + (d.Rs.compact ? 2.0 : 4.0)) + z.rows * w +
              z.nc * w * (coarse_next ? 1.0 : 4.0),
          k_restrict<T>, dim3(nblocks(d.Rs.nrows)), dim3(BLK), 0, s, Rs, d.rs_lg, dn.A.nrows,
          (const T*)rl, nn.b.template as<T>(), (const T*)level_invd(t, h, l + 1),
          nn.d.template as<T>(), nn.x.template as<T>(), (T)chs[l + 1].c0, coarse_next ? 0 : 1,
          (const int*)done);
  SellD<T> P = d.P.view<T>(n.pval.template as<T>());
  if (b16) P.v16 = n.p16.template as<uint16_t>();
  if (coarse_next && dn.A.nrows <= COARSE_FOLD_MAX) {
    // the dense coarsest solve is folded into the prolongation
    klaunch(t, KC_PROLONG_COARSE, l, w,
            z.pnnz * (wt + z.pcol) + z.nc * z.nc * 8.0 + z.nc * w + 2.0 * z.rows * w,
            k_prolong_coarse<T>, dim3(g), dim3(BLK), 0, s, P, (int)dn.A.nrows,
            (const double*)h.W.template as<double>(), (const T*)nn.b.template as<T>(), cur,
            (const int*)done);
  } else {
    T* xc = nullptr;
    vcycle<T>(t, h, s, l + 1, nn.b.template as<T>(), nn.x.template as<T>(),
              nn.xs.template as<T>(), &xc, chs, done, false);
    klaunch(t, KC_PROLONG, l, w, z.pnnz * (wt + z.pcol) + z.nc * w + 2.0 * z.rows * w,
            k_prolong<T>, dim3(g), dim3(BLK), 0, s, P, (const T*)xc, cur, (const int*)done);
  }
  for (int k = 0; k < m; k++) {
    double c1 = k == 0 ? 0.0 : ch.c1[k];
    double c2 = k == 0 ? ch.c0 : ch.c2[k];
    int fuse = (fuse_dot_at_fine && k == m - 1) ? 1 : 0;
    bool wd = k != m - 1;
    klaunch(t, KC_CHEB, l, w, cheb_bytes(t, l, w, k != 0, wd, a16), cheb_kernel<T>(l),
            dim3(gs), dim3(BLK), 0, s, Asm, lg, nr, (const T*)cur, oth, bl, (const T*)dd,
            wd ? dd : (T*)nullptr, invd, dg, (T)c1, (T)c2, (const int*)done, fuse, bm_ptr(t),
            part);
    if (fuse) finish(t, s, FIN_RZ, g);
    std::swap(cur, oth);
  }
  *xout = cur;
}

// Refills level l's split-copy values in hierarchy h from the level's
// operator, as the numeric refresh does (after the copy was re-laid).
template <typename T>
void refill_split(dpf_topology* t, Hier<T>& h, int l, hipStream_t s) {
  if (!h.allocated) return;
  LevelDev& d = *t->lev[l];
  LevelNum<T>& n = h.lev[l];
  constexpr bool is32 = std::is_same<T, float>::value;
  n.asval.alloc(d.As.entries * sizeof(T));
  if (is32) n.as16.alloc(d.As.entries * sizeof(uint16_t));
  const double w = sizeof(T);
  const LvlSize z = lvl_size(t, l);
  klaunch(t, KC_RGATHER, l, (int)w, z.nnz * (4.0 + 2.0 * w + (is32 ? 4.0 : 0.0)), k_rgather<T>,
          dim3(nblocks(d.As.entries)), dim3(BLK), 0, s, d.As.entries, d.as_src.as<int32_t>(),
          (const T*)level_val(t, h, l), n.asval.template as<T>(),
          (const uint16_t*)n.a16.template as<uint16_t>(), n.as16.template as<uint16_t>());
}

// Re-lays level l's split copy with 2^lg lanes a row (1 <= lg <= 6): host
// split, compact upload, values refilled in every allocated hierarchy.
// Captured iteration graphs hold the old buffers and are dropped.
void set_level_lanes(dpf_topology* t, int l, int lg, hipStream_t s) {
  LevelHost& lh = t->H.levels[l];
  LevelDev& d = *t->lev[l];
  if (d.as_lg == lg) return;
  // work queued by earlier calls, on any stream, may still read the old copy
  HIPTRY(hipDeviceSynchronize());
  int got = 0;
  split_rows(lh.A, got, lh.As, lh.as_src, lg);
  lh.as_lg = got;
  d.As.up_compact(lh.As);
  upload(d.as_src, lh.as_src);
  d.as_lg = got;
  refill_split(t, t->h64, l, s);
  refill_split(t, t->h32, l, s);
  for (auto& kv : t->iter_graphs) (void)hipGraphExecDestroy(kv.second);
  t->iter_graphs.clear();
}

// Mean time in microseconds of one level-l smoothing sweep in the V-cycle's
// configuration for precision T (d read, not written), over reps launches
// back to back after one untimed launch, on zeroed level vectors.
template <typename T>
double time_level_sweep(dpf_topology* t, Hier<T>& h, int l, hipStream_t s, int reps) {
  LevelDev& d = *t->lev[l];
  LevelNum<T>& n = h.lev[l];
  constexpr bool is32 = std::is_same<T, float>::value;
  const bool a16 = is32 && t->vbf16 && d.as_lg > 0;
  SellD<T> Asm = d.As.view<T>(n.asval.template as<T>());
  const T* dg = nullptr;
  if (a16) {
    Asm.v16 = n.as16.template as<uint16_t>();
    dg = n.dg.template as<T>();
  }
  const int64_t nr = d.A.nrows;
  for (DBuf* v : {&n.x, &n.xs, &n.b, &n.d}) HIPTRY(hipMemsetAsync(v->p, 0, nr * sizeof(T), s));
  const unsigned gs = nblocks(Asm.nrows);
  auto kern = &k_cheb<T, false>;
  auto launch = [&]() {
    hipLaunchKernelGGL(kern, dim3(gs), dim3(BLK), 0, s, Asm, d.as_lg, nr,
                       (const T*)n.x.template as<T>(), n.xs.template as<T>(),
                       (const T*)n.b.template as<T>(), (const T*)n.d.template as<T>(),
                       (T*)nullptr, (const T*)level_invd(t, h, l), dg, (T)0.5, (T)0.5,
                       (const int*)t->zero_flag.as<int>(), 0, bm_ptr(t),
                       t->partial.as<double>());
  };
  hipEvent_t e0, e1;
  HIPTRY(hipEventCreate(&e0));
  HIPTRY(hipEventCreate(&e1));
  launch();
  HIPTRY(hipEventRecord(e0, s));
  for (int k = 0; k < reps; k++) launch();
  HIPTRY(hipEventRecord(e1, s));
  HIPTRY(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPTRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPTRY(hipGetLastError());
  return 1e3 * (double)ms / reps;
}

// Lanes-per-row autotuner: for every coarse level with a split copy, time its
// smoothing sweep with 2, 4, ..., 32 lanes a row (on the hierarchy h the
// V-cycle runs in, after its numeric refresh) and keep the fastest layout.
template <typename T>
void tune_lanes(dpf_topology* t, Hier<T>& h, hipStream_t s) {
  const int L = t->levels();
  t->lane_us.assign(L, std::array<double, 7>{});
  for (int l = 1; l < L; l++) {
    if (t->lev[l]->as_lg <= 0) continue;
    if (l == L - 1 && t->H.coarse_dense) continue;  // solved densely, never swept
    int best = t->lev[l]->as_lg;
    double best_us = 1e300;
    for (int lg = 1; lg <= 5; lg++) {
      set_level_lanes(t, l, lg, s);
      double us = time_level_sweep(t, h, l, s, 20);
      t->lane_us[l][lg] = us;
      if (us < best_us) {
        best_us = us;
        best = lg;
      }
    }
    set_level_lanes(t, l, best, s);
  }
  HIPTRY(hipStreamSynchronize(s));
  t->lanes_fixed = true;
}

struct SolveResult {
  int status;  // SolveScalars.done
  int iters;
  double rr, rr0;
};

// PCG on the fine level: solves A x = b (b length nf, device); x ends in
// hk.cx.  The Krylov vectors, the operator A and its SpMV are in TK (from
// hk); the V-cycle preconditioner runs in TP on hp's hierarchy.  TK == TP
// (hk and hp the same hierarchy) is the plain FP64 or FP32 solve; TK =
// double, TP = float is the FP64 Krylov method with an FP32 V-cycle, whose
// input is the FP64 residual rounded to FP32 (in hp.cr).  Blocking (polls the
// device convergence flag through the host mirror).
template <typename TK, typename TP>
SolveResult pcg(dpf_topology* t, Hier<TK>& hk, Hier<TP>& hp, hipStream_t s, const TK* b,
                double reltol, double abstol, int maxiter, const std::vector<Cheb>& chs,
                int abs_target = 0, const double* b64 = nullptr) {
  constexpr bool split = !std::is_same<TK, TP>::value;
  const Cheb& ch = chs[0];
  LevelDev& d0 = *t->lev[0];
  LevelNum<TP>& n0 = hp.lev[0];
  int64_t nf = t->H.nf;
  unsigned g = nblocks(nf);
  double* part = t->partial.as<double>();
  SolveScalars* sc = t->scal.as<SolveScalars>();
  int* done = t->gdone.as<int>();
  TK* x = hk.cx.template as<TK>();
  TK* r = hk.cr.template as<TK>();
  TK* q = hk.cq.template as<TK>();
  TK* pb[2] = {hk.cp0.template as<TK>(), hk.cp1.template as<TK>()};
  TP* rp = split ? hp.cr.template as<TP>() : nullptr;  // V-cycle input when split
  const TP* vb = split ? (const TP*)rp : (const TP*)(void*)r;
  const TP* invd = level_invd(t, hp, 0);
  SellD<TK> A = d0.A.view<TK>(level_val(t, hk, 0));
  // the zero-guess step's d is read only by further pre-smoothing sweeps, or
  // by the longer sweep that stands in for a dense solve on a one-level
  // hierarchy
  bool need_d = ch.c1.size() > 1 || (t->levels() == 1 && !t->H.coarse_dense);
  TP* dd = need_d ? n0.d.template as<TP>() : nullptr;
  TP* z0 = n0.x.template as<TP>();
  TP* zs = n0.xs.template as<TP>();
  int seq = ++t->seq;
  // algorithmic bytes of the CG kernels (Krylov values wk, V-cycle values wp)
  const double wk = sizeof(TK), wp = sizeof(TP), rows = (double)nf;
  const LvlSize z0s = lvl_size(t, 0);
  const double vec_p = (split ? wp : 0.0) + (dd ? wp : 0.0) + 2.0 * wp;  // rp, d0, invd, z0
  const double b_start = rows * (3.0 * wk + vec_p) + 16.0 * g;
  const double b_pq = z0s.nnz * wk + z0s.cols + rows * (wp + 3.0 * wk) + 16.0 * g;
  const double b_xr = rows * (6.0 * wk + vec_p) + 16.0 * g;
  t->kiter = 0;
  t->kphase = 0;
  klaunch(t, KC_CG_START, 0, (int)wk, b_start + (b64 ? rows * (8.0 - wk) : 0.0),
          k_cg_start<TK, TP>, dim3(g), dim3(BLK), 0, s, nf, b64 ? (const TK*)nullptr : b, b64,
          (const SolveScalars*)sc, r, x, rp, invd, dd, z0, (TP)ch.c0, bm_ptr(t), part);
  finish(t, s, FIN_START, g, reltol, abstol, maxiter, abs_target, seq);
  TP* z = nullptr;
  t->kphase = 1;
  vcycle<TP>(t, hp, s, 0, vb, z0, zs, &z, chs, done, true);
  int par = 0;
  // one iteration's launches (iteration it, for the timing records); the
  // V-cycle's output buffer is the same on every call, so z does not change
  // from one iteration to the next
  // (in two parts: the CG update up to the convergence test, then the V-cycle).
  // A single system's convergence test (FIN_XR) rides on block 0 of the
  // V-cycle's first kernel, the finest level's residual (finest smoother of
  // degree 1), when the V-cycle follows the update in the same queue: one
  // launch less an iteration.  A batch, and an update queued without its
  // V-cycle (the forecast's), keep the k_finish launch.
  const bool fold_xr = t->nm == 1 && t->levels() >= 2 && ch.c1.size() == 1;
  FoldFin fxr = no_fold();
  if (fold_xr) {
    fxr.partial = part;
    fxr.sc = sc;
    fxr.gdone = done;
    fxr.hm = t->hm_dev;
    fxr.nblk = (int)g;
    fxr.mode = FIN_XR;
  }
  auto body_a = [&](int pr, int it, bool xr_finish) {
    t->kiter = it;
    t->kphase = 0;
    klaunch(t, KC_CG_PQ, 0, (int)wk, b_pq, k_cg_pq<TK, TP>, dim3(g), dim3(BLK), 0, s, A,
            (const TP*)z, (const TK*)pb[pr], pb[pr ^ 1], q, (const int*)done,
            (const SolveScalars*)sc, bm_ptr(t), part);
    finish(t, s, FIN_PQ, g);
    klaunch(t, KC_CG_XR, 0, (int)wk, b_xr, k_cg_xr<TK, TP>, dim3(g), dim3(BLK), 0, s, nf, x, r,
            (const TK*)pb[pr ^ 1], (const TK*)q, rp, invd, dd, z0, (TP)ch.c0, (const int*)done,
            (const SolveScalars*)sc, bm_ptr(t), part);
    if (xr_finish) finish(t, s, FIN_XR, g);
  };
  auto body_b = [&](bool folded) {
    TP* zz = nullptr;
    t->kphase = 1;
    vcycle<TP>(t, hp, s, 0, vb, z0, zs, &zz, chs, done, true, folded ? &fxr : nullptr);
    if (zz != z) throw Err{DPF_EHIP, "V-cycle output buffer moved between iterations"};
  };
  auto body = [&](int pr, int it) {
    body_a(pr, it, !fold_xr);
    body_b(fold_xr);
  };
  bool graphs = t->use_graphs && !t->timing && !t->ktiming;
  std::string gkey;
  if (graphs) {
    // everything the captured launches take as arguments and can differ
    // between calls: types, the smoother's coefficients, the buffers
    char buf[160];
    std::snprintf(buf, sizeof(buf), "%zu/%zu/%d/%p/%p/%p/%p/%p/", sizeof(TK), sizeof(TP),
                  (int)t->vbf16, (void*)z, (const void*)vb, (void*)dd, (void*)A.val, (void*)x);
    gkey = buf;
    auto num = [&](double v) {  // exact (hexadecimal) text of a coefficient
      std::snprintf(buf, sizeof(buf), "%a,", v);
      gkey += buf;
    };
    for (const Cheb& c : chs) {
      num(c.c0);
      for (size_t k = 0; k < c.c1.size(); k++) {
        num(c.c1[k]);
        num(c.c2[k]);
      }
      gkey += ";";
    }
  }
  auto graph_for = [&](int pr) -> hipGraphExec_t {
    std::string key = gkey + std::to_string(pr);
    auto it = t->iter_graphs.find(key);
    if (it != t->iter_graphs.end()) return it->second;
    HIPTRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    hipGraph_t gr = nullptr;
    try {
      body(pr, 0);
    } catch (...) {
      (void)hipStreamEndCapture(s, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    HIPTRY(hipStreamEndCapture(s, &gr));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPTRY(e);
    t->iter_graphs[key] = ex;
    return ex;
  };
  SolveResult res{0, 0, 0, 0};
  // Keep one iteration queued beyond the last one whose convergence test the
  // device has published: the host enqueues iteration k+1 while the GPU runs
  // iteration k's V-cycle, no launch outlives convergence, and the stream is
  // never drained to poll.
  HostMirror* m = t->hm_host;
  int launched = 0;
  double t0 = now_wall_ms();
  unsigned spins = 0;
  // Convergence forecast (single systems, launched iterations): from rr of
  // the last two published iterations; an iteration forecast to reach the
  // stop threshold is queued without its V-cycle, which follows only once its
  // convergence test has come back negative.  A converged solve then leaves
  // no V-cycle of no-op launches behind; a wrong forecast costs one host
  // round trip.
  int seen = -1;
  double rr_a = -1.0, rr_b = -1.0;
  bool half = false;  // the last queued iteration's V-cycle is not queued
  while (true) {
    int its = 0, dn = 0;
    if (mirror_load(&m->seq) == seq) {
      dn = mirror_load(&m->done);
      its = mirror_load(&m->iters);
    }
    if (dn != 0) {
      res.status = dn;
      res.iters = its;
      res.rr = mirror_load(&m->rr);
      break;
    }
    if (its != seen && its > 0) {
      seen = its;
      rr_a = rr_b;
      rr_b = mirror_load(&m->rr);
    }
    if (half) {
      if (its >= launched) {  // its test came back and the solve goes on
        body_b(false);
        half = false;
        HIPTRY(hipGetLastError());
        continue;
      }
    } else if (launched < maxiter && launched <= its) {
      bool forecast = false;
      if (!graphs && t->nm == 1 && rr_a > 0.0 && rr_b > 0.0 && rr_b < rr_a) {
        const double stop = mirror_load(&m->rr_stop);
        forecast = stop > 0.0 && rr_b * (rr_b / rr_a) <= stop;
      }
      if (graphs)
        HIPTRY(hipGraphLaunch(graph_for(par), s));
      else if (forecast)
        body_a(par, launched + 1, true);
      else
        body(par, launched + 1);
      half = forecast;
      par ^= 1;
      launched++;
      HIPTRY(hipGetLastError());
      continue;
    }
    if ((++spins & 1023) == 0) {
      hipError_t qs = hipStreamQuery(s);
      if (qs != hipSuccess && qs != hipErrorNotReady)
        throw Err{DPF_EHIP, std::string("stream error in the Krylov loop: ") + hipGetErrorString(qs)};
      if (qs == hipSuccess && mirror_load(&m->seq) == seq && mirror_load(&m->done) == 0 &&
          mirror_load(&m->iters) < launched)
        throw Err{DPF_EHIP, "Krylov loop made no progress on the device"};
      if (now_wall_ms() - t0 > 300000.0) throw Err{DPF_EHIP, "timed out waiting for the GPU"};
    }
    __builtin_ia32_pause();
  }
  t->kiter = -1;
  t->kphase = 0;
  harvest_timing(t, res.iters);
  return res;
}

}  // namespace

// ====================================================================== C-ABI

extern "C" {

const char* dpf_last_error(void) { return g_err.c_str(); }

void dpf_opts_init(dpf_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->reltol = 1e-10;
  o->abstol = 0.0;
  o->maxiter = 2000;
  o->precision = DPF_FP64;
  o->inner_reltol = 1e-5;
  o->max_outer = 20;
  o->cheb_degree = 0;
  o->memory = DPF_HOST;
  o->stream = nullptr;
}

}  // extern "C"

namespace {

// Shared by dpf_topology_create and dpf_batch_topology_create.  members:
// cell offsets of the packed members (empty for a single system).
// Staging of the allocation-free entry point dpf_hm_forward_noalloc: exp(logK)
// per cell, DPF_NOALLOC_MAX_CELLS doubles (32 MiB) in every GPU's copy of the
// code object, held until the process ends.  Host code never passes this
// name to a kernel (on the host it is only a stand-in for the symbol): it
// uses the device address hipGetSymbolAddress returns for the current
// device, kept per device in g_scratch.
static __device__ __align__(256) double scratch[DPF_NOALLOC_MAX_CELLS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The address of `scratch` on device dev, looked up once per device with dev
// current (the caller makes it current).
double* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) throw Err{DPF_EINVAL, "device index out of range"};
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Err{DPF_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_scratch[dev] = (double*)ptr;
  }
  return g_scratch[dev];
}
}  // namespace

int create_impl(int64_t n, int64_t nedges, const int64_t* neighbors,
                const double* areasoverlengths, int64_t ndirichlet,
                const int64_t* dirichletnodes, const std::vector<int64_t>& cell0,
                dpf_topology** out) {
  auto t = std::make_unique<dpf_topology>();
  try {
    std::string e = build_topology(t->H, n, nedges, neighbors, areasoverlengths, ndirichlet,
                                   dirichletnodes);
    if (!e.empty()) {
      set_err(e);
      return e.find("singular") != std::string::npos ? DPF_ESINGULAR : DPF_EINVAL;
    }
    if (!cell0.empty()) {
      // members: free rows are numbered in cell order, so each member owns a
      // contiguous range of them
      t->nm = (int)cell0.size() - 1;
      t->m_cell0 = cell0;
      t->m_row0.assign(t->nm + 1, 0);
      for (int m = 0; m < t->nm; m++) {
        int64_t cnt = 0;
        for (int64_t i = cell0[m]; i < cell0[m + 1]; i++) cnt += t->H.isfree[i];
        t->m_row0[m + 1] = t->m_row0[m] + cnt;
        if (cnt < BLK) {
          set_err("batch member " + std::to_string(m + 1) + " has fewer than " +
                  std::to_string(BLK) + " free cells");
          return DPF_EINVAL;
        }
      }
    }
    e = build_hierarchy(t->H, 25, 128, 512, t->nm > 1 ? t->m_row0 : std::vector<int64_t>());
    if (!e.empty()) {
      set_err(e);
      return DPF_EINVAL;
    }
    HIPTRY(hipGetDevice(&t->device));
    // the allocation-free entry point's staging address on this device
    (void)scratch_on(t->device);
    HIPTRY(hipStreamCreateWithFlags(&t->own_stream, hipStreamNonBlocking));
    for (auto& ev : t->ev_setup) HIPTRY(hipEventCreate(&ev));
    for (auto& ev : t->ev_null) HIPTRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPTRY(hipEventCreateWithFlags(&t->ev_call, hipEventDisableTiming));
    // room for a few forward states (K_e, dh, Q, heads, multipliers) and the
    // per-call temporaries of the host-memory interface
    size_t state_bytes = (size_t)(t->H.ne + 3 * t->H.n + 2 * t->H.nf) * sizeof(double);
    t->pool = std::make_shared<BufPool>(std::max<size_t>(size_t(1) << 30, 4 * state_bytes));
    TopologyHost& H = t->H;
    upload(t->edges, H.edges);
    upload(t->aol, H.aol);
    upload(t->n2f, H.n2f);
    upload(t->f2n, H.f2n);
    upload(t->inc_ptr, H.inc_ptr);
    upload(t->inc_edge, H.inc_edge);
    upload(t->inc_other, H.inc_other);
    // the fine operator goes to the GPU in the offset-slice layout when the
    // hierarchy has coarse levels (a one-level hierarchy inverts it densely,
    // which walks rows through rowlen)
    FineLayout F;
    bool fine_offsets = H.levels.size() >= 2;
    if (fine_offsets) {
      build_fine_layout(H, F);
      t->fine_offset_slices = F.offset_slices;
      t->fine_explicit_nnz = F.explicit_nnz;
    } else {
      t->fine_explicit_nnz = H.levels[0].A.nnz;
    }
    upload(t->inc_slot, fine_offsets ? F.inc_slot : H.inc_slot);
    int64_t maxrows = 0;
    for (size_t l = 0; l < H.levels.size(); l++) {
      auto d = std::make_unique<LevelDev>();
      LevelHost& lh = H.levels[l];
      bool fl = l == 0 && fine_offsets;
      if (fl)
        d->A.up_fine(lh.A, F);
      else
        d->A.up(lh.A);
      upload(d->diag_slot, fl ? F.diag_slot : lh.diag_slot);
      d->as_lg = lh.as_lg;
      if (lh.as_lg > 0) {
        d->As.up_compact(lh.As);
        upload(d->as_src, lh.as_src);
        maxrows = std::max(maxrows, lh.As.nrows);
      }
      maxrows = std::max(maxrows, lh.A.nrows);
      if (!lh.agg.empty()) {
        d->transfer = true;
        d->P.up_compact(lh.P);
        d->Rs.up_compact(lh.Rs);
        d->rs_lg = lh.rs_lg;
        d->nc = lh.R.nrows;
        upload(d->p_slot, fl ? F.p_slot : lh.p_slot);
        upload(d->p_self, lh.p_self);
        upload(d->rs_src, lh.rs_src);
        upload(d->ap_off, lh.ap_off);
        upload(d->ap_a, fl ? F.ap_a : lh.ap_a);
        upload(d->ap_p, lh.ap_p);
        upload(d->rap_off, lh.rap_off);
        upload(d->rap_dest, lh.rap_dest);
        upload(d->rap_mdst, lh.rap_mdst);
        upload(d->rap_msrc, lh.rap_msrc);
        d->n_mirror = (int64_t)lh.rap_mdst.size();
        d->n_rap = (int64_t)lh.rap_dest.size();
        d->ap_lg = lh.ap_lg;
        d->rap_lg = lh.rap_lg;
        upload(d->rap_p, lh.rap_p);
        d->ap_nnz = lh.ap_rp.back();
        GalerkinPlan G;
        std::string ge = build_galerkin_plan(H, (int)l, fl ? &F : nullptr, G);
        if (!ge.empty()) throw Err{DPF_EINVAL, ge};
        d->ap_rows = G.ap_rows;
        d->rap_rows = G.rap_rows;
        d->apr_lg = G.ap_lg;
        d->ap_w = G.ap_w;
        d->rap_w = G.rap_w;
        if (G.ap_rows) {
          d->ap_prods = (int64_t)G.apc.size();
          upload(d->ap_rp, lh.ap_rp);
          upload(d->apc_ptr, G.apc_ptr);
          upload(d->apc, G.apc);
        }
        if (G.rap_rows) {
          d->rap_prods = (int64_t)G.rcd.size();
          upload(d->rcd_ptr, G.rcd_ptr);
          upload(d->rcd, G.rcd);
          upload(d->rnp, G.rnp);
        } else {
          upload(d->rap_ap, lh.rap_ap);
        }
        maxrows = std::max(maxrows, lh.Rs.nrows);
      }
      t->lev.push_back(std::move(d));
    }
    maxrows = std::max<int64_t>(maxrows, std::max(H.n, H.ne));
    t->max_grid = nblocks(maxrows) + 8;
    t->partial.alloc(2 * t->max_grid * sizeof(double));
    t->partial2.alloc(2 * t->max_grid * sizeof(double));
    setup_members(t.get());
    t->scal.alloc(t->nm * sizeof(SolveScalars));
    HIPTRY(hipMemset(t->scal.p, 0, t->nm * sizeof(SolveScalars)));
    t->gdone.alloc(sizeof(int));
    HIPTRY(hipMemset(t->gdone.p, 0, sizeof(int)));
    t->zero_flag.alloc(sizeof(int));
    HIPTRY(hipMemset(t->zero_flag.p, 0, sizeof(int)));
    HIPTRY(hipHostMalloc((void**)&t->scal_host, sizeof(SolveScalars), hipHostMallocDefault));
    HIPTRY(hipHostMalloc((void**)&t->hm_host, sizeof(HostMirror), hipHostMallocCoherent));
    std::memset(t->hm_host, 0, sizeof(HostMirror));
    t->hm_host->seq = -1;
    t->hm_host->outer_seq = -1;
    HIPTRY(hipHostGetDevicePointer((void**)&t->hm_dev, t->hm_host, 0));
    HIPTRY(hipHostMalloc((void**)&t->hm_outer_host, t->nm * sizeof(double),
                         hipHostMallocCoherent));
    std::memset(t->hm_outer_host, 0, t->nm * sizeof(double));
    HIPTRY(hipHostGetDevicePointer((void**)&t->hm_outer_dev, t->hm_outer_host, 0));
    HIPTRY(hipHostMalloc((void**)&t->dots_host, 3 * sizeof(double), hipHostMallocCoherent));
    std::memset(t->dots_host, 0, 3 * sizeof(double));
    HIPTRY(hipHostGetDevicePointer((void**)&t->dots_dev, t->dots_host, 0));
    t->A64.alloc(t->lev[0]->A.entries * sizeof(double));
    t->invd64.alloc(H.nf * sizeof(double));
    t->b64.alloc(H.nf * sizeof(double));
    // pooled: a forward that keeps a state hands these buffers to the state
    // and takes fresh ones from the pool
    for (DBuf* b : {&t->Ke, &t->dh, &t->Q}) b->pool = t->pool.get();
    t->Ke.alloc(H.ne * sizeof(double));
    t->dh.alloc(H.n * sizeof(double));
    t->Q.alloc(H.n * sizeof(double));
    HIPTRY(hipDeviceSynchronize());
  } catch (const Err& e) {
    set_err(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return DPF_ENOMEM;
  }
  *out = t.release();
  return DPF_OK;
}

}  // namespace

extern "C" {

int dpf_topology_create(int64_t n, int64_t nedges, const int64_t* neighbors,
                        const double* areasoverlengths, int64_t ndirichlet,
                        const int64_t* dirichletnodes, dpf_topology** out) {
  if (!out) {
    set_err("out is NULL");
    return DPF_EINVAL;
  }
  *out = nullptr;
  if ((nedges > 0 && (!neighbors || !areasoverlengths)) || (ndirichlet > 0 && !dirichletnodes)) {
    set_err("NULL array argument");
    return DPF_EINVAL;
  }
  return create_impl(n, nedges, neighbors, areasoverlengths, ndirichlet, dirichletnodes, {}, out);
}

int dpf_batch_topology_create(int32_t nmembers, const int64_t* ncells, const int64_t* nedges,
                              const int64_t* neighbors, const double* areasoverlengths,
                              const int64_t* ndirichlet, const int64_t* dirichletnodes,
                              dpf_topology** out) {
  if (!out) {
    set_err("out is NULL");
    return DPF_EINVAL;
  }
  *out = nullptr;
  if (nmembers <= 0 || !ncells || !nedges || !ndirichlet) {
    set_err("invalid member description");
    return DPF_EINVAL;
  }
  try {
    std::vector<int64_t> cell0(nmembers + 1, 0), edge0(nmembers + 1, 0), d0(nmembers + 1, 0);
    for (int m = 0; m < nmembers; m++) {
      if (ncells[m] <= 0 || nedges[m] < 0 || ndirichlet[m] < 0) {
        set_err("member " + std::to_string(m + 1) + " has a negative or empty size");
        return DPF_EINVAL;
      }
      cell0[m + 1] = cell0[m] + ncells[m];
      edge0[m + 1] = edge0[m] + nedges[m];
      d0[m + 1] = d0[m] + ndirichlet[m];
    }
    if ((edge0[nmembers] > 0 && (!neighbors || !areasoverlengths)) ||
        (d0[nmembers] > 0 && !dirichletnodes)) {
      set_err("NULL array argument");
      return DPF_EINVAL;
    }
    // pack: member-local 1-based indices -> global 1-based indices
    std::vector<int64_t> nb(2 * edge0[nmembers]), dn(d0[nmembers]);
    for (int m = 0; m < nmembers; m++) {
      for (int64_t e = edge0[m]; e < edge0[m + 1]; e++)
        for (int k = 0; k < 2; k++) {
          int64_t v = neighbors[2 * e + k];
          if (v < 1 || v > ncells[m]) {
            set_err("member " + std::to_string(m + 1) + ": neighbors refer to a cell outside 1.." +
                    std::to_string(ncells[m]));
            return DPF_EINVAL;
          }
          nb[2 * e + k] = v + cell0[m];
        }
      for (int64_t k = d0[m]; k < d0[m + 1]; k++) {
        int64_t v = dirichletnodes[k];
        if (v < 1 || v > ncells[m]) {
          set_err("member " + std::to_string(m + 1) + ": dirichletnodes outside 1.." +
                  std::to_string(ncells[m]));
          return DPF_EINVAL;
        }
        dn[k] = v + cell0[m];
      }
    }
    int rc = create_impl(cell0[nmembers], edge0[nmembers], nb.data(), areasoverlengths,
                         d0[nmembers], dn.data(), cell0, out);
    if (rc == DPF_OK) (*out)->m_edge0 = edge0;
    return rc;
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return DPF_ENOMEM;
  }
}

int dpf_member_stats(const dpf_topology* t, double* relres, int32_t nmembers) {
  if (!t || !relres || nmembers != t->nm) {
    set_err("invalid argument (nmembers must equal the topology's member count)");
    return DPF_EINVAL;
  }
  for (int m = 0; m < t->nm; m++)
    relres[m] = m < (int)t->m_relres.size() ? t->m_relres[m] : -1.0;
  return DPF_OK;
}

int dpf_topology_destroy(dpf_topology* t) {
  if (!t) return DPF_OK;
  {
    std::lock_guard<std::mutex> g(t->mu);
    // every call's work done, on whatever stream it ran
    if (t->own_stream && t->call_seen) (void)hipStreamWaitEvent(t->own_stream, t->ev_call, 0);
    if (t->own_stream) (void)hipStreamSynchronize(t->own_stream);
    for (auto& ev : t->ev_setup)
      if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : t->ev_null)
      if (ev) (void)hipEventDestroy(ev);
    if (t->ev_call) (void)hipEventDestroy(t->ev_call);
    if (t->scal_host) (void)hipHostFree(t->scal_host);
    if (t->hm_host) (void)hipHostFree(t->hm_host);
    if (t->hm_outer_host) (void)hipHostFree(t->hm_outer_host);
    if (t->dots_host) (void)hipHostFree(t->dots_host);
    for (auto& kv : t->iter_graphs) (void)hipGraphExecDestroy(kv.second);
    if (t->own_stream) (void)hipStreamDestroy(t->own_stream);
  }
  delete t;
  return DPF_OK;
}

int dpf_topology_tuning(const dpf_topology* t, int32_t precision, double* ms, int32_t* iters,
                        int32_t n) {
  if (!t || !ms || !iters || precision < DPF_FP64 || precision > DPF_MIXED_PC) {
    set_err("invalid argument");
    return DPF_EINVAL;
  }
  for (int k = 0; k < n && k < 7; k++) {
    bool have = t->tuned[precision];
    ms[k] = have ? t->tune_ms[precision][k] : 0.0;
    iters[k] = have ? t->tune_iters[precision][k] : 0;
  }
  return DPF_OK;
}

int dpf_topology_set_smoother(dpf_topology* t, int32_t precision, int32_t deg_fine,
                              int32_t deg_second, int32_t deg_deeper, double ratio) {
  if (!t || precision < DPF_FP64 || precision > DPF_MIXED_PC || deg_fine < 1 || deg_fine > 16 ||
      deg_second < 1 || deg_second > 16 || deg_deeper < 1 || deg_deeper > 16 ||
      !(ratio > 1.0) || !std::isfinite(ratio)) {
    set_err("need a precision code, degrees in 1..16 and an interval ratio > 1");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  t->deg_fine[precision] = deg_fine;
  t->deg_coarse[precision] = deg_second;
  t->deg_deep[precision] = deg_deeper;
  t->ratio_tuned[precision] = ratio;
  for (int k = 0; k < 7; k++) {
    t->tune_ms[precision][k] = 0.0;
    t->tune_iters[precision][k] = 0;
  }
  t->tuned[precision] = true;
  return DPF_OK;
}

int dpf_topology_set_damping(dpf_topology* t, double omega_fine, double omega_coarse) {
  if (!t || !(omega_fine > 0.0 && omega_fine < 4.0) || !(omega_coarse > 0.0 && omega_coarse < 4.0)) {
    set_err("need a topology and damping factors in (0, 4)");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  t->omega_fine = omega_fine;
  t->omega_coarse = omega_coarse;
  // the hierarchies refreshed so far used the old damping: kept states
  // refresh again before their adjoint solves
  t->version++;
  return DPF_OK;
}

int dpf_topology_set_tuning(dpf_topology* t, int32_t precision, int32_t deg_fine,
                            int32_t deg_coarse, double ratio) {
  return dpf_topology_set_smoother(t, precision, deg_fine, deg_coarse, deg_coarse, ratio);
}

int dpf_topology_lanes(const dpf_topology* t, int32_t* lg, double* us, int32_t nlevels) {
  if (!t || !lg) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  const int L = t->levels();
  for (int l = 0; l < nlevels; l++) {
    lg[l] = l < L ? t->lev[l]->as_lg : 0;
    if (us)
      for (int k = 0; k < 7; k++)
        us[7 * l + k] = (l < L && l < (int)t->lane_us.size()) ? t->lane_us[l][k] : 0.0;
  }
  return DPF_OK;
}

int dpf_topology_set_lanes(dpf_topology* t, int32_t level, int32_t lg) {
  if (!t) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  const int L = t->levels();
  if (level < 1 || level >= L || t->lev[level]->as_lg <= 0 || lg < 1 || lg > 6) {
    set_err("need a level that has a split copy (1 <= level < levels, lanes exponent > 0) "
            "and a lanes exponent in 1..6");
    return DPF_EINVAL;
  }
  int prev = -1;
  try {
    // the topology's device is made current for the re-layout and the
    // caller's current device restored afterwards
    HIPTRY(hipGetDevice(&prev));
    HIPTRY(hipSetDevice(t->device));
    set_level_lanes(t, level, lg, t->own_stream);
    HIPTRY(hipStreamSynchronize(t->own_stream));
    HIPTRY(hipSetDevice(prev));
  } catch (const Err& e) {
    if (prev >= 0) (void)hipSetDevice(prev);
    set_err(e.msg);
    return e.code;
  }
  t->lanes_fixed = true;
  return DPF_OK;
}

int dpf_topology_info(const dpf_topology* t, int64_t* info, int32_t ninfo) {
  if (!t || !info) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  const TopologyHost& H = t->H;
  int64_t nnz_all = 0;
  for (auto& l : H.levels) nnz_all += l.A.nnz;
  int64_t v[21] = {H.n,
                  H.ne,
                  H.nf,
                  H.levels[0].A.nnz,
                  (int64_t)H.levels.size(),
                  H.levels.back().A.nrows,
                  (int64_t)(1000.0 * (double)nnz_all / (double)H.levels[0].A.nnz),
                  (int64_t)(H.setup_ms * 1000.0),
                  t->deg_fine[0],
                  t->deg_coarse[0],
                  t->deg_fine[1],
                  t->deg_coarse[1],
                  t->fine_offset_slices,
                  t->deg_fine[2],
                  t->deg_coarse[2],
                  (int64_t)t->ratio_tuned[0],
                  (int64_t)t->ratio_tuned[1],
                  (int64_t)t->ratio_tuned[2],
                  t->deg_deep[0],
                  t->deg_deep[1],
                  t->deg_deep[2]};
  for (int i = 0; i < ninfo && i < 21; i++) info[i] = v[i];
  return DPF_OK;
}

}  // extern "C"

namespace {

// Copy-in helper: host pointer -> staging device buffer, or pass the device pointer.
const double* stage_in(dpf_topology* t, const dpf_opts* o, DBuf& buf, const double* p,
                       int64_t count, hipStream_t s) {
  if (!p) return nullptr;
  if (o->memory == DPF_DEVICE) return p;
  buf.alloc(count * sizeof(double));
  HIPTRY(hipMemcpyAsync(buf.p, p, count * sizeof(double), hipMemcpyHostToDevice, s));
  return buf.as<double>();
}

void copy_dev(void* dst, const void* src, size_t bytes, hipStream_t s) {
  HIPTRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// Solve A x = b64 (device, FP64, nf) with the current hierarchy; x -> xout (device FP64).
// Returns 0 / DPF_ENOCONV / DPF_EBREAKDOWN and fills stats.
int solve_system(dpf_topology* t, hipStream_t s, const double* b, double* xout,
                 const dpf_opts* o, dpf_stats* st, const double* x0 = nullptr,
                 bool steady_check = false) {
  int64_t nf = t->H.nf;
  t->timing = (o->flags & DPF_FLAG_SWEEP_TIMING) != 0;
  t->use_graphs = (o->flags & DPF_FLAG_GRAPH) != 0;
  t->vbf16 = (o->flags & DPF_FLAG_FP32_STORAGE) == 0;
  const LvlSize z0s = lvl_size(t, 0);
  t->sweep_ms = 0;
  t->sweep_n = 0;
  int pidx = o->precision;
  int L = t->levels();
  // degree m0 on the finest level, m1 on the second, md on every deeper one
  auto make_chs = [&](int m0, int m1, int md, double ratio) {
    std::vector<Cheb> v(L);
    for (int l = 0; l < L; l++) v[l].make(l == 0 ? m0 : l == 1 ? m1 : md, 1.0, ratio);
    return v;
  };
  // smoothers: explicit degree (every level, default interval), or the
  // topology's degrees and interval for this precision: the fixed default,
  // pinned ones, or with DPF_FLAG_AUTOTUNE the ones tuned on its first such
  // solve in this precision
  const bool autotune = o->cheb_degree == 0 && (o->flags & DPF_FLAG_AUTOTUNE) != 0;
  bool tune = autotune && !t->tuned[pidx];
  std::vector<Cheb> chs =
      o->cheb_degree > 0
          ? make_chs(o->cheb_degree, o->cheb_degree, o->cheb_degree, t->cheb_ratio)
          : make_chs(t->deg_fine[pidx], t->deg_coarse[pidx], t->deg_deep[pidx],
                     t->ratio_tuned[pidx]);
  double* part = t->partial.as<double>();
  SolveScalars* sc = t->scal.as<SolveScalars>();
  LevelDev& d0 = *t->lev[0];
  SellD<double> A64 = d0.A.view<double>(t->A64.as<double>());
  unsigned g = nblocks(nf);
  int status = DPF_OK;
  int total_it = 0, outer = 0;
  t->x64.alloc(nf * sizeof(double));
  t->r64.alloc(nf * sizeof(double));
  // the members' ||r64|| into the pinned mirror (first: also ||b|| and the
  // refinement target), waited for; returns the members above their target
  auto publish_outer = [&](int first) -> int {
    int oseq = ++t->seq;
    finish(t, s, FIN_OUTER, g, o->reltol, o->abstol, 0, first, oseq);
    HIPTRY(hipGetLastError());
    wait_published(t, s, &t->hm_host->outer_seq, oseq);
    return mirror_load(&t->hm_host->active);
  };
  // r64 = b - A x64 (b alone when first: x64 = 0); every member's ||r64||
  // lands in the pinned mirror.  first: also take it as the member's ||b||
  // and set its refinement target.  warm: x64 holds the initial guess x0,
  // which is first replaced by its A-norm-optimal multiple.  Returns the
  // number of members whose residual is above their target.
  auto residual = [&](int first, bool warm = false) -> int {
    klaunch(t, KC_OUTER_RESIDUAL, 0, 8,
            first ? 16.0 * z0s.rows + 16.0 * g
                  : z0s.nnz * 8.0 + z0s.cols + 24.0 * z0s.rows + 16.0 * g,
            k_outer_residual, dim3(g), dim3(BLK), 0, s, A64,
                       first ? (const double*)nullptr : (const double*)t->x64.as<double>(), b,
                       t->r64.as<double>(), (const SolveScalars*)sc, bm_ptr(t), part);
    if (warm) {
      double* part2 = t->partial2.as<double>();
      klaunch(t, KC_REFINE_VEC, 0, 8, 24.0 * z0s.rows, k_warm_dots, dim3(g), dim3(BLK), 0, s, nf,
                         (const double*)t->x64.as<double>(), b, (const double*)t->r64.as<double>(),
                         (const SolveScalars*)sc, bm_ptr(t), part, part2);
      finish(t, s, FIN_WXB, g);
      finish(t, s, FIN_WXR, g, 0, 0, 0, 0, 0, part2);
      klaunch(t, KC_REFINE_VEC, 0, 8, 40.0 * z0s.rows, k_warm_apply, dim3(g), dim3(BLK), 0, s, nf,
              t->x64.as<double>(), b,
                         t->r64.as<double>(), (const SolveScalars*)sc, bm_ptr(t), part);
    }
    return publish_outer(first);
  };
  auto all_finite = [&]() {
    for (int m = 0; m < t->nm; m++)
      if (!std::isfinite(mirror_load(&t->hm_outer_host[m]))) return false;
    return true;
  };
  // every member works on b 2^-e, max|b| = f 2^e with f in [1/2, 1): exact,
  // and keeps the FP64 norms of b and of the residuals clear of underflow
  // and overflow (an adjoint that decays over many time steps reaches
  // |b| ~ 1e-200 long before it reaches zero)
  klaunch(t, KC_REFINE_VEC, 0, 8, 8.0 * z0s.rows, k_amax, dim3(g), dim3(BLK), 0, s, nf, b,
          bm_ptr(t), part);
  finish(t, s, FIN_AMAX, g);
  klaunch(t, KC_REFINE_VEC, 0, 8, 8.0 * z0s.rows, k_zero_f64, dim3(g), dim3(BLK), 0, s, nf,
          t->x64.as<double>());
  int active = residual(1);
  // a steady system whose assembly met a zero conductance: singular if that
  // cut leaves a component without a Dirichlet cell (host check on K_e)
  if (steady_check && mirror_load(&t->hm_host->cut)) {
    std::vector<double> K(t->H.ne);
    HIPTRY(hipMemcpy(K.data(), t->Ke.p, t->H.ne * sizeof(double), hipMemcpyDeviceToHost));
    std::string e = floating_component(t->H, K.data());
    if (!e.empty()) throw Err{DPF_ESINGULAR, e};
  }
  std::vector<double> bnorm(t->nm);
  for (int m = 0; m < t->nm; m++) bnorm[m] = mirror_load(&t->hm_outer_host[m]);
  bool finite = all_finite();
  if (finite && x0) {
    klaunch(t, KC_REFINE_VEC, 0, 8, 16.0 * z0s.rows, k_scale_copy, dim3(g), dim3(BLK), 0, s, nf,
            x0, t->x64.as<double>(),
                       (const SolveScalars*)sc, bm_ptr(t), 0);
    active = residual(0, true);
    finite = all_finite();
  }
  if (!finite) {
    if (st) std::memset(st, 0, sizeof(*st));
    throw Err{DPF_EINVAL, "non-finite values in the system (NaN or Inf in K, Q, dirichleths or "
                          "the initial guess)"};
  }
  // Iterative refinement on x64: each step solves A d = r for the current FP64
  // residual of every member still above its target (FP64 CG straight to the
  // target in DPF_FP64 and DPF_MIXED_PC modes; FP32 CG to inner_reltol in
  // DPF_MIXED mode), then x64 += d.
  int budget = o->maxiter;
  // FP64 Krylov: one CG to the full tolerance; a restart only if rounding
  // left the true residual just above it
  int max_outer = o->precision == DPF_MIXED ? o->max_outer : 3;
  if (o->precision == DPF_MIXED) ensure_hier(t, t->h32);
  if (o->precision == DPF_MIXED_PC) {
    ensure_hier(t, t->h32);
    ensure_hier(t, t->h64);  // its fine-level Krylov vectors
  }
  // autotuner, part 1: lanes per row of the coarse split copies (once per
  // topology, as the layout serves every precision; unless pinned), timed in
  // the precision the V-cycle runs in on the freshly refreshed hierarchy
  if (autotune && !t->lanes_fixed) {
    bool ktiming_on = t->ktiming, timing_on = t->timing;
    t->ktiming = t->timing = false;
    if (o->precision == DPF_FP64)
      tune_lanes(t, t->h64, s);
    else
      tune_lanes(t, t->h32, s);
    t->ktiming = ktiming_on;
    t->timing = timing_on;
  }
  // one Krylov solve of the current residual with the given smoothers (the
  // FP32 inner solve of DPF_MIXED to relative tolerance inner)
  auto krylov = [&](const std::vector<Cheb>& cc, double inner) -> SolveResult {
    if (o->precision == DPF_FP64)
      return pcg<double, double>(t, t->h64, t->h64, s, t->r64.as<double>(), 0.0, 0.0, budget,
                                 cc, 1);
    if (o->precision == DPF_MIXED_PC)
      return pcg<double, float>(t, t->h64, t->h32, s, t->r64.as<double>(), 0.0, 0.0, budget,
                                cc, 1);
    return pcg<float, float>(t, t->h32, t->h32, s, (const float*)nullptr, inner, 0.0, budget,
                             cc, 0, (const double*)t->r64.as<double>());
  };
  while (active > 0) {
    if (outer >= max_outer || budget <= 0) {
      status = DPF_ENOCONV;
      break;
    }
    SolveResult r;
    if (tune) {
      // Autotuner: time each (finest, second, deeper levels) Chebyshev degree
      // triple on this very system at the default interval, then two other
      // intervals at the best triple, and keep the fastest; the result is fixed
      // for the life of the topology, so its later solves are bitwise
      // reproducible (another process may time differently: it repeats a tuned
      // configuration by pinning it, dpf_topology_set_smoother and
      // dpf_topology_set_lanes).  An untimed
      // solve first brings the kernels' code and the clocks up, and the FP32
      // trials of DPF_MIXED solve to inner_reltol^2 rather than inner_reltol:
      // the whole refinement's worth of convergence, where a cheaper V-cycle
      // that converges more slowly loses what it seems to win over the first
      // few digits.
      bool graphs_on = t->use_graphs, ktiming_on = t->ktiming, timing_on = t->timing;
      t->use_graphs = false;  // trials use launches: a capture would distort their times
      t->ktiming = t->timing = false;  // nor do the trials enter the timing records
      static const int cand[][3] = {{1, 1, 2}, {1, 1, 3}, {1, 1, 4}, {1, 2, 2}, {1, 2, 3}};
      static const double ratios[] = {30.0, 8.0};
      double trial_tol = std::max(o->inner_reltol * o->inner_reltol, 1e-12);
      double best = 1e300;
      (void)krylov(make_chs(1, 1, 2, t->cheb_ratio), trial_tol);
      auto trial = [&](int m0, int m1, int md, double ratio, int slot) {
        std::vector<Cheb> cc = make_chs(m0, m1, md, ratio);
        HIPTRY(hipStreamSynchronize(s));
        double ta = now_ms();
        SolveResult tr = krylov(cc, trial_tol);
        HIPTRY(hipStreamSynchronize(s));
        double el = now_ms() - ta;
        if (slot >= 0) {
          t->tune_ms[pidx][slot] = tr.status == 1 ? el : -1.0;
          t->tune_iters[pidx][slot] = tr.iters;
        }
        if (tr.status == 1 && el < best) {
          best = el;
          t->deg_fine[pidx] = m0;
          t->deg_coarse[pidx] = m1;
          t->deg_deep[pidx] = md;
          t->ratio_tuned[pidx] = ratio;
        }
      };
      for (int c = 0; c < 5; c++) trial(cand[c][0], cand[c][1], cand[c][2], t->cheb_ratio, c);
      int m0 = t->deg_fine[pidx], m1 = t->deg_coarse[pidx], md = t->deg_deep[pidx];
      for (int k = 0; k < 2; k++) trial(m0, m1, md, ratios[k], 5 + k);
      t->tuned[pidx] = true;
      t->use_graphs = graphs_on;
      t->ktiming = ktiming_on;
      t->timing = timing_on;
      tune = false;
      chs = make_chs(t->deg_fine[pidx], t->deg_coarse[pidx], t->deg_deep[pidx],
                     t->ratio_tuned[pidx]);
    }
    r = krylov(chs, o->inner_reltol);
    if (o->precision != DPF_MIXED)
      klaunch(t, KC_REFINE_VEC, 0, 8, 24.0 * z0s.rows, k_axpy_f64, dim3(g), dim3(BLK), 0, s, nf,
              1.0, (const double*)t->h64.cx.as<double>(), t->x64.as<double>());
    else
      klaunch(t, KC_REFINE_VEC, 0, 8, 20.0 * z0s.rows, k_refine, dim3(g), dim3(BLK), 0, s, nf,
              t->x64.as<double>(), (const float*)t->h32.cx.as<float>(), (const SolveScalars*)sc,
              bm_ptr(t));
    total_it += r.iters;
    budget -= r.iters;
    outer++;
    if (r.status == 3) {
      status = DPF_EBREAKDOWN;
      break;
    }
    active = residual(0);
    if (!all_finite()) {
      status = DPF_EBREAKDOWN;
      break;
    }
  }
  double relres = 0.0;
  t->m_relres.resize(t->nm);
  for (int m = 0; m < t->nm; m++) {
    double rm = mirror_load(&t->hm_outer_host[m]);
    t->m_relres[m] = bnorm[m] > 0 ? rm / bnorm[m] : rm;
    relres = std::max(relres, t->m_relres[m]);
  }
  if (!(relres == relres)) relres = INFINITY;

  klaunch(t, KC_REFINE_VEC, 0, 8, 16.0 * z0s.rows, k_scale_copy, dim3(g), dim3(BLK), 0, s, nf,
                     (const double*)t->x64.as<double>(), xout, (const SolveScalars*)sc, bm_ptr(t),
                     1);
  if (st) {
    st->iterations = total_it;
    st->outer = outer;
    st->converged = (status == DPF_OK);
    st->levels = t->levels();
    st->relres = relres;
    st->sweep_ms = t->sweep_ms;
    st->sweep_launches = t->sweep_n;
  }
  t->timing = false;
  t->use_graphs = false;
  return status;
}

// Assemble + numeric refresh for the conductivities already in t->Ke.
void prepare(dpf_topology* t, hipStream_t s, int precision, const double* Q, bool want_b,
             double tscale = 1.0, const double* svec = nullptr) {
  assemble(t, s, precision != DPF_FP64, Q, want_b, tscale, svec);
  t->version++;
  if (precision != DPF_FP64) {
    refresh<float>(t, t->h32, s);
    t->h32.version = t->version;
  } else {
    ensure_hier(t, t->h64);
    refresh<double>(t, t->h64, s);
    t->h64.version = t->version;
  }
  t->assembled_for = t->version;
}

int check_opts(const dpf_opts* o) {
  if (!o) return DPF_OK;
  if (o->precision != DPF_FP64 && o->precision != DPF_MIXED && o->precision != DPF_MIXED_PC)
    return DPF_EINVAL;
  if (o->memory != DPF_HOST && o->memory != DPF_DEVICE) return DPF_EINVAL;
  if (!(o->reltol >= 0) || !(o->abstol >= 0) || o->maxiter <= 0) return DPF_EINVAL;
  if (o->precision == DPF_MIXED && (!(o->inner_reltol > 0 && o->inner_reltol < 1) || o->max_outer <= 0))
    return DPF_EINVAL;
  if (o->cheb_degree < 0 || o->cheb_degree > 16) return DPF_EINVAL;
  return DPF_OK;
}

template <typename F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Err& e) {
    set_err(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return DPF_ENOMEM;
  }
}

}  // namespace

extern "C" {

int dpf_residuals(dpf_topology* t, const double* hfree, const double* Ks,
                  const double* dirichleths, const double* Qs, const double* Ss,
                  const double* V, double* r_out, const dpf_opts* opts) {
  if (!t || !hfree || !Ks || !dirichleths || !Qs || !Ss || !V || !r_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, nf = t->H.nf, ne = t->H.ne;
    const double* dh = stage_in(t, o, t->stage_n, dirichleths, n, s);
    const double* Q = stage_in(t, o, t->stage_n2, Qs, n, s);
    const double* K = stage_in(t, o, t->stage_e, Ks, ne, s);
    const double* h = stage_in(t, o, t->stage_f, hfree, nf, s);
    DBuf ssb, vb;
    const double* ss = stage_in(t, o, ssb, Ss, n, s);
    const double* vv = stage_in(t, o, vb, V, n, s);
    double* r = r_out;
    if (o->memory == DPF_HOST) {
      t->stage_f2.alloc(nf * sizeof(double));
      r = t->stage_f2.as<double>();
    }
    hipLaunchKernelGGL(k_residuals, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                       t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                       t->inc_other.as<int32_t>(), t->n2f.as<int32_t>(), K,
                       t->aol.as<double>(), h, dh, Q, ss, vv, r);
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(r_out, r, nf * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return DPF_OK;
  });
}

// CSR patterns of the four Jacobians and the incidence -> value maps
// (host), built once per topology.
static void build_jac_maps(dpf_topology* t) {
  if (t->jac_maps) return;
  const TopologyHost& H = t->H;
  const LevelHost& L0 = H.levels[0];
  const int64_t nf = H.nf, ninc = H.inc_ptr.back();
  for (int k = 0; k < 4; k++) {
    t->jac_rp[k].assign(nf + 1, 0);
    t->jac_cols[k].clear();
  }
  for (int k = 0; k < 3; k++) t->jac_pos[k].assign(ninc, -1);
  t->jac_dpos.assign(nf, 0);
  std::vector<int32_t> dn;
  for (int64_t j = 0; j < nf; j++) {
    const int32_t i = H.f2n[j];
    const int64_t p0 = H.inc_ptr[i], p1 = H.inc_ptr[i + 1];
    // h: the fine pattern's row (columns ascending)
    const int64_t b0 = (int64_t)t->jac_cols[0].size();
    for (int k = 0; k < L0.A.rowlen[j]; k++) t->jac_cols[0].push_back(L0.A.at(j, k));
    t->jac_dpos[j] = (int32_t)(b0 + L0.diag_slot[j]);
    for (int64_t p = p0; p < p1; p++)
      if (H.inc_slot[p] != NO_SLOT) t->jac_pos[0][p] = (int32_t)(b0 + H.inc_slot[p]);
    t->jac_rp[0][j + 1] = (int64_t)t->jac_cols[0].size();
    // Ks: one entry per incident edge, in edge order
    const int64_t b1 = (int64_t)t->jac_cols[1].size();
    for (int64_t p = p0; p < p1; p++) {
      t->jac_pos[1][p] = (int32_t)(b1 + (p - p0));
      t->jac_cols[1].push_back(H.inc_edge[p]);
    }
    t->jac_rp[1][j + 1] = (int64_t)t->jac_cols[1].size();
    // dirichleths: the Dirichlet neighbours, ascending, each once
    dn.clear();
    for (int64_t p = p0; p < p1; p++)
      if (!H.isfree[H.inc_other[p]]) dn.push_back(H.inc_other[p]);
    std::sort(dn.begin(), dn.end());
    dn.erase(std::unique(dn.begin(), dn.end()), dn.end());
    const int64_t b2 = (int64_t)t->jac_cols[2].size();
    for (int64_t p = p0; p < p1; p++) {
      int32_t o = H.inc_other[p];
      if (!H.isfree[o])
        t->jac_pos[2][p] = (int32_t)(b2 + (std::lower_bound(dn.begin(), dn.end(), o) - dn.begin()));
    }
    t->jac_cols[2].insert(t->jac_cols[2].end(), dn.begin(), dn.end());
    t->jac_rp[2][j + 1] = (int64_t)t->jac_cols[2].size();
    // Qs: the row's own cell
    t->jac_cols[3].push_back(i);
    t->jac_rp[3][j + 1] = j + 1;
  }
  t->jac_maps = true;
}

int dpf_jacobian_pattern(dpf_topology* t, int32_t which, int64_t* nnz, int64_t* rowptr,
                         int32_t* cols) {
  if (!t || !nnz || which < DPF_JAC_H || which > DPF_JAC_QS) {
    set_err("need a topology, a Jacobian kind (DPF_JAC_*) and nnz");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  try {
    build_jac_maps(t);
  } catch (const std::bad_alloc&) {
    set_err("host allocation failed");
    return DPF_ENOMEM;
  }
  const auto& rp = t->jac_rp[which];
  const auto& cc = t->jac_cols[which];
  *nnz = (int64_t)cc.size();
  if (rowptr) std::memcpy(rowptr, rp.data(), rp.size() * sizeof(int64_t));
  if (cols && !cc.empty()) std::memcpy(cols, cc.data(), cc.size() * sizeof(int32_t));
  return DPF_OK;
}

int dpf_jacobian_values(dpf_topology* t, int32_t which, const double* hfree, const double* Ks,
                        const double* dirichleths, const double* Ss, const double* V,
                        double* vals, const dpf_opts* opts) {
  if (!t || !Ss || !V || !vals || which < DPF_JAC_H || which > DPF_JAC_QS ||
      ((which == DPF_JAC_H || which == DPF_JAC_DIRICHLETHS) && !Ks) ||
      (which == DPF_JAC_KS && (!hfree || !dirichleths))) {
    set_err("NULL argument or unknown Jacobian kind (DPF_JAC_H and DPF_JAC_DIRICHLETHS read "
            "Ks; DPF_JAC_KS reads hfree and dirichleths; every kind reads Ss and V)");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    build_jac_maps(t);
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    if (!t->jac_dev) {
      for (int k = 0; k < 4; k++) upload(t->jac_rp_d[k], t->jac_rp[k]);
      for (int k = 0; k < 3; k++) upload(t->jac_pos_d[k], t->jac_pos[k]);
      upload(t->jac_dpos_d, t->jac_dpos);
      t->jac_dev = true;
    }
    const int64_t n = t->H.n, nf = t->H.nf, ne = t->H.ne;
    const int64_t nnz = (int64_t)t->jac_cols[which].size();
    const double* K = which == DPF_JAC_KS || which == DPF_JAC_QS
                          ? nullptr
                          : stage_in(t, o, t->stage_e, Ks, ne, s);
    const double* h = which == DPF_JAC_KS ? stage_in(t, o, t->stage_f, hfree, nf, s) : nullptr;
    const double* dh = which == DPF_JAC_KS ? stage_in(t, o, t->stage_n, dirichleths, n, s)
                                           : nullptr;
    DBuf ssb, vb, outb;
    const double* ss = stage_in(t, o, ssb, Ss, n, s);
    const double* vv = stage_in(t, o, vb, V, n, s);
    double* out = vals;
    if (o->memory == DPF_HOST) {
      outb.alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
      out = outb.as<double>();
    }
    hipLaunchKernelGGL(k_jacobian, dim3(nblocks(nf)), dim3(BLK), 0, s, (int)which, nf,
                       t->f2n.as<int32_t>(), t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                       t->inc_other.as<int32_t>(), t->n2f.as<int32_t>(),
                       t->jac_rp_d[which].as<int64_t>(),
                       which < 3 ? t->jac_pos_d[which].as<int32_t>() : (int32_t*)nullptr,
                       t->jac_dpos_d.as<int32_t>(), K, t->aol.as<double>(), h, dh, ss, vv, out);
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST && nnz > 0)
      HIPTRY(hipMemcpyAsync(vals, out, nnz * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return DPF_OK;
  });
}

}  // extern "C"

namespace {

// logK -> K_e (geometric = false: harmonic mean, true: geometric mean)
int kmap_forward(dpf_topology* t, bool geometric, const double* logK, double* Ke,
                 const dpf_opts* opts) {
  if (!t || !logK || !Ke) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    const double* lk = stage_in(t, o, t->stage_n, logK, t->H.n, s);
    double* out = Ke;
    if (o->memory == DPF_HOST) {
      t->stage_e.alloc(t->H.ne * sizeof(double));
      out = t->stage_e.as<double>();
    }
    ktiming_begin(t, o);
    const double nn = (double)t->H.n, ne = (double)t->H.ne;
    DBuf kb(t->pool.get(), s);  // back to the pool ordered after k_hm_forward
    if (geometric) {
      klaunch(t, KC_KMAP, -1, 8, ne * 16.0 + nn * 8.0, k_geo_forward, dim3(nblocks(t->H.ne)),
              dim3(BLK), 0, s, t->H.ne, t->edges.as<int2>(), lk, out);
    } else {
      kb.alloc(t->H.n * sizeof(double));
      klaunch(t, KC_KMAP, -1, 8, nn * 16.0, k_exp, dim3(nblocks(t->H.n)), dim3(BLK), 0, s,
              t->H.n, lk, kb.as<double>());
      klaunch(t, KC_KMAP, -1, 8, ne * 16.0 + nn * 8.0, k_hm_forward, dim3(nblocks(t->H.ne)),
              dim3(BLK), 0, s, t->H.ne, t->edges.as<int2>(), (const double*)kb.as<double>(),
              out);
    }
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(Ke, out, t->H.ne * sizeof(double), hipMemcpyDeviceToHost, s));
    // host memory: K_e has reached the caller's array when the call returns;
    // device memory: the call returns with its work queued
    if (o->memory == DPF_HOST) HIPTRY(hipStreamSynchronize(s));
    ktiming_end(t);
    return DPF_OK;
  });
}

int kmap_pullback(dpf_topology* t, bool geometric, const double* logK, const double* Ke,
                  const double* dKe, double* dlogK, const dpf_opts* opts) {
  if (!t || (!geometric && !logK) || !Ke || !dKe || !dlogK) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    const double* lk = geometric ? nullptr : stage_in(t, o, t->stage_n, logK, t->H.n, s);
    const double* ke = stage_in(t, o, t->stage_e, Ke, t->H.ne, s);
    DBuf dkb(t->pool.get(), s);
    const double* dk = stage_in(t, o, dkb, dKe, t->H.ne, s);
    double* out = dlogK;
    if (o->memory == DPF_HOST) {
      t->stage_n2.alloc(t->H.n * sizeof(double));
      out = t->stage_n2.as<double>();
    }
    ktiming_begin(t, o);
    const double nn = (double)t->H.n, ninc = (double)t->H.inc_ptr.back();
    DBuf kb(t->pool.get(), s);
    if (geometric) {
      klaunch(t, KC_KMAP, -1, 8, ninc * 4.0 + ninc * 16.0 + nn * 16.0, k_geo_pullback,
              dim3(nblocks(t->H.n)), dim3(BLK), 0, s, t->H.n, t->inc_ptr.as<int64_t>(),
              t->inc_edge.as<int32_t>(), ke, dk, out);
    } else {
      kb.alloc(t->H.n * sizeof(double));
      klaunch(t, KC_KMAP, -1, 8, nn * 16.0, k_exp, dim3(nblocks(t->H.n)), dim3(BLK), 0, s,
              t->H.n, lk, kb.as<double>());
      // per incidence entry: edge, other cell, K_e and dK_e; per cell k, dlogK
      klaunch(t, KC_KMAP, -1, 8, ninc * 24.0 + nn * 24.0, k_hm_pullback, dim3(nblocks(t->H.n)),
              dim3(BLK), 0, s, t->H.n, t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
              t->inc_other.as<int32_t>(), (const double*)kb.as<double>(), ke, dk, out);
    }
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST) {
      HIPTRY(hipMemcpyAsync(dlogK, out, t->H.n * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPTRY(hipStreamSynchronize(s));
    }
    ktiming_end(t);
    return DPF_OK;
  });
}

}  // namespace

extern "C" {

int dpf_hm_forward(dpf_topology* t, const double* logK, double* Ke, const dpf_opts* opts) {
  return kmap_forward(t, false, logK, Ke, opts);
}

int dpf_hm_pullback(dpf_topology* t, const double* logK, const double* Ke, const double* dKe,
                    double* dlogK, const dpf_opts* opts) {
  return kmap_pullback(t, false, logK, Ke, dKe, dlogK, opts);
}

int dpf_geo_forward(dpf_topology* t, const double* logK, double* Ke, const dpf_opts* opts) {
  return kmap_forward(t, true, logK, Ke, opts);
}

int dpf_hm_forward_noalloc(dpf_topology* t, const double* logK, double* Ke,
                           const dpf_opts* opts) {
  if (!t || !logK || !Ke || !opts) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (opts->memory != DPF_DEVICE) {
    set_err("dpf_hm_forward_noalloc takes device memory only (opts->memory = DPF_DEVICE)");
    return DPF_EINVAL;
  }
  if (t->H.n > DPF_NOALLOC_MAX_CELLS) {
    set_err(std::to_string(t->H.n) + " cells exceed the allocation-free staging of " +
            std::to_string(DPF_NOALLOC_MAX_CELLS) + " cells (use dpf_hm_forward)");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  try {
    // the address was looked up when the topology was created on its device
    double* kb = g_scratch[t->device];
    if (!kb) throw Err{DPF_EHIP, "no staging address for the topology's device"};
    CallStream cs(t, opts);
    hipStream_t s = cs.s;
    const int64_t n = t->H.n, ne = t->H.ne;
    // the same two kernels as dpf_hm_forward, with the staging in `scratch`;
    // k_exp writes every entry k_hm_forward reads, so nothing is cleared
    hipLaunchKernelGGL(k_exp, dim3(nblocks(n)), dim3(BLK), 0, s, n, logK, kb);
    hipLaunchKernelGGL(k_hm_forward, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                       (const int2*)t->edges.as<int2>(), (const double*)kb, Ke);
    HIPTRY(hipGetLastError());
  } catch (const Err& e) {
    set_err(e.msg);
    return e.code;
  }
  return DPF_OK;
}

int dpf_geo_pullback(dpf_topology* t, const double* Ke, const double* dKe, double* dlogK,
                     const dpf_opts* opts) {
  return kmap_pullback(t, true, nullptr, Ke, dKe, dlogK, opts);
}

int dpf_steady_forward(dpf_topology* t, const double* Ks, const double* dirichleths,
                       const double* Qs, double* h_out, const dpf_opts* opts, dpf_stats* stats,
                       dpf_state** state_out) {
  if (state_out) *state_out = nullptr;
  if (!t || !Ks || !dirichleths || !Qs || !h_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    // inputs -> topology-owned device buffers
    if (o->memory == DPF_DEVICE) {
      copy_dev(t->Ke.p, Ks, ne * sizeof(double), s);
      copy_dev(t->dh.p, dirichleths, n * sizeof(double), s);
      copy_dev(t->Q.p, Qs, n * sizeof(double), s);
    } else {
      HIPTRY(hipMemcpyAsync(t->Ke.p, Ks, ne * sizeof(double), hipMemcpyHostToDevice, s));
      HIPTRY(hipMemcpyAsync(t->dh.p, dirichleths, n * sizeof(double), hipMemcpyHostToDevice, s));
      HIPTRY(hipMemcpyAsync(t->Q.p, Qs, n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    // a component without a Dirichlet cell (e.g. dirichletnodes empty)
    // makes the steady operator singular
    if (!t->H.floating.empty()) throw Err{DPF_ESINGULAR, t->H.floating};
    ktiming_begin(t, o);
    mirror_store(&t->hm_host->cut, 0);
    HIPTRY(hipEventRecord(t->ev_setup[0], s));
    prepare(t, s, o->precision, t->Q.as<double>(), true);
    HIPTRY(hipEventRecord(t->ev_setup[1], s));
    std::unique_ptr<dpf_state> st;
    double* hfree;
    DBuf tmp(t->pool.get(), s);
    if (state_out) {
      st = std::make_unique<dpf_state>(t);
      st->release_on(s);  // should the state be freed here: behind this call's work
      st->hfree.alloc(nf * sizeof(double));
      hfree = st->hfree.as<double>();
    } else {
      tmp.alloc(nf * sizeof(double));
      hfree = tmp.as<double>();
    }
    int status = solve_system(t, s, t->b64.as<double>(), hfree, o, stats, nullptr, true);
    double* hdev = h_out;
    if (o->memory == DPF_HOST) {
      t->stage_n2.alloc(n * sizeof(double));
      hdev = t->stage_n2.as<double>();
    }
    klaunch(t, KC_SCATTER, -1, 8, (double)n * 20.0, k_scatter_h, dim3(nblocks(n)), dim3(BLK), 0,
            s, n, t->n2f.as<int32_t>(), (const double*)hfree, t->dh.as<double>(), hdev);
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(h_out, hdev, n * sizeof(double), hipMemcpyDeviceToHost, s));
    if (st) {
      st->version = t->version;
      st->precision = o->precision;
      // the state takes this call's K_e, dirichleths and Q buffers as they
      // are (no copies; the kernels queued on them keep valid memory) and the
      // topology takes fresh ones from the shared pool for its next call
      for (auto pr : {std::make_pair(&st->Ke, &t->Ke), std::make_pair(&st->dh, &t->dh),
                      std::make_pair(&st->Q, &t->Q)}) {
        std::swap(pr.first->p, pr.second->p);
        std::swap(pr.first->bytes, pr.second->bytes);
      }
      t->Ke.alloc_on(ne * sizeof(double), s);
      t->dh.alloc_on(n * sizeof(double), s);
      t->Q.alloc_on(n * sizeof(double), s);
    }
    // host memory: the heads have reached the caller's array when the call
    // returns; device memory: the call returns with its work queued (the
    // solve loop has already seen it converge), and the next call on the
    // topology, or the state's release, is ordered after it on the GPU
    if (o->memory == DPF_HOST) HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      // (the setup's events completed before the solve converged)
      float setup = 0.f;
      HIPTRY(hipEventElapsedTime(&setup, t->ev_setup[0], t->ev_setup[1]));
      stats->total_ms = now_ms() - t0;
      stats->setup_ms = setup;
      stats->solve_ms = stats->total_ms - setup;
    }
    ktiming_end(t);
    if (status == DPF_ENOCONV) set_err("AMG-PCG did not reach the tolerance within maxiter");
    if (status == DPF_EBREAKDOWN) set_err("AMG-PCG breakdown (non-SPD operator or NaN input)");
    if (state_out && status == DPF_OK) {
      st->release_on(nullptr, false);  // dpf_state_destroy orders its release
      *state_out = st.release();
    }
    return status;
  });
}

int dpf_steady_pullback(dpf_state* st, const double* delta, double* dKs, double* ddh,
                        double* dQs, const dpf_opts* opts, dpf_stats* stats) {
  if (!st || !st->topo || !delta) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  dpf_topology* t = st->topo;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    // the hierarchy must hold this state's conductivities in the requested precision
    ktiming_begin(t, o);
    bool fresh = (t->version == st->version) &&
                 ((o->precision != DPF_FP64) ? t->h32.version == st->version
                                              : t->h64.version == st->version);
    if (!fresh) {
      copy_dev(t->Ke.p, st->Ke.p, ne * sizeof(double), s);
      copy_dev(t->dh.p, st->dh.p, n * sizeof(double), s);
      prepare(t, s, o->precision, nullptr, false);
      st->version = t->version;
    }
    const double* dl = stage_in(t, o, t->stage_n, delta, n, s);
    st->lam.alloc_on(nf * sizeof(double), s);
    t->stage_f3.alloc(nf * sizeof(double));
    double* rhs = t->stage_f3.as<double>();
    klaunch(t, KC_SCATTER, -1, 8, (double)nf * 20.0, k_gather_free, dim3(nblocks(nf)), dim3(BLK),
            0, s, nf, t->f2n.as<int32_t>(), dl, rhs);
    double t1 = now_ms();
    int status = solve_system(t, s, rhs, st->lam.as<double>(), o, stats);
    double t2 = now_ms();
    const double* lam = st->lam.as<double>();
    double* dK = dKs;
    double* dd = ddh;
    double* dq = dQs;
    DBuf bk(t->pool.get()), bd(t->pool.get()), bq(t->pool.get());
    if (o->memory == DPF_HOST) {
      if (dKs) { bk.alloc(ne * sizeof(double)); dK = bk.as<double>(); }
      if (ddh) { bd.alloc(n * sizeof(double)); dd = bd.as<double>(); }
      if (dQs) { bq.alloc(n * sizeof(double)); dq = bq.as<double>(); }
    }
    // per edge: the pair, aol, the free maps of its cells and the write;
    // per free cell its head and multiplier
    if (dK)
      klaunch(t, KC_GRADIENT, -1, 8, (double)ne * 32.0 + (double)nf * 16.0, k_edge_grad,
              dim3(nblocks(ne)), dim3(BLK), 0, s, ne, t->edges.as<int2>(), t->n2f.as<int32_t>(),
              t->aol.as<double>(), (const double*)st->hfree.as<double>(),
              (const double*)st->dh.as<double>(), lam, 1.0, 0, dK);
    if (dd || dq)
      klaunch(t, KC_GRADIENT, -1, 8, (double)n * 28.0 + (double)nf * 8.0, k_ddh_dq,
              dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                         t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                         t->inc_other.as<int32_t>(), (const double*)st->Ke.as<double>(),
                         t->aol.as<double>(), lam, dl, dd, dq);
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST) {
      if (dKs) HIPTRY(hipMemcpyAsync(dKs, dK, ne * sizeof(double), hipMemcpyDeviceToHost, s));
      if (ddh) HIPTRY(hipMemcpyAsync(ddh, dd, n * sizeof(double), hipMemcpyDeviceToHost, s));
      if (dQs) HIPTRY(hipMemcpyAsync(dQs, dq, n * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPTRY(hipStreamSynchronize(s));
    }
    if (stats) {
      stats->setup_ms = t1 - t0;
      stats->solve_ms = t2 - t1;
      stats->total_ms = now_ms() - t0;
    }
    ktiming_end(t);
    if (status == DPF_ENOCONV) set_err("adjoint AMG-PCG did not reach the tolerance within maxiter");
    if (status == DPF_EBREAKDOWN) set_err("adjoint AMG-PCG breakdown");
    return status;
  });
}

int dpf_kernel_table(dpf_topology* t, dpf_kernel_record* out, int32_t cap, int32_t* count,
                     int32_t reset) {
  if (!t || !count || cap < 0 || (cap > 0 && !out)) {
    set_err("invalid argument");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  int32_t k = 0;
  for (const auto& kv : t->ktable) {
    if (k < cap) {
      dpf_kernel_record& r = out[k];
      std::memset(&r, 0, sizeof(r));
      std::snprintf(r.kernel, sizeof(r.kernel), "%s", kcat_name[std::get<0>(kv.first)]);
      r.level = std::get<1>(kv.first);
      r.value_bytes = std::get<2>(kv.first);
      r.launches = kv.second.launches;
      r.ms = kv.second.ms;
      r.bytes = kv.second.bytes;
    }
    k++;
  }
  *count = k;
  if (reset) t->ktable.clear();
  return DPF_OK;
}

int dpf_state_destroy(dpf_state* st) {
  if (!st) return DPF_OK;
  if (st->topo) {
    // the state's buffers go back to the pool ordered after every call made
    // so far on its topology (a call in DPF_DEVICE mode may return before its
    // work is done): on the topology's stream, behind the last call's end
    dpf_topology* t = st->topo;
    std::lock_guard<std::mutex> g(t->mu);
    bool ordered = !t->call_seen || hipStreamWaitEvent(t->own_stream, t->ev_call, 0) == hipSuccess;
    if (ordered)
      st->release_on(t->own_stream);
    else
      (void)hipStreamSynchronize(t->own_stream);
    delete st;
    return DPF_OK;
  }
  delete st;
  return DPF_OK;
}

}  // extern "C"

namespace {

template <typename T>
double profile_one(dpf_topology* t, Hier<T>& h, hipStream_t s, int kernel, int reps,
                   double* bytes) {
  LevelDev& d0 = *t->lev[0];
  LevelNum<T>& n0 = h.lev[0];
  int64_t nf = t->H.nf, nnz = t->H.levels[0].A.nnz;
  // column bytes only where the layout stores columns
  double cb = 4.0 * (double)t->fine_explicit_nnz;
  unsigned g = nblocks(nf);
  const int* nodone = t->zero_flag.as<int>();
  double* part = t->partial.as<double>();
  SolveScalars* sc = t->scal.as<SolveScalars>();
  SellD<T> A = d0.A.view<T>(level_val(t, h, 0));
  double w = (double)sizeof(T);
  Cheb ch;
  ch.make(2, 1.0, t->cheb_ratio);
  hipEvent_t e0, e1;
  HIPTRY(hipEventCreate(&e0));
  HIPTRY(hipEventCreate(&e1));
  auto launch = [&](int k) {
    T* xa = (k & 1) ? n0.xs.template as<T>() : n0.x.template as<T>();
    T* xb = (k & 1) ? n0.x.template as<T>() : n0.xs.template as<T>();
    if (kernel == 0) {
      // as in the V-cycle with finest degree 1: the post-smoothing sweep after
      // the coarse correction (c1 = 0: no d read, d dead: not written), with
      // the fused r.z partials
      hipLaunchKernelGGL((k_cheb<T, true>), dim3(g), dim3(BLK), 0, s, A, 0, nf, (const T*)xa, xb,
                         (const T*)h.cr.template as<T>(), (const T*)nullptr, (T*)nullptr,
                         (const T*)level_invd(t, h, 0), (const T*)nullptr, (T)0, (T)ch.c0, nodone,
                         1, bm_ptr(t), part);
    } else if (kernel == 1) {
      hipLaunchKernelGGL((k_cg_pq<T, T>), dim3(g), dim3(BLK), 0, s, A, (const T*)xa,
                         (const T*)h.cp0.template as<T>(), h.cp1.template as<T>(),
                         h.cq.template as<T>(), nodone, (const SolveScalars*)sc, bm_ptr(t), part);
    } else {
      hipLaunchKernelGGL((k_residual<double, true>), dim3(g), dim3(BLK), 0, s,
                         d0.A.view<double>(t->A64.as<double>()), 0, nf,
                         (const double*)t->x64.as<double>(),
                         (const double*)t->b64.as<double>(), (const double*)nullptr,
                         t->r64.as<double>(), nodone, no_fold());
    }
  };
  launch(0);  // warm
  HIPTRY(hipEventRecord(e0, s));
  for (int k = 0; k < reps; k++) launch(k);
  HIPTRY(hipEventRecord(e1, s));
  HIPTRY(hipGetLastError());
  HIPTRY(hipEventSynchronize(e1));
  float ms = 0;
  HIPTRY(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (kernel == 0)  // values (+ columns), reads of xin, b, invd, the write of xout
    *bytes = (double)nnz * w + cb + 4.0 * (double)nf * w;
  else if (kernel == 1)
    *bytes = (double)nnz * w + cb + 4.0 * (double)nf * w;
  else
    *bytes = (double)nnz * 8.0 + cb + 3.0 * (double)nf * 8.0;
  return (double)ms / reps;
}

}  // namespace

namespace {

// Grid-wide barrier of a cooperative launch (every workgroup resident):
// thread 0 of each taking-part workgroup arrives on *count; the last arrival
// resets it and moves *gen on, the others wait for *gen to move.  Agent-scope
// fences on both sides make each workgroup's earlier writes visible to every
// other's later reads, across the XCDs' separate L2 caches.  Each wait is
// bounded (spin_limit polls): one that runs out sets *err and the barrier
// returns false in every thread of its workgroup, so no wave waits forever.
__device__ bool grid_barrier(unsigned* count, unsigned* gen, unsigned nparts, unsigned& epoch,
                             int spin_limit, int* err) {
  __shared__ int ok_sh;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = 1;
    const unsigned g = epoch;
    __threadfence();
    unsigned prev = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (prev == nparts - 1) {
      __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      int spins = 0;
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > spin_limit) {
          ok = 0;
          __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    epoch = g + 1u;
    ok_sh = ok;
  }
  __syncthreads();
  return ok_sh != 0;
}

// Barrier cost probe: rounds of (write own word, barrier, read the next
// workgroup's word and check it, barrier).  stride 8: only the workgroups
// with blockIdx % 8 == 0 take part -- all on one XCD when the dispatcher
// deals workgroups to the 8 XCDs in turn; stride 1: every workgroup.
__global__ void __launch_bounds__(256) k_barrier_probe(unsigned* count, unsigned* gen, int rounds,
                                                       int stride, int* err, int* data) {
  if (blockIdx.x % stride != 0) return;
  const unsigned me = blockIdx.x / stride, np = gridDim.x / stride;
  unsigned epoch = 0;
  for (int k = 0; k < rounds; k++) {
    if (threadIdx.x == 0) data[me] = k * 4096 + (int)me;
    if (!grid_barrier(count, gen, np, epoch, 1 << 17, err)) return;
    if (threadIdx.x == 0) {
      const unsigned o = (me + 1) % np;
      if (data[o] != k * 4096 + (int)o)
        __hip_atomic_fetch_or(err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!grid_barrier(count, gen, np, epoch, 1 << 17, err)) return;
  }
}

}  // namespace

extern "C" int dpf_profile_grid_barrier(int32_t nparts, int32_t one_xcd, int32_t rounds,
                                        double* us, int32_t* errors) {
  if (!us || !errors || nparts < 1 || nparts > 4096 || rounds < 1 || rounds > 100000) {
    set_err("need 1 <= nparts <= 4096, 1 <= rounds <= 100000 and output pointers");
    return DPF_EINVAL;
  }
  try {
    int dev = 0, coop = 0;
    HIPTRY(hipGetDevice(&dev));
    HIPTRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    if (!coop) throw Err{DPF_EINVAL, "the device does not support cooperative launches"};
    const int stride = one_xcd ? 8 : 1;
    DBuf words;
    words.alloc((size_t)(4 + nparts) * sizeof(int));
    int* base = words.as<int>();
    unsigned* count = (unsigned*)base;
    unsigned* gen = (unsigned*)base + 1;
    int* err = base + 2;
    int* data = base + 4;
    hipStream_t s = nullptr;
    HIPTRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto run = [&](int nr) -> hipError_t {
      HIPTRY(hipMemsetAsync(base, 0, (size_t)(4 + nparts) * sizeof(int), s));
      void* args[] = {&count, &gen, &nr, (void*)&stride, &err, &data};
      return hipLaunchCooperativeKernel((const void*)k_barrier_probe,
                                        dim3((unsigned)(nparts * stride)), dim3(256), args, 0, s);
    };
    hipError_t le = run(std::min(rounds, 10));
    float ms = 0.f;
    if (le == hipSuccess) {
      HIPTRY(hipEventCreate(&e0));
      HIPTRY(hipEventCreate(&e1));
      HIPTRY(hipMemsetAsync(base, 0, (size_t)(4 + nparts) * sizeof(int), s));
      void* args[] = {&count, &gen, &rounds, (void*)&stride, &err, &data};
      HIPTRY(hipEventRecord(e0, s));
      le = hipLaunchCooperativeKernel((const void*)k_barrier_probe,
                                      dim3((unsigned)(nparts * stride)), dim3(256), args, 0, s);
      HIPTRY(hipEventRecord(e1, s));
      HIPTRY(hipEventSynchronize(e1));
      HIPTRY(hipEventElapsedTime(&ms, e0, e1));
    }
    int herr = 0;
    if (le == hipSuccess) HIPTRY(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    HIPTRY(hipStreamSynchronize(s));
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    if (le != hipSuccess) {
      (void)hipGetLastError();
      throw Err{DPF_EINVAL, std::string("cooperative launch refused: ") + hipGetErrorString(le)};
    }
    *us = 1e3 * (double)ms / (2.0 * rounds);
    *errors = herr;
  } catch (const Err& e) {
    set_err(e.msg);
    return e.code;
  }
  return DPF_OK;
}

extern "C" int dpf_profile_kernel(dpf_topology* t, int32_t precision, int32_t kernel,
                                  int32_t reps, double* avg_ms, double* bytes,
                                  const dpf_opts* opts) {
  if (!t || !avg_ms || !bytes || reps <= 0 || kernel < 0 || kernel > 2) {
    set_err("invalid argument");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    if (precision != DPF_FP64) {
      if (!t->h32.allocated || t->h32.version == 0) throw Err{DPF_EINVAL, "no FP32 solve yet"};
      *avg_ms = profile_one<float>(t, t->h32, s, kernel, reps, bytes);
    } else {
      if (!t->h64.allocated || t->h64.version == 0) throw Err{DPF_EINVAL, "no FP64 solve yet"};
      if (kernel == 2 && (!t->x64.p || !t->r64.p)) throw Err{DPF_EINVAL, "no solve yet"};
      *avg_ms = profile_one<double>(t, t->h64, s, kernel, reps, bytes);
    }
    return DPF_OK;
  });
}

// ====================================================================== transient

namespace {

// (S + dt A) and its hierarchy for the state's K, S and step length dt (the
// Dirichlet part of b -> t->b64, independent of dt)
void prepare_be(dpf_topology* t, hipStream_t s, int precision, dpf_state* st, double dt) {
  copy_dev(t->Ke.p, st->Ke.p, t->H.ne * sizeof(double), s);
  copy_dev(t->dh.p, st->dh.p, t->H.n * sizeof(double), s);
  prepare(t, s, precision, nullptr, true, dt, st->svec.as<double>());
  st->version = t->version;
  st->cur_dt = dt;
}

// The hierarchy holds this state's (S + dt A) in the requested precision.
bool hier_fresh(dpf_topology* t, dpf_state* st, int precision, double dt) {
  return t->version == st->version && st->cur_dt == dt &&
         (precision != DPF_FP64 ? t->h32.version == st->version : t->h64.version == st->version);
}

void ensure_be(dpf_topology* t, hipStream_t s, int precision, dpf_state* st, double dt) {
  if (!hier_fresh(t, st, precision, dt)) prepare_be(t, s, precision, st, dt);
}

// Error text for a failed time step: which step, which phase, how far it got.
void step_failure(const char* what, int step, int code, const dpf_stats& ss) {
  char buf[256];
  std::snprintf(buf, sizeof(buf), "%s %d did not converge (%s after %d iterations in %d "
                "refinement steps, relative residual %.3e)", what, step,
                code == DPF_EBREAKDOWN ? "breakdown" : "iteration limit", ss.iterations, ss.outer,
                ss.relres);
  set_err(buf);
}

// u_{n} from u_{n-1}: one backward-Euler step of length dts[n-1] (warm-started
// at u_{n-1})
int be_step(dpf_topology* t, hipStream_t s, dpf_state* st, int n, const double* uprev,
            double* unext, const dpf_opts* o, int& iters) {
  int64_t nf = t->H.nf;
  double dt = st->dts[n - 1];
  ensure_be(t, s, o->precision, st, dt);
  t->stage_f3.alloc(nf * sizeof(double));
  double* rhs = t->stage_f3.as<double>();
  hipLaunchKernelGGL(k_be_rhs, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                     (const double*)st->svec.as<double>(), uprev,
                     (const double*)st->Q.as<double>(), st->qs[n], dt,
                     (const double*)t->b64.as<double>(), rhs);
  dpf_stats ss;
  std::memset(&ss, 0, sizeof(ss));
  int code = solve_system(t, s, rhs, unext, o, &ss, uprev);
  iters += ss.iterations;
  if (code != DPF_OK) step_failure("backward-Euler step", n, code, ss);
  return code;
}

// Host or device array of count doubles -> host vector (the step lengths and
// Q scales steer the host loop).
void host_copy(std::vector<double>& dst, const double* src, int64_t count, const dpf_opts* o,
               hipStream_t s) {
  dst.resize(count);
  if (o->memory == DPF_DEVICE) {
    HIPTRY(hipMemcpyAsync(dst.data(), src, count * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
  } else {
    std::memcpy(dst.data(), src, count * sizeof(double));
  }
}

int transient_forward_impl(dpf_topology* t, const double* Ks, const double* dirichleths,
                           const double* Qs, const double* Ss, const double* V, const double* u0,
                           int32_t nsteps, const double* dts, double dt_const,
                           const double* qscale, int32_t checkpoint_every, double* u_final,
                           double* states_out, const dpf_opts* opts, dpf_stats* stats,
                           dpf_state** state_out) {
  if (state_out) *state_out = nullptr;
  if (!t || !Ks || !dirichleths || !Qs || !Ss || !V || !u0 || !qscale || !u_final ||
      nsteps <= 0 || (!dts && !(dt_const > 0 && std::isfinite(dt_const)))) {
    set_err("NULL argument or non-positive nsteps/dt");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    auto st = std::make_unique<dpf_state>(t);
    st->kind = 1;
    st->precision = o->precision;
    st->nsteps = nsteps;
    if (dts) {
      host_copy(st->dts, dts, nsteps, o, s);
      for (int k = 0; k < nsteps; k++)
        if (!(st->dts[k] > 0) || !std::isfinite(st->dts[k]))
          throw Err{DPF_EINVAL, "step length " + std::to_string(k + 1) +
                                    " is not positive and finite"};
    } else {
      st->dts.assign(nsteps, dt_const);
    }
    host_copy(st->qs, qscale, nsteps + 1, o, s);
    st->ck = checkpoint_every > 0 ? std::min<int>(checkpoint_every, nsteps) : nsteps;
    st->Ke.alloc(ne * sizeof(double));
    st->dh.alloc(n * sizeof(double));
    st->Q.alloc(n * sizeof(double));
    st->svec.alloc(nf * sizeof(double));
    DBuf ssb(t->pool.get()), vb(t->pool.get()), u0b(t->pool.get());
    const double *ss_d, *v_d, *u0_d;
    if (o->memory == DPF_DEVICE) {
      copy_dev(st->Ke.p, Ks, ne * sizeof(double), s);
      copy_dev(st->dh.p, dirichleths, n * sizeof(double), s);
      copy_dev(st->Q.p, Qs, n * sizeof(double), s);
      ss_d = Ss;
      v_d = V;
      u0_d = u0;
    } else {
      HIPTRY(hipMemcpyAsync(st->Ke.p, Ks, ne * sizeof(double), hipMemcpyHostToDevice, s));
      HIPTRY(hipMemcpyAsync(st->dh.p, dirichleths, n * sizeof(double), hipMemcpyHostToDevice, s));
      HIPTRY(hipMemcpyAsync(st->Q.p, Qs, n * sizeof(double), hipMemcpyHostToDevice, s));
      ss_d = stage_in(t, o, ssb, Ss, n, s);
      v_d = stage_in(t, o, vb, V, n, s);
      u0_d = stage_in(t, o, u0b, u0, nf, s);
    }
    hipLaunchKernelGGL(k_storage, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                       ss_d, v_d, st->svec.as<double>());
    HIPTRY(hipStreamSynchronize(s));
    prepare_be(t, s, o->precision, st.get(), st->dts[0]);
    double t1 = now_ms();
    int nck = nsteps / st->ck + 1;
    st->ckpt.alloc((size_t)nck * nf * sizeof(double));
    double* ck = st->ckpt.as<double>();
    copy_dev(ck, u0_d, nf * sizeof(double), s);
    DBuf ua(t->pool.get()), ub(t->pool.get());
    ua.alloc(nf * sizeof(double));
    ub.alloc(nf * sizeof(double));
    copy_dev(ua.p, u0_d, nf * sizeof(double), s);
    if (states_out) {
      if (o->memory == DPF_DEVICE)
        copy_dev(states_out, u0_d, nf * sizeof(double), s);
      else
        HIPTRY(hipMemcpyAsync(states_out, u0_d, nf * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    double* cur = ua.as<double>();
    double* nxt = ub.as<double>();
    int iters = 0, status = DPF_OK;
    for (int k = 1; k <= nsteps; k++) {
      int code = be_step(t, s, st.get(), k, cur, nxt, o, iters);
      if (code != DPF_OK) {
        status = code;
        break;
      }
      std::swap(cur, nxt);
      if (k % st->ck == 0) copy_dev(ck + (size_t)(k / st->ck) * nf, cur, nf * sizeof(double), s);
      if (states_out) {
        double* dst = states_out + (size_t)k * nf;
        if (o->memory == DPF_DEVICE)
          copy_dev(dst, cur, nf * sizeof(double), s);
        else
          HIPTRY(hipMemcpyAsync(dst, cur, nf * sizeof(double), hipMemcpyDeviceToHost, s));
      }
    }
    if (o->memory == DPF_DEVICE)
      copy_dev(u_final, cur, nf * sizeof(double), s);
    else
      HIPTRY(hipMemcpyAsync(u_final, cur, nf * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = iters;
      stats->outer = nsteps;
      stats->converged = status == DPF_OK;
      stats->levels = t->levels();
      stats->relres = 0;
      stats->setup_ms = t1 - t0;
      stats->solve_ms = now_ms() - t1;
      stats->total_ms = now_ms() - t0;
    }
    if (status != DPF_OK) return status;
    if (state_out) *state_out = st.release();
    return (int)DPF_OK;
  });
}

}  // namespace

extern "C" int dpf_transient_forward(dpf_topology* t, const double* Ks, const double* dirichleths,
                                     const double* Qs, const double* Ss, const double* V,
                                     const double* u0, int32_t nsteps, double dt,
                                     const double* qscale, int32_t checkpoint_every,
                                     double* u_final, double* states_out, const dpf_opts* opts,
                                     dpf_stats* stats, dpf_state** state_out) {
  return transient_forward_impl(t, Ks, dirichleths, Qs, Ss, V, u0, nsteps, nullptr, dt, qscale,
                                checkpoint_every, u_final, states_out, opts, stats, state_out);
}

extern "C" int dpf_transient_forward_dts(dpf_topology* t, const double* Ks,
                                         const double* dirichleths, const double* Qs,
                                         const double* Ss, const double* V, const double* u0,
                                         int32_t nsteps, const double* dts, const double* qscale,
                                         int32_t checkpoint_every, double* u_final,
                                         double* states_out, const dpf_opts* opts,
                                         dpf_stats* stats, dpf_state** state_out) {
  if (!dts) {
    if (state_out) *state_out = nullptr;
    set_err("dts is NULL");
    return DPF_EINVAL;
  }
  return transient_forward_impl(t, Ks, dirichleths, Qs, Ss, V, u0, nsteps, dts, 0.0, qscale,
                                checkpoint_every, u_final, states_out, opts, stats, state_out);
}

extern "C" int dpf_transient_pullback_seeds(dpf_state* st, int32_t nseed,
                                            const int32_t* seed_steps, const double* seeds,
                                            double* dKs, double* ddirichleths, double* dQs,
                                            double* du0, const dpf_opts* opts,
                                            dpf_stats* stats) {
  if (!st || !st->topo || st->kind != 1 || nseed < 0 || (nseed > 0 && (!seed_steps || !seeds))) {
    set_err("invalid argument (state must come from dpf_transient_forward)");
    return DPF_EINVAL;
  }
  for (int i = 0; i < nseed; i++)
    if (seed_steps[i] < 0 || seed_steps[i] > st->nsteps || (i > 0 && seed_steps[i] <= seed_steps[i - 1])) {
      set_err("seed_steps must be strictly ascending step indices in 0..nsteps");
      return DPF_EINVAL;
    }
  dpf_topology* t = st->topo;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    BufPool* bp = t->pool.get();
    // the seeds on the device, and per step the seed row (null: no seed)
    DBuf seedb(bp);
    const double* sd = seeds;
    if (nseed > 0 && o->memory == DPF_HOST) {
      seedb.alloc((size_t)nseed * nf * sizeof(double));
      HIPTRY(hipMemcpyAsync(seedb.p, seeds, (size_t)nseed * nf * sizeof(double),
                            hipMemcpyHostToDevice, s));
      sd = seedb.as<double>();
    }
    int N = st->nsteps, C = st->ck;
    std::vector<const double*> gk(N + 1, nullptr);
    for (int i = 0; i < nseed; i++) gk[seed_steps[i]] = sd + (size_t)i * nf;
    int smax = nseed > 0 ? seed_steps[nseed - 1] : 0;  // no state after it carries a seed
    // outputs (device), and the two accumulated multiplier sums:
    // macc = sum_n dt_n mu_n (-> ddirichleths), qacc = sum_n dt_n q_n mu_n (-> dQs)
    DBuf bk(bp), bd(bp), bq(bp), bu(bp), macc(bp), qacc(bp);
    auto out = [&](DBuf& b, double* user, int64_t cnt) -> double* {
      if (!user) return nullptr;
      if (o->memory == DPF_DEVICE) return user;
      b.alloc(cnt * sizeof(double));
      return b.as<double>();
    };
    double* dK = out(bk, dKs, ne);
    double* dd = out(bd, ddirichleths, n);
    double* dQ = out(bq, dQs, n);
    double* du = out(bu, du0, nf);
    if (dd) macc.alloc(nf * sizeof(double));
    if (dQ) qacc.alloc(nf * sizeof(double));
    DBuf seg(bp), mua(bp), mub(bp), rhsb(bp);
    seg.alloc((size_t)(C + 1) * nf * sizeof(double));
    mua.alloc(nf * sizeof(double));
    mub.alloc(nf * sizeof(double));
    rhsb.alloc(nf * sizeof(double));
    double* segp = seg.as<double>();
    double* mu_next = mua.as<double>();
    double* mu = mub.as<double>();
    double* rhs = rhsb.as<double>();
    bool have_next = false, first_acc = true;
    int iters = 0, status = DPF_OK;
    for (int c = smax >= 1 ? (smax - 1) / C : -1; c >= 0 && status == DPF_OK; c--) {
      int a = c * C, b = std::min(smax, (c + 1) * C);
      // u_{a..b} into the segment buffer: u_a and, when b closes a full
      // segment, u_b from the checkpoints; the states in between recomputed
      // (bitwise what the forward produced: same steps, operators and warm
      // starts)
      copy_dev(segp, st->ckpt.as<double>() + (size_t)c * nf, nf * sizeof(double), s);
      bool b_stored = b == (c + 1) * C;
      for (int k = a + 1; k < (b_stored ? b : b + 1) && status == DPF_OK; k++)
        status = be_step(t, s, st, k, segp + (size_t)(k - a - 1) * nf,
                         segp + (size_t)(k - a) * nf, o, iters);
      if (b_stored)
        copy_dev(segp + (size_t)(b - a) * nf, st->ckpt.as<double>() + (size_t)(c + 1) * nf,
                 nf * sizeof(double), s);
      for (int k = b; k > a && status == DPF_OK; k--) {
        double dt = st->dts[k - 1];
        ensure_be(t, s, o->precision, st, dt);
        // (S + dt_k A) mu_k = dJ/du_k + S mu_{k+1}
        hipLaunchKernelGGL(k_be_adj_rhs, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                           (const double*)st->svec.as<double>(),
                           have_next ? (const double*)mu_next : (const double*)nullptr, gk[k],
                           rhs);
        dpf_stats ss;
        std::memset(&ss, 0, sizeof(ss));
        status = solve_system(t, s, rhs, mu, o, &ss, have_next ? mu_next : nullptr);
        iters += ss.iterations;
        if (status != DPF_OK) {
          step_failure("adjoint backward-Euler step", k, status, ss);
          break;
        }
        const double* uk = segp + (size_t)(k - a) * nf;
        if (dK)
          hipLaunchKernelGGL(k_edge_grad, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                             t->edges.as<int2>(), t->n2f.as<int32_t>(), t->aol.as<double>(), uk,
                             (const double*)st->dh.as<double>(), (const double*)mu, dt,
                             first_acc ? 0 : 1, dK);
        if (dd || dQ)
          hipLaunchKernelGGL(k_adj_acc, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                             (const double*)mu, dt, dt * st->qs[k], first_acc ? 0 : 1,
                             dd ? macc.as<double>() : (double*)nullptr,
                             dQ ? qacc.as<double>() : (double*)nullptr);
        first_acc = false;
        std::swap(mu, mu_next);
        have_next = true;
      }
    }
    HIPTRY(hipGetLastError());
    if (status == DPF_OK) {
      if (first_acc) {  // no step carried a multiplier: every sum is zero
        if (dK) hipLaunchKernelGGL(k_zero_f64, dim3(nblocks(ne)), dim3(BLK), 0, s, ne, dK);
        if (dd) hipLaunchKernelGGL(k_zero_f64, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, macc.as<double>());
        if (dQ) hipLaunchKernelGGL(k_zero_f64, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, qacc.as<double>());
      }
      // dirichleths: sum over the free neighbours of K_e aol_e macc (0 on free
      // cells); Qs: qacc scattered to the free cells (0 on Dirichlet cells)
      if (dd)
        hipLaunchKernelGGL(k_ddh_dq, dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                           t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                           t->inc_other.as<int32_t>(), (const double*)st->Ke.as<double>(),
                           t->aol.as<double>(), (const double*)macc.as<double>(),
                           (const double*)nullptr, dd, (double*)nullptr);
      if (dQ)
        hipLaunchKernelGGL(k_ddh_dq, dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                           t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                           t->inc_other.as<int32_t>(), (const double*)st->Ke.as<double>(),
                           t->aol.as<double>(), (const double*)qacc.as<double>(),
                           (const double*)nullptr, (double*)nullptr, dQ);
      // u_0: dJ/du_0 + S mu_1
      if (du)
        hipLaunchKernelGGL(k_be_adj_rhs, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                           (const double*)st->svec.as<double>(),
                           have_next ? (const double*)mu_next : (const double*)nullptr, gk[0],
                           du);
      HIPTRY(hipGetLastError());
      if (o->memory == DPF_HOST) {
        if (dKs) HIPTRY(hipMemcpyAsync(dKs, dK, ne * sizeof(double), hipMemcpyDeviceToHost, s));
        if (ddirichleths)
          HIPTRY(hipMemcpyAsync(ddirichleths, dd, n * sizeof(double), hipMemcpyDeviceToHost, s));
        if (dQs) HIPTRY(hipMemcpyAsync(dQs, dQ, n * sizeof(double), hipMemcpyDeviceToHost, s));
        if (du0) HIPTRY(hipMemcpyAsync(du0, du, nf * sizeof(double), hipMemcpyDeviceToHost, s));
      }
    }
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = iters;
      stats->outer = smax;
      stats->converged = status == DPF_OK;
      stats->levels = t->levels();
      stats->relres = 0;
      stats->setup_ms = 0;
      stats->solve_ms = now_ms() - t0;
      stats->total_ms = now_ms() - t0;
    }
    return status;
  });
}

extern "C" int dpf_transient_pullback(dpf_state* st, const double* dJ_du_final, double* dKs,
                                      double* dQs, const dpf_opts* opts, dpf_stats* stats) {
  if (!st || st->kind != 1 || !dJ_du_final || !dKs) {
    set_err("invalid argument (state must come from dpf_transient_forward)");
    return DPF_EINVAL;
  }
  int32_t step = st->nsteps;
  return dpf_transient_pullback_seeds(st, 1, &step, dJ_du_final, dKs, nullptr, dQs, nullptr,
                                      opts, stats);
}

// ====================================================================== two-phase (IMPES)

namespace {

Fluid make_fluid(const double* fl) {
  Fluid f;
  f.vw = fl[0];
  f.vo = fl[1];
  f.swc = fl[2];
  f.sor = fl[3];
  f.span = (1.0 - f.swc) - f.sor;
  return f;
}

int twophase_checks(const dpf_topology* t) {
  if (t->nm > 1) {
    set_err("the two-phase calls take a single-system topology");
    return DPF_EINVAL;
  }
  return DPF_OK;
}

// With t->Ke = K_a + K_b, t->dh and t->Q in place: pressure P (device, n) of
// the transmissivity system, shifted to mean zero if asked, and fluxes Vn
// (device, ne).
// warm (nullable, nf): an initial guess for the free pressures when
// have_warm (the previous IMPES step's), overwritten with this step's.
int tpfa_core(dpf_topology* t, hipStream_t s, const dpf_opts* o, dpf_stats* st, double* P,
              double* Vn, int zero_mean, double* warm = nullptr, bool have_warm = false) {
  int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
  prepare(t, s, o->precision, t->Q.as<double>(), true);
  DBuf hf(t->pool.get());
  hf.alloc(nf * sizeof(double));
  int status = solve_system(t, s, t->b64.as<double>(), hf.as<double>(), o, st,
                            have_warm ? (const double*)warm : nullptr);
  if (warm) copy_dev(warm, hf.p, nf * sizeof(double), s);
  hipLaunchKernelGGL(k_scatter_h, dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                     (const double*)hf.as<double>(), t->dh.as<double>(), P);
  if (zero_mean) {
    unsigned nb = nblocks(n);
    DBuf part(t->pool.get());
    part.alloc((nb + 1) * sizeof(double));
    double* pp = part.as<double>();
    hipLaunchKernelGGL(k_tp_blocksum, dim3(nb), dim3(BLK), 0, s, n, (const double*)P, pp);
    hipLaunchKernelGGL(k_tp_mean, dim3(1), dim3(64), 0, s, n, (int64_t)nb, (const double*)pp,
                       pp + nb);
    hipLaunchKernelGGL(k_tp_sub, dim3(nb), dim3(BLK), 0, s, n, P, (const double*)(pp + nb));
  }
  hipLaunchKernelGGL(k_tp_flux, dim3(nblocks(ne)), dim3(BLK), 0, s, ne, t->edges.as<int2>(),
                     t->aol.as<double>(), (const double*)t->Ke.as<double>(), (const double*)P,
                     Vn);
  HIPTRY(hipGetLastError());
  if (status == DPF_ENOCONV) set_err("AMG-PCG did not reach the tolerance within maxiter");
  if (status == DPF_EBREAKDOWN) set_err("AMG-PCG breakdown (non-SPD operator or NaN input)");
  return status;
}

// Upstream transport over time T of S (device, n, in place), as
// src/twophase.jl:208-235.  Returns the number of CFL sub-steps.
int64_t upstream_core(dpf_topology* t, hipStream_t s, double* S, const Fluid& f,
                      const double* Q, double T, const double* P, const double* Vn,
                      const double* pv) {
  int64_t n = t->H.n;
  unsigned nb = nblocks(n);
  DBuf part(t->pool.get());
  part.alloc(nb * sizeof(double));
  hipLaunchKernelGGL(k_tp_cfl, dim3(nb), dim3(BLK), 0, s, n, t->edges.as<int2>(),
                     t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(), Vn, Q, pv,
                     part.as<double>());
  HIPTRY(hipGetLastError());
  std::vector<double> hp(nb);
  HIPTRY(hipMemcpyAsync(hp.data(), part.p, nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  // minimum over the blocks, NaN-propagating like Julia's minimum
  double pm = INFINITY;
  for (double v : hp) pm = (v != v || pm != pm) ? v + pm : std::min(pm, v);
  double cfl = (f.span / 3.0) * pm;
  double steps = std::ceil(T / cfl);
  if (!(steps == steps) || std::fabs(steps) > 1e9)
    throw Err{DPF_EINVAL, "CFL sub-step count T/cfl = " + std::to_string(T / cfl) +
                              " is not finite or exceeds 1e9 (no inflow, zero pore volume or "
                              "non-finite fluxes?)"};
  int64_t nts = (int64_t)steps;
  if (nts <= 0) return 0;
  double tau = T / (double)nts;
  // upwind-signed flux weights per incidence entry, and f at S
  int64_t ninc = t->H.inc_ptr.back();
  DBuf wb(t->pool.get()), other(t->pool.get()), f0(t->pool.get()), f1(t->pool.get());
  wb.alloc(std::max<int64_t>(ninc, 1) * sizeof(double));
  other.alloc(n * sizeof(double));
  f0.alloc(n * sizeof(double));
  f1.alloc(n * sizeof(double));
  hipLaunchKernelGGL(k_tp_weights, dim3(nb), dim3(BLK), 0, s, n, t->inc_ptr.as<int64_t>(),
                     t->inc_edge.as<int32_t>(), t->inc_other.as<int32_t>(), P, Vn,
                     wb.as<double>());
  hipLaunchKernelGGL(k_tp_frac, dim3(nb), dim3(BLK), 0, s, n, (const double*)S, f,
                     f0.as<double>());
  double* cur = S;
  double* oth = other.as<double>();
  double* fc = f0.as<double>();
  double* fo = f1.as<double>();
  double lo = f.swc, hi = 1.0 - f.sor;
  for (int64_t k = 0; k < nts; k++) {
    hipLaunchKernelGGL(k_tp_upstream, dim3(nb), dim3(BLK), 0, s, n, t->inc_ptr.as<int64_t>(),
                       t->inc_other.as<int32_t>(), (const double*)wb.as<double>(),
                       (const double*)cur, (const double*)fc, oth, fo, Q, pv, tau, f, lo, hi);
    std::swap(cur, oth);
    std::swap(fc, fo);
  }
  HIPTRY(hipGetLastError());
  if (cur != S) copy_dev(S, cur, n * sizeof(double), s);
  return nts;
}

// Host or device input -> device pointer (staged through buf for DPF_HOST).
const double* dev_in(dpf_topology* t, const dpf_opts* o, DBuf& buf, const double* p,
                     int64_t count, hipStream_t s) {
  return stage_in(t, o, buf, p, count, s);
}

}  // namespace

extern "C" {

int dpf_twophase_tpfa(dpf_topology* t, const double* K, const double* dirichleths,
                      const double* Qs, double* P, double* Vn, int32_t zero_mean,
                      const dpf_opts* opts, dpf_stats* stats) {
  if (!t || !K || !dirichleths || !Qs || !P || !Vn) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne;
    DBuf kb(t->pool.get()), pb(t->pool.get()), vb(t->pool.get());
    const double* Kd = dev_in(t, o, kb, K, n, s);
    hipLaunchKernelGGL(k_tp_ksum, dim3(nblocks(ne)), dim3(BLK), 0, s, ne, t->edges.as<int2>(),
                       Kd, t->Ke.as<double>());
    if (o->memory == DPF_DEVICE) {
      copy_dev(t->dh.p, dirichleths, n * sizeof(double), s);
      copy_dev(t->Q.p, Qs, n * sizeof(double), s);
    } else {
      HIPTRY(hipMemcpyAsync(t->dh.p, dirichleths, n * sizeof(double), hipMemcpyHostToDevice, s));
      HIPTRY(hipMemcpyAsync(t->Q.p, Qs, n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    double* Pd = P;
    double* Vd = Vn;
    if (o->memory == DPF_HOST) {
      pb.alloc(n * sizeof(double));
      vb.alloc(ne * sizeof(double));
      Pd = pb.as<double>();
      Vd = vb.as<double>();
    }
    int status = tpfa_core(t, s, o, stats, Pd, Vd, zero_mean);
    if (o->memory == DPF_HOST) {
      HIPTRY(hipMemcpyAsync(P, Pd, n * sizeof(double), hipMemcpyDeviceToHost, s));
      HIPTRY(hipMemcpyAsync(Vn, Vd, ne * sizeof(double), hipMemcpyDeviceToHost, s));
    }
    HIPTRY(hipStreamSynchronize(s));
    if (stats) stats->total_ms = now_ms() - t0;
    return status;
  });
}

int dpf_twophase_upstream(dpf_topology* t, double* S, const double* fluid, const double* Qs,
                          double T, const double* P, const double* Vn, const double* volumes,
                          int64_t* nsub, const dpf_opts* opts) {
  if (nsub) *nsub = 0;
  if (!t || !S || !fluid || !Qs || !P || !Vn || !volumes) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne;
    DBuf sb(t->pool.get()), qb(t->pool.get()), pb(t->pool.get()), vb(t->pool.get()),
        wb(t->pool.get());
    double* Sd = S;
    if (o->memory == DPF_HOST) {
      sb.alloc(n * sizeof(double));
      Sd = sb.as<double>();
      HIPTRY(hipMemcpyAsync(Sd, S, n * sizeof(double), hipMemcpyHostToDevice, s));
    }
    const double* Qd = dev_in(t, o, qb, Qs, n, s);
    const double* Pd = dev_in(t, o, pb, P, n, s);
    const double* Vd = dev_in(t, o, vb, Vn, ne, s);
    const double* Wd = dev_in(t, o, wb, volumes, n, s);
    int64_t k = upstream_core(t, s, Sd, make_fluid(fluid), Qd, T, Pd, Vd, Wd);
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(S, Sd, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (nsub) *nsub = k;
    return DPF_OK;
  });
}

int dpf_solvetwophase(dpf_topology* t, const double* h0, const double* S0, const double* K,
                      const double* dirichleths, const double* Qs, const double* volumes,
                      const double* fluid, double dt, int32_t nt, int32_t zero_mean,
                      double* P_out, double* S_out, double* P_all, double* S_all,
                      int64_t* nsub_total, dpf_state** state_out, const dpf_opts* opts,
                      dpf_stats* stats) {
  if (nsub_total) *nsub_total = 0;
  if (state_out) *state_out = nullptr;
  if (!t || !h0 || !S0 || !K || !dirichleths || !Qs || !volumes || !fluid || !P_out ||
      !S_out || nt < 0) {
    set_err(nt < 0 ? "nt must be >= 0" : "NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne;
    bool host = o->memory == DPF_HOST;
    auto out = [&](double* dst, const double* src, int64_t count) {
      if (host)
        HIPTRY(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToHost, s));
      else
        copy_dev(dst, src, count * sizeof(double), s);
    };
    DBuf kb(t->pool.get()), qb(t->pool.get()), wb(t->pool.get()), sb(t->pool.get()),
        pb(t->pool.get()), vb(t->pool.get()), mb(t->pool.get());
    const double* Kd = dev_in(t, o, kb, K, n, s);
    const double* Qd = dev_in(t, o, qb, Qs, n, s);
    const double* Wd = dev_in(t, o, wb, volumes, n, s);
    sb.alloc(n * sizeof(double));
    pb.alloc(n * sizeof(double));
    vb.alloc(ne * sizeof(double));
    mb.alloc(n * sizeof(double));
    double* Sd = sb.as<double>();
    double* Pd = pb.as<double>();
    double* Vd = vb.as<double>();
    HIPTRY(hipMemcpyAsync(Sd, S0, n * sizeof(double),
                          host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(Pd, h0, n * sizeof(double),
                          host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    HIPTRY(hipMemcpyAsync(t->dh.p, dirichleths, n * sizeof(double),
                          host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    copy_dev(t->Q.p, Qd, n * sizeof(double), s);
    Fluid f = make_fluid(fluid);
    int64_t subs = 0;
    int status = DPF_OK;
    int iters = 0;
    std::unique_ptr<dpf_state> st;
    if (state_out) {
      st = std::make_unique<dpf_state>(t);
      st->kind = 2;
      st->precision = o->precision;
      st->nsteps = nt;
      st->dt = dt;
      st->tp_zero_mean = zero_mean;
      std::memcpy(st->tp_fluid, fluid, sizeof(st->tp_fluid));
      size_t nn = (size_t)std::max<int64_t>((int64_t)nt * n, 1) * sizeof(double);
      size_t ee = (size_t)std::max<int64_t>((int64_t)nt * ne, 1) * sizeof(double);
      st->tp_S.alloc(nn);
      st->tp_P.alloc(nn);
      st->tp_Ksum.alloc(ee);
      st->tp_Vn.alloc(ee);
      st->tp_K.alloc(n * sizeof(double));
      st->tp_pv.alloc(n * sizeof(double));
      st->Q.alloc(n * sizeof(double));
      st->dh.alloc(n * sizeof(double));
      copy_dev(st->tp_K.p, Kd, n * sizeof(double), s);
      copy_dev(st->tp_pv.p, Wd, n * sizeof(double), s);
      copy_dev(st->Q.p, Qd, n * sizeof(double), s);
      copy_dev(st->dh.p, t->dh.p, n * sizeof(double), s);
    }
    // each pressure solve starts from the previous step's pressures
    DBuf hwarm(t->pool.get());
    hwarm.alloc(t->H.nf * sizeof(double));
    for (int32_t step = 0; step < nt && status == DPF_OK; step++) {
      // total mobility times K, arithmetic mean per edge, pressure and fluxes,
      // then the saturation transport over dt
      hipLaunchKernelGGL(k_tp_kmob, dim3(nblocks(n)), dim3(BLK), 0, s, n, (const double*)Sd,
                         Kd, f, mb.as<double>());
      hipLaunchKernelGGL(k_tp_ksum, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                         t->edges.as<int2>(), (const double*)mb.as<double>(),
                         t->Ke.as<double>());
      status = tpfa_core(t, s, o, stats, Pd, Vd, zero_mean, hwarm.as<double>(), step > 0);
      if (stats) iters += stats->iterations;
      if (status != DPF_OK) break;
      if (st) {
        copy_dev(st->tp_S.as<double>() + (int64_t)step * n, Sd, n * sizeof(double), s);
        copy_dev(st->tp_P.as<double>() + (int64_t)step * n, Pd, n * sizeof(double), s);
        copy_dev(st->tp_Ksum.as<double>() + (int64_t)step * ne, t->Ke.p, ne * sizeof(double), s);
        copy_dev(st->tp_Vn.as<double>() + (int64_t)step * ne, Vd, ne * sizeof(double), s);
      }
      int64_t k = upstream_core(t, s, Sd, f, Qd, dt, Pd, Vd, Wd);
      if (st) st->tp_nts.push_back(k);
      subs += k;
      if (P_all) out(P_all + (int64_t)step * n, Pd, n);
      if (S_all) out(S_all + (int64_t)step * n, Sd, n);
    }
    out(P_out, Pd, n);
    out(S_out, Sd, n);
    HIPTRY(hipStreamSynchronize(s));
    if (nsub_total) *nsub_total = subs;
    if (stats) {
      stats->iterations = iters;
      stats->total_ms = now_ms() - t0;
    }
    if (st && status == DPF_OK) {
      st->version = t->version;
      *state_out = st.release();
    }
    return status;
  });
}

}  // extern "C"

namespace {

// Per incidence entry the same edge's entry in the other cell's list, and per
// edge its entries in its first and second cell's lists.
void tp_build_maps(dpf_topology* t) {
  if (t->tp_maps) return;
  const TopologyHost& H = t->H;
  int64_t n = H.n, ne = H.ne;
  int64_t ninc = H.inc_ptr[n];
  std::vector<int32_t> mirror(std::max<int64_t>(ninc, 1), 0);
  std::vector<int32_t> epos(2 * std::max<int64_t>(ne, 1), 0);
  for (int64_t c = 0; c < n; c++)
    for (int64_t p = H.inc_ptr[c]; p < H.inc_ptr[c + 1]; p++) {
      int32_t e = H.inc_edge[p];
      epos[2 * e + (H.edges[2 * e] == c ? 0 : 1)] = (int32_t)p;
    }
  for (int64_t c = 0; c < n; c++)
    for (int64_t p = H.inc_ptr[c]; p < H.inc_ptr[c + 1]; p++) {
      int32_t e = H.inc_edge[p];
      mirror[p] = H.edges[2 * e] == c ? epos[2 * e + 1] : epos[2 * e];
    }
  upload(t->tp_mirror, mirror);
  upload(t->tp_epos, epos);
  t->tp_maps = true;
}

}  // namespace

extern "C" int dpf_solvetwophase_pullback(dpf_state* st, const double* dP_final,
                                          const double* dS_final, double* dK_out,
                                          double* dS0_out, const dpf_opts* opts,
                                          dpf_stats* stats) {
  if (!st || !st->topo || !dK_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (st->kind != 2) {
    set_err("the state does not come from dpf_solvetwophase");
    return DPF_EINVAL;
  }
  dpf_topology* t = st->topo;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    int64_t ninc = t->H.inc_ptr.back();
    int nt = st->nsteps;
    bool host = o->memory == DPF_HOST;
    tp_build_maps(t);
    Fluid f = make_fluid(st->tp_fluid);
    double lo = f.swc, hi = 1.0 - f.sor;
    int64_t maxnts = 0;
    for (int64_t k : st->tp_nts) maxnts = std::max(maxnts, k);
    BufPool* pl = t->pool.get();
    DBuf dS(pl), dK(pl), seedb(pl), w(pl), dabs(pl), du(pl), dR(pl), dR2(pl), dVn(pl), dKsum(pl),
        dP(pl), rhs(pl), lam(pl), hf(pl), part(pl), tS(pl), tf(pl);
    unsigned nb = nblocks(n);
    dS.alloc(n * sizeof(double));
    dK.alloc(n * sizeof(double));
    w.alloc(std::max<int64_t>(ninc, 1) * sizeof(double));
    dabs.alloc(std::max<int64_t>(ninc, 1) * sizeof(double));
    du.alloc(n * sizeof(double));
    dR.alloc(n * sizeof(double));
    dR2.alloc(n * sizeof(double));
    dVn.alloc(ne * sizeof(double));
    dKsum.alloc(ne * sizeof(double));
    dP.alloc(n * sizeof(double));
    rhs.alloc(nf * sizeof(double));
    lam.alloc(nf * sizeof(double));
    hf.alloc(nf * sizeof(double));
    part.alloc((nb + 1) * sizeof(double));
    tS.alloc(std::max<int64_t>(maxnts, 1) * n * sizeof(double));
    tf.alloc(std::max<int64_t>(maxnts, 1) * n * sizeof(double));
    if (dS_final)
      HIPTRY(hipMemcpyAsync(dS.p, dS_final, n * sizeof(double),
                            host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    else
      HIPTRY(hipMemsetAsync(dS.p, 0, n * sizeof(double), s));
    HIPTRY(hipMemsetAsync(dK.p, 0, n * sizeof(double), s));
    const double* seedP = dP_final ? stage_in(t, o, seedb, dP_final, n, s) : nullptr;
    const int64_t* ip = t->inc_ptr.as<int64_t>();
    const int32_t* ie = t->inc_edge.as<int32_t>();
    const int32_t* io = t->inc_other.as<int32_t>();
    const double* Qd = st->Q.as<double>();
    const double* pv = st->tp_pv.as<double>();
    int status = DPF_OK, iters = 0;
    for (int step = nt - 1; step >= 0 && status == DPF_OK; step--) {
      const double* Sp = st->tp_S.as<double>() + (int64_t)step * n;
      const double* Pt = st->tp_P.as<double>() + (int64_t)step * n;
      const double* Vt = st->tp_Vn.as<double>() + (int64_t)step * ne;
      const double* Kt = st->tp_Ksum.as<double>() + (int64_t)step * ne;
      int64_t nts = st->tp_nts[step];
      HIPTRY(hipMemsetAsync(dabs.p, 0, std::max<int64_t>(ninc, 1) * sizeof(double), s));
      if (nts > 0) {
        double tau = st->dt / (double)nts;
        hipLaunchKernelGGL(k_tp_weights, dim3(nb), dim3(BLK), 0, s, n, ip, ie, io, Pt, Vt,
                           w.as<double>());
        // the sub-steps' input states and fractional flows, recomputed
        double* S0k = tS.as<double>();
        double* f0k = tf.as<double>();
        copy_dev(S0k, Sp, n * sizeof(double), s);
        hipLaunchKernelGGL(k_tp_frac, dim3(nb), dim3(BLK), 0, s, n, (const double*)S0k, f,
                           f0k);
        for (int64_t k = 0; k + 1 < nts; k++)
          hipLaunchKernelGGL(k_tp_upstream, dim3(nb), dim3(BLK), 0, s, n, ip, io,
                             (const double*)w.as<double>(), (const double*)(S0k + k * n),
                             (const double*)(f0k + k * n), S0k + (k + 1) * n, f0k + (k + 1) * n,
                             Qd, pv, tau, f, lo, hi);
        // launch j: part 2 of sub-step j (j < nts) and part 1 of sub-step
        // j - 1 (j > 0), for j = nts .. 0; dR ping-pongs between two buffers
        double* dRs[2] = {dR.as<double>(), dR2.as<double>()};
        int cur = 0;  // buffer holding dR of the sub-step part 2 reads
        for (int64_t j = nts; j >= 0; j--) {
          int do_b = j < nts, do_a = j > 0;
          const double* Sb = S0k + (do_b ? j : 0) * n;
          const double* Sa = S0k + (do_a ? j - 1 : 0) * n;
          const double* fa = f0k + (do_a ? j - 1 : 0) * n;
          hipLaunchKernelGGL(k_tp_adj, dim3(nb), dim3(BLK), 0, s, n, ip, io,
                             t->tp_mirror.as<int32_t>(), (const double*)w.as<double>(), Qd, pv, f,
                             tau, lo, hi, do_b, do_a, Sb, (const double*)dRs[cur], Sa, fa,
                             du.as<double>(), dRs[cur ^ 1], dabs.as<double>(), dS.as<double>());
          cur ^= 1;
        }
      }
      hipLaunchKernelGGL(k_tp_flux_adj, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                         t->edges.as<int2>(), t->tp_epos.as<int2>(), t->aol.as<double>(), Pt, Vt,
                         (const double*)dabs.as<double>(), dVn.as<double>(), dKsum.as<double>());
      hipLaunchKernelGGL(k_tp_dP, dim3(nb), dim3(BLK), 0, s, n, t->edges.as<int2>(), ip, ie,
                         t->aol.as<double>(), Kt, (const double*)dVn.as<double>(),
                         step == nt - 1 ? seedP : (const double*)nullptr, dP.as<double>());
      if (st->tp_zero_mean) {
        double* pp = part.as<double>();
        hipLaunchKernelGGL(k_tp_blocksum, dim3(nb), dim3(BLK), 0, s, n,
                           (const double*)dP.as<double>(), pp);
        hipLaunchKernelGGL(k_tp_mean, dim3(1), dim3(64), 0, s, n, (int64_t)nb,
                           (const double*)pp, pp + nb);
        hipLaunchKernelGGL(k_tp_sub, dim3(nb), dim3(BLK), 0, s, n, dP.as<double>(),
                           (const double*)(pp + nb));
      }
      // the pressure solve's adjoint at this step's operator
      copy_dev(t->Ke.p, Kt, ne * sizeof(double), s);
      copy_dev(t->dh.p, st->dh.p, n * sizeof(double), s);
      prepare(t, s, o->precision, nullptr, false);
      hipLaunchKernelGGL(k_gather_free, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                         t->f2n.as<int32_t>(), (const double*)dP.as<double>(), rhs.as<double>());
      // warm-started from the later step's multiplier (lam holds it)
      status = solve_system(t, s, rhs.as<double>(), lam.as<double>(), o, stats,
                            step < nt - 1 ? (const double*)lam.as<double>() : nullptr);
      if (stats) iters += stats->iterations;
      hipLaunchKernelGGL(k_gather_free, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                         t->f2n.as<int32_t>(), Pt, hf.as<double>());
      hipLaunchKernelGGL(k_edge_grad, dim3(nblocks(ne)), dim3(BLK), 0, s, ne, t->edges.as<int2>(),
                         t->n2f.as<int32_t>(), t->aol.as<double>(), (const double*)hf.as<double>(),
                         Pt, (const double*)lam.as<double>(), 1.0, 1, dKsum.as<double>());
      hipLaunchKernelGGL(k_tp_kmob_adj, dim3(nb), dim3(BLK), 0, s, n, ip, ie,
                         (const double*)dKsum.as<double>(), Sp,
                         (const double*)st->tp_K.as<double>(), f, dK.as<double>(),
                         dS.as<double>());
      HIPTRY(hipGetLastError());
    }
    auto out = [&](double* dst, const DBuf& src) {
      HIPTRY(hipMemcpyAsync(dst, src.p, n * sizeof(double),
                            host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
    };
    out(dK_out, dK);
    if (dS0_out) out(dS0_out, dS);
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      stats->iterations = iters;
      stats->total_ms = now_ms() - t0;
    }
    if (status == DPF_ENOCONV) set_err("adjoint AMG-PCG did not reach the tolerance within maxiter");
    if (status == DPF_EBREAKDOWN) set_err("adjoint AMG-PCG breakdown");
    return status;
  });
}

// ====================================================================== steady Richards

namespace {

// z = V-cycle(b) on the FP64 hierarchy (the Picard operator's), Chebyshev
// degrees 1 on the finest level and 2 below, interval ratio 15.
void rich_precond(dpf_topology* t, hipStream_t s, const std::vector<Cheb>& chs, const double* b,
                  double* z) {
  int64_t nf = t->H.nf;
  LevelNum<double>& n0 = t->h64.lev[0];
  double* x0 = n0.x.as<double>();
  hipLaunchKernelGGL(k_rich_pre, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, b,
                     (const double*)level_invd(t, t->h64, 0), chs[0].c0, n0.d.as<double>(), x0);
  double* out = nullptr;
  vcycle<double>(t, t->h64, s, 0, b, x0, n0.xs.as<double>(), &out, chs, t->zero_flag.as<int>(),
                 false);
  copy_dev(z, out, nf * sizeof(double), s);
}

// The same V-cycle on the FP32 hierarchy (b rounded to FP32, z widened back).
void rich_precond32(dpf_topology* t, hipStream_t s, const std::vector<Cheb>& chs,
                    const double* b, double* z) {
  int64_t nf = t->H.nf;
  LevelNum<float>& n0 = t->h32.lev[0];
  float* b32 = t->h32.cr.as<float>();
  float* x0 = n0.x.as<float>();
  hipLaunchKernelGGL(k_rich_pre32, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, b,
                     (const float*)level_invd(t, t->h32, 0), (float)chs[0].c0, b32,
                     n0.d.as<float>(), x0);
  float* out = nullptr;
  vcycle<float>(t, t->h32, s, 0, (const float*)b32, x0, n0.xs.as<float>(), &out, chs,
                t->zero_flag.as<int>(), false);
  hipLaunchKernelGGL(k_widen, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, (const float*)out, z);
}

// Up to three dot products x_k . y_k of length n in one pass, finished on the
// device in a fixed order; the finishing kernel stores the three sums in the
// topology's pinned host buffer (one synchronisation, no copy).  Calls on one
// topology are serialised by its mutex, so the buffer is not shared.
struct Dots3 {
  DBuf part;
  double h[3];
  unsigned grid;
  double *host, *dev;
  Dots3(dpf_topology* t, int64_t n) : part(t->pool.get()), host(t->dots_host), dev(t->dots_dev) {
    grid = std::max<unsigned>(1, std::min<unsigned>(nblocks(n), DOT_GRID));
    part.alloc(3 * (size_t)grid * sizeof(double));
  }
  void run(hipStream_t s, int64_t n, const double* x0, const double* y0, const double* x1 = nullptr,
           const double* y1 = nullptr, const double* x2 = nullptr, const double* y2 = nullptr) {
    hipLaunchKernelGGL(k_dots3, dim3(grid), dim3(BLK), 0, s, n, x0, y0, x1, y1, x2, y2,
                       part.as<double>());
    finish(s);
  }
  // sums the partials a kernel of `grid` blocks left in part
  void finish(hipStream_t s) {
    hipLaunchKernelGGL(k_dots3_fin, dim3(1), dim3(BLK), 0, s, (int64_t)grid,
                       (const double*)part.as<double>(), dev);
    HIPTRY(hipStreamSynchronize(s));
    for (int k = 0; k < 3; k++) h[k] = mirror_load(&host[k]);
  }
};

// Right-preconditioned BiCGStab for A x = b (apply(x, y): y = A x;
// prec(b, z): z ~ A^-1 b; or, with jac given, the Jacobi preconditioner
// z = b / jac fused into the vector updates and prec unused), vectors of
// length nf, x from 0 (guess: from the x passed in), to ||b - A x|| <= tol
// ||b||.  Three synchronisations per iteration: r^.v; s.s, t.s and t.t
// together; r.r and r^.r together (fused with the x and r updates).  Returns
// 0, DPF_ENOCONV or DPF_EBREAKDOWN; *iters counts the iterations.
template <class Apply, class Prec>
int bicgstab(dpf_topology* t, hipStream_t s, int64_t nf, Apply apply, Prec prec, const double* b,
             double* x, double tol, int maxit, int* iters, bool guess = false,
             const double* jac = nullptr) {
  BufPool* pl = t->pool.get();
  DBuf r(pl), rh(pl), p(pl), v(pl), ph(pl), sv(pl), sh(pl), tv(pl);
  for (DBuf* d : {&r, &rh, &p, &v, &ph, &sv, &sh, &tv}) d->alloc(nf * sizeof(double));
  Dots3 dots(t, nf);
  size_t bytes = nf * sizeof(double);
  dim3 g(nblocks(nf)), bl(BLK);
  if (guess) {
    // r = b - A x
    apply((const double*)x, r.as<double>());
    hipLaunchKernelGGL(k_rich_axpby, g, bl, 0, s, nf, 1.0, b, -1.0, r.as<double>(), 0.0,
                       (const double*)nullptr);
  } else {
    HIPTRY(hipMemsetAsync(x, 0, bytes, s));
    copy_dev(r.p, b, bytes, s);
  }
  copy_dev(rh.p, r.p, bytes, s);
  HIPTRY(hipMemsetAsync(p.p, 0, bytes, s));
  HIPTRY(hipMemsetAsync(v.p, 0, bytes, s));
  dots.run(s, nf, b, b, r.as<double>(), r.as<double>());
  double bn = std::sqrt(dots.h[0]);
  *iters = 0;
  if (!(bn > 0.0)) {
    if (bn == 0.0) HIPTRY(hipMemsetAsync(x, 0, bytes, s));
    return bn == 0.0 ? DPF_OK : DPF_EBREAKDOWN;
  }
  if (std::sqrt(dots.h[1]) <= tol * bn) return DPF_OK;
  double rho = 1.0, alpha = 1.0, omega = 1.0, rho1 = dots.h[1];
  for (int it = 1; it <= maxit; it++) {
    if (rho1 == 0.0 || !std::isfinite(rho1)) return DPF_EBREAKDOWN;
    double beta = (rho1 / rho) * (alpha / omega);
    // p = r + beta (p - omega v), ph = M^-1 p
    if (jac) {
      hipLaunchKernelGGL(k_bicg_jacobi, g, bl, 0, s, nf, (const double*)r.as<double>(), beta,
                         -beta * omega, (const double*)v.as<double>(), jac, p.as<double>(),
                         ph.as<double>());
    } else {
      hipLaunchKernelGGL(k_rich_axpby, g, bl, 0, s, nf, 1.0, (const double*)r.as<double>(),
                         beta, p.as<double>(), -beta * omega, (const double*)v.as<double>());
      prec((const double*)p.as<double>(), ph.as<double>());
    }
    apply((const double*)ph.as<double>(), v.as<double>());
    dots.run(s, nf, rh.as<double>(), v.as<double>());
    double rv = dots.h[0];
    if (rv == 0.0 || !std::isfinite(rv)) return DPF_EBREAKDOWN;
    alpha = rho1 / rv;
    // s = r - alpha v, then t = A M^-1 s (also at the last iteration: one
    // synchronisation fewer per iteration)
    if (jac) {
      hipLaunchKernelGGL(k_bicg_jacobi, g, bl, 0, s, nf, (const double*)r.as<double>(), 0.0,
                         -alpha, (const double*)v.as<double>(), jac, sv.as<double>(),
                         sh.as<double>());
    } else {
      hipLaunchKernelGGL(k_lin2, g, bl, 0, s, nf, -alpha, (const double*)v.as<double>(), 1.0,
                         (const double*)r.as<double>(), sv.as<double>());
      prec((const double*)sv.as<double>(), sh.as<double>());
    }
    apply((const double*)sh.as<double>(), tv.as<double>());
    dots.run(s, nf, sv.as<double>(), sv.as<double>(), tv.as<double>(), sv.as<double>(),
             tv.as<double>(), tv.as<double>());
    double ss = dots.h[0], ts = dots.h[1], tt = dots.h[2];
    *iters = it;
    if (std::sqrt(ss) <= tol * bn) {
      hipLaunchKernelGGL(k_rich_axpby, g, bl, 0, s, nf, alpha, (const double*)ph.as<double>(),
                         1.0, x, 0.0, (const double*)nullptr);
      return DPF_OK;
    }
    if (tt == 0.0 || !std::isfinite(tt)) return DPF_EBREAKDOWN;
    omega = ts / tt;
    // x += alpha ph + omega sh ; r = s - omega t ; r.r and r^.r
    hipLaunchKernelGGL(k_bicg_xr, dim3(dots.grid), bl, 0, s, nf, alpha,
                       (const double*)ph.as<double>(), omega, (const double*)sh.as<double>(),
                       (const double*)sv.as<double>(), (const double*)tv.as<double>(),
                       (const double*)rh.as<double>(), x, r.as<double>(),
                       dots.part.as<double>());
    dots.finish(s);
    if (std::sqrt(dots.h[0]) <= tol * bn) return DPF_OK;
    if (omega == 0.0) return DPF_EBREAKDOWN;
    rho = rho1;
    rho1 = dots.h[1];
  }
  return DPF_ENOCONV;
}

// BiCGStab on a Newton matrix of the fine layout; the AMG preconditioner runs
// in FP32 when fp32 (the hierarchy prepared in that precision).
int rich_bicgstab(dpf_topology* t, hipStream_t s, const std::vector<Cheb>& chs,
                  SellD<double> M, const double* b, double* x, double tol, int maxit,
                  int* iters, bool fp32) {
  int64_t nf = t->H.nf;
  auto apply = [&](const double* xx, double* yy) {
    hipLaunchKernelGGL(k_rich_spmv, dim3(nblocks(nf)), dim3(BLK), 0, s, M, xx, yy);
  };
  auto prec = [&](const double* bb, double* zz) {
    if (fp32)
      rich_precond32(t, s, chs, bb, zz);
    else
      rich_precond(t, s, chs, bb, zz);
  };
  return bicgstab(t, s, nf, apply, prec, b, x, tol, maxit, iters);
}

// The inputs of one Richards problem on the device.
struct RichIn {
  const double *K, *dpsi, *z, *al, *N, *Q;
};

// Edge fluxes F (and dFa, dFb) and the residual r at psi (free); returns
// (max |r|, ||r||_2).
std::pair<double, double> rich_residual(dpf_topology* t, hipStream_t s, const RichIn& in,
                                        const double* psi, double* F, double* dFa, double* dFb,
                                        double* r, DBuf& part) {
  int64_t ne = t->H.ne, nf = t->H.nf;
  hipLaunchKernelGGL(k_rich_edge_flux, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                     t->edges.as<int2>(), t->n2f.as<int32_t>(), psi, in.dpsi, in.z, in.al, in.N,
                     in.K, t->aol.as<double>(), F, dFa, dFb);
  unsigned nb = nblocks(nf);
  hipLaunchKernelGGL(k_rich_residual, dim3(nb), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                     t->edges.as<int2>(), t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                     (const double*)F, in.Q, (const double*)nullptr, r, part.as<double>());
  std::vector<double> hp(2 * nb);
  HIPTRY(hipMemcpyAsync(hp.data(), part.p, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  double mx = 0.0, sq = 0.0;
  for (unsigned b = 0; b < nb; b++) {
    mx = (hp[2 * b] != hp[2 * b] || hp[2 * b] > mx) ? hp[2 * b] : mx;
    sq += hp[2 * b + 1];
  }
  return {mx, std::sqrt(sq)};
}

// Newton matrix M (and MT) at psi: Picard operator (and its AMG hierarchy) for
// the preconditioner, then the kr-slope terms.
void rich_matrices(dpf_topology* t, hipStream_t s, const RichIn& in, const double* psi,
                   const double* dFa, const double* dFb, double* M, double* MT, bool fp32) {
  int64_t ne = t->H.ne, nf = t->H.nf;
  hipLaunchKernelGGL(k_rich_picard, dim3(nblocks(ne)), dim3(BLK), 0, s, ne, t->edges.as<int2>(),
                     t->n2f.as<int32_t>(), psi, in.dpsi, in.z, in.al, in.N, in.K,
                     t->Ke.as<double>());
  prepare(t, s, fp32 ? DPF_MIXED_PC : DPF_FP64, nullptr, false);
  LevelDev& d0 = *t->lev[0];
  size_t bytes = d0.A.entries * sizeof(double);
  if (M) HIPTRY(hipMemsetAsync(M, 0, bytes, s));
  if (MT) HIPTRY(hipMemsetAsync(MT, 0, bytes, s));
  hipLaunchKernelGGL(k_rich_newton_matrix, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                     t->f2n.as<int32_t>(), t->n2f.as<int32_t>(), t->edges.as<int2>(),
                     t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                     t->inc_other.as<int32_t>(), t->tp_mirror.as<int32_t>(),
                     t->inc_slot.as<uint16_t>(), d0.diag_slot.as<uint16_t>(),
                     d0.A.view<double>(M ? M : MT), dFa, dFb, (const double*)nullptr, M, MT);
  HIPTRY(hipGetLastError());
}

std::vector<Cheb> rich_chebs(dpf_topology* t) {
  std::vector<Cheb> v(t->levels());
  // degree 1 on the finest and the second level, 3 below (the BiCGStab
  // iteration of the Newton steps gains more from the deeper levels'
  // smoothing than the steady CG does: 33 + 14 against 36 + 15 iterations
  // with degree 2 below at 2M cells)
  for (int l = 0; l < t->levels(); l++) v[l].make(l <= 1 ? 1 : 3, 1.0, 15.0);
  return v;
}

}  // namespace

extern "C" int dpf_richards_steady(dpf_topology* t, const double* psi0, const double* Ks,
                                   const double* dirichletpsis, const double* z,
                                   const double* alphas, const double* Ns, const double* Qs,
                                   double* psi_out, const dpf_opts* opts, dpf_stats* stats,
                                   dpf_state** state_out) {
  if (state_out) *state_out = nullptr;
  if (!t || !psi0 || !Ks || !dirichletpsis || !z || !alphas || !Ns || !Qs || !psi_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  if (t->levels() < 2) {
    set_err("the Richards solver needs a multilevel hierarchy (a larger mesh)");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    bool fp32 = o->precision != DPF_FP64;  // FP32 V-cycle preconditioner
    t->timing = false;
    tp_build_maps(t);
    BufPool* pl = t->pool.get();
    std::unique_ptr<dpf_state> st = std::make_unique<dpf_state>(t);
    st->kind = 3;
    st->precision = DPF_FP64;
    // inputs, kept in the state for the pullback
    auto keep = [&](DBuf& d, const double* src, int64_t count) -> const double* {
      d.alloc(count * sizeof(double));
      HIPTRY(hipMemcpyAsync(d.p, src, count * sizeof(double),
                            o->memory == DPF_HOST ? hipMemcpyHostToDevice
                                                  : hipMemcpyDeviceToDevice, s));
      return d.as<double>();
    };
    RichIn in;
    in.K = keep(st->Ke, Ks, ne);
    in.dpsi = keep(st->dh, dirichletpsis, n);
    in.z = keep(st->r_z, z, n);
    in.al = keep(st->r_al, alphas, ne);
    in.N = keep(st->r_N, Ns, ne);
    in.Q = keep(st->Q, Qs, n);
    DBuf p0(pl), F(pl), dFa(pl), dFb(pl), r(pl), rn(pl), part(pl), dx(pl), xn(pl), Mv(pl),
        out(pl);
    const double* P0 = stage_in(t, o, p0, psi0, n, s);
    st->hfree.alloc(nf * sizeof(double));
    double* x = st->hfree.as<double>();
    hipLaunchKernelGGL(k_gather_free, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                       t->f2n.as<int32_t>(), P0, x);
    for (DBuf* d : {&F, &dFa, &dFb}) d->alloc(ne * sizeof(double));
    for (DBuf* d : {&r, &rn, &dx, &xn}) d->alloc(nf * sizeof(double));
    part.alloc(2 * nblocks(nf) * sizeof(double));
    LevelDev& d0 = *t->lev[0];
    Mv.alloc(d0.A.entries * sizeof(double));
    std::vector<Cheb> chs = rich_chebs(t);
    auto res = rich_residual(t, s, in, x, F.as<double>(), dFa.as<double>(), dFb.as<double>(),
                             r.as<double>(), part);
    double r0 = res.first;
    double tol = std::max(o->reltol * r0, o->abstol);
    int newton = 0, lin = 0, status = DPF_OK;
    while (res.first > tol) {
      if (newton >= 100 || lin >= o->maxiter * 50) {
        status = DPF_ENOCONV;
        break;
      }
      if (!std::isfinite(res.first)) {
        status = DPF_EBREAKDOWN;
        break;
      }
      rich_matrices(t, s, in, x, dFa.as<double>(), dFb.as<double>(), Mv.as<double>(), nullptr,
                    fp32);
      int its = 0;
      int ls = rich_bicgstab(t, s, chs, d0.A.view<double>(Mv.as<double>()), r.as<double>(),
                             dx.as<double>(), 1e-10, o->maxiter, &its, fp32);
      lin += its;
      if (ls == DPF_EBREAKDOWN) {
        status = DPF_EBREAKDOWN;
        break;
      }
      // M = -dr/dpsi, so the Newton step is +dx; backtrack on ||r||_2
      double lam = 1.0;
      std::pair<double, double> rnew;
      while (true) {
        copy_dev(xn.p, x, nf * sizeof(double), s);
        hipLaunchKernelGGL(k_rich_axpby, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, lam,
                           (const double*)dx.as<double>(), 1.0, xn.as<double>(), 0.0,
                           (const double*)nullptr);
        rnew = rich_residual(t, s, in, xn.as<double>(), F.as<double>(), dFa.as<double>(),
                             dFb.as<double>(), rn.as<double>(), part);
        if (rnew.second < (1.0 - 1e-4 * lam) * res.second || lam < 1e-6) break;
        lam *= 0.5;
      }
      copy_dev(x, xn.p, nf * sizeof(double), s);
      copy_dev(r.p, rn.p, nf * sizeof(double), s);
      res = rnew;
      newton++;
    }
    double* pdev = psi_out;
    if (o->memory == DPF_HOST) {
      out.alloc(n * sizeof(double));
      pdev = out.as<double>();
    }
    hipLaunchKernelGGL(k_scatter_h, dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                       (const double*)x, in.dpsi, pdev);
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(psi_out, pdev, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->iterations = lin;
      stats->outer = newton;
      stats->converged = status == DPF_OK;
      stats->levels = t->levels();
      stats->relres = r0 > 0 ? res.first / r0 : res.first;
      stats->total_ms = now_ms() - t0;
    }
    if (status == DPF_ENOCONV) set_err("Richards Newton iteration did not converge");
    if (status == DPF_EBREAKDOWN) set_err("Richards Newton iteration broke down (non-finite "
                                          "residual or a failed linear solve)");
    if (state_out && status == DPF_OK) *state_out = st.release();
    return status;
  });
}

extern "C" int dpf_richards_pullback(dpf_state* st, const double* delta, double* dKs,
                                     double* ddirichletpsis, double* dalphas, double* dNs,
                                     double* dQs, const dpf_opts* opts, dpf_stats* stats) {
  if (!st || !st->topo || !delta) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (st->kind != 3) {
    set_err("the state does not come from dpf_richards_steady");
    return DPF_EINVAL;
  }
  dpf_topology* t = st->topo;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  if (check_opts(o)) {
    set_err("invalid options");
    return DPF_EINVAL;
  }
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double t0 = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    bool fp32 = o->precision != DPF_FP64;  // FP32 V-cycle preconditioner
    bool host = o->memory == DPF_HOST;
    t->timing = false;
    tp_build_maps(t);
    BufPool* pl = t->pool.get();
    RichIn in{st->Ke.as<double>(), st->dh.as<double>(), st->r_z.as<double>(),
              st->r_al.as<double>(), st->r_N.as<double>(), st->Q.as<double>()};
    const double* x = st->hfree.as<double>();
    DBuf F(pl), dFa(pl), dFb(pl), MTv(pl), db(pl), rhs(pl), mu(pl), ga(pl), gb(pl),
        oK(pl), oA(pl), oN(pl), oD(pl), oQ(pl);
    for (DBuf* d : {&F, &dFa, &dFb, &ga, &gb}) d->alloc(ne * sizeof(double));
    LevelDev& d0 = *t->lev[0];
    MTv.alloc(d0.A.entries * sizeof(double));
    hipLaunchKernelGGL(k_rich_edge_flux, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                       t->edges.as<int2>(), t->n2f.as<int32_t>(), x, in.dpsi, in.z, in.al, in.N,
                       in.K, t->aol.as<double>(), F.as<double>(), dFa.as<double>(),
                       dFb.as<double>());
    rich_matrices(t, s, in, x, dFa.as<double>(), dFb.as<double>(), nullptr, MTv.as<double>(),
                  fp32);
    const double* dl = stage_in(t, o, db, delta, n, s);
    rhs.alloc(nf * sizeof(double));
    mu.alloc(nf * sizeof(double));
    hipLaunchKernelGGL(k_gather_free, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                       t->f2n.as<int32_t>(), dl, rhs.as<double>());
    // lambda = J^-T delta = -(M^T)^-1 delta
    int its = 0;
    std::vector<Cheb> chs = rich_chebs(t);
    int status = rich_bicgstab(t, s, chs, d0.A.view<double>(MTv.as<double>()), rhs.as<double>(),
                               mu.as<double>(), 1e-12, o->maxiter, &its, fp32);
    hipLaunchKernelGGL(k_rich_axpby, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, 0.0,
                       (const double*)mu.as<double>(), -1.0, mu.as<double>(), 0.0,
                       (const double*)nullptr);
    auto dev_out = [&](DBuf& b, double* user, int64_t count) -> double* {
      if (!user) return nullptr;
      if (!host) return user;
      b.alloc(count * sizeof(double));
      return b.as<double>();
    };
    double* k_ = dev_out(oK, dKs, ne);
    double* a_ = dev_out(oA, dalphas, ne);
    double* n_ = dev_out(oN, dNs, ne);
    double* d_ = dev_out(oD, ddirichletpsis, n);
    double* q_ = dev_out(oQ, dQs, n);
    hipLaunchKernelGGL(k_rich_tangents, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                       t->edges.as<int2>(), t->n2f.as<int32_t>(), x, in.dpsi, in.z, in.al, in.N,
                       in.K, t->aol.as<double>(), (const double*)mu.as<double>(), k_, a_, n_,
                       ga.as<double>(), gb.as<double>());
    hipLaunchKernelGGL(k_rich_cell_tangents, dim3(nblocks(n)), dim3(BLK), 0, s, n,
                       t->n2f.as<int32_t>(), t->edges.as<int2>(), t->inc_ptr.as<int64_t>(),
                       t->inc_edge.as<int32_t>(), (const double*)ga.as<double>(),
                       (const double*)gb.as<double>(), (const double*)mu.as<double>(), dl, d_,
                       q_);
    HIPTRY(hipGetLastError());
    if (host) {
      auto back = [&](double* user, double* dev, int64_t count) {
        if (user)
          HIPTRY(hipMemcpyAsync(user, dev, count * sizeof(double), hipMemcpyDeviceToHost, s));
      };
      back(dKs, k_, ne);
      back(dalphas, a_, ne);
      back(dNs, n_, ne);
      back(ddirichletpsis, d_, n);
      back(dQs, q_, n);
    }
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->iterations = its;
      stats->converged = status == DPF_OK;
      stats->total_ms = now_ms() - t0;
    }
    if (status == DPF_ENOCONV) set_err("adjoint BiCGStab did not reach the tolerance");
    if (status == DPF_EBREAKDOWN) set_err("adjoint BiCGStab broke down");
    return status;
  });
}

extern "C" int dpf_richards_residuals(dpf_topology* t, const double* psi_free, const double* Ks,
                                      const double* dirichletpsis, const double* z,
                                      const double* alphas, const double* Ns, const double* Qs,
                                      const double* Ss, const double* V, double* r_out,
                                      const dpf_opts* opts) {
  if (!t || !psi_free || !Ks || !dirichletpsis || !z || !alphas || !Ns || !Qs || !Ss || !V ||
      !r_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne, nf = t->H.nf;
    BufPool* pl = t->pool.get();
    DBuf b1(pl), b2(pl), b3(pl), b4(pl), b5(pl), b6(pl), b7(pl), b8(pl), b9(pl), F(pl), sv(pl),
        r(pl), part(pl);
    const double* x = stage_in(t, o, b1, psi_free, nf, s);
    RichIn in{stage_in(t, o, b2, Ks, ne, s), stage_in(t, o, b3, dirichletpsis, n, s),
              stage_in(t, o, b4, z, n, s),   stage_in(t, o, b5, alphas, ne, s),
              stage_in(t, o, b6, Ns, ne, s), stage_in(t, o, b7, Qs, n, s)};
    const double* ss = stage_in(t, o, b8, Ss, n, s);
    const double* vv = stage_in(t, o, b9, V, n, s);
    F.alloc(ne * sizeof(double));
    sv.alloc(n * sizeof(double));
    part.alloc(2 * nblocks(nf) * sizeof(double));
    double* rd = r_out;
    if (o->memory == DPF_HOST) {
      r.alloc(nf * sizeof(double));
      rd = r.as<double>();
    }
    hipLaunchKernelGGL(k_rich_mul, dim3(nblocks(n)), dim3(BLK), 0, s, n, ss, vv, sv.as<double>());
    hipLaunchKernelGGL(k_rich_edge_flux, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                       t->edges.as<int2>(), t->n2f.as<int32_t>(), x, in.dpsi, in.z, in.al, in.N,
                       in.K, t->aol.as<double>(), F.as<double>(), (double*)nullptr,
                       (double*)nullptr);
    hipLaunchKernelGGL(k_rich_residual, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                       t->f2n.as<int32_t>(), t->edges.as<int2>(), t->inc_ptr.as<int64_t>(),
                       t->inc_edge.as<int32_t>(), (const double*)F.as<double>(), in.Q,
                       (const double*)sv.as<double>(), rd, part.as<double>());
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(r_out, rd, nf * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return DPF_OK;
  });
}

// ====================================================================== advection-dispersion

namespace {

// Device copies of one transport problem's inputs (kept in st when given).
struct TrBufs {
  DBuf v, D, ud, Q, V, c;
  explicit TrBufs(BufPool* p) : v(p), D(p), ud(p), Q(p), V(p), c(p) {}
};

TrIn tr_inputs(dpf_topology* t, const dpf_opts* o, hipStream_t s, DBuf& v, DBuf& D, DBuf& ud,
               DBuf& Q, DBuf& V, DBuf& c, const double* vx, const double* vy, const double* vz,
               const double* Ds, const double* dirichletus, const double* Qs,
               const double* volumes, const double* cx, const double* cy, const double* cz,
               int upwind) {
  int64_t n = t->H.n, ne = t->H.ne;
  auto kind = o->memory == DPF_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  auto put = [&](DBuf& d, std::initializer_list<const double*> src, int64_t count) {
    d.alloc(std::max<int64_t>(count * (int64_t)src.size(), 1) * sizeof(double));
    int64_t k = 0;
    for (const double* p : src)
      HIPTRY(hipMemcpyAsync(d.as<double>() + (k++) * count, p, count * sizeof(double), kind, s));
  };
  put(v, {vx, vy, vz}, ne);
  put(D, {Ds}, ne);
  put(ud, {dirichletus}, n);
  put(Q, {Qs}, n);
  put(V, {volumes}, n);
  put(c, {cx, cy, cz}, n);
  TrIn in;
  in.vx = v.as<double>();
  in.vy = in.vx + ne;
  in.vz = in.vy + ne;
  in.D = D.as<double>();
  in.ud = ud.as<double>();
  in.Q = Q.as<double>();
  in.V = V.as<double>();
  in.cx = c.as<double>();
  in.cy = in.cx + n;
  in.cz = in.cy + n;
  in.upwind = upwind;
  return in;
}

void tr_rows(dpf_topology* t, hipStream_t s, const TrIn& in, const double* x, double a, double b,
             double g, double* y) {
  int64_t n = t->H.n;
  hipLaunchKernelGGL(k_tr_rows, dim3(nblocks(n)), dim3(BLK), 0, s, n, t->n2f.as<int32_t>(),
                     t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                     t->inc_other.as<int32_t>(), t->aol.as<double>(), in, x, a, b, g, y);
}

}  // namespace

// M = I - dt J on the free cells' fine pattern (and/or its transpose MT), its
// diagonal dg and the free rows' constant part c0 (nullable).
void tr_matrix(dpf_topology* t, hipStream_t s, const TrIn& in, double dt, double* M, double* MT,
               double* dg, double* c0) {
  tp_build_maps(t);
  LevelDev& d0 = *t->lev[0];
  int64_t nf = t->H.nf;
  size_t bytes = d0.A.entries * sizeof(double);
  if (M) HIPTRY(hipMemsetAsync(M, 0, bytes, s));
  if (MT) HIPTRY(hipMemsetAsync(MT, 0, bytes, s));
  hipLaunchKernelGGL(k_tr_matrix, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                     t->n2f.as<int32_t>(), t->inc_ptr.as<int64_t>(), t->inc_edge.as<int32_t>(),
                     t->inc_other.as<int32_t>(), t->tp_mirror.as<int32_t>(),
                     t->inc_slot.as<uint16_t>(), d0.diag_slot.as<uint16_t>(),
                     d0.A.view<double>(M ? M : MT), t->aol.as<double>(), in, dt, M, MT, dg, c0);
  HIPTRY(hipGetLastError());
}

extern "C" int dpf_transport_residuals(dpf_topology* t, const double* u, const double* vx,
                                       const double* vy, const double* vz, const double* Ds,
                                       const double* dirichletus, const double* Qs,
                                       const double* volumes, const double* cx,
                                       const double* cy, const double* cz, int32_t use_upwind,
                                       double* r_out, const dpf_opts* opts) {
  if (!t || !u || !vx || !vy || !vz || !Ds || !dirichletus || !Qs || !volumes || !cx || !cy ||
      !cz || !r_out) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n;
    BufPool* pl = t->pool.get();
    TrBufs b(pl);
    DBuf ub(pl), rb(pl);
    TrIn in = tr_inputs(t, o, s, b.v, b.D, b.ud, b.Q, b.V, b.c, vx, vy, vz, Ds, dirichletus, Qs,
                        volumes, cx, cy, cz, use_upwind);
    const double* ud = stage_in(t, o, ub, u, n, s);
    double* rd = r_out;
    if (o->memory == DPF_HOST) {
      rb.alloc(n * sizeof(double));
      rd = rb.as<double>();
    }
    tr_rows(t, s, in, ud, 0.0, 1.0, 1.0, rd);
    HIPTRY(hipGetLastError());
    if (o->memory == DPF_HOST)
      HIPTRY(hipMemcpyAsync(r_out, rd, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    return DPF_OK;
  });
}

extern "C" int dpf_transport_steps(dpf_topology* t, const double* u0, const double* vx,
                                   const double* vy, const double* vz, const double* Ds,
                                   const double* dirichletus, const double* Qs,
                                   const double* volumes, const double* cx, const double* cy,
                                   const double* cz, int32_t use_upwind, double t0, double t1,
                                   int32_t nsteps, double* u_final, double* u_all,
                                   const dpf_opts* opts, dpf_stats* stats,
                                   dpf_state** state_out) {
  if (state_out) *state_out = nullptr;
  if (!t || !u0 || !vx || !vy || !vz || !Ds || !dirichletus || !Qs || !volumes || !cx || !cy ||
      !cz || !u_final || nsteps < 1 || !(t1 > t0)) {
    set_err(nsteps < 1 || !(t1 > t0) ? "need nsteps >= 1 and t1 > t0" : "NULL argument");
    return DPF_EINVAL;
  }
  if (int rc = twophase_checks(t)) return rc;
  if (t->H.nf < t->H.n && (t1 - t0) / nsteps == 1.0) {
    // a Dirichlet row of backward Euler is (1 - dt) u' = u - dt ud
    set_err("a step of exactly 1 makes the Dirichlet rows of backward Euler singular");
    return DPF_EINVAL;
  }
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double tstart = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n;
    bool host = o->memory == DPF_HOST;
    BufPool* pl = t->pool.get();
    auto st = std::make_unique<dpf_state>(t);
    st->kind = 4;
    st->nsteps = nsteps;
    st->dt = (t1 - t0) / nsteps;
    st->tr_upwind = use_upwind;
    TrIn in = tr_inputs(t, o, s, st->tr_v, st->tr_D, st->tr_ud, st->tr_Q, st->tr_V, st->tr_c, vx,
                        vy, vz, Ds, dirichletus, Qs, volumes, cx, cy, cz, use_upwind);
    double dt = st->dt;
    st->tr_u.alloc((size_t)(nsteps + 1) * n * sizeof(double));
    double* U = st->tr_u.as<double>();
    HIPTRY(hipMemcpyAsync(U, u0, n * sizeof(double),
                          host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    int64_t nf = t->H.nf;
    LevelDev& d0 = *t->lev[0];
    DBuf M(pl), dg(pl), c0(pl), rhs(pl), x(pl);
    M.alloc(d0.A.entries * sizeof(double));
    for (DBuf* d : {&dg, &c0, &rhs, &x}) d->alloc(std::max<int64_t>(nf, 1) * sizeof(double));
    tr_matrix(t, s, in, dt, M.as<double>(), nullptr, dg.as<double>(), c0.as<double>());
    SellD<double> Mv = d0.A.view<double>(M.as<double>());
    auto apply = [&](const double* xx, double* yy) {
      hipLaunchKernelGGL(k_rich_spmv, dim3(nblocks(nf)), dim3(BLK), 0, s, Mv, xx, yy);
    };
    // the Jacobi preconditioner (diagonal dg) is fused into BiCGStab's updates
    auto prec = [](const double*, double*) {};
    int status = DPF_OK, iters = 0;
    for (int k = 0; k < nsteps && status == DPF_OK; k++) {
      const double* uk = U + (int64_t)k * n;
      double* un = U + (int64_t)(k + 1) * n;
      // Dirichlet cells in place; the free cells' system starts from u_k
      hipLaunchKernelGGL(k_tr_step_in, dim3(nblocks(n)), dim3(BLK), 0, s, n,
                         t->n2f.as<int32_t>(), in.ud, (const double*)c0.as<double>(), dt, uk, un,
                         rhs.as<double>(), x.as<double>());
      int it = 0;
      status = bicgstab(t, s, nf, apply, prec, (const double*)rhs.as<double>(), x.as<double>(),
                        std::max(o->reltol, 1e-14), o->maxiter, &it, true,
                        (const double*)dg.as<double>());
      hipLaunchKernelGGL(k_tr_step_out, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                         t->f2n.as<int32_t>(), (const double*)x.as<double>(), un);
      iters += it;
    }
    auto out = [&](double* dst, const double* src, int64_t count) {
      HIPTRY(hipMemcpyAsync(dst, src, count * sizeof(double),
                            host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
    };
    out(u_final, U + (int64_t)nsteps * n, n);
    if (u_all) out(u_all, U, (int64_t)(nsteps + 1) * n);
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->iterations = iters;
      stats->outer = nsteps;
      stats->converged = status == DPF_OK;
      stats->total_ms = now_ms() - tstart;
    }
    if (status == DPF_ENOCONV) set_err("BiCGStab of a backward-Euler step did not converge");
    if (status == DPF_EBREAKDOWN) set_err("BiCGStab of a backward-Euler step broke down");
    if (state_out && status == DPF_OK) *state_out = st.release();
    return status;
  });
}

extern "C" int dpf_transport_pullback(dpf_state* st, const double* dJ_du_final, double* dvx,
                                      double* dvy, double* dvz, double* dDs, double* dQs,
                                      const dpf_opts* opts, dpf_stats* stats) {
  if (!st || !st->topo || !dJ_du_final) {
    set_err("NULL argument");
    return DPF_EINVAL;
  }
  if (st->kind != 4) {
    set_err("the state does not come from dpf_transport_steps");
    return DPF_EINVAL;
  }
  dpf_topology* t = st->topo;
  dpf_opts dflt;
  dpf_opts_init(&dflt);
  const dpf_opts* o = opts ? opts : &dflt;
  std::lock_guard<std::mutex> g(t->mu);
  return guarded([&]() {
    double tstart = now_ms();
    CallStream cs(t, o);
    hipStream_t s = cs.s;
    int64_t n = t->H.n, ne = t->H.ne;
    bool host = o->memory == DPF_HOST;
    BufPool* pl = t->pool.get();
    TrIn in;
    in.vx = st->tr_v.as<double>();
    in.vy = in.vx + ne;
    in.vz = in.vy + ne;
    in.D = st->tr_D.as<double>();
    in.ud = st->tr_ud.as<double>();
    in.Q = st->tr_Q.as<double>();
    in.V = st->tr_V.as<double>();
    in.cx = st->tr_c.as<double>();
    in.cy = in.cx + n;
    in.cz = in.cy + n;
    in.upwind = st->tr_upwind;
    double dt = st->dt;
    // (I - dt J)^T is block diagonal over free and Dirichlet cells, and the
    // gradients read the adjoint at free cells only: the adjoint steps solve
    // the free block's transpose
    int64_t nf = t->H.nf;
    LevelDev& d0 = *t->lev[0];
    DBuf MT(pl), lamn(pl), lam(pl), mu(pl), dg(pl), gv(pl), gD(pl), gQ(pl);
    MT.alloc(d0.A.entries * sizeof(double));
    lamn.alloc(n * sizeof(double));
    for (DBuf* d : {&lam, &mu, &dg}) d->alloc(std::max<int64_t>(nf, 1) * sizeof(double));
    gv.alloc(3 * ne * sizeof(double));
    gD.alloc(ne * sizeof(double));
    gQ.alloc(n * sizeof(double));
    HIPTRY(hipMemsetAsync(gv.p, 0, 3 * ne * sizeof(double), s));
    HIPTRY(hipMemsetAsync(gD.p, 0, ne * sizeof(double), s));
    HIPTRY(hipMemsetAsync(gQ.p, 0, n * sizeof(double), s));
    HIPTRY(hipMemcpyAsync(lamn.p, dJ_du_final, n * sizeof(double),
                          host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, s));
    tr_matrix(t, s, in, dt, nullptr, MT.as<double>(), dg.as<double>(), nullptr);
    SellD<double> MTv = d0.A.view<double>(MT.as<double>());
    hipLaunchKernelGGL(k_tr_gather, dim3(nblocks(nf)), dim3(BLK), 0, s, nf, t->f2n.as<int32_t>(),
                       (const double*)lamn.as<double>(), lam.as<double>());
    auto apply = [&](const double* xx, double* yy) {
      hipLaunchKernelGGL(k_rich_spmv, dim3(nblocks(nf)), dim3(BLK), 0, s, MTv, xx, yy);
    };
    // the Jacobi preconditioner (diagonal dg) is fused into BiCGStab's updates
    auto prec = [](const double*, double*) {};
    int status = DPF_OK, iters = 0;
    double* gvx = gv.as<double>();
    // each adjoint step starts from the previous one (mu holds lam)
    copy_dev(mu.p, lam.p, nf * sizeof(double), s);
    for (int k = st->nsteps; k >= 1 && status == DPF_OK; k--) {
      int it = 0;
      status = bicgstab(t, s, nf, apply, prec, (const double*)lam.as<double>(), mu.as<double>(),
                        std::max(o->reltol, 1e-14), o->maxiter, &it, true,
                        (const double*)dg.as<double>());
      iters += it;
      const double* uk = st->tr_u.as<double>() + (int64_t)k * n;
      hipLaunchKernelGGL(k_tr_grad_edges, dim3(nblocks(ne)), dim3(BLK), 0, s, ne,
                         t->edges.as<int2>(), t->n2f.as<int32_t>(), t->aol.as<double>(), in, uk,
                         (const double*)mu.as<double>(), dt, gvx, gvx + ne, gvx + 2 * ne,
                         gD.as<double>());
      hipLaunchKernelGGL(k_tr_grad_q, dim3(nblocks(nf)), dim3(BLK), 0, s, nf,
                         t->f2n.as<int32_t>(), in, (const double*)mu.as<double>(), dt,
                         gQ.as<double>());
      copy_dev(lam.p, mu.p, nf * sizeof(double), s);
    }
    HIPTRY(hipGetLastError());
    auto out = [&](double* dst, const double* src, int64_t count) {
      if (!dst) return;
      HIPTRY(hipMemcpyAsync(dst, src, count * sizeof(double),
                            host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
    };
    out(dvx, gvx, ne);
    out(dvy, gvx + ne, ne);
    out(dvz, gvx + 2 * ne, ne);
    out(dDs, gD.as<double>(), ne);
    out(dQs, gQ.as<double>(), n);
    HIPTRY(hipStreamSynchronize(s));
    if (stats) {
      std::memset(stats, 0, sizeof(*stats));
      stats->iterations = iters;
      stats->converged = status == DPF_OK;
      stats->total_ms = now_ms() - tstart;
    }
    if (status == DPF_ENOCONV) set_err("adjoint BiCGStab did not converge");
    if (status == DPF_EBREAKDOWN) set_err("adjoint BiCGStab broke down");
    return status;
  });
}
