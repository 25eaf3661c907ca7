/*
 * dpfgpu.h -- C-ABI of libdpfgpu.so, the MI355X (gfx950) groundwater solve
 * and adjoint behind DPFEHM's groundwater interface.
 *
 * Plain pointers and sizes only.  Index arguments follow the reference's
 * 1-based convention (Julia Pair{Int,Int} neighbors reinterpret as 2*E
 * int64 without copying).  Every call returns DPF_OK or an error code;
 * dpf_last_error() gives the message of the calling thread's last failure.
 *
 * Array arguments of compute calls are host pointers when opts->memory is
 * DPF_HOST (the library stages them through device buffers) and device
 * pointers when it is DPF_DEVICE (work is queued on opts->stream; when that
 * is NULL, on the topology's own stream, ordered after the work already on
 * the legacy NULL stream when the call starts and ahead of the NULL-stream
 * work queued after it returns, as if the call ran on the NULL stream).
 * Calls on one topology are serialised by the library, and their work runs
 * in call order on the GPU whatever streams they name.  With DPF_HOST the
 * outputs are in place when a call returns; with DPF_DEVICE the steady
 * forward and adjoint solves (which wait for their own convergence) and the
 * logK maps return with the rest of their work queued, and the outputs are
 * ready once the call's stream has reached it (synchronise it, or queue the
 * consumer on it).
 */
#ifndef DPFGPU_H
#define DPFGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpf_topology dpf_topology; /* mesh + Dirichlet set + AMG symbolic setup */
typedef struct dpf_state dpf_state;       /* what a pullback needs from its forward call */

enum {
  DPF_OK = 0,
  DPF_EINVAL = 1,      /* bad argument (the reference would throw) */
  DPF_EHIP = 2,        /* HIP runtime failure */
  DPF_ENOCONV = 3,     /* iteration cap reached before the tolerance */
  DPF_EBREAKDOWN = 4,  /* CG breakdown: matrix or preconditioner not SPD, or NaN */
  DPF_ESINGULAR = 5,   /* singular system: a free cell without edges, or (steady
                          solves) a connected component that no edge of nonzero
                          conductance K_e aol_e joins to a Dirichlet cell, e.g.
                          dirichletnodes empty */
  DPF_ENOMEM = 6
};
/* DPF_FP64: FP64 AMG-PCG.  DPF_MIXED: FP32 AMG-PCG inner solves inside FP64
 * iterative refinement.  DPF_MIXED_PC: FP64 CG (vectors, SpMV, scalars)
 * preconditioned by the FP32 V-cycle -- no refinement restarts, so the FP64
 * iteration count, with the V-cycle's traffic halved. */
enum { DPF_FP64 = 0, DPF_MIXED = 1, DPF_MIXED_PC = 2 };
enum { DPF_HOST = 0, DPF_DEVICE = 1 };
/* flags: DPF_FLAG_SWEEP_TIMING times every finest-level smoothing sweep with
 * a hipEvent pair and reports the sum in dpf_stats.sweep_ms / sweep_launches
 * (adds two event records per sweep).  DPF_FLAG_GRAPH replays each PCG
 * iteration (SpMV and updates, the reductions' finishing kernels, the whole
 * V-cycle: ~20 launches) as one hipGraph, captured on the topology's first
 * iteration with the same precision and smoother and kept until it is
 * destroyed; the kernels and their order are those of the launched form, so
 * results are bitwise equal.  Ignored together with sweep or kernel timing
 * and during the autotuner's trials.  The stream must not be the legacy NULL
 * stream.  DPF_FLAG_KERNEL_TIMING times every launch of dpf_steady_forward,
 * dpf_steady_pullback and the logK maps the same way and adds the durations
 * and algorithmic bytes to the topology's table (dpf_kernel_table); launches
 * that the converged solve turned into no-ops are left out.
 * The FP32 V-cycle (DPF_MIXED, DPF_MIXED_PC) computes in FP32 and reads the
 * prolongators, restrictions and coarse-level operators as bf16 values (each
 * coarse row's rounding error folded into an FP32 diagonal, so row sums are
 * kept); DPF_FLAG_FP32_STORAGE makes it read FP32 values instead.
 * DPF_FLAG_AUTOTUNE (with cheb_degree 0): the topology's first solve in a
 * precision times the lanes per row of the coarse split copies (once per
 * topology) and the Chebyshev degrees and interval, and keeps the fastest
 * for the topology's life (dpf_topology_lanes, dpf_topology_tuning report
 * the trials).  Without it the smoother and layout are fixed defaults
 * (degrees 1, 1 and 2, ratio 15, ~8 entries a lane) or what was pinned, so
 * results do not depend on timing. */
enum { DPF_FLAG_SWEEP_TIMING = 1, DPF_FLAG_GRAPH = 2, DPF_FLAG_KERNEL_TIMING = 4,
       DPF_FLAG_FP32_STORAGE = 8, DPF_FLAG_AUTOTUNE = 16 };

typedef struct {
  double reltol;       /* stop when ||b - A h|| <= max(reltol ||b||, abstol), FP64 residual */
  double abstol;
  int32_t maxiter;     /* PCG iteration cap, summed over refinement steps */
  int32_t precision;   /* DPF_FP64, DPF_MIXED or DPF_MIXED_PC (above) */
  double inner_reltol; /* DPF_MIXED: relative tolerance of each FP32 inner solve; a step
                          whose residual is within 10/inner_reltol of the target solves
                          to half the target instead */
  int32_t max_outer;   /* DPF_MIXED: refinement step cap */
  int32_t cheb_degree; /* Chebyshev smoother degree on every level; 0 = the
                          topology's degrees: default (1 on the finest and the
                          second level, 2 below), pinned, or autotuned with
                          DPF_FLAG_AUTOTUNE */
  int32_t memory;      /* DPF_HOST or DPF_DEVICE */
  int32_t flags;       /* DPF_FLAG_* bits, 0 by default */
  void *stream;        /* hipStream_t for DPF_DEVICE calls, NULL = NULL-stream order */
} dpf_opts;

typedef struct {
  int32_t iterations;  /* PCG iterations, all refinement steps */
  int32_t outer;       /* refinement steps (1 for DPF_FP64) */
  int32_t converged;   /* 1 if the FP64 residual met the tolerance */
  int32_t levels;      /* AMG levels */
  double relres;       /* final ||b - A h|| / ||b|| in FP64 */
  double setup_ms;     /* assembly + numeric AMG refresh (GPU) */
  double solve_ms;     /* Krylov solve(s) */
  double total_ms;     /* the whole call, wall clock */
  double sweep_ms;     /* with DPF_FLAG_SWEEP_TIMING: summed finest-level sweep time */
  int32_t sweep_launches;
  int32_t reserved;
} dpf_stats;

/* Message of the last failing call on this thread ("" if none). */
const char *dpf_last_error(void);

/* Defaults: reltol 1e-10, abstol 0, maxiter 2000, DPF_FP64, inner_reltol 1e-5,
 * max_outer 20, cheb_degree 0 (the topology's degrees), DPF_HOST, stream NULL,
 * flags 0. */
void dpf_opts_init(dpf_opts *opts);

/* Topology cache.  Replaces the per-call work the reference repeats inside
 * every generated function: getfreenodes (src/utilities.jl:10-24, effective
 * src/twophase.jl:15-20), the COO -> sparse() pattern build of the
 * NonlinearEquations-generated Jacobians (src/groundwater.jl:1,79) and the
 * AMG setup of amg_solver (src/groundwater.jl:45-49) -- here the
 * conductivity-independent part of a smoothed-aggregation hierarchy.
 *   neighbors: 2*nedges int64, pairs (a, b), 1-based;
 *   areasoverlengths: nedges;  dirichletnodes: ndirichlet, 1-based. */
int dpf_topology_create(int64_t n, int64_t nedges, const int64_t *neighbors,
                        const double *areasoverlengths, int64_t ndirichlet,
                        const int64_t *dirichletnodes, dpf_topology **out);
int dpf_topology_destroy(dpf_topology *topo);

/* Batched ensembles (no counterpart in the reference, whose ensembles are
 * independent calls, e.g. pmap in examples/scott_example/ex.jl:10): the
 * members' meshes packed into one block-diagonal topology.  Member m has
 * ncells[m] cells, nedges[m] edges and ndirichlet[m] Dirichlet cells; the
 * neighbors (2 per edge), areasoverlengths and dirichletnodes arrays hold the
 * members one after another, with 1-based indices local to the member.
 * Every other entry point then works on the packed arrays (cells, edges and
 * free cells concatenated member by member) in one set of launches, while
 * each member keeps its own CG scalars and its own convergence test against
 * reltol.  Every member needs at least 256 free cells. */
int dpf_batch_topology_create(int32_t nmembers, const int64_t *ncells, const int64_t *nedges,
                              const int64_t *neighbors, const double *areasoverlengths,
                              const int64_t *ndirichlet, const int64_t *dirichletnodes,
                              dpf_topology **out);

/* Relative FP64 residual ||b - A h|| / ||b|| of every member after the last
 * solve on the topology (nmembers must equal its member count; 1 for a
 * single system). */
int dpf_member_stats(const dpf_topology *topo, double *relres, int32_t nmembers);

/* info[0..20] = n, nedges, nfree, nnz(A), levels, coarse rows,
 * operator-complexity*1000, host symbolic setup time in microseconds, the
 * Chebyshev degrees (finest level, second level) in use for DPF_FP64 and for
 * DPF_MIXED (the default 1 and 1 unless tuned or pinned), the fine operator's
 * 64-row slices stored as column offsets (no column indices), the degrees
 * for DPF_MIXED_PC, then the Chebyshev interval ratio r (interval
 * [lmax/r, lmax]; default 15) for DPF_FP64, DPF_MIXED and DPF_MIXED_PC, then
 * the degree of the levels below the second (default 2) for the same three. */
int dpf_topology_info(const dpf_topology *topo, int64_t *info, int32_t ninfo);

/* The Chebyshev autotuner's trials for one precision (DPF_FP64, DPF_MIXED or
 * DPF_MIXED_PC), zeros unless a solve with DPF_FLAG_AUTOTUNE tuned it:
 * ms[k] / iters[k] for k < min(n, 7) = wall time and PCG iterations of trial
 * k, which solves the first refinement step's system with (finest, second,
 * deeper levels) degrees (1,1,2), (1,1,3), (1,1,4), (1,2,2), (1,2,3) at
 * interval ratio 15 (k = 0..4), then with the best of these at ratios 30 and
 * 8 (k = 5, 6); -1 ms marks a trial that did not converge.  No counterpart in
 * the reference (its AMG smoother is fixed). */
int dpf_topology_tuning(const dpf_topology *topo, int32_t precision, double *ms,
                        int32_t *iters, int32_t n);

/* Fixes the smoother of one precision on this topology, in place of the
 * autotuner (which then does not run for it): Chebyshev degree deg_fine on
 * the finest level and deg_coarse on every level below (1..16), interval
 * [lmax/ratio, lmax], ratio > 1.  For runs that must not depend on timing,
 * e.g. to repeat a tuned configuration under a profiler.  No counterpart in
 * the reference. */
int dpf_topology_set_tuning(dpf_topology *topo, int32_t precision, int32_t deg_fine,
                            int32_t deg_coarse, double ratio);

/* As dpf_topology_set_tuning with three degrees: deg_fine on the finest
 * level, deg_second on the second level, deg_deeper on every level below it
 * (each 1..16; the default smoother is 1, 1, 2 at ratio 15).  No counterpart
 * in the reference. */
int dpf_topology_set_smoother(dpf_topology *topo, int32_t precision, int32_t deg_fine,
                              int32_t deg_second, int32_t deg_deeper, double ratio);

/* Damping of the AMG prolongator smoothing P = (I - omega D_l1^-1 A) T on
 * the finest level (omega_fine) and on every coarser level (omega_coarse),
 * each in (0, 4); default 1.5 and 2, which suit spatially correlated
 * conductivity fields.  Fields that are rough from cell to cell can converge
 * in fewer iterations with 4/3 on both.  Applies from the next numeric
 * refresh; states kept by earlier forward solves refresh again before their
 * adjoint solves.  No counterpart in the reference. */
int dpf_topology_set_damping(dpf_topology *topo, double omega_fine, double omega_coarse);

/* Lanes per row of each level's row-split copy, the layout the coarse
 * levels' smoothing sweeps and residuals read (2^lg[l] lanes share row r;
 * lg[l] = 0: level l has no split copy), for levels 0..nlevels-1; us (may be
 * NULL, else 7*nlevels) gets the autotuner's sweep time in microseconds of
 * level l with 2^k lanes a row at us[7*l + k], 0 where not timed.  The
 * default layout gives each lane about 8 entries; a solve with
 * DPF_FLAG_AUTOTUNE times 2..32 lanes a row on every split level first and
 * keeps the fastest.  No counterpart in the reference. */
int dpf_topology_lanes(const dpf_topology *topo, int32_t *lg, double *us, int32_t nlevels);

/* Fixes level's split copy at 2^lg lanes a row (1 <= lg <= 6; level >= 1 must
 * have a split copy), in place of the autotuner (which then leaves the
 * layout alone).  The row sums' order follows the layout, so two processes
 * reproduce each other's results bit for bit when they pin the same lanes
 * and smoother.  No counterpart in the reference. */
int dpf_topology_set_lanes(dpf_topology *topo, int32_t level, int32_t lg);

/* groundwater_residuals(h, Ks, neighbors, areasoverlengths, dirichletnodes,
 * dirichleths, Qs, specificstorage, volumes), src/groundwater.jl:1-38.
 *   hfree: nfree;  Ks: nedges;  dirichleths, Qs, specificstorage, volumes: n;
 *   r_out: nfree. */
int dpf_residuals(dpf_topology *topo, const double *hfree, const double *Ks,
                  const double *dirichleths, const double *Qs, const double *specificstorage,
                  const double *volumes, double *r_out, const dpf_opts *opts);

/* Jacobians of groundwater_residuals as sparse matrices over the free cells:
 * the groundwater_h, groundwater_Ks, groundwater_dirichleths and
 * groundwater_Qs that the @NonlinearEquations.equations macro generates
 * (src/groundwater.jl:1 for groundwater_residuals at :1-38), the f_gw_u /
 * f_gw_p blocks of examples/gw_transient_3d/inputdeck.jl:82-91.  CSR, rows =
 * free cells in order, columns 0-based and ascending, contributions to one
 * entry summed (as sparse() sums duplicates):
 *   DPF_JAC_H            nfree x nfree   dr_j/dh_k       (reads Ks, Ss, V)
 *   DPF_JAC_KS           nfree x nedges  dr_j/dKs_e      (reads hfree, dirichleths, Ss, V)
 *   DPF_JAC_DIRICHLETHS  nfree x n       dr_j/ddirichleths_i (reads Ks, Ss, V)
 *   DPF_JAC_QS           nfree x n       dr_j/dQs_i      (reads Ss, V)
 * The pattern is structural (an entry per edge term, even where K_e = 0);
 * self-edges (a == b), whose terms vanish, add none.
 * dpf_jacobian_pattern: *nnz, and where not NULL rowptr (nfree + 1) and cols
 * (nnz), host arrays.  dpf_jacobian_values: the nnz values in pattern order
 * (arrays in opts->memory, like dpf_residuals). */
enum { DPF_JAC_H = 0, DPF_JAC_KS = 1, DPF_JAC_DIRICHLETHS = 2, DPF_JAC_QS = 3 };
int dpf_jacobian_pattern(dpf_topology *topo, int32_t which, int64_t *nnz, int64_t *rowptr,
                         int32_t *cols);
int dpf_jacobian_values(dpf_topology *topo, int32_t which, const double *hfree,
                        const double *Ks, const double *dirichleths, const double *Ss,
                        const double *V, double *vals, const dpf_opts *opts);

/* Edge conductivity from cell log-conductivity, K_e = hm(exp(logK_a),
 * exp(logK_b)) with DPFEHM.hm (src/richards.jl:50-52), and its pullback
 * dlogK = (dK_e/dlogK)' dK (the scatter Zygote performs for the user-side
 * map of examples/gw_steadystate_3d/inputdeck.jl:51), as a cell gather.
 *   logK: n;  Ke, dKe: nedges;  dlogK: n. */
int dpf_hm_forward(dpf_topology *topo, const double *logK, double *Ke, const dpf_opts *opts);

/* Allocation-free dpf_hm_forward.  No counterpart in the reference: for
 * callers that cannot allocate (code under hipGraph stream capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernels and
 * results as dpf_hm_forward (bit for bit), but its exp(logK) staging lives in
 * a fixed device array of DPF_NOALLOC_MAX_CELLS doubles (32 MiB per GPU, held
 * until the process ends) instead of a pooled buffer, and the call neither
 * allocates, frees nor synchronises: it queues two kernels on opts->stream
 * (or the topology's stream) and returns.  Device memory only
 * (opts->memory = DPF_DEVICE, else DPF_EINVAL); a topology of more than
 * DPF_NOALLOC_MAX_CELLS cells is refused with DPF_EINVAL.  Every call on one
 * GPU shares that GPU's staging array: calls that may overlap on one GPU
 * (different streams, or different host threads) race and must be ordered
 * by the caller. */
#define DPF_NOALLOC_MAX_CELLS 4194304
int dpf_hm_forward_noalloc(dpf_topology *topo, const double *logK, double *Ke,
                           const dpf_opts *opts);
int dpf_hm_pullback(dpf_topology *topo, const double *logK, const double *Ke,
                    const double *dKe, double *dlogK, const dpf_opts *opts);

/* The geometric-mean map of examples/gw_steadystate_3d/inputdeck.jl:51,
 * K_e = exp(0.5 (logK_a + logK_b)), and its pullback
 * dlogK_c = sum over the edges of c of dK_e K_e / 2 (a cell gather).
 * Self-edges (a == b) are not incident to their cell and contribute nothing
 * to dlogK (nor, for the same reason, to dpf_hm_pullback); the residual gives
 * them zero flux.
 *   logK: n;  Ke, dKe: nedges;  dlogK: n. */
int dpf_geo_forward(dpf_topology *topo, const double *logK, double *Ke, const dpf_opts *opts);
int dpf_geo_pullback(dpf_topology *topo, const double *Ke, const double *dKe, double *dlogK,
                     const dpf_opts *opts);

/* Two-phase flow, IMPES (src/twophase.jl).  Single-system topologies only.
 * The pressure system is singular when the topology has no Dirichlet cell;
 * a caller then creates the topology with one pinned cell (head 0) and
 * passes zero_mean = 1, which shifts P to mean zero afterwards (the
 * reference's AMG solve of the singular system fixes the constant through
 * its own iterates; fluxes and saturations do not depend on it).
 *
 * tpfa (src/twophase.jl:189-198, with solvepressure :48-56 and its
 * transmissivity equations :23-44): cell values K (n) -> K_e = 0.5 (K_a +
 * K_b), pressure P (n; Dirichlet cells hold dirichleths) of the TPFA system
 * with transmissibility 2 aol_e K_e, solved by AMG-PCG to opts->reltol, and
 * Darcy fluxes Vn (nedges) = (P_a - P_b) 2 aol_e K_e. */
int dpf_twophase_tpfa(dpf_topology *topo, const double *K, const double *dirichleths,
                      const double *Qs, double *P, double *Vn, int32_t zero_mean,
                      const dpf_opts *opts, dpf_stats *stats);

/* upstream (src/twophase.jl:208-235): explicit upwind transport of the water
 * saturation S (n, updated in place) over time T, in ceil(T / cfl) equal
 * sub-steps, cfl = ((1 - swc - sor) / 3) min(pv / (inflow + max(Q, 0))),
 * pv = volumes; each sub-step adds the saturationcalc residual
 * (src/twophase.jl:85-112) times dtx = (T / Nts) / pv and max(Q, 0) dtx, then
 * clamps to [swc, 1 - sor].  fluid = {vw, vo, swc, sor}
 * (relativeperm, src/twophase.jl:201-206), 4 doubles in host memory in every
 * memory mode.  *nsub (may be NULL) = Nts.  opts: memory and stream only. */
int dpf_twophase_upstream(dpf_topology *topo, double *S, const double *fluid, const double *Qs,
                          double T, const double *P, const double *Vn, const double *volumes,
                          int64_t *nsub, const dpf_opts *opts);

/* solvetwophase (src/twophase.jl:238-264): nt IMPES steps from (h0, S0); each
 * step forms the total mobility Mt(S) K, runs tpfa, then upstream over dt
 * (fluid: 4 host doubles, as for dpf_twophase_upstream).
 * P_out, S_out (n): the last step's (h0 and S0 when nt = 0); P_all, S_all
 * (nt x n, step-major, may be NULL): every step's, the reference's everyStep
 * output.  *nsub_total (may be NULL) = CFL sub-steps over all steps; stats:
 * PCG iterations summed over the pressure solves.  state_out (may be NULL)
 * receives the run's tape for dpf_solvetwophase_pullback (per step: the
 * starting saturation, the edge factors, pressure and fluxes; release with
 * dpf_state_destroy). */
int dpf_solvetwophase(dpf_topology *topo, const double *h0, const double *S0, const double *K,
                      const double *dirichleths, const double *Qs, const double *volumes,
                      const double *fluid, double dt, int32_t nt, int32_t zero_mean,
                      double *P_out, double *S_out, double *P_all, double *S_all,
                      int64_t *nsub_total, dpf_state **state_out, const dpf_opts *opts,
                      dpf_stats *stats);

/* Steady Richards equation (src/richards.jl:97-205), FP64 (with precision
 * DPF_MIXED or DPF_MIXED_PC the preconditioner's V-cycle runs in FP32; the
 * residuals, Newton matrices and BiCGStab stay FP64).  Per-edge van
 * Genuchten parameters alphas, Ns (nedges), the cell elevation z (n, the
 * third row of the reference's coords), Dirichlet pressure heads
 * dirichletpsis (n).  Single-system topologies with at least two AMG levels.
 *
 * richards_residuals (src/richards.jl:97-117): r (nfree) at the free heads
 * psi_free (nfree), with storage Ss and volumes V (n). */
int dpf_richards_residuals(dpf_topology *topo, const double *psi_free, const double *Ks,
                           const double *dirichletpsis, const double *z, const double *alphas,
                           const double *Ns, const double *Qs, const double *Ss, const double *V,
                           double *r_out, const dpf_opts *opts);

/* richards_steadystate (src/richards.jl:156-179; Ss = V = 1): Newton's
 * method from psi0 (n) with backtracking on ||r||_2, until max|r| <=
 * max(reltol max|r(psi0)|, abstol).  Each Newton system M dpsi = r, M =
 * -dr/dpsi (the Picard operator K_e hm(kr_a, kr_b) plus the kr-slope terms,
 * nonsymmetric), is solved by BiCGStab to 1e-10 preconditioned by the AMG
 * V-cycle of the Picard operator.  psi_out (n): free heads, Dirichlet cells
 * hold dirichletpsis.  stats: iterations = BiCGStab iterations, outer = Newton
 * steps, relres = final max|r| / max|r(psi0)|.  state_out (may be NULL) keeps
 * what the pullback needs. */
int dpf_richards_steady(dpf_topology *topo, const double *psi0, const double *Ks,
                        const double *dirichletpsis, const double *z, const double *alphas,
                        const double *Ns, const double *Qs, double *psi_out,
                        const dpf_opts *opts, dpf_stats *stats, dpf_state **state_out);

/* The rrule's pullback (src/richards.jl:185-203): lambda = J^-T delta[free]
 * (BiCGStab on the transposed Newton matrix, same preconditioner), then
 * dKs, dalphas, dNs (nedges), ddirichletpsis, dQs (n) = -(dr/d.)^T lambda
 * (+ delta on the Dirichlet cells); any output may be NULL.  (The tangent in
 * psi0 is zero.) */
int dpf_richards_pullback(dpf_state *state, const double *delta, double *dKs,
                          double *ddirichletpsis, double *dalphas, double *dNs, double *dQs,
                          const dpf_opts *opts, dpf_stats *stats);

/* Advection-dispersion (src/transport.jl:2-33), FP64, single-system
 * topologies (their Dirichlet cells are the transport's).  Per-edge velocity
 * components vx, vy, vz and dispersion Ds (nedges); cell coordinates cx, cy,
 * cz, volumes, Qs, dirichletus (n); use_upwind 1 (the reference's default) or
 * 0 (downwind).
 *
 * transport_residuals: r (n) at u (n), Dirichlet rows u - dirichletus. */
int dpf_transport_residuals(dpf_topology *topo, const double *u, const double *vx,
                            const double *vy, const double *vz, const double *Ds,
                            const double *dirichletus, const double *Qs, const double *volumes,
                            const double *cx, const double *cy, const double *cz,
                            int32_t use_upwind, double *r_out, const dpf_opts *opts);

/* du/dt = transport_residuals(u) from t0 to t1 by nsteps fixed backward-Euler
 * steps (the reference integrates with DifferentiableBackwardEuler.steps,
 * adaptive, test/transport.jl:30-33); each step solves (I - dt J) u' = u +
 * dt r(0) by Jacobi-preconditioned BiCGStab to opts->reltol.  u_final (n);
 * u_all ((nsteps + 1) x n, may be NULL) every state.  state_out (may be NULL)
 * keeps the trajectory for dpf_transport_pullback.  dt = 1 makes the Dirichlet
 * rows singular, as in the reference's equations. */
int dpf_transport_steps(dpf_topology *topo, const double *u0, const double *vx, const double *vy,
                        const double *vz, const double *Ds, const double *dirichletus,
                        const double *Qs, const double *volumes, const double *cx,
                        const double *cy, const double *cz, int32_t use_upwind, double t0,
                        double t1, int32_t nsteps, double *u_final, double *u_all,
                        const dpf_opts *opts, dpf_stats *stats, dpf_state **state_out);

/* Gradient of an objective of u_final (dJ_du_final, n) through those steps
 * (the discrete adjoint; test/transport.jl:50-70 takes the same gradient with
 * Zygote): dvx, dvy, dvz, dDs (nedges), dQs (n), any of them NULL.  Upwind
 * choices are held fixed. */
int dpf_transport_pullback(dpf_state *state, const double *dJ_du_final, double *dvx, double *dvy,
                           double *dvz, double *dDs, double *dQs, const dpf_opts *opts,
                           dpf_stats *stats);

/* Reverse mode of dpf_solvetwophase: the gradient that test/twophase.jl:192-205
 * takes with Zygote through the rrules of solvepressure and saturationcalc
 * (src/twophase.jl:59-80,117-183).  Given dJ/dP and dJ/dS of the last step
 * (either may be NULL = 0), dK (n) = dJ/dK and dS0 (n, may be NULL) = dJ/dS0.
 * Sub-step counts, upwind choices and the clamp's activity are held fixed (they
 * are piecewise constant in K; the clamp passes the tangent where
 * swc <= u <= 1 - sor).  One adjoint pressure solve per step, at that step's
 * operator; the sub-steps are recomputed from the tape. */
int dpf_solvetwophase_pullback(dpf_state *state, const double *dP_final, const double *dS_final,
                               double *dK, double *dS0, const dpf_opts *opts, dpf_stats *stats);

/* groundwater_steadystate(Ks, neighbors, areasoverlengths, dirichletnodes,
 * dirichleths, Qs; kwargs...) (src/groundwater.jl:75-83) and the forward
 * half of its rrule (src/groundwater.jl:85-92): assemble A = dr/dh and
 * b = -r(0) at Ss = V = 1, solve with AMG-preconditioned CG, scatter.
 *   Ks: nedges;  dirichleths, Qs: n;  h_out: n.
 * If state_out is non-NULL it receives a state for dpf_steady_pullback
 * (release with dpf_state_destroy).  stats may be NULL. */
int dpf_steady_forward(dpf_topology *topo, const double *Ks, const double *dirichleths,
                       const double *Qs, double *h_out, const dpf_opts *opts,
                       dpf_stats *stats, dpf_state **state_out);

/* The rrule pullback (src/groundwater.jl:93-107): lambda = A'^-1 delta[free]
 * through the same hierarchy (A is symmetric), then
 *   dKs = -(dr/dKs)' lambda,  ddirichleths = -(dr/ddh)' lambda + delta .* !free,
 *   dQs = -(dr/dQs)' lambda.
 *   delta: n;  dKs: nedges;  ddirichleths, dQs: n.  Output pointers may be
 * NULL to skip that tangent. */
int dpf_steady_pullback(dpf_state *state, const double *delta, double *dKs,
                        double *ddirichleths, double *dQs, const dpf_opts *opts,
                        dpf_stats *stats);

int dpf_state_destroy(dpf_state *state);

/* Transient groundwater flow by backward Euler: the reference integrates
 * du/dt = f(u) = -groundwater_residuals(u, Ks, ..., specificstorage,
 * volumes) with DifferentiableBackwardEuler.steps, parameters
 * p = [Ks; dirichleths] (examples/gw_transient_3d/inputdeck.jl:60-92,
 * test/theis.jl:61-106).  Step n (1..nsteps) of length dt_n solves
 * (S + dt_n A) u_n = S u_{n-1} + dt_n b(t_n), S = diag(Ss V) on free cells,
 * with Q(t_n) = qscale[n] Qs, by AMG-PCG warm-started at u_{n-1}.
 *   Ks: nedges;  dirichleths, Qs, specificstorage, volumes: n;  u0, u_final:
 *   nfree;  qscale: nsteps + 1;  states_out: NULL or (nsteps + 1) * nfree
 *   (row k = u_k, like the columns of the reference's steps matrix).
 * dpf_transient_forward takes one step length dt; dpf_transient_forward_dts
 * a caller-given sequence dts (nsteps lengths > 0, e.g. the steps an
 * adaptive integrator chose; a hierarchy refresh runs whenever the length
 * changes).  qscale and dts follow opts->memory.
 * The state keeps every checkpoint_every-th state (0 = only u_0) for the
 * pullback, which recomputes the states in between. */
int dpf_transient_forward(dpf_topology *topo, const double *Ks, const double *dirichleths,
                          const double *Qs, const double *specificstorage,
                          const double *volumes, const double *u0, int32_t nsteps, double dt,
                          const double *qscale, int32_t checkpoint_every, double *u_final,
                          double *states_out, const dpf_opts *opts, dpf_stats *stats,
                          dpf_state **state_out);
int dpf_transient_forward_dts(dpf_topology *topo, const double *Ks, const double *dirichleths,
                              const double *Qs, const double *specificstorage,
                              const double *volumes, const double *u0, int32_t nsteps,
                              const double *dts, const double *qscale, int32_t checkpoint_every,
                              double *u_final, double *states_out, const dpf_opts *opts,
                              dpf_stats *stats, dpf_state **state_out);

/* Discrete adjoint of a transient run: the reverse mode the rrule of
 * DifferentiableBackwardEuler.steps provides for the gradient in
 * p = [Ks; dirichleths] (examples/gw_transient_3d/inputdeck.jl:63,86-91,
 * 103-104), and in Qs and the initial state.  The objective may depend on
 * any states: seeds holds dJ/du_k (nfree each, row i for step
 * seed_steps[i]; seed_steps: nseed strictly ascending indices in 0..nsteps,
 * host memory; seeds follow opts->memory).  With M_n = S + dt_n A:
 *   M_n mu_n = dJ/du_n + S mu_{n+1}  (n = N..1, mu_{N+1} = 0),
 *   dKs_e          = -sum_n dt_n aol_e (u~n_a - u~n_b)(mu~n_a - mu~n_b),
 *   ddirichleths_o = sum_n dt_n sum_{e = (o, free i)} Ks_e aol_e mu_n[i]
 *                    (Dirichlet cells; 0 on free cells),
 *   dQs_i          = sum_n dt_n qscale[n] mu_n[i]  (0 on Dirichlet cells),
 *   du0            = dJ/du_0 + S mu_1.
 *   dKs: nedges;  ddirichleths, dQs: n;  du0: nfree; any output may be NULL.
 * The reverse sweep starts at the last seeded step. */
int dpf_transient_pullback_seeds(dpf_state *state, int32_t nseed, const int32_t *seed_steps,
                                 const double *seeds, double *dKs, double *ddirichleths,
                                 double *dQs, double *du0, const dpf_opts *opts,
                                 dpf_stats *stats);

/* dpf_transient_pullback_seeds for an objective of the final state alone:
 *   dJ_du_final: nfree;  dKs: nedges;  dQs: n or NULL. */
int dpf_transient_pullback(dpf_state *state, const double *dJ_du_final, double *dKs,
                           double *dQs, const dpf_opts *opts, dpf_stats *stats);

/* The kernel table of calls made with DPF_FLAG_KERNEL_TIMING (no counterpart
 * in the reference): one record per kernel class, AMG level (-1 where none
 * applies) and value size (4 or 8 bytes), with the launches, their summed
 * duration (ms, from dispatch timestamps) and summed algorithmic bytes (the
 * bytes a launch must move: operator values and the columns it stores,
 * vectors read and written once each, index maps of the gathers).
 * *count = records in the table; at most cap are written to out.  reset != 0
 * empties the table afterwards. */
typedef struct {
  char kernel[32];
  int32_t level;
  int32_t value_bytes;
  int64_t launches;
  double ms;
  double bytes;
} dpf_kernel_record;
int dpf_kernel_table(dpf_topology *topo, dpf_kernel_record *out, int32_t cap, int32_t *count,
                     int32_t reset);

/* Kernel timing for roofline reporting (no counterpart in the reference).
 * Launches one kernel of the current hierarchy `reps` times back to back
 * between two hipEvents on opts->stream (or the topology stream) and returns
 * the mean duration per launch and the algorithmic bytes one launch moves:
 *   kernel 0: finest-level Chebyshev sweep as the V-cycle runs it at finest
 *             degree 1: the post-smoothing sweep after the coarse correction,
 *             x' = x + c2 D^-1 (b - A x) (c1 = 0: no d read, none written),
 *             with the fused b.x' block partials
 *             bytes = nnz w + cols + 4 nfree w     (w = 8 FP64, 4 FP32)
 *   kernel 1: CG direction + SpMV  p' = z + beta p, q = A p'
 *             bytes = nnz w + cols + 4 nfree w
 *   kernel 2: FP64 refinement residual  r = b - A x
 *             bytes = nnz 8 + cols + 3 nfree 8
 * where cols = 4 x (nonzeros of the fine operator's slices that store column
 * indices; regular-grid slices store column offsets and no indices).
 * precision selects the FP64 or the FP32 (mixed) hierarchy; it must have been
 * used by a solve on this topology. */
int dpf_profile_kernel(dpf_topology *topo, int32_t precision, int32_t kernel, int32_t reps,
                       double *avg_ms, double *bytes, const dpf_opts *opts);

/* Cost of a grid-wide barrier inside one cooperative launch (no counterpart
 * in the reference): nparts workgroups -- on every XCD, or with one_xcd only
 * those the dispatcher places on one XCD -- run `rounds` rounds of (write a
 * word, barrier, read a neighbour's word, barrier); us = mean time per
 * barrier, errors = 0, or bit 1 when a bounded wait ran out, bit 2 when a
 * read missed the neighbour's write.  The measurement behind the choice of
 * kernel boundaries over a persistent V-cycle kernel. */
int dpf_profile_grid_barrier(int32_t nparts, int32_t one_xcd, int32_t rounds, double *us,
                             int32_t *errors);

#ifdef __cplusplus
}
#endif

#endif /* DPFGPU_H */
