/*
 * dpfgpu_inspect.h -- read-only view of libdpfgpu's host symbolic setup.
 *
 * No counterpart in the reference.  It builds the same topology cache and
 * AMG patterns as dpf_topology_create, on the host only (no GPU needed), and
 * exposes the index arrays the GPU kernels consume, so that they can be
 * checked against independent sparse algebra on a machine without a GPU.
 *
 * Array names and element types (level-independent ones ignore `level`):
 *   n2f, f2n, inc_edge, inc_other: int32   inc_ptr: int64   inc_slot: uint16
 *   A.sptr, P.sptr, R.sptr, Rs.sptr, As.sptr, ap_rp, ap_off, rap_off: int64
 *   A.col, A.rowlen, P.col, P.rowlen, R.col, R.rowlen, Rs.col, Rs.rowlen, As.col,
 *   As.rowlen, agg, r_src, rs_src, as_src, ap_col, ap_a, ap_p, rap_dest, rap_p, rap_ap: int32
 *   diag_slot, p_slot, p_self: uint16
 * dpf_inspect_size(h, level, "rs_lg" | "as_lg") returns log2 of the lanes per
 * row of the split restriction / split level operator (no array);
 * "G.ap_rows" | "G.rap_rows" 1 when the level computes A*P / R*(A*P) by rows
 * (dpf_host.h GalerkinPlan), "G.ap_lg" log2 of the lanes per fine row of the
 * former; the plan's arrays are
 *   G.apc_ptr, G.rcd_ptr: int64   G.apc: uint8   G.rcd, G.rnp: uint16
 */
#ifndef DPFGPU_INSPECT_H
#define DPFGPU_INSPECT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dpf_inspect dpf_inspect;

const char *dpf_inspect_last_error(void);
int dpf_inspect_create(int64_t n, int64_t nedges, const int64_t *neighbors,
                       const double *areasoverlengths, int64_t ndirichlet,
                       const int64_t *dirichletnodes, int32_t max_levels, int32_t coarse_max,
                       dpf_inspect **out);
void dpf_inspect_destroy(dpf_inspect *h);
int32_t dpf_inspect_levels(const dpf_inspect *h);
/* The check dpf_steady_forward makes before it solves: 0, or DPF_ESINGULAR
 * (5) with the message in dpf_inspect_last_error() when some connected
 * component of the mesh holds no Dirichlet cell, through every edge or, with
 * Ks (nedges) given, through the edges with Ks[e] != 0. */
int dpf_inspect_check_steady(const dpf_inspect *h, const double *Ks);
int64_t dpf_inspect_size(const dpf_inspect *h, int32_t level, const char *name);
int dpf_inspect_get(const dpf_inspect *h, int32_t level, const char *name, void *out,
                    int64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* DPFGPU_INSPECT_H */
