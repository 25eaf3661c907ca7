// Read-only inspection of the host symbolic setup (topology maps and AMG
// patterns) without touching the GPU.  Used by the CPU test-suite to check
// the index maps the GPU kernels consume against independent sparse algebra.
#include <cstring>
#include <string>

#include "../../include/dpfgpu_inspect.h"
#include "dpf_host.h"

using namespace dpf;

struct dpf_inspect {
  TopologyHost H;
  FineLayout F;  // the fine operator's GPU layout (hierarchies of two or more levels)
  std::vector<GalerkinPlan> G;  // per level: the row-wise Galerkin plan the GPU uses
};

static thread_local std::string g_ierr;

namespace {
template <typename T>
int put(const std::vector<T>& v, void* out, int64_t cap) {
  if ((int64_t)v.size() > cap) return -1;
  if (!v.empty()) std::memcpy(out, v.data(), v.size() * sizeof(T));
  return 0;
}
}  // namespace

extern "C" {

const char* dpf_inspect_last_error(void) { return g_ierr.c_str(); }

int dpf_inspect_create(int64_t n, int64_t nedges, const int64_t* neighbors,
                       const double* areasoverlengths, int64_t ndirichlet,
                       const int64_t* dirichletnodes, int32_t max_levels, int32_t coarse_max,
                       dpf_inspect** out) {
  *out = nullptr;
  auto* h = new dpf_inspect();
  std::string e = build_topology(h->H, n, nedges, neighbors, areasoverlengths, ndirichlet,
                                 dirichletnodes);
  if (e.empty()) e = build_hierarchy(h->H, max_levels, coarse_max, 512);
  if (e.empty() && h->H.levels.size() >= 2) build_fine_layout(h->H, h->F);
  h->G.resize(h->H.levels.size());
  for (size_t l = 0; e.empty() && l + 1 < h->H.levels.size(); l++)
    e = build_galerkin_plan(h->H, (int)l, l == 0 ? &h->F : nullptr, h->G[l]);
  if (!e.empty()) {
    g_ierr = e;
    delete h;
    return 1;
  }
  *out = h;
  return 0;
}

void dpf_inspect_destroy(dpf_inspect* h) { delete h; }

int32_t dpf_inspect_levels(const dpf_inspect* h) { return (int32_t)h->H.levels.size(); }

// The steady solve's singularity check: 0, or 5 (DPF_ESINGULAR) with the
// message in dpf_inspect_last_error when a connected component -- through
// every edge, or through the edges with Ks[e] != 0 when Ks is given -- holds
// no Dirichlet cell.
int dpf_inspect_check_steady(const dpf_inspect* h, const double* Ks) {
  std::string e = h->H.floating;
  if (e.empty() && Ks) e = floating_component(h->H, Ks);
  if (e.empty()) return 0;
  g_ierr = e;
  return 5;
}

}  // extern "C"

namespace {

// name -> (size, copy) over the topology / level arrays
template <typename Fn>
int with_array(const dpf_inspect* h, int32_t level, const std::string& s, Fn&& f) {
  const TopologyHost& T = h->H;
  // the fine operator's GPU layout ("F." names; empty for one-level hierarchies)
  if (s == "F.sptr") return f(h->F.sptr);
  if (s == "F.col") return f(h->F.col);
  if (s == "F.soff") return f(h->F.soff);
  if (s == "F.inc_slot") return f(h->F.inc_slot);
  if (s == "F.diag_slot") return f(h->F.diag_slot);
  if (s == "F.p_slot") return f(h->F.p_slot);
  if (s == "F.ap_a") return f(h->F.ap_a);
  if (s == "n2f") return f(T.n2f);
  if (s == "f2n") return f(T.f2n);
  if (s == "inc_ptr") return f(T.inc_ptr);
  if (s == "inc_edge") return f(T.inc_edge);
  if (s == "inc_other") return f(T.inc_other);
  if (s == "inc_slot") return f(T.inc_slot);
  if (level < 0 || level >= (int32_t)T.levels.size()) return -1;
  const LevelHost& L = T.levels[level];
  if (s == "A.sptr") return f(L.A.sptr);
  if (s == "A.col") return f(L.A.col);
  if (s == "A.rowlen") return f(L.A.rowlen);
  if (s == "diag_slot") return f(L.diag_slot);
  if (s == "agg") return f(L.agg);
  if (s == "P.sptr") return f(L.P.sptr);
  if (s == "P.col") return f(L.P.col);
  if (s == "P.rowlen") return f(L.P.rowlen);
  if (s == "p_slot") return f(L.p_slot);
  if (s == "p_self") return f(L.p_self);
  if (s == "R.sptr") return f(L.R.sptr);
  if (s == "R.col") return f(L.R.col);
  if (s == "R.rowlen") return f(L.R.rowlen);
  if (s == "r_src") return f(L.r_src);
  if (s == "Rs.sptr") return f(L.Rs.sptr);
  if (s == "Rs.col") return f(L.Rs.col);
  if (s == "Rs.rowlen") return f(L.Rs.rowlen);
  if (s == "rs_src") return f(L.rs_src);
  if (s == "As.sptr") return f(L.As.sptr);
  if (s == "As.col") return f(L.As.col);
  if (s == "As.rowlen") return f(L.As.rowlen);
  if (s == "as_src") return f(L.as_src);
  if (s == "ap_rp") return f(L.ap_rp);
  if (s == "ap_col") return f(L.ap_col);
  if (s == "ap_off") return f(L.ap_off);
  if (s == "ap_a") return f(L.ap_a);
  if (s == "ap_p") return f(L.ap_p);
  if (s == "rap_off") return f(L.rap_off);
  if (s == "rap_p") return f(L.rap_p);
  if (s == "rap_dest") return f(L.rap_dest);
  if (s == "rap_ap") return f(L.rap_ap);
  if (s == "rap_mdst") return f(L.rap_mdst);
  if (s == "rap_msrc") return f(L.rap_msrc);
  // the level's row-wise Galerkin plan ("G." names; empty where it uses lists)
  const GalerkinPlan& G = h->G[level];
  if (s == "G.apc_ptr") return f(G.apc_ptr);
  if (s == "G.apc") return f(G.apc);
  if (s == "G.rcd_ptr") return f(G.rcd_ptr);
  if (s == "G.rcd") return f(G.rcd);
  if (s == "G.rnp") return f(G.rnp);
  return -1;
}

}  // namespace

extern "C" {

// Sizes (element counts) of the arrays dpf_inspect_get can return; the
// pseudo-names "rs_lg" / "as_lg" return log2 of the lanes per row of the
// split restriction / split level operator, "G.ap_rows" / "G.rap_rows" 1 when
// the level computes A*P / R*(A*P) row-wise, "G.ap_lg" log2 of the lanes per
// fine row of the row-wise A*P.
int64_t dpf_inspect_size(const dpf_inspect* h, int32_t level, const char* name) {
  std::string s(name);
  if (s == "rs_lg" || s == "as_lg") {
    if (level < 0 || level >= (int32_t)h->H.levels.size()) return -1;
    return s == "rs_lg" ? h->H.levels[level].rs_lg : h->H.levels[level].as_lg;
  }
  if (s == "G.ap_rows" || s == "G.rap_rows" || s == "G.ap_lg") {
    if (level < 0 || level >= (int32_t)h->G.size()) return -1;
    const GalerkinPlan& G = h->G[level];
    return s == "G.ap_rows" ? G.ap_rows : (s == "G.rap_rows" ? G.rap_rows : G.ap_lg);
  }
  int64_t out = -1;
  with_array(h, level, s, [&](const auto& v) {
    out = (int64_t)v.size();
    return 0;
  });
  return out;
}

int dpf_inspect_get(const dpf_inspect* h, int32_t level, const char* name, void* out,
                    int64_t cap) {
  return with_array(h, level, std::string(name),
                    [&](const auto& v) { return put(v, out, cap); });
}

}  // extern "C"
