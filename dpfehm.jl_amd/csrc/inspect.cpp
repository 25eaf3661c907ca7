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

// Sizes (element counts) of the arrays dpf_inspect_get can return.
int64_t dpf_inspect_size(const dpf_inspect* h, int32_t level, const char* name) {
  const TopologyHost& T = h->H;
  std::string s(name);
  if (s == "n2f") return (int64_t)T.n2f.size();
  if (s == "f2n") return (int64_t)T.f2n.size();
  if (s == "inc_ptr") return (int64_t)T.inc_ptr.size();
  if (s == "inc_edge") return (int64_t)T.inc_edge.size();
  if (s == "inc_other") return (int64_t)T.inc_other.size();
  if (s == "inc_slot") return (int64_t)T.inc_slot.size();
  if (level < 0 || level >= (int32_t)T.levels.size()) return -1;
  const LevelHost& L = T.levels[level];
  if (s == "A.sptr") return (int64_t)L.A.sptr.size();
  if (s == "A.col") return (int64_t)L.A.col.size();
  if (s == "A.rowlen") return (int64_t)L.A.rowlen.size();
  if (s == "diag_slot") return (int64_t)L.diag_slot.size();
  if (s == "agg") return (int64_t)L.agg.size();
  if (s == "P.sptr") return (int64_t)L.P.sptr.size();
  if (s == "P.col") return (int64_t)L.P.col.size();
  if (s == "P.rowlen") return (int64_t)L.P.rowlen.size();
  if (s == "p_slot") return (int64_t)L.p_slot.size();
  if (s == "p_self") return (int64_t)L.p_self.size();
  if (s == "R.sptr") return (int64_t)L.R.sptr.size();
  if (s == "R.col") return (int64_t)L.R.col.size();
  if (s == "R.rowlen") return (int64_t)L.R.rowlen.size();
  if (s == "r_src") return (int64_t)L.r_src.size();
  if (s == "ap_rp") return (int64_t)L.ap_rp.size();
  if (s == "ap_col") return (int64_t)L.ap_col.size();
  if (s == "ap_pp") return (int64_t)L.ap_pp.size();
  if (s == "ap_map") return (int64_t)L.ap_map.size();
  if (s == "rap_pp") return (int64_t)L.rap_pp.size();
  if (s == "rap_map") return (int64_t)L.rap_map.size();
  return -1;
}

int dpf_inspect_get(const dpf_inspect* h, int32_t level, const char* name, void* out,
                    int64_t cap) {
  const TopologyHost& T = h->H;
  std::string s(name);
  if (s == "n2f") return put(T.n2f, out, cap);
  if (s == "f2n") return put(T.f2n, out, cap);
  if (s == "inc_ptr") return put(T.inc_ptr, out, cap);
  if (s == "inc_edge") return put(T.inc_edge, out, cap);
  if (s == "inc_other") return put(T.inc_other, out, cap);
  if (s == "inc_slot") return put(T.inc_slot, out, cap);
  if (level < 0 || level >= (int32_t)T.levels.size()) return -1;
  const LevelHost& L = T.levels[level];
  if (s == "A.sptr") return put(L.A.sptr, out, cap);
  if (s == "A.col") return put(L.A.col, out, cap);
  if (s == "A.rowlen") return put(L.A.rowlen, out, cap);
  if (s == "diag_slot") return put(L.diag_slot, out, cap);
  if (s == "agg") return put(L.agg, out, cap);
  if (s == "P.sptr") return put(L.P.sptr, out, cap);
  if (s == "P.col") return put(L.P.col, out, cap);
  if (s == "P.rowlen") return put(L.P.rowlen, out, cap);
  if (s == "p_slot") return put(L.p_slot, out, cap);
  if (s == "p_self") return put(L.p_self, out, cap);
  if (s == "R.sptr") return put(L.R.sptr, out, cap);
  if (s == "R.col") return put(L.R.col, out, cap);
  if (s == "R.rowlen") return put(L.R.rowlen, out, cap);
  if (s == "r_src") return put(L.r_src, out, cap);
  if (s == "ap_rp") return put(L.ap_rp, out, cap);
  if (s == "ap_col") return put(L.ap_col, out, cap);
  if (s == "ap_pp") return put(L.ap_pp, out, cap);
  if (s == "ap_map") return put(L.ap_map, out, cap);
  if (s == "rap_pp") return put(L.rap_pp, out, cap);
  if (s == "rap_map") return put(L.rap_map, out, cap);
  return -1;
}

}  // extern "C"
