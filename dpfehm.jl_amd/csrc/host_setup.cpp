// Host-side topology cache and smoothed-aggregation symbolic setup.
//
// Topology: free-cell maps of getfreenodes (src/utilities.jl:10-24), the
// cell -> edge incidence in edge-list order (the order the residual of
// src/groundwater.jl:11-20 visits edges in), and the pattern of the free-cell
// Jacobian dr/dh with a slot map from every (cell, incident edge) to the
// nonzero it feeds, so that assembly on the GPU is a deterministic row
// gather with no atomics.
//
// Hierarchy: smoothed aggregation on the matrix graph (all couplings strong),
// the pattern of P = (I - w D^-1 A) T, R = P', A*P and R*A*P, with product
// slot maps so the GPU recomputes every level's values for new
// conductivities without any symbolic work.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "dpf_host.h"

namespace dpf {

void SellHost::build(int64_t nr, int64_t nc, const std::vector<int64_t>& rp,
                     const std::vector<int32_t>& ci, bool square) {
  nrows = nr;
  ncols = nc;
  int64_t ns = (nr + SELL_C - 1) / SELL_C;
  sptr.assign(ns + 1, 0);
  rowlen.assign(nr, 0);
  std::vector<int> width(ns, 0);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < ns; s++) {
    int w = 0;
    for (int64_t r = s * SELL_C; r < std::min<int64_t>(nr, (s + 1) * SELL_C); r++) {
      int len = (int)(rp[r + 1] - rp[r]);
      rowlen[r] = len;
      w = std::max(w, len);
    }
    width[s] = w;
  }
  max_rowlen = 0;
  for (int64_t s = 0; s < ns; s++) {
    max_rowlen = std::max(max_rowlen, width[s]);
    sptr[s + 1] = sptr[s] + (int64_t)width[s] * SELL_C;
  }
  nnz = rp[nr] - rp[0];
  col.assign(sptr[ns], 0);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < ns; s++) {
    int w = width[s];
    for (int l = 0; l < SELL_C; l++) {
      int64_t r = s * SELL_C + l;
      int32_t pad = (square && r < nr) ? (int32_t)r : 0;
      for (int k = 0; k < w; k++) {
        int64_t e = sptr[s] + (int64_t)k * SELL_C + l;
        col[e] = (r < nr && k < rowlen[r]) ? ci[rp[r] + k] : pad;
      }
    }
  }
}

std::string build_topology(TopologyHost& T, int64_t n, int64_t ne, const int64_t* nb,
                           const double* aol, int64_t nd, const int64_t* dn) {
  if (n <= 0) return "n must be positive";
  if (n >= INT32_MAX || ne >= INT32_MAX) return "too many cells or edges for 32-bit indices";
  if (ne < 0 || nd < 0) return "negative size";
  T.n = n;
  T.ne = ne;
  T.edges.resize(2 * ne);
  T.aol.assign(aol, aol + ne);
  for (int64_t e = 0; e < ne; e++) {
    int64_t a = nb[2 * e], b = nb[2 * e + 1];
    if (a < 1 || a > n || b < 1 || b > n)
      return "neighbors[" + std::to_string(e + 1) + "] refers to a cell outside 1.." +
             std::to_string(n);
    T.edges[2 * e] = (int32_t)(a - 1);
    T.edges[2 * e + 1] = (int32_t)(b - 1);
  }
  T.isfree.assign(n, 1);
  for (int64_t k = 0; k < nd; k++) {
    if (dn[k] < 1 || dn[k] > n)
      return "dirichletnodes[" + std::to_string(k + 1) + "] outside 1.." + std::to_string(n);
    T.isfree[dn[k] - 1] = 0;
  }
  T.n2f.assign(n, -1);
  T.f2n.clear();
  for (int64_t i = 0; i < n; i++)
    if (T.isfree[i]) {
      T.n2f[i] = (int32_t)T.f2n.size();
      T.f2n.push_back((int32_t)i);
    }
  T.nf = (int64_t)T.f2n.size();
  T.nd = n - T.nf;
  if (T.nf == 0) return "every cell is a Dirichlet cell";

  // incidence, edges in list order per cell
  T.inc_ptr.assign(n + 1, 0);
  for (int64_t e = 0; e < ne; e++) {
    int32_t a = T.edges[2 * e], b = T.edges[2 * e + 1];
    if (a == b) continue;
    T.inc_ptr[a + 1]++;
    T.inc_ptr[b + 1]++;
  }
  for (int64_t i = 0; i < n; i++) T.inc_ptr[i + 1] += T.inc_ptr[i];
  int64_t ninc = T.inc_ptr[n];
  T.inc_edge.resize(ninc);
  T.inc_other.resize(ninc);
  T.inc_slot.assign(ninc, NO_SLOT);
  std::vector<int64_t> pos(T.inc_ptr.begin(), T.inc_ptr.end() - 1);
  for (int64_t e = 0; e < ne; e++) {
    int32_t a = T.edges[2 * e], b = T.edges[2 * e + 1];
    if (a == b) continue;
    T.inc_edge[pos[a]] = (int32_t)e;
    T.inc_other[pos[a]++] = b;
    T.inc_edge[pos[b]] = (int32_t)e;
    T.inc_other[pos[b]++] = a;
  }

  // fine pattern: row j = {j} U {free neighbours}, sorted unique
  int64_t nf = T.nf;
  std::vector<int64_t> rp(nf + 1, 0);
  std::vector<int32_t> ci;
  ci.reserve(ninc + nf);
  std::vector<int32_t> row;
  for (int64_t j = 0; j < nf; j++) {
    int32_t i = T.f2n[j];
    row.clear();
    row.push_back((int32_t)j);
    for (int64_t p = T.inc_ptr[i]; p < T.inc_ptr[i + 1]; p++) {
      int32_t o = T.inc_other[p];
      if (T.isfree[o]) row.push_back(T.n2f[o]);
    }
    std::sort(row.begin(), row.end());
    row.erase(std::unique(row.begin(), row.end()), row.end());
    if (row.size() >= NO_SLOT) return "a cell has too many distinct neighbours";
    ci.insert(ci.end(), row.begin(), row.end());
    rp[j + 1] = (int64_t)ci.size();
  }
  for (int64_t j = 0; j < nf; j++) {
    int32_t i = T.f2n[j];
    if (T.inc_ptr[i + 1] == T.inc_ptr[i])
      return "free cell " + std::to_string(i + 1) + " has no edges: the system is singular";
  }
  T.floating = floating_component(T, nullptr);
  T.levels.assign(1, LevelHost());
  LevelHost& L0 = T.levels[0];
  L0.A.build(nf, nf, rp, ci, true);
  L0.diag_slot.resize(nf);
  for (int64_t j = 0; j < nf; j++) {
    int64_t s = rp[j], t = rp[j + 1];
    for (int64_t k = s; k < t; k++)
      if (ci[k] == j) L0.diag_slot[j] = (uint16_t)(k - s);
    int32_t i = T.f2n[j];
    for (int64_t p = T.inc_ptr[i]; p < T.inc_ptr[i + 1]; p++) {
      int32_t o = T.inc_other[p];
      if (!T.isfree[o]) continue;
      int32_t c = T.n2f[o];
      auto it = std::lower_bound(ci.begin() + s, ci.begin() + t, c);
      T.inc_slot[p] = (uint16_t)(it - (ci.begin() + s));
    }
  }
  return "";
}

namespace {

// Standard SA aggregation on the graph of A (pattern only), over the rows
// [r0, r1) (a range closed under the graph: a whole system or one batch
// member): roots whose whole neighbourhood is free start an aggregate;
// leftovers join a neighbouring phase-1 aggregate; the rest form new
// aggregates with their unaggregated neighbours.  Aggregates are numbered
// from `base`.  Returns how many were made.
int64_t aggregate_range(const SellHost& A, int64_t r0, int64_t r1, int64_t base,
                        std::vector<int32_t>& agg) {
  int64_t nc = 0;
  for (int64_t i = r0; i < r1; i++) agg[i] = -1;
  for (int64_t i = r0; i < r1; i++) {
    if (agg[i] != -1) continue;
    bool ok = true;
    for (int s = 0; s < A.rowlen[i] && ok; s++) {
      int32_t j = A.at(i, s);
      if (j != i && agg[j] != -1) ok = false;
    }
    if (!ok) continue;
    for (int s = 0; s < A.rowlen[i]; s++) agg[A.at(i, s)] = (int32_t)(base + nc);
    agg[i] = (int32_t)(base + nc);
    nc++;
  }
  std::vector<int32_t> agg1(agg.begin() + r0, agg.begin() + r1);
  for (int64_t i = r0; i < r1; i++) {
    if (agg[i] != -1) continue;
    for (int s = 0; s < A.rowlen[i]; s++) {
      int32_t j = A.at(i, s);
      if (agg1[j - r0] != -1) {
        agg[i] = agg1[j - r0];
        break;
      }
    }
  }
  for (int64_t i = r0; i < r1; i++) {
    if (agg[i] != -1) continue;
    agg[i] = (int32_t)(base + nc);
    for (int s = 0; s < A.rowlen[i]; s++) {
      int32_t j = A.at(i, s);
      if (agg[j] == -1) agg[j] = (int32_t)(base + nc);
    }
    nc++;
  }
  return nc;
}

// Aggregates every member range independently (in parallel); member m's
// aggregates are numbered after those of members < m, so each member also
// owns a contiguous row range on the next level (returned in next_row0).
int64_t aggregate(const SellHost& A, const std::vector<int64_t>& row0, std::vector<int32_t>& agg,
                  std::vector<int64_t>& next_row0) {
  int64_t nm = (int64_t)row0.size() - 1;
  agg.assign(A.nrows, -1);
  std::vector<int64_t> cnt(nm, 0);
  // aggregate each member from 0, then shift by the members before it
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t m = 0; m < nm; m++) cnt[m] = aggregate_range(A, row0[m], row0[m + 1], 0, agg);
  next_row0.assign(nm + 1, 0);
  for (int64_t m = 0; m < nm; m++) next_row0[m + 1] = next_row0[m] + cnt[m];
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < nm; m++)
    for (int64_t i = row0[m]; i < row0[m + 1]; i++) agg[i] += (int32_t)next_row0[m];
  return next_row0[nm];
}

// Row-wise unique-column accumulator keyed by a stamp array.
struct RowSet {
  std::vector<int64_t> stamp;
  std::vector<int32_t> slot;
  std::vector<int32_t> cols;
  explicit RowSet(int64_t ncols) : stamp(ncols, -1), slot(ncols, 0) {}
  void begin() { cols.clear(); }
  void add(int32_t c, int64_t row) {
    if (stamp[c] != row) {
      stamp[c] = row;
      cols.push_back(c);
    }
  }
  // sort and record the slot of every column; the stamp stays set for lookups
  void finish() {
    std::sort(cols.begin(), cols.end());
    for (size_t k = 0; k < cols.size(); k++) slot[cols[k]] = (int32_t)k;
  }
};

// Builds a sorted, duplicate-free row pattern in parallel.  gen(i, rs)
// adds row i's candidate columns to rs (rs.add(c, i)); after(i, rs) may use
// rs.slot for the row's columns.  Returns false if a row exceeds maxlen.
template <typename Gen, typename After>
bool par_pattern(int64_t nrows, int64_t ncols, Gen gen, After after, std::vector<int64_t>& rp,
                 std::vector<int32_t>& ci, size_t maxlen) {
  rp.assign(nrows + 1, 0);
  bool bad = false;
#pragma omp parallel
  {
    RowSet rs(ncols);
#pragma omp for schedule(dynamic, 2048)
    for (int64_t i = 0; i < nrows; i++) {
      rs.begin();
      gen(i, rs);
      rp[i + 1] = (int64_t)rs.cols.size();
      if (rs.cols.size() >= maxlen) bad = true;
    }
  }
  if (bad) return false;
  for (int64_t i = 0; i < nrows; i++) rp[i + 1] += rp[i];
  ci.resize(rp[nrows]);
#pragma omp parallel
  {
    RowSet rs(ncols);
#pragma omp for schedule(dynamic, 2048)
    for (int64_t i = 0; i < nrows; i++) {
      rs.begin();
      gen(i, rs);
      rs.finish();
      std::copy(rs.cols.begin(), rs.cols.end(), ci.begin() + rp[i]);
      after(i, rs);
    }
  }
  return true;
}

// Exclusive prefix sum of counts stored at [1..n] of v (v[0] = 0).
void prefix(std::vector<int64_t>& v) {
  for (size_t i = 1; i < v.size(); i++) v[i] += v[i - 1];
}

std::string build_transfer(LevelHost& L, LevelHost& next, int64_t nc) {
  const SellHost& A = L.A;
  int64_t n = A.nrows;
  // P pattern: aggregates seen from each row (including the row's own)
  {
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    L.p_slot.assign(A.entries(), NO_SLOT);
    L.p_self.resize(n);
    bool ok = par_pattern(
        n, nc,
        [&](int64_t i, RowSet& rs) {
          rs.add(L.agg[i], i);
          for (int s = 0; s < A.rowlen[i]; s++) rs.add(L.agg[A.at(i, s)], i);
        },
        [&](int64_t i, RowSet& rs) {
          for (int s = 0; s < A.rowlen[i]; s++)
            L.p_slot[A.idx(i, s)] = (uint16_t)rs.slot[L.agg[A.at(i, s)]];
          L.p_self[i] = (uint16_t)rs.slot[L.agg[i]];
        },
        rp, ci, NO_SLOT);
    if (!ok) return "prolongator row too long";
    L.P.build(n, nc, rp, ci, false);
  }
  const SellHost& P = L.P;
  // R = P' (columns of R = fine rows, ascending by construction)
  {
    std::vector<int64_t> rp(nc + 1, 0);
    for (int64_t i = 0; i < n; i++)
      for (int s = 0; s < P.rowlen[i]; s++) rp[P.at(i, s) + 1]++;
    prefix(rp);
    std::vector<int32_t> ci(rp[nc]);
    std::vector<int64_t> src(rp[nc]);
    std::vector<int64_t> pos(rp.begin(), rp.end() - 1);
    for (int64_t i = 0; i < n; i++)
      for (int s = 0; s < P.rowlen[i]; s++) {
        int32_t I = P.at(i, s);
        ci[pos[I]] = (int32_t)i;
        src[pos[I]++] = P.idx(i, s);
      }
    for (int64_t I = 0; I < nc; I++)
      if (rp[I + 1] - rp[I] >= NO_SLOT) return "restriction row too long";
    L.R.build(nc, n, rp, ci, false);
    L.r_src.assign(L.R.entries(), -1);
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < nc; I++)
      for (int t = 0; t < L.R.rowlen[I]; t++) L.r_src[L.R.idx(I, t)] = (int32_t)src[rp[I] + t];
  }
  const SellHost& R = L.R;
  // R split over G lanes per row (sub-row I*G+g: entries g, g+G, ...)
  {
    double avg = nc > 0 ? (double)R.nnz / (double)nc : 1.0;
    int lg = 0;
    while ((1 << lg) * 4 < avg && lg < 6) lg++;
    L.rs_lg = lg;
    int G = 1 << lg;
    int64_t ns = nc * G;
    std::vector<int64_t> rp(ns + 1, 0);
    for (int64_t I = 0; I < nc; I++)
      for (int g = 0; g < G; g++) rp[I * G + g + 1] = (R.rowlen[I] - g + G - 1) / G;
    prefix(rp);
    std::vector<int32_t> ci(rp[ns]);
    std::vector<int32_t> src(rp[ns]);
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < nc; I++)
      for (int g = 0; g < G; g++) {
        int64_t q = rp[I * G + g];
        for (int t = g; t < R.rowlen[I]; t += G, q++) {
          ci[q] = R.at(I, t);
          src[q] = L.r_src[R.idx(I, t)];
        }
      }
    L.Rs.build(ns, n, rp, ci, false);
    L.rs_src.assign(L.Rs.entries(), -1);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < ns; r++)
      for (int k = 0; k < L.Rs.rowlen[r]; k++) L.rs_src[L.Rs.idx(r, k)] = src[rp[r] + k];
  }
  // A*P pattern; products grouped by destination nonzero
  {
    bool ok = par_pattern(
        n, nc,
        [&](int64_t i, RowSet& rs) {
          for (int s = 0; s < A.rowlen[i]; s++) {
            int32_t k = A.at(i, s);
            for (int m = 0; m < P.rowlen[k]; m++) rs.add(P.at(k, m), i);
          }
        },
        [&](int64_t, RowSet&) {}, L.ap_rp, L.ap_col, NO_SLOT);
    if (!ok) return "A*P row too long";
    int64_t nap = L.ap_rp[n];
    if (nap >= INT32_MAX) return "A*P too large for 32-bit indices";
    L.ap_off.assign(nap + 1, 0);
    // products of row i in enumeration order: (s, m); counts per destination
#pragma omp parallel
    {
      std::vector<int32_t> slot_of(nc, 0);
#pragma omp for schedule(dynamic, 2048)
      for (int64_t i = 0; i < n; i++) {
        for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++)
          slot_of[L.ap_col[a]] = (int32_t)(a - L.ap_rp[i]);
        for (int s = 0; s < A.rowlen[i]; s++) {
          int32_t k = A.at(i, s);
          for (int m = 0; m < P.rowlen[k]; m++) L.ap_off[L.ap_rp[i] + slot_of[P.at(k, m)] + 1]++;
        }
      }
    }
    prefix(L.ap_off);
    int64_t nprod = L.ap_off[nap];
    if (nprod >= INT32_MAX) return "too many A*P products";
    L.ap_a.resize(nprod);
    L.ap_p.resize(nprod);
#pragma omp parallel
    {
      std::vector<int32_t> slot_of(nc, 0);
      std::vector<int64_t> fill;
#pragma omp for schedule(dynamic, 2048)
      for (int64_t i = 0; i < n; i++) {
        int64_t a0 = L.ap_rp[i], a1 = L.ap_rp[i + 1];
        fill.assign(L.ap_off.begin() + a0, L.ap_off.begin() + a1);
        for (int64_t a = a0; a < a1; a++) slot_of[L.ap_col[a]] = (int32_t)(a - a0);
        for (int s = 0; s < A.rowlen[i]; s++) {
          int32_t k = A.at(i, s);
          int64_t ea = A.idx(i, s);
          for (int m = 0; m < P.rowlen[k]; m++) {
            int64_t q = fill[slot_of[P.at(k, m)]]++;
            L.ap_a[q] = (int32_t)ea;
            L.ap_p[q] = (int32_t)P.idx(k, m);
          }
        }
      }
    }
  }
  // R*(A*P): the next level's pattern, then products grouped by its SELL
  // entries (lists numbered row-major over every slot, padding included)
  {
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    bool ok = par_pattern(
        nc, nc,
        [&](int64_t I, RowSet& rs) {
          rs.add((int32_t)I, I);
          for (int t = 0; t < R.rowlen[I]; t++) {
            int32_t i = R.at(I, t);
            for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++) rs.add(L.ap_col[a], I);
          }
        },
        [&](int64_t, RowSet&) {}, rp, ci, NO_SLOT);
    if (!ok) return "coarse operator row too long";
    next = LevelHost();
    next.A.build(nc, nc, rp, ci, true);
    next.diag_slot.resize(nc);
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < nc; I++) {
      auto b = ci.begin() + rp[I], e = ci.begin() + rp[I + 1];
      next.diag_slot[I] = (uint16_t)(std::lower_bound(b, e, (int32_t)I) - b);
    }
    const SellHost& C = next.A;
    // one contribution list per entry on or above the diagonal, in row-major
    // order (row I's such entries are its slots diag_slot[I] ..)
    std::vector<int64_t> up_base(nc + 1, 0);
    for (int64_t I = 0; I < nc; I++) up_base[I + 1] = up_base[I] + (rp[I + 1] - rp[I]) - next.diag_slot[I];
    int64_t nent = up_base[nc];
    L.rap_dest.resize(nent);
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < nc; I++)
      for (int64_t u = up_base[I]; u < up_base[I + 1]; u++)
        L.rap_dest[u] = (int32_t)C.idx(I, (int)(next.diag_slot[I] + (u - up_base[I])));
    L.rap_off.assign(nent + 1, 0);
#pragma omp parallel
    {
      std::vector<int32_t> slot_of(nc, 0);
#pragma omp for schedule(dynamic, 512)
      for (int64_t I = 0; I < nc; I++) {
        for (int64_t k = rp[I]; k < rp[I + 1]; k++) slot_of[ci[k]] = (int32_t)(k - rp[I]);
        int64_t ub = up_base[I] - next.diag_slot[I];
        for (int t = 0; t < R.rowlen[I]; t++) {
          int32_t i = R.at(I, t);
          for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++)
            if (L.ap_col[a] >= I) L.rap_off[ub + slot_of[L.ap_col[a]] + 1]++;
        }
      }
    }
    prefix(L.rap_off);
    int64_t nprod = L.rap_off[nent];
    if (nprod >= INT32_MAX) return "too many R*A*P products";
    L.rap_p.resize(nprod);
    L.rap_ap.resize(nprod);
#pragma omp parallel
    {
      std::vector<int32_t> slot_of(nc, 0);
      std::vector<int64_t> fill;
#pragma omp for schedule(dynamic, 512)
      for (int64_t I = 0; I < nc; I++) {
        int64_t b0 = up_base[I], b1 = up_base[I + 1];
        int ds = next.diag_slot[I];
        fill.assign(L.rap_off.begin() + b0, L.rap_off.begin() + b1);
        for (int64_t k = rp[I]; k < rp[I + 1]; k++) slot_of[ci[k]] = (int32_t)(k - rp[I]) - ds;
        for (int t = 0; t < R.rowlen[I]; t++) {
          int32_t i = R.at(I, t);
          int32_t pe = L.r_src[R.idx(I, t)];
          for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++) {
            if (L.ap_col[a] < I) continue;  // below the diagonal: mirrored
            int64_t q = fill[slot_of[L.ap_col[a]]]++;
            L.rap_p[q] = pe;
            L.rap_ap[q] = (int32_t)a;
          }
        }
      }
    }
    // entries below the diagonal and the positions of their mirrors (the
    // pattern is symmetric: column I of row J exists whenever column J of
    // row I does)
    std::vector<int64_t> mo(nc + 1, 0);  // per row: its entries below the diagonal
#pragma omp parallel for schedule(static)
    for (int64_t I = 0; I < nc; I++) {
      int64_t c = 0;
      for (int64_t k = rp[I]; k < rp[I + 1] && ci[k] < I; k++) c++;
      mo[I + 1] = c;
    }
    prefix(mo);
    L.rap_mdst.assign(mo[nc], 0);
    L.rap_msrc.assign(mo[nc], 0);
    bool sym = true;
#pragma omp parallel for schedule(dynamic, 512) reduction(&& : sym)
    for (int64_t I = 0; I < nc; I++)
      for (int64_t k = rp[I]; k < rp[I] + (mo[I + 1] - mo[I]); k++) {
        int32_t J = ci[k];
        auto b = ci.begin() + rp[J], e = ci.begin() + rp[J + 1];
        auto it = std::lower_bound(b, e, (int32_t)I);
        if (it == e || *it != (int32_t)I) {
          sym = false;
          continue;
        }
        int64_t m = mo[I] + (k - rp[I]);
        L.rap_mdst[m] = (int32_t)C.idx(I, (int)(k - rp[I]));
        L.rap_msrc[m] = (int32_t)C.idx(J, (int)(it - b));
      }
    if (!sym) return "coarse operator pattern is not symmetric";
    auto lanes = [](int64_t nprod_, int64_t nlists) {
      double avg = nlists > 0 ? (double)nprod_ / (double)nlists : 1.0;
      int lg = 0;
      while ((1 << lg) * 4 < avg && lg < 5) lg++;
      return lg;
    };
    L.rap_lg = lanes(nprod, nent);
    L.ap_lg = lanes(L.ap_off.back(), (int64_t)L.ap_col.size());
  }
  return "";
}

}  // namespace

std::string floating_component(const TopologyHost& T, const double* Ks) {
  const int64_t n = T.n;
  std::vector<int32_t> up(n);
  for (int64_t i = 0; i < n; i++) up[i] = (int32_t)i;
  auto find = [&](int32_t a) {
    while (up[a] != a) {
      up[a] = up[up[a]];
      a = up[a];
    }
    return a;
  };
  for (int64_t e = 0; e < T.ne; e++) {
    // a zero conductance K_e aol_e connects nothing
    if (T.aol[e] * (Ks ? Ks[e] : 1.0) == 0.0) continue;
    int32_t a = find(T.edges[2 * e]), b = find(T.edges[2 * e + 1]);
    if (a != b) up[std::max(a, b)] = std::min(a, b);
  }
  std::vector<uint8_t> anchored(n, 0);
  std::vector<int64_t> size(n, 0);
  for (int64_t i = 0; i < n; i++) {
    int32_t r = find((int32_t)i);
    size[r]++;
    if (!T.isfree[i]) anchored[r] = 1;
  }
  for (int64_t i = 0; i < n; i++) {
    int32_t r = find((int32_t)i);
    if (!anchored[r])
      return "free cell " + std::to_string(i + 1) + " lies in a component of " +
             std::to_string(size[r]) + " cells that " +
             (Ks ? "no edge with K_e aol_e != 0 connects to a Dirichlet cell"
                 : "holds no Dirichlet cell") +
             ": the steady system is singular";
  }
  return "";
}

// Row-split copy of a SELL matrix: sub-row r*G + g holds the entries
// g, g + G, ... of row r, G = 2^lg: force_lg when it is given (0..6), else
// chosen so that a lane sees ~8 entries.
void split_rows(const SellHost& A, int& lg, SellHost& S, std::vector<int32_t>& src,
                int force_lg) {
  double avg = A.nrows > 0 ? (double)A.nnz / (double)A.nrows : 1.0;
  lg = 0;
  // ~8 entries a lane: one batch of eight loads in flight per lane
  if (force_lg >= 0)
    lg = std::min(force_lg, 6);
  else
    while ((1 << lg) * 8 < avg && lg < 6) lg++;
  int G = 1 << lg;
  int64_t ns = A.nrows * G;
  std::vector<int64_t> rp(ns + 1, 0);
  std::vector<int32_t> ci, si;
  ci.reserve(A.nnz);
  si.reserve(A.nnz);
  for (int64_t r = 0; r < A.nrows; r++)
    for (int g = 0; g < G; g++) {
      for (int t = g; t < A.rowlen[r]; t += G) {
        ci.push_back(A.at(r, t));
        si.push_back((int32_t)A.idx(r, t));
      }
      rp[r * G + g + 1] = (int64_t)ci.size();
    }
  S.build(ns, A.ncols, rp, ci, false);
  src.assign(S.entries(), -1);
  for (int64_t r = 0; r < ns; r++)
    for (int k = 0; k < S.rowlen[r]; k++) src[S.idx(r, k)] = si[rp[r] + k];
}

void build_fine_layout(const TopologyHost& T, FineLayout& F) {
  const LevelHost& L = T.levels[0];
  const SellHost& A = L.A;
  const int64_t nr = A.nrows, ns = A.nslices();
  // slice widths and offset sets
  std::vector<int> width(ns, 0);
  F.soff.assign(9 * ns, 0);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < ns; s++) {
    int64_t r0 = s * SELL_C, r1 = std::min<int64_t>(nr, r0 + SELL_C);
    int ws = (int)((A.sptr[s + 1] - A.sptr[s]) / SELL_C);
    width[s] = ws;
    if (r1 - r0 < SELL_C) continue;
    std::vector<int32_t> offs;
    bool ok = true;
    for (int64_t r = r0; r < r1 && ok; r++)
      for (int k = 0; k < A.rowlen[r] && ok; k++) {
        int32_t o = (int32_t)(A.at(r, k) - r);
        if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
          offs.push_back(o);
          ok = offs.size() <= 8;
        }
      }
    if (!ok || offs.empty()) continue;
    std::sort(offs.begin(), offs.end());
    for (int32_t o : offs)
      if (r0 + o < 0 || r1 - 1 + o >= A.ncols) ok = false;
    if (!ok) continue;
    F.soff[9 * s] = (int32_t)offs.size();
    for (size_t k = 0; k < offs.size(); k++) F.soff[9 * s + 1 + k] = offs[k];
    width[s] = (int)offs.size();
  }
  F.sptr.assign(ns + 1, 0);
  F.offset_slices = 0;
  F.explicit_nnz = 0;
  for (int64_t s = 0; s < ns; s++) {
    F.sptr[s + 1] = F.sptr[s] + (int64_t)width[s] * SELL_C;
    if (F.soff[9 * s]) {
      F.offset_slices++;
    } else {
      for (int64_t r = s * SELL_C; r < std::min<int64_t>(nr, (s + 1) * SELL_C); r++)
        F.explicit_nnz += A.rowlen[r];
    }
  }
  // host entry -> device entry (-1 for host padding), and the slot map per row
  std::vector<int64_t> emap(A.entries(), -1);
  F.col.assign(F.sptr[ns], 0);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < ns; s++) {
    const int32_t* so = &F.soff[9 * s];
    for (int64_t r = s * SELL_C; r < std::min<int64_t>(nr, (s + 1) * SELL_C); r++) {
      int lane = (int)(r % SELL_C);
      if (so[0]) {
        int k2 = 0;
        for (int k = 0; k < A.rowlen[r]; k++) {
          int32_t o = (int32_t)(A.at(r, k) - r);
          while (so[1 + k2] != o) k2++;  // ascending in both
          emap[A.idx(r, k)] = F.sptr[s] + (int64_t)k2 * SELL_C + lane;
        }
      } else {
        int ws = (int)((F.sptr[s + 1] - F.sptr[s]) / SELL_C);
        for (int k = 0; k < ws; k++) {
          int64_t e = F.sptr[s] + (int64_t)k * SELL_C + lane;
          F.col[e] = A.at(r, k);
          if (k < A.rowlen[r]) emap[A.idx(r, k)] = e;
        }
      }
    }
    if (!so[0])  // rows past the end of the matrix (last slice): padding columns
      for (int64_t r = nr; r < (s + 1) * SELL_C; r++) {
        int ws = (int)((F.sptr[s + 1] - F.sptr[s]) / SELL_C);
        for (int k = 0; k < ws; k++)
          F.col[F.sptr[s] + (int64_t)k * SELL_C + (r % SELL_C)] = A.at(r, k);
      }
  }
  auto dslot = [&](int64_t r, int k) -> uint16_t {
    if (k == NO_SLOT) return NO_SLOT;
    int64_t e = emap[A.idx(r, k)];
    return (uint16_t)((e - F.sptr[r / SELL_C] - r % SELL_C) / SELL_C);
  };
  F.diag_slot.resize(nr);
  for (int64_t r = 0; r < nr; r++) F.diag_slot[r] = dslot(r, L.diag_slot[r]);
  F.inc_slot.assign(T.inc_slot.size(), NO_SLOT);
  for (int64_t i = 0; i < T.n; i++) {
    int32_t j = T.n2f[i];
    if (j < 0) continue;
    for (int64_t p = T.inc_ptr[i]; p < T.inc_ptr[i + 1]; p++) F.inc_slot[p] = dslot(j, T.inc_slot[p]);
  }
  if (!L.p_slot.empty()) {
    F.p_slot.assign(F.sptr[ns], NO_SLOT);
    for (int64_t e = 0; e < A.entries(); e++)
      if (emap[e] >= 0) F.p_slot[emap[e]] = L.p_slot[e];
  }
  F.ap_a.resize(L.ap_a.size());
#pragma omp parallel for schedule(static)
  for (int64_t q = 0; q < (int64_t)L.ap_a.size(); q++) F.ap_a[q] = (int32_t)emap[L.ap_a[q]];
}

std::string build_galerkin_plan(const TopologyHost& T, int level, const FineLayout* F,
                                GalerkinPlan& G) {
  G = GalerkinPlan();
  if (level < 0 || level + 1 >= (int)T.levels.size()) return "";
  const LevelHost& L = T.levels[level];
  if (L.agg.empty() || L.ap_rp.empty()) return "";
  const SellHost& A = L.A;
  const SellHost& P = L.P;
  const SellHost& R = L.R;
  const SellHost& Rs = L.Rs;
  const LevelHost& Nx = T.levels[level + 1];
  const SellHost& C = Nx.A;
  const int64_t n = A.nrows, nc = R.nrows;
  if (F && (int64_t)F->sptr.size() != A.nslices() + 1) return "fine layout does not match level 0";
  int apw = 0;
  for (int64_t i = 0; i < n; i++) apw = std::max(apw, (int)(L.ap_rp[i + 1] - L.ap_rp[i]));
  // row-wise A*P: enough rows to fill the GPU, and rows short enough for the
  // per-lane accumulators in LDS
  if (n < 8192 || apw > 32) return "";
  G.ap_rows = true;
  G.ap_w = apw;
  // lanes per fine row: about 8 A slots a lane (level 0's offset slices: one)
  int lg = 0;
  if (!F) {
    double avg = n > 0 ? (double)A.nnz / (double)n : 1.0;
    while ((1 << lg) * 8 < avg && lg < 3) lg++;
  }
  G.ap_lg = lg;
  const int GL = 1 << lg;
  const int64_t nsub = n * GL, nss = (nsub + SELL_C - 1) / SELL_C;
  // the A slots the GPU layout walks for sub-row (i, g): f(column, present)
  auto walk = [&](int64_t i, int g, auto&& f) {
    if (F) {
      const int64_t sl = i / SELL_C;
      const int lane = (int)(i % SELL_C);
      const int32_t* so = &F->soff[9 * sl];
      if (so[0]) {
        for (int k = 0; k < so[0]; k++) {
          int32_t c = (int32_t)i + so[1 + k];
          bool present = false;
          for (int m = 0; m < A.rowlen[i]; m++)
            if (A.at(i, m) == c) present = true;
          f(c, present);
        }
      } else {
        int w = (int)((F->sptr[sl + 1] - F->sptr[sl]) / SELL_C);
        for (int k = 0; k < w; k++) f(F->col[F->sptr[sl] + (int64_t)k * SELL_C + lane], k < A.rowlen[i]);
      }
    } else {
      for (int k = g; k < A.rowlen[i]; k += GL) f(A.at(i, k), true);
    }
  };
  std::vector<int32_t> nprod(nsub, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < nsub; r++) {
    int c = 0;
    walk(r / GL, (int)(r % GL), [&](int32_t j, bool) { c += P.rowlen[j]; });
    nprod[r] = c;
  }
  G.apc_ptr.assign(nss + 1, 0);
  for (int64_t sl = 0; sl < nss; sl++) {
    int wc = 0;
    for (int64_t r = sl * SELL_C; r < std::min<int64_t>(nsub, (sl + 1) * SELL_C); r++)
      wc = std::max(wc, nprod[r]);
    G.apc_ptr[sl + 1] = G.apc_ptr[sl] + (int64_t)wc * SELL_C;
  }
  G.apc.assign(G.apc_ptr[nss], 255);
#pragma omp parallel
  {
    std::vector<int32_t> slot_of(P.ncols, -1);
#pragma omp for schedule(dynamic, 64)
    for (int64_t sl = 0; sl < nss; sl++)
      for (int64_t r = sl * SELL_C; r < std::min<int64_t>(nsub, (sl + 1) * SELL_C); r++) {
        const int64_t i = r / GL;
        for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++)
          slot_of[L.ap_col[a]] = (int32_t)(a - L.ap_rp[i]);
        int64_t q = 0;
        const int64_t base = G.apc_ptr[sl] + r % SELL_C;
        walk(i, (int)(r % GL), [&](int32_t j, bool present) {
          for (int t = 0; t < P.rowlen[j]; t++, q++)
            G.apc[base + q * SELL_C] = present ? (uint8_t)slot_of[P.at(j, t)] : (uint8_t)255;
        });
        for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++) slot_of[L.ap_col[a]] = -1;
      }
  }
  // row-wise R*(A*P) over Rs's sub-rows: enough coarse rows, at most 8 entries
  // a sub-row, short upper parts and AP rows (a product code holds the
  // entry in 3 bits, the AP entry in 5 and the destination in 8)
  int rw = 0;
  for (int64_t I = 0; I < nc; I++) rw = std::max(rw, C.rowlen[I] - (int)Nx.diag_slot[I]);
  int rsw = 0;
  for (int64_t r = 0; r < Rs.nrows; r++) rsw = std::max(rsw, Rs.rowlen[r]);
  if (nc < 8192 || rw > 32 || apw > 32 || rsw > 8) return "";
  G.rap_rows = true;
  G.rap_w = rw;
  const int rG = 1 << L.rs_lg;
  const int64_t rsub = Rs.nrows, rss = Rs.nslices();
  // per sub-row its products, sorted by destination and, within one, by entry
  // (ascending fine row): every destination's products form one run
  std::vector<std::vector<uint16_t>> prods(rsub);
#pragma omp parallel
  {
    std::vector<int32_t> slot_of(nc, -1);
#pragma omp for schedule(dynamic, 256)
    for (int64_t r = 0; r < rsub; r++) {
      const int64_t I = r / rG;
      const int ds = Nx.diag_slot[I];
      for (int k = ds; k < C.rowlen[I]; k++) slot_of[C.at(I, k)] = k - ds;
      std::vector<uint16_t>& v = prods[r];
      for (int t = 0; t < Rs.rowlen[r]; t++) {
        const int32_t i = Rs.at(r, t);
        for (int64_t a = L.ap_rp[i]; a < L.ap_rp[i + 1]; a++) {
          if (L.ap_col[a] < I) continue;
          const int k = (int)(a - L.ap_rp[i]);
          v.push_back((uint16_t)(t | (k << 3) | (slot_of[L.ap_col[a]] << 8)));
        }
      }
      std::stable_sort(v.begin(), v.end(), [](uint16_t x, uint16_t y) { return (x >> 8) < (y >> 8); });
      for (int k = ds; k < C.rowlen[I]; k++) slot_of[C.at(I, k)] = -1;
    }
  }
  G.rnp.assign(rsub, 0);
  G.rcd_ptr.assign(rss + 1, 0);
  for (int64_t sl = 0; sl < rss; sl++) {
    size_t wc = 0;
    for (int64_t r = sl * SELL_C; r < std::min<int64_t>(rsub, (sl + 1) * SELL_C); r++) {
      wc = std::max(wc, prods[r].size());
      if (prods[r].size() > 65535) return "row-wise R*(A*P): too many products in a lane";
      G.rnp[r] = (uint16_t)prods[r].size();
    }
    G.rcd_ptr[sl + 1] = G.rcd_ptr[sl] + (int64_t)wc * SELL_C;
  }
  G.rcd.assign(G.rcd_ptr[rss], 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < rsub; r++)
    for (size_t q = 0; q < prods[r].size(); q++)
      G.rcd[G.rcd_ptr[r / SELL_C] + (int64_t)q * SELL_C + r % SELL_C] = prods[r][q];
  return "";
}

std::string build_hierarchy(TopologyHost& T, int max_levels, int coarse_max, int dense_max,
                            const std::vector<int64_t>& member_row0) {
  auto t0 = std::chrono::steady_clock::now();
  T.levels.resize(1);
  std::vector<int64_t> row0 = member_row0;
  if (row0.size() < 2) row0 = {0, T.levels[0].A.nrows};
  while ((int)T.levels.size() < max_levels && T.levels.back().A.nrows > coarse_max) {
    LevelHost& L = T.levels.back();
    std::vector<int64_t> next_row0;
    int64_t nc = aggregate(L.A, row0, L.agg, next_row0);
    row0 = next_row0;
    if (nc <= 0 || (double)nc > 0.85 * (double)L.A.nrows) {
      L.agg.clear();
      break;
    }
    LevelHost next;
    std::string err = build_transfer(L, next, nc);
    if (!err.empty()) return err;
    T.levels.push_back(std::move(next));
  }
  T.coarse_dense = T.levels.back().A.nrows <= dense_max;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t l = 1; l < T.levels.size(); l++)
    split_rows(T.levels[l].A, T.levels[l].as_lg, T.levels[l].As, T.levels[l].as_src);
  T.setup_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return "";
}

}  // namespace dpf
