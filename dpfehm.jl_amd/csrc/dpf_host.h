// Host-side topology cache and smoothed-aggregation (SA) symbolic setup.
//
// Everything here depends only on the mesh topology (neighbors,
// areasoverlengths, dirichletnodes), never on the conductivities, so it is
// built once per topology and reused for every solve: the per-solve work
// (assembly, prolongator smoothing, Galerkin products) runs on the GPU
// through the index maps built here.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace dpf {

// Sliced ELLPACK with 64-row slices (one row per lane of a 64-wide wavefront),
// column-major inside a slice: entry s of row r lives at
//   sptr[r / 64] + s * 64 + r % 64.
// Rows shorter than their slice's width are padded with value 0 and a valid
// column (the row itself for square matrices, 0 otherwise).
constexpr int SELL_C = 64;
constexpr uint16_t NO_SLOT = 0xFFFF;

struct SellHost {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> sptr;    // nslices + 1
  std::vector<int32_t> col;     // sptr.back() entries
  std::vector<int32_t> rowlen;  // nrows
  int64_t nnz = 0;
  int max_rowlen = 0;
  int64_t entries() const { return sptr.empty() ? 0 : sptr.back(); }
  int64_t nslices() const { return sptr.empty() ? 0 : (int64_t)sptr.size() - 1; }
  int64_t idx(int64_t r, int s) const {
    return sptr[r / SELL_C] + (int64_t)s * SELL_C + (r % SELL_C);
  }
  int32_t at(int64_t r, int s) const { return col[idx(r, s)]; }
  // rp/ci: row lists with sorted unique columns.
  void build(int64_t nrows, int64_t ncols, const std::vector<int64_t>& rp,
             const std::vector<int32_t>& ci, bool square);
};

// One level of the AMG hierarchy.  A is the level operator; the transfer
// data (agg .. rap_map) connect it to the next coarser level and are empty
// on the coarsest level.
struct LevelHost {
  SellHost A;
  std::vector<uint16_t> diag_slot;  // per row: slot of the diagonal entry
  // A with each row split over 2^as_lg lanes for the smoothing sweeps
  // (coarse levels, whose rows are long); as_lg = 0: the sweeps use A itself
  int as_lg = 0;
  SellHost As;
  std::vector<int32_t> as_src;      // per As entry: index of the A entry, -1 = padding
  std::vector<int32_t> agg;         // per row: aggregate (= coarse row)
  SellHost P;                       // prolongator, nrows x ncoarse
  std::vector<uint16_t> p_slot;     // per A entry: slot of agg(col) in P's row
  std::vector<uint16_t> p_self;     // per row: slot of agg(row) in P's row
  SellHost R;                       // restriction = P', ncoarse x nrows
  std::vector<int32_t> r_src;       // per R entry: index of the P entry, -1 = padding
  // R with each row split over G lanes (sub-row I*G+g holds the entries
  // t = g, g+G, ... of row I), for the V-cycle's restriction
  int rs_lg = 0;                    // log2 G
  SellHost Rs;
  std::vector<int32_t> rs_src;      // per Rs entry: index of the P entry, -1 = padding
  // A*P (CSR) and its products grouped by destination nonzero:
  //   AP[d] = sum_{q in [ap_off[d], ap_off[d+1])} A.val[ap_a[q]] * P.val[ap_p[q]]
  std::vector<int64_t> ap_rp;
  std::vector<int32_t> ap_col;
  std::vector<int64_t> ap_off;
  std::vector<int32_t> ap_a, ap_p;
  // next level's A.  The Galerkin operator of a symmetric A is symmetric, so
  // only its entries with column >= row are summed, one contribution list
  // each, lists in row-major order of (row, slot) so that neighbouring
  // threads share fine rows:
  //   Ac[rap_dest[k]] = sum_{q in [rap_off[k], rap_off[k+1])} P.val[rap_p[q]] * AP[rap_ap[q]]
  // and each entry below the diagonal is copied from its mirror:
  //   Ac[rap_mdst[m]] = Ac[rap_msrc[m]]
  // (SELL padding is never written: zero from allocation on)
  std::vector<int64_t> rap_off;
  std::vector<int32_t> rap_dest, rap_p, rap_ap;
  std::vector<int32_t> rap_mdst, rap_msrc;
  int ap_lg = 0, rap_lg = 0;  // log2 lanes per contribution list on the GPU
};

struct TopologyHost {
  int64_t n = 0, ne = 0, nd = 0, nf = 0;
  std::vector<int32_t> edges;       // 2*ne, 0-based
  std::vector<double> aol;
  std::vector<uint8_t> isfree;
  std::vector<int32_t> n2f;         // -1 on Dirichlet cells
  std::vector<int32_t> f2n;
  // cell -> incident edges (all cells, edges in list order; self-edges,
  // which contribute nothing, are left out)
  std::vector<int64_t> inc_ptr;     // n + 1
  std::vector<int32_t> inc_edge;
  std::vector<int32_t> inc_other;   // the edge's other cell
  std::vector<uint16_t> inc_slot;   // free cell, free other: slot of the
                                    // other's column in this cell's A row
  std::vector<LevelHost> levels;    // levels[0].A has the fine pattern
  bool coarse_dense = false;        // coarsest level solved by a dense inverse
  double setup_ms = 0;              // host time of the symbolic setup
  // non-empty when some connected component of the mesh holds no Dirichlet
  // cell (e.g. dirichletnodes is empty): the steady operator is singular
  // there, while a transient one (storage on the diagonal) is not
  std::string floating;
};

// GPU layout of the fine operator.  A 64-row slice whose nonzeros all sit at
// no more than 8 distinct offsets o = col - row (with row + o inside the
// matrix for every lane) is an offset slice: slot k of every lane holds the
// entry at offset soff[9s + 1 + k] (0 where the row has none), and no column
// is stored; soff[9s] = the slice's width (0: an explicit slice, stored as in
// the host layout).  Regular grids make almost every slice an offset slice,
// which halves the FP32 operator's bytes.  Entries keep their ascending
// column order inside a row, so every accumulation over a row's entries runs
// in the same order as on the host layout.  The maps that address fine
// entries (inc_slot, diag_slot, p_slot, ap_a) are translated to the layout.
struct FineLayout {
  std::vector<int64_t> sptr;      // nslices + 1
  std::vector<int32_t> col;       // explicit slices' columns (offset slices: unused)
  std::vector<int32_t> soff;      // 9 per slice
  std::vector<uint16_t> inc_slot, diag_slot, p_slot;
  std::vector<int32_t> ap_a;
  int64_t offset_slices = 0, explicit_nnz = 0;
};
void build_fine_layout(const TopologyHost& T, FineLayout& F);

// Row-wise Galerkin products of one level with transfer operators (the GPU
// kernels k_ap_rows and k_rap_rows), an alternative to the product lists of
// LevelHost that reads each operator row once instead of gathering two
// values through two 32-bit indices per product.
//  * A*P: each fine row i is spread over G = 2^ap_lg lanes (sub-row i G + g);
//    lane g walks the row's A slots s = g, g + G, ... in the order of the GPU
//    layout (the fine offset layout on level 0: every slot of an offset slice,
//    the absent ones too; else the row's entries) and, for the slot's column
//    j, the entries t < rowlen(j) of P's row j, adding each product into AP
//    entry apc[.] of the row -- 255 for the products of a slot the row does
//    not hold (dropped).  apc is a SELL layout over the sub-rows (apc_ptr;
//    product q of sub-row r at apc_ptr[r/64] + 64 q + r%64).  The G lanes'
//    partial rows are then summed in a fixed pattern and the row's entries
//    written in A*P's CSR order (ap_rp, ap_col).
//  * The next level's entries on and above the diagonal: each coarse row is
//    spread over the lanes of R's row-split copy (Rs, 2^rs_lg lanes a row; at
//    most 8 entries a lane).  A lane's products AP(i_t, k) P(i_t, I), over its
//    entries t (fine row i_t) and the AP entries k of row i_t with column >=
//    I, are coded t | k << 3 | d << 8 (d: the destination among the coarse
//    row's entries from its diagonal on) and sorted by d, then t, so that
//    every destination's products form one run that the lane sums in a
//    register, in ascending fine row (rcd, a SELL layout over Rs's sub-rows,
//    rcd_ptr; rnp: products per sub-row).  The lanes' partial rows are then
//    summed in a fixed pattern.
struct GalerkinPlan {
  bool ap_rows = false, rap_rows = false;
  int ap_lg = 0;            // log2 lanes per fine row of the row-wise A*P
  int ap_w = 0, rap_w = 0;  // widest AP row, widest upper part of a coarse row
  std::vector<int64_t> apc_ptr, rcd_ptr;
  std::vector<uint8_t> apc;
  std::vector<uint16_t> rcd, rnp;
};
// fine: level 0's GPU layout (null: the level's operator is in the host
// layout, walked through its row lengths).  Returns "" or an error message.
std::string build_galerkin_plan(const TopologyHost& T, int level, const FineLayout* fine,
                                GalerkinPlan& G);

// Both return "" on success or an error message.
std::string build_topology(TopologyHost& T, int64_t n, int64_t ne, const int64_t* nb,
                           const double* aol, int64_t nd, const int64_t* dn);
// member_row0: fine-row offsets of the members of a batch (empty: one
// system); members are aggregated independently so that they never share an
// aggregate and each owns a contiguous row range on every level.
std::string build_hierarchy(TopologyHost& T, int max_levels, int coarse_max, int dense_max,
                            const std::vector<int64_t>& member_row0 = {});
void split_rows(const SellHost& A, int& lg, SellHost& S, std::vector<int32_t>& src,
                int force_lg = -1);
// "" when every connected component of the mesh -- through the edges with
// conductance Ks[e] aol[e] != 0 (aol[e] != 0 when Ks is null) -- holds a Dirichlet
// cell, else a message naming the lowest-numbered free cell of the first
// component without one and the component's size.
std::string floating_component(const TopologyHost& T, const double* Ks);

}  // namespace dpf
