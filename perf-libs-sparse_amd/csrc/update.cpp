// update.cpp — value updates written into an optimized device plan in place.
//
// perflibs_spmat_update_{s,d} changes values of existing entries
// (src/matvec.cpp:530-562); the reference writes them straight into its
// optimized SELL layout (src/sell_c_sigma.cpp:966-1049), so a solver loop can
// update and execute without re-optimizing.  Here the host arrays are updated
// first (the CSR / BSR search decides which entry an update names); this file
// then finds where each updated entry lives in the device layout of the plan
// and overwrites just those values with one small scatter launch:
//   CSR-vector / CSR-adaptive: the device CSR is the operand itself, slot =
//     the entry's position j in it;
//   SELL-C-sigma: row r sits at permuted position pi = in2permd[r], chunk
//     pi / C, lane pi % C, and its t-th entry (CSR order) at cs + t*C + lane;
//   column panel: entries of (tile, panel) segments keep row order, so the
//     slot is the segment start plus the number of earlier entries of the same
//     tile and panel (one pass over each touched tile);
//   BSR built from CSR: block (r/b, c/b) of the kept block structure, value at
//     (r%b)*b + c%b, holding the sum of the row's entries at column c (the
//     block build sums duplicates);
//   native BSR: the device values are the handle's, slot = j.
// Plans of a transposed operand (TRANS / CONJTRANS hints) keep that operand's
// structure, and each hit is first moved to it: row c of the transpose lists
// the original rows in ascending order, so the entry (r, c) sits at the lower
// bound of r in that row (plus its rank among duplicate copies).  CSR-family
// plans of a BSR handle (whose expansion drops zero values, so an update can
// change its structure) are not patched: they are rebuilt by the next
// execution.
#include <algorithm>
#include <numeric>

#include "plsparse.hpp"

namespace pls {

namespace {

// Slots of hits in a column-panel plan, NOTRANS operand = the handle's CSR.
void panel_slots(const DevPanel &P, const int32_t *rp, const int32_t *ci,
                 const std::vector<ValueHit> &hits, std::vector<int64_t> &slot) {
  const int32_t base = rp[0];
  const int64_t np = P.npanels;
  const int32_t pw = P.pw_log2;
  const auto &t0 = P.h_tile_row0;
  std::vector<size_t> order(hits.size());
  std::iota(order.begin(), order.end(), size_t(0));
  std::vector<int32_t> tile(hits.size());
  for (size_t h = 0; h < hits.size(); ++h)
    tile[h] = int32_t(std::upper_bound(t0.begin(), t0.end(), hits[h].r) - t0.begin()) - 1;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    if (tile[a] != tile[b]) return tile[a] < tile[b];
    if (hits[a].r != hits[b].r) return hits[a].r < hits[b].r;
    return hits[a].j < hits[b].j;
  });
  std::vector<int64_t> cnt(size_t(np), 0);
  size_t q = 0;
  while (q < order.size()) {
    const int32_t t = tile[order[q]];
    std::fill(cnt.begin(), cnt.end(), 0);
    const int32_t row_end = t0[size_t(t) + 1];
    for (int32_t row = t0[size_t(t)]; row < row_end && q < order.size() && tile[order[q]] == t;
         ++row) {
      const int64_t j0 = int64_t(rp[row]) - base, j1 = int64_t(rp[row + 1]) - base;
      for (int64_t j = j0; j < j1; ++j) {
        const int64_t p = int64_t(ci[j] - base) >> pw;
        while (q < order.size() && tile[order[q]] == t && hits[order[q]].r == row &&
               hits[order[q]].j == j) {
          slot[order[q]] = (P.h_seg[size_t(t * np + p)] & ~int64_t(3)) + cnt[size_t(p)];
          ++q;
        }
        ++cnt[size_t(p)];
      }
    }
    while (q < order.size() && tile[order[q]] == t) ++q;  // unmatched: slot stays -1
  }
}

}  // namespace

bool patch_plan(perflibs_spmat_top_t *A, const std::vector<ValueHit> &hits) {
  Plan &pl = A->plan;
  if (!pl.csr && !pl.sell && !pl.bsr && !pl.panel) return true;  // nothing on the device yet
  if (hits.empty()) return true;
  const bool trans = pl.op != PERFLIBS_SPARSE_OPERATION_NOTRANS;
  const bool csr_host = A->fmt == Fmt::Csr || A->fmt == Fmt::Scs;
  const bool native_bsr = !trans && pl.bsr && !pl.bsr->from_csr && A->fmt == Fmt::Bsr;
  if (!csr_host && !native_bsr) return false;
  if (trans && pl.h_trp.empty()) return false;

  const size_t nh = hits.size();
  // The operand the plan was built from: the handle's CSR (NOTRANS) or its
  // transpose, and every hit in operand coordinates (row, column, position).
  const int32_t *orp = csr_host ? A->csr.rp : nullptr, *oci = csr_host ? A->csr.ci : nullptr;
  std::vector<ValueHit> oh(hits);
  if (trans) {
    orp = pl.h_trp.data();
    oci = pl.h_tci.data();
    const int32_t base = A->csr.rp[0];
    for (size_t h = 0; h < nh; ++h) {
      const int32_t r = hits[h].r, c = hits[h].c;
      // the transposed row c lists original rows in ascending order and the
      // copies of one (r, c) entry in their original order: the hit is the
      // k-th copy of column c in row r
      int64_t kth = 0;
      for (int64_t j = int64_t(A->csr.rp[r]) - base; j < hits[h].j; ++j)
        kth += A->csr.ci[j] - base == c;
      const int32_t *lo = std::lower_bound(oci + orp[c], oci + orp[c + 1], r);
      oh[h] = {c, r, int64_t(lo - oci) + kth};
    }
  }
  std::vector<int64_t> slot(nh, -1);
  // value written per hit, in the handle dtype
  std::vector<double> vd;
  std::vector<float> vf;
  const bool dbl = A->dt == DType::F64;
  if (dbl) vd.resize(nh);
  else vf.resize(nh);
  const void *hv = csr_host ? A->csr.vals : A->bsr.vals;
  for (size_t h = 0; h < nh; ++h) {
    if (dbl) vd[h] = static_cast<const double *>(hv)[hits[h].j];
    else vf[h] = static_cast<const float *>(hv)[hits[h].j];
  }
  int64_t count = 0;  // values in the device array
  VStore vs = VStore::F64;
  void *dvals = nullptr;
  if (pl.csr) {
    for (size_t h = 0; h < nh; ++h) slot[h] = oh[h].j;
    count = pl.csr->nnz;
    vs = pl.csr->vs;
    dvals = pl.csr->vals.get();
  } else if (pl.sell) {
    const DevSell &S = *pl.sell;
    const int32_t base = orp[0];
    for (size_t h = 0; h < nh; ++h) {
      const int64_t pi = S.h_in2permd[size_t(oh[h].r)];
      const int64_t c = pi / S.C, lane = pi % S.C;
      const int64_t t = oh[h].j - (int64_t(orp[oh[h].r]) - base);
      slot[h] = S.h_cs[size_t(c)] + t * S.C + lane;
    }
    count = S.padded;
    vs = S.vs;
    dvals = S.vals.get();
  } else if (pl.panel) {
    panel_slots(*pl.panel, orp, oci, oh, slot);
    count = int64_t(pl.panel->vals.bytes() / (pl.panel->vs == VStore::F64 ? 8
                                               : pl.panel->vs == VStore::F32 ? 4 : 2));
    vs = pl.panel->vs;
    dvals = pl.panel->vals.get();
  } else if (native_bsr) {
    for (size_t h = 0; h < nh; ++h) slot[h] = hits[h].j;
    count = pl.bsr->nnzb * pl.bsr->b * pl.bsr->b;
    vs = pl.bsr->vs;
    dvals = pl.bsr->vals.get();
  } else {  // BSR blocks built from the handle's CSR
    const DevBsr &B = *pl.bsr;
    const int32_t base = A->csr.rp[0];
    const int64_t bb = int64_t(B.b) * B.b;
    for (size_t h = 0; h < nh; ++h) {
      const int32_t orow = oh[h].r, ocol = oh[h].c, rb = orow / B.b, cb = ocol / B.b;
      int64_t k = -1;
      for (int32_t q = B.h_rp[size_t(rb)]; q < B.h_rp[size_t(rb) + 1]; ++q)
        if (B.h_ci[size_t(q)] == cb) {
          k = q;
          break;
        }
      if (k < 0) return false;
      slot[h] = k * bb + int64_t(orow % B.b) * B.b + ocol % B.b;
      // the block value is the sum of every stored copy of the entry, added
      // in the operand's order from zero as the block build does (for a
      // transposed operand the copies keep the original row's order)
      const int32_t r = hits[h].r, c = hits[h].c;
      double sd = 0.0;
      float sf = 0.0f;
      for (int64_t j = int64_t(A->csr.rp[r]) - base; j < int64_t(A->csr.rp[r + 1]) - base; ++j)
        if (A->csr.ci[j] - base == c) {
          if (dbl) sd += static_cast<const double *>(A->csr.vals)[j];
          else sf += static_cast<const float *>(A->csr.vals)[j];
        }
      if (dbl) vd[h] = sd;
      else vf[h] = sf;
    }
    count = B.nnzb * bb;
    vs = B.vs;
    dvals = B.vals.get();
  }
  for (size_t h = 0; h < nh; ++h)
    if (slot[h] < 0 || slot[h] >= count) return false;  // never scatter outside the array
  // Hits that name the same slot carry the same (final) value, so the
  // scatter's write order does not matter.
  const void *src = dbl ? (const void *)vd.data() : (const void *)vf.data();
  if ((vs == VStore::F64) != dbl) return false;  // FP64 values are stored as FP64 only
  DevBuf dslot, dval;
  hipError_t e = dslot.upload(slot.data(), sizeof(int64_t) * nh);
  if (e == hipSuccess) e = dval.upload(src, (dbl ? 8 : 4) * nh);
  if (e == hipSuccess) e = plk::patch_vals(vs, dvals, dslot.as<int64_t>(), dval.get(), int64_t(nh), stream());
  // clustered SpMM blocks keep a block-order copy of the CSR plan's values
  if (e == hipSuccess && pl.csr) e = refresh_block_values(*pl.csr);
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  return e == hipSuccess;
}

}  // namespace pls
