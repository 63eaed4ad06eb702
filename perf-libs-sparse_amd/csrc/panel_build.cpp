// panel_build.cpp — host construction of the column-panel CSR layout.
//
// Rows are cut into wave tiles (contiguous, nnz-balanced, at most `rcap`
// rows so one wave's row accumulators fit in LDS).  Columns are cut into
// panels of 2^pw_log2 columns.  The non-zeros of each (tile, panel) pair are
// stored contiguously, tile-major then panel-major, each as
//   idx = (row - tile_first_row) << pw_log2 | (col - panel_first_col)
// plus its value, in row order (and the row's own order inside a row), so a
// wave can sweep the panels in column order with only one panel's slice of x
// live at a time and reduce rows by segmented sums over sorted keys.
#include <algorithm>
#include <cstring>
#include <thread>

#include "plsparse.hpp"

namespace pls {

namespace {
template <class F> void parallel_chunks(int64_t n, F &&fn) {
  unsigned hw = std::thread::hardware_concurrency();
  const int64_t nt = std::min<int64_t>(hw ? hw : 1, 16);
  if (n < 64 || nt <= 1) {
    fn(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t step = (n + nt - 1) / nt;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t lo = t * step, hi = std::min(n, lo + step);
    if (lo >= hi) break;
    th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
  }
  for (auto &t : th) t.join();
}
}  // namespace

bool panel_build(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci, const void *vals,
                 DType dt, int32_t want_tiles, int32_t rcap, int32_t pw_log2, PanelHost &P) {
  if (m <= 0 || rcap <= 0 || pw_log2 < 1 || pw_log2 > 30) return false;
  // the row field must hold rows 0..rcap-1 and the all-ones padding marker
  if (int64_t(rcap) >= (int64_t(1) << (32 - pw_log2)) || rcap > 65536) return false;
  P = PanelHost{};
  P.m = m;
  P.n = n;
  P.pw_log2 = pw_log2;
  P.npanels = std::max<int32_t>(1, int32_t((int64_t(n) + (int64_t(1) << pw_log2) - 1) >> pw_log2));
  if (P.npanels > 65536) return false;  // the kernel keys a step's panels in 16 bits
  const int32_t base = rp[0];
  const int64_t nnz = int64_t(rp[m]) - base;

  // Wave tiles: cut where the running cost (one unit per non-zero and one per
  // row, for the accumulator clear and the y update) crosses k/want of the
  // total, so exactly want_tiles tiles come out unless the row cap splits one.
  const int64_t want = std::max<int32_t>(want_tiles, 1);
  const int64_t total = nnz + m;
  auto cost = [&](int32_t i) { return int64_t(rp[i]) - base + i; };
  P.tile_row0.push_back(0);
  int32_t prev = 0;
  for (int64_t k = 1; k <= want && prev < m; ++k) {
    int32_t b = m;
    if (k < want) {
      const int64_t target = total * k / want;
      int32_t lo = prev, hi = m;  // first row i >= prev with cost(i) >= target
      while (lo < hi) {
        const int32_t mid = lo + (hi - lo) / 2;
        if (cost(mid) >= target) hi = mid;
        else lo = mid + 1;
      }
      b = lo;
    }
    while (b - prev > rcap) {
      prev += rcap;
      P.tile_row0.push_back(prev);
    }
    if (b > prev) {
      P.tile_row0.push_back(b);
      prev = b;
    }
  }
  // whole workgroups of 4 wave tiles
  while ((P.tile_row0.size() - 1) % 4 != 0) P.tile_row0.push_back(m);
  P.ntiles = int32_t(P.tile_row0.size() - 1);

  const int64_t np = P.npanels;
  std::vector<int64_t> count(size_t(P.ntiles) * np, 0);
  parallel_chunks(P.ntiles, [&](int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; ++t)
      for (int32_t row = P.tile_row0[t]; row < P.tile_row0[t + 1]; ++row)
        for (int32_t j = rp[row] - base; j < rp[row + 1] - base; ++j)
          ++count[size_t(t * np + ((ci[j] - base) >> pw_log2))];
  });
  // Segments are padded to whole quads (4 entries) so a lane can load four
  // consecutive entries with one 16-byte index load; seg[k] holds the padded
  // start of segment k (a multiple of 4) plus, in its low 2 bits, how many of
  // that segment's entries are padding.  A padding entry has value 0 and the
  // packed index with an all-ones row field and column 0 of its panel.
  P.seg.assign(size_t(P.ntiles) * np + 1, 0);
  int64_t at = 0;
  for (size_t k = 0; k < count.size(); ++k) {
    const int64_t padded = (count[k] + 3) & ~int64_t(3);
    P.seg[k] = at | (padded - count[k]);
    at += padded;
  }
  P.seg.back() = at;
  const size_t vsz = dt == DType::F64 ? 8 : 4;
  P.idx.assign(size_t(std::max<int64_t>(at, 4)), 0u);
  P.vals.assign(size_t(std::max<int64_t>(at, 4)) * vsz, uint8_t(0));
  parallel_chunks(P.ntiles, [&](int64_t lo, int64_t hi) {
    std::vector<int64_t> fill(static_cast<size_t>(np));
    for (int64_t t = lo; t < hi; ++t) {
      for (int64_t p = 0; p < np; ++p) fill[size_t(p)] = P.seg[size_t(t * np + p)] & ~int64_t(3);
      const int32_t row0 = P.tile_row0[t];
      for (int32_t row = row0; row < P.tile_row0[t + 1]; ++row)
        for (int32_t j = rp[row] - base; j < rp[row + 1] - base; ++j) {
          const int32_t c = ci[j] - base;
          const int64_t p = c >> pw_log2;
          const int64_t e = fill[size_t(p)]++;
          P.idx[size_t(e)] = (uint32_t(row - row0) << pw_log2) | uint32_t(c - (int32_t(p) << pw_log2));
          std::memcpy(&P.vals[size_t(e) * vsz], static_cast<const uint8_t *>(vals) + size_t(j) * vsz, vsz);
        }
      const uint32_t pad_idx = ~((1u << pw_log2) - 1u);
      for (int64_t p = 0; p < np; ++p) {  // padding entries
        const size_t k = size_t(t * np + p);
        const int64_t npad = P.seg[k] & 3;
        for (int64_t e = fill[size_t(p)]; e < fill[size_t(p)] + npad; ++e) P.idx[size_t(e)] = pad_idx;
      }
    }
  });
  return true;
}

}  // namespace pls
