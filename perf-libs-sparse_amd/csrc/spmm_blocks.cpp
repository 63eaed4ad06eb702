// spmm_blocks.cpp — row blocks of the column-compressed SpMM kernel
// (k_spmm_lds): which rows share a workgroup, the distinct columns each block
// lists once, and every non-zero's position in its block's list.
//
// Two cuts are made and the one with the shorter lists kept:
//  * consecutive: rows in order, a block closed before the row that would
//    exceed a limit;
//  * clustered: a block grows from its first unassigned row by repeatedly
//    adding the row that brings the fewest columns not yet listed (ties: the
//    row adjacent to the most rows already in the block, then the lowest row
//    index), so the rows of a block share their columns.  On the 27-point
//    stencil of a 128^3 grid a block becomes a compact patch of neighbouring
//    grid points instead of a segment of one grid line: 6.2 instead of 10.4
//    listed columns per row with lists of 128, i.e. 40% fewer B rows staged
//    in LDS per product (k = 128 FP64 product 3.39 -> 2.83 ms measured on
//    the symmetrically permuted matrix before the kernel took row lists).
// A clustered block's rows are listed in `brow` (the kernel reads them from
// there); consecutive blocks leave it empty.
#include <algorithm>
#include <thread>

#include "plsparse.hpp"

namespace pls {

namespace {
struct LbHost {
  std::vector<int32_t> row0{0};  // block b: entries [row0[b], row0[b+1]) of the row order
  std::vector<int32_t> uoff{0};  // block b: list entries [uoff[b], uoff[b+1]) of ucol
  std::vector<int32_t> ucol;     // distinct columns of each block, in list order
  std::vector<int32_t> brow;     // row order (empty: consecutive rows)
  std::vector<uint8_t> lidx;     // list position of every non-zero (CSR order)
  int64_t direct = 0;            // non-zeros of rows without a list
  int64_t cost() const { return int64_t(ucol.size()) + direct; }
};

template <class F> void parallel_ranges(int64_t n, int64_t min_per_range, int64_t max_ranges, F &&fn) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t nt =
      std::max<int64_t>(1, std::min<int64_t>({hw ? hw : 1, max_ranges, n / min_per_range}));
  if (nt <= 1) {
    fn(0, int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t step = (n + nt - 1) / nt;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t lo = t * step, hi = std::min(n, lo + step);
    if (lo >= hi) break;
    th.emplace_back([&fn, t, lo, hi] { fn(int(t), lo, hi); });
  }
  for (auto &t : th) t.join();
}

// Greedy cut of consecutive rows into blocks of at most umax distinct
// columns, lb_nnz_per_ucol(umax) * umax non-zeros and kLbRows rows; a row that
// alone exceeds a limit becomes a block without a column list.
void cut_consecutive(const Operand &o, int32_t umax, LbHost &h) {
  const int32_t nnz_max = lb_nnz_per_ucol(umax) * umax;
  const int32_t m = o.m;
  const int32_t *rp = o.prp, *ci = o.pci;
  const int32_t base = rp[0];
  const int64_t nnz = int64_t(rp[m]) - base;
  const size_t ncols = size_t(std::max(o.n, 1));
  h.lidx.assign(size_t(std::max<int64_t>(nnz, 1)), 0);
  // inblk[c] == bid: column c is in the current block's list at pos[c];
  // seen[c] == pass: column c was counted by the current counting pass
  std::vector<int32_t> inblk(ncols, -1), pos(ncols, 0), seen(ncols, -1);
  int32_t bid = 0, pass = 0, u = 0, rows = 0;
  int64_t bnnz = 0;
  auto close = [&](int32_t r) {  // the current block ends before row r
    h.row0.push_back(r);
    h.uoff.push_back(int32_t(h.ucol.size()));
    ++bid;
    u = rows = 0;
    bnnz = 0;
  };
  auto count_fresh = [&](int64_t j0, int64_t j1) {  // distinct columns not in the list
    ++pass;
    int32_t f = 0;
    for (int64_t j = j0; j < j1; ++j) {
      const int32_t c = ci[j] - base;
      if (inblk[size_t(c)] != bid && seen[size_t(c)] != pass) {
        seen[size_t(c)] = pass;
        ++f;
      }
    }
    return f;
  };
  for (int32_t r = 0; r < m; ++r) {
    const int64_t j0 = int64_t(rp[r]) - base, j1 = int64_t(rp[r + 1]) - base;
    int32_t fresh = count_fresh(j0, j1);
    if (rows > 0 && !(u + fresh <= umax && bnnz + (j1 - j0) <= nnz_max && rows < kLbRows)) {
      close(r);
      fresh = count_fresh(j0, j1);
    }
    if (fresh > umax || j1 - j0 > nnz_max) {
      close(r + 1);  // rows == 0: the row is a block of its own without a list
      h.direct += j1 - j0;
      continue;
    }
    for (int64_t j = j0; j < j1; ++j) {
      const int32_t c = ci[j] - base;
      if (inblk[size_t(c)] != bid) {
        inblk[size_t(c)] = bid;
        pos[size_t(c)] = u++;
        h.ucol.push_back(c);
      }
      h.lidx[size_t(j)] = uint8_t(pos[size_t(c)]);
    }
    bnnz += j1 - j0;
    ++rows;
  }
  if (rows > 0) close(m);
}

// Columns whose rows are not followed when scoring candidates (a column shared
// by very many rows says little about which rows belong together, and following
// it would make the candidate sets large).
constexpr int32_t kDegreeCap = 256;

// Rows of each column of the operand's pattern (0-based, rows ascending).
void rows_by_column(const Operand &o, std::vector<int32_t> &cp, std::vector<int32_t> &cr) {
  const int32_t m = o.m, n = o.n, base = o.prp[0];
  const int64_t nnz = int64_t(o.prp[m]) - base;
  cp.assign(size_t(n) + 1, 0);
  cr.resize(size_t(std::max<int64_t>(nnz, 1)));
  for (int64_t j = 0; j < nnz; ++j) ++cp[size_t(o.pci[j] - base) + 1];
  for (int32_t c = 0; c < n; ++c) cp[size_t(c) + 1] += cp[size_t(c)];
  std::vector<int32_t> next(cp.begin(), cp.end() - 1);
  for (int32_t r = 0; r < m; ++r)
    for (int64_t j = int64_t(o.prp[r]) - base; j < int64_t(o.prp[r + 1]) - base; ++j)
      cr[size_t(next[size_t(o.pci[j] - base)]++)] = r;
}

// Over up to 256 evenly spaced non-empty rows: the mean of the largest
// fraction of a row's non-zeros whose columns one other row shares (near 0 for
// scattered columns: no block can share columns, so clustering is skipped; 2/3
// on the 27-point stencil), and the fraction of rows whose best partner lies
// within m / 32 rows of it (row ranges can then be clustered independently).
struct Overlap {
  double shared = 0, near = 0;
};
Overlap sampled_overlap(const Operand &o, const std::vector<int32_t> &cp,
                        const std::vector<int32_t> &cr) {
  const int32_t m = o.m, base = o.prp[0];
  std::vector<int32_t> cnt(size_t(m), 0), hit;
  double sum = 0;
  int32_t ns = 0, near = 0;
  const int32_t step = std::max(1, m / 256);
  for (int32_t r = 0; r < m; r += step) {
    const int64_t j0 = int64_t(o.prp[r]) - base, j1 = int64_t(o.prp[r + 1]) - base;
    if (j1 == j0) continue;
    int32_t best = 0, partner = r;
    for (int64_t j = j0; j < j1; ++j) {
      const int32_t c = o.pci[j] - base;
      if (cp[size_t(c) + 1] - cp[size_t(c)] > kDegreeCap) continue;
      for (int32_t q = cp[size_t(c)]; q < cp[size_t(c) + 1]; ++q) {
        const int32_t s = cr[size_t(q)];
        if (s == r) continue;
        if (cnt[size_t(s)]++ == 0) hit.push_back(s);
        if (cnt[size_t(s)] > best) {
          best = cnt[size_t(s)];
          partner = s;
        }
      }
    }
    for (int32_t s : hit) cnt[size_t(s)] = 0;
    hit.clear();
    sum += double(best) / double(j1 - j0);
    near += std::abs(int64_t(partner) - r) <= m / 32;
    ++ns;
  }
  Overlap v;
  if (ns) {
    v.shared = sum / ns;
    v.near = double(near) / ns;
  }
  return v;
}

// Column -> list position of the block being grown (open addressing; at most
// kLbUcolsLarge columns at a quarter load).
class ColMap {
 public:
  ColMap() : key_(kCap, -1), val_(kCap, 0) { used_.reserve(kCap); }
  int32_t find(int32_t c) const {
    for (uint32_t i = hash(c);; i = (i + 1) & (kCap - 1)) {
      if (key_[i] == c) return val_[i];
      if (key_[i] < 0) return -1;
    }
  }
  void insert(int32_t c, int32_t v) {
    uint32_t i = hash(c);
    while (key_[i] >= 0) i = (i + 1) & (kCap - 1);
    key_[i] = c;
    val_[i] = v;
    used_.push_back(i);
  }
  void clear() {
    for (uint32_t i : used_) key_[i] = -1;
    used_.clear();
  }

 private:
  static constexpr uint32_t kCap = 4 * kLbUcolsLarge;
  static uint32_t hash(int32_t c) { return (uint32_t(c) * 2654435761u) >> 22; }  // 10 bits
  std::vector<int32_t> key_, val_;
  std::vector<uint32_t> used_;
};
static_assert(4 * kLbUcolsLarge == 1024, "ColMap hashes to 10 bits");

// One range of rows clustered on its own (ranges run in parallel).
struct RangeBlocks {
  std::vector<int32_t> order, nrow, nu, ucol;
  int64_t direct = 0;
  bool ok = true;
};

// Clustered cut over up to `ranges` row ranges clustered in parallel; false
// when the candidate scans exceed their work budget (48 steps per non-zero of a
// range: the greedy then gives up and the consecutive cut is used).
bool cut_clustered(const Operand &o, int32_t umax, const std::vector<int32_t> &cp,
                   const std::vector<int32_t> &cr, int64_t ranges, LbHost &h) {
  const int32_t nnz_max = lb_nnz_per_ucol(umax) * umax;
  const int32_t m = o.m, base = o.prp[0];
  const int32_t *rp = o.prp, *ci = o.pci;
  const int64_t nnz = int64_t(rp[m]) - base;
  h.lidx.assign(size_t(std::max<int64_t>(nnz, 1)), 0);
  // per row, written only by the range that owns the row: assigned to a
  // block; listed columns shared with the block; rows of the block it is
  // adjacent to (its index appears among their columns)
  std::vector<uint8_t> assigned(size_t(m), 0);
  std::vector<int32_t> ov(size_t(m), 0), adj(size_t(m), 0);
  std::vector<RangeBlocks> parts(16);
  auto len = [&](int32_t r) { return rp[r + 1] - rp[r]; };
  parallel_ranges(m, 4096, ranges, [&](int t, int64_t lo64, int64_t hi64) {
    const int32_t lo = int32_t(lo64), hi = int32_t(hi64);
    RangeBlocks &P = parts[size_t(t)];
    const int64_t budget = 48 * (int64_t(rp[hi]) - rp[lo]) + 65536;
    int64_t work = 0;
    ColMap cm;
    std::vector<int32_t> touched, empties;
    auto own = [&](int32_t r) { return r >= lo && r < hi && !assigned[size_t(r)]; };
    auto note = [&](int32_t r) {
      if (ov[size_t(r)] == 0 && adj[size_t(r)] == 0) touched.push_back(r);
    };
    auto flush_empties = [&] {
      if (empties.empty()) return;
      P.order.insert(P.order.end(), empties.begin(), empties.end());
      P.nrow.push_back(int32_t(empties.size()));
      P.nu.push_back(0);
      empties.clear();
    };
    for (int32_t seed = lo; seed < hi; ++seed) {
      if (assigned[size_t(seed)]) continue;
      const int32_t sl = len(seed);
      if (sl == 0) {  // empty rows share blocks of their own (C = beta * C)
        assigned[size_t(seed)] = 1;
        empties.push_back(seed);
        if (int32_t(empties.size()) == kLbRows) flush_empties();
        continue;
      }
      cm.clear();
      int32_t fresh = 0;  // distinct columns, counted up to umax + 1
      for (int64_t j = int64_t(rp[seed]) - base; j < int64_t(rp[seed + 1]) - base && fresh <= umax; ++j)
        if (cm.find(ci[j] - base) < 0) cm.insert(ci[j] - base, fresh++);
      cm.clear();
      if (fresh > umax || sl > nnz_max) {  // a block of its own without a list
        assigned[size_t(seed)] = 1;
        P.order.push_back(seed);
        P.nrow.push_back(1);
        P.nu.push_back(0);
        P.direct += sl;
        continue;
      }
      int32_t u = 0, rows = 0, cur = seed;
      int64_t bnnz = 0;
      touched.clear();
      for (;;) {
        assigned[size_t(cur)] = 1;
        P.order.push_back(cur);
        ++rows;
        bnnz += len(cur);
        for (int64_t j = int64_t(rp[cur]) - base; j < int64_t(rp[cur + 1]) - base; ++j) {
          const int32_t c = ci[j] - base;
          if (c < m && own(c)) {
            note(c);
            ++adj[size_t(c)];
          }
          int32_t p = cm.find(c);
          if (p < 0) {
            p = u++;
            cm.insert(c, p);
            P.ucol.push_back(c);
            const int32_t q0 = cp[size_t(c)], q1 = cp[size_t(c) + 1];
            if (q1 - q0 <= kDegreeCap) {
              work += q1 - q0;
              for (int32_t q = q0; q < q1; ++q) {
                const int32_t r = cr[size_t(q)];
                if (!own(r)) continue;
                note(r);
                ++ov[size_t(r)];
              }
            }
          }
          h.lidx[size_t(j)] = uint8_t(p);
        }
        if (rows >= kLbRows) break;
        // the candidate bringing the fewest unlisted columns (len - shared is
        // an upper bound on them, so the limits hold), then the most adjacent
        // block rows, then the lowest index
        int32_t best = -1, bnew = 0, badj = 0;
        work += int64_t(touched.size());
        for (const int32_t r : touched) {
          if (assigned[size_t(r)]) continue;
          const int32_t l = len(r), nw = l - ov[size_t(r)];
          if (l == 0 || u + nw > umax || bnnz + l > nnz_max) continue;
          const int32_t a = adj[size_t(r)];
          if (best < 0 || nw < bnew || (nw == bnew && (a > badj || (a == badj && r < best)))) {
            best = r;
            bnew = nw;
            badj = a;
          }
        }
        if (best < 0) break;
        cur = best;
      }
      for (const int32_t r : touched) ov[size_t(r)] = adj[size_t(r)] = 0;
      P.nrow.push_back(rows);
      P.nu.push_back(u);
      if (work > budget) {
        P.ok = false;
        return;
      }
    }
    flush_empties();
  });
  for (const RangeBlocks &P : parts)
    if (!P.ok) return false;
  h.brow.reserve(size_t(m));
  for (const RangeBlocks &P : parts) {
    h.brow.insert(h.brow.end(), P.order.begin(), P.order.end());
    h.ucol.insert(h.ucol.end(), P.ucol.begin(), P.ucol.end());
    for (size_t b = 0; b < P.nrow.size(); ++b) {
      h.row0.push_back(h.row0.back() + P.nrow[b]);
      h.uoff.push_back(h.uoff.back() + P.nu[b]);
    }
    h.direct += P.direct;
  }
  return int32_t(h.brow.size()) == m;
}

// Clustered blocks also keep their non-zeros in block order: first non-zero
// of each entry of the row order (prp), list positions, the CSR position of
// each (epos) and the values gathered through it on the device.
hipError_t upload_blocks(const Operand &o, const LbHost &h, int32_t umax, DevCsr &d) {
  d.lb_brow.reset();
  d.lb_prp.reset();
  d.lb_epos.reset();
  d.lb_val.reset();
  hipError_t e = d.lb_row0.upload(h.row0.data(), sizeof(int32_t) * h.row0.size());
  if (e == hipSuccess) e = d.lb_uoff.upload(h.uoff.data(), sizeof(int32_t) * h.uoff.size());
  if (e == hipSuccess)
    e = d.lb_ucol.upload(h.ucol.data(), sizeof(int32_t) * std::max<size_t>(h.ucol.size(), 1));
  // list positions and block-order values end with one padding entry (the
  // kernel stages them in aligned pairs); the CSR plan's values carry theirs
  if (h.brow.empty()) {
    std::vector<uint8_t> lidx(h.lidx);
    lidx.push_back(0);
    if (e == hipSuccess) e = d.lb_lidx.upload(lidx.data(), lidx.size());
  } else {
    const int32_t m = o.m, base = o.prp[0];
    const size_t nnz = h.lidx.size();
    std::vector<int32_t> prp(size_t(m) + 1, 0), epos(nnz + 1, 0);
    std::vector<uint8_t> lidx(nnz + 1, 0);
    for (int32_t i = 0; i < m; ++i) {
      const int32_t r = h.brow[size_t(i)];
      prp[size_t(i) + 1] = prp[size_t(i)] + (o.prp[r + 1] - o.prp[r]);
    }
    parallel_ranges(m, 65536, 16, [&](int, int64_t lo, int64_t hi) {
      for (int64_t i = lo; i < hi; ++i) {
        const int32_t r = h.brow[size_t(i)], j0 = o.prp[r] - base;
        const int32_t q0 = prp[size_t(i)], len = prp[size_t(i) + 1] - q0;
        for (int32_t t = 0; t < len; ++t) {
          epos[size_t(q0 + t)] = j0 + t;
          lidx[size_t(q0 + t)] = h.lidx[size_t(j0 + t)];
        }
      }
    });
    const size_t es = d.vs == VStore::F64 ? 8 : d.vs == VStore::F32 ? 4 : 2;
    if (e == hipSuccess) e = d.lb_lidx.upload(lidx.data(), lidx.size());
    if (e == hipSuccess) e = d.lb_brow.upload(h.brow.data(), sizeof(int32_t) * size_t(m));
    if (e == hipSuccess) e = d.lb_prp.upload(prp.data(), sizeof(int32_t) * prp.size());
    // the padding entry gathers the CSR plan's own padding value (zero)
    epos[nnz] = int32_t(nnz);
    if (e == hipSuccess) e = d.lb_epos.upload(epos.data(), sizeof(int32_t) * (nnz + 1));
    if (e == hipSuccess) e = d.lb_val.alloc(es * (nnz + 1));
    if (e == hipSuccess)
      e = plk::gather_vals(d.vs, d.lb_val.get(), d.vals.get(), d.lb_epos.as<int32_t>(),
                           int64_t(nnz + 1), stream());
  }
  d.lb_umax = umax;
  d.lb_state = e == hipSuccess ? 1 : 0;
  return e;
}
}  // namespace

// Lists of at most 128 columns keep the kernel's LDS small enough for six
// workgroups per CU (measured 1.3x faster than 256 on the 27-point stencil at
// k = 32 and 128); lists of 256 serve operands whose rows alone have more
// distinct columns (e.g. 8 x 8 block rows).  The blocks are used when their
// lists plus the non-zeros of list-less rows are at most 60% of nnz (a B row
// loaded into LDS then feeds 1.7 non-zeros or more); otherwise the direct
// kernels, which read B for every non-zero, run.
namespace {
// The blocks the plan uses (true), or the last cut tried when no list length
// limit gives enough reuse (false); umax = the limit of the returned cut.
bool choose_blocks(const Operand &o, int rows, LbHost &out, int32_t &umax_out) {
  const int64_t nnz = int64_t(o.prp[o.m]) - o.prp[0];
  std::vector<int32_t> cp, cr;
  Overlap ov;
  // the column -> rows structure and the sampled overlap are built on a thread
  // of their own while the consecutive cut runs
  const bool may_cluster = rows != 1 && o.m >= 2;
  std::thread pre;
  if (may_cluster)
    pre = std::thread([&] {
      rows_by_column(o, cp, cr);
      ov = sampled_overlap(o, cp, cr);
    });
  int overlap = -1;     // 1: rows share enough columns for clustering to pay
  int64_t ranges = 16;  // row ranges clustered in parallel
  bool used = false;
  for (const int32_t umax : {kLbUcolsSmall, kLbUcolsLarge}) {
    LbHost h;
    cut_consecutive(o, umax, h);
    if (may_cluster && double(h.direct) < 0.4 * double(nnz)) {
      if (overlap < 0) {
        pre.join();
        overlap = rows == 2 || ov.shared >= 0.25 ? 1 : 0;
        // rows whose partners are far apart (a scattered row order) are
        // clustered as one range when that stays quick
        if (ov.near < 0.5 && o.m <= (1 << 18)) ranges = 1;
      }
      LbHost c;
      if (overlap == 1 && cut_clustered(o, umax, cp, cr, ranges, c) &&
          (rows == 2 || double(c.cost()) < 0.9 * double(h.cost())))
        h = std::move(c);
    }
    used = double(h.cost()) <= 0.6 * double(nnz);
    out = std::move(h);
    umax_out = umax;
    if (used) break;
  }
  if (pre.joinable()) pre.join();
  return used;
}
}  // namespace

hipError_t build_spmm_blocks(const Operand &o, DevCsr &d, int rows) {
  LbHost h;
  int32_t umax = 0;
  const bool used = choose_blocks(o, rows, h, umax);
  d.lb_nblocks = int32_t(h.row0.size() - 1);
  d.lb_ucols = int64_t(h.ucol.size());
  d.lb_umax = umax;
  d.lb_clustered = !h.brow.empty();
  if (used) return upload_blocks(o, h, umax, d);
  d.lb_state = -1;
  return hipSuccess;
}

int spmm_blocks_host(const Operand &o, int rows, int32_t *nblocks, int32_t *list_max,
                     int64_t *list_total, int32_t *clustered, int32_t *row_order,
                     int32_t *block_starts) {
  LbHost h;
  int32_t umax = 0;
  const bool used = choose_blocks(o, rows, h, umax);
  const int32_t nb = int32_t(h.row0.size() - 1);
  if (nblocks) *nblocks = nb;
  if (list_max) *list_max = umax;
  if (list_total) *list_total = int64_t(h.ucol.size());
  if (clustered) *clustered = h.brow.empty() ? 0 : 1;
  if (row_order)
    for (int32_t i = 0; i < o.m; ++i) row_order[i] = h.brow.empty() ? i : h.brow[size_t(i)];
  if (block_starts) std::copy(h.row0.begin(), h.row0.end(), block_starts);
  return used ? 1 : -1;
}

hipError_t refresh_block_values(DevCsr &d) {
  if (d.lb_state != 1 || !d.lb_clustered || !d.lb_val.get()) return hipSuccess;
  return plk::gather_vals(d.vs, d.lb_val.get(), d.vals.get(), d.lb_epos.as<int32_t>(), d.nnz + 1,
                          stream());
}

}  // namespace pls
