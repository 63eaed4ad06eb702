// sell_build.cpp — SELL-C-sigma layout construction at optimize time.
//
// Follows the reference's csr2scs (src/sell_c_sigma.cpp:145-336) and
// gen_data_vectors (src/sell_c_sigma.cpp:584-705) so the device layout is the
// same matrix object the reference builds: rows sorted inside sigma-row windows
// by (length, row index) descending, chunk width = longest row of the chunk,
// chunk offsets the exclusive prefix of width*C, per-chunk minimum column,
// column offsets stored in the fewest of 1/2/4 bytes that hold the largest
// per-chunk column range, column-major slices, zero padding.
//
// The sort key packs (length, row) into one 64-bit word (length in the high
// half), so a descending integer sort is exactly the reference's
// std::greater<> on std::pair<int,int>; window padding is the key 0 = (0, 0).
#include <algorithm>
#include <climits>
#include <cstring>
#include <functional>
#include <thread>

#include "plsparse.hpp"

namespace pls {

namespace {
template <class F> void parallel_for(int64_t n, F &&fn) {
  unsigned hw = std::thread::hardware_concurrency();
  int64_t nt = std::min<int64_t>(hw ? hw : 1, 16);
  if (n < 4096 || nt <= 1) {
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

bool sell_structure(int32_t m, const int32_t *rp, const int32_t *ci, int32_t C,
                    int32_t sigma, SellHost &s) {
  if (m <= 0 || C <= 0 || C > 64 || sigma < 0 || sigma % C != 0) return false;
  s = SellHost{};
  s.C = C;
  s.sigma = sigma;
  const int64_t nC = (int64_t(m) + C - 1) / C;
  s.nC = nC;
  const bool sorted = sigma != 0;
  const int64_t win = sorted ? sigma : C;
  const int64_t nwin = (int64_t(m) + win - 1) / win;
  const int32_t base = rp[0];

  std::vector<uint64_t> key(size_t(nwin * win), 0);
  parallel_for(m, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i)
      key[i] = (uint64_t(uint32_t(rp[i + 1] - rp[i])) << 32) | uint32_t(i);
  });
  if (sorted)
    parallel_for(nwin, [&](int64_t lo, int64_t hi) {
      for (int64_t w = lo; w < hi; ++w)
        std::sort(key.begin() + w * win, key.begin() + (w + 1) * win,
                  std::greater<uint64_t>());
    });
  auto len_of = [](uint64_t k) { return int64_t(k >> 32); };
  auto row_of = [](uint64_t k) { return int64_t(k & 0xffffffffu); };

  s.cl.assign(nC, 0);
  s.cs.assign(nC, 0);
  const int64_t chunks_per_win = win / C;
  for (int64_t c = 0; c < nC; ++c) {
    const int64_t w = c / chunks_per_win, r0 = (c * C) % win;
    if (sorted) {
      s.cl[c] = len_of(key[w * win + r0]);
    } else {  // unsorted: widest row of the whole window (= the chunk)
      int64_t mx = 0;
      for (int64_t k = 0; k < win; ++k) mx = std::max(mx, len_of(key[w * win + k]));
      s.cl[c] = mx;
    }
  }
  for (int64_t c = 1; c < nC; ++c) s.cs[c] = s.cs[c - 1] + s.cl[c - 1] * C;

  // Per-row column range (empty rows: [INT32_MAX, 0]).
  std::vector<int32_t> rmin(m), rmax(m);
  parallel_for(m, [&](int64_t lo, int64_t hi) {
    for (int64_t i = lo; i < hi; ++i) {
      int32_t mn = INT32_MAX, mx = 0;
      for (int32_t j = rp[i] - base; j < rp[i + 1] - base; ++j) {
        const int32_t c = ci[j] - base;
        mn = std::min(mn, c);
        mx = std::max(mx, c);
      }
      rmin[i] = mn;
      rmax[i] = mx;
    }
  });

  s.permd2in.assign(nC * C, -1);
  s.in2permd.assign(nC * C, -1);
  s.cmin.assign(nC, 0);
  s.chunk_min_nnz.assign(nC, 0);
  std::vector<int32_t> crange(nC, 0);
  parallel_for(nC, [&](int64_t lo, int64_t hi) {
    for (int64_t c = lo; c < hi; ++c) {
      const int64_t w = c / chunks_per_win, r0 = (c * C) % win;
      const int64_t rows = std::min<int64_t>(C, int64_t(m) - c * C);
      int32_t mn = INT32_MAX, mx = 0, cnt = 0;
      for (int64_t r = 0; r < rows; ++r) {
        const uint64_t k = key[w * win + r0 + r];
        const int64_t in = row_of(k);
        cnt += int32_t(len_of(k));
        mn = std::min(mn, rmin[in]);
        mx = std::max(mx, rmax[in]);
        s.in2permd[in] = c * C + r;
        s.permd2in[c * C + r] = in;
      }
      s.chunk_min_nnz[c] = cnt;
      s.cmin[c] = mn;
      crange[c] = mx - mn;  // negative for a chunk of empty rows
    }
  });
  int32_t range = 0;
  for (int64_t c = 0; c < nC; ++c) range = std::max(range, crange[c]);
  int bits = 0;
  while (range) {
    range >>= 1;
    ++bits;
  }
  int bytes = bits / 8 + 1;
  if (bytes == 3) bytes = 4;
  if (bytes > 4) bytes = 8;  // unreachable with int32 columns
  s.col_bytes = bytes;

  s.n_full_chunks = int32_t(s.permd2in[nC * C - 1] < 0 ? nC - 1 : nC);
  int32_t pad = 0;
  while (pad < C && s.permd2in[nC * C - pad - 1] < 0) ++pad;
  s.ntidyup = pad == 0 ? 0 : C - pad;
  s.padded = s.cs[nC - 1] + s.cl[nC - 1] * C;
  return true;
}

// Padding efficiency of a SELL-C-sigma layout from its chunk widths and
// chunk_min_nnz.  beta is get_beta (src/sell_c_sigma.cpp:369-381): non-zeros
// over the slots of real rows, cl * C per full chunk and cl * ntidyup for a
// partial last chunk.  The per-thread figures follow get_beta_threads
// (:389-452) over the nnz-division chunk split of gen_parallel_decomp
// (:493-552), including its threshold recurrence vals_next += (chunks of the
// thread) * C + vals_per_thread.  Where that recurrence would open a thread
// past the last one (the reference asserts there) the remaining chunks stay
// with the last thread.  64-bit arithmetic throughout.
SellBeta sell_padding_beta(const std::vector<int32_t> &cl, const std::vector<int64_t> &cmn,
                           int32_t C, int32_t m, int32_t nthreads) {
  SellBeta out;
  const int64_t nC = int64_t(cl.size());
  if (nC == 0 || C <= 0 || nthreads <= 0) return out;
  std::vector<int64_t> cs(size_t(nC) + 1, 0);
  for (int64_t c = 0; c < nC; ++c) cs[c + 1] = cs[c] + int64_t(cl[c]) * C;
  int64_t nnz = 0;
  for (int64_t v : cmn) nnz += v;
  const int64_t ntidyup = m % C, n_full = m / C;
  const int64_t last_w = cl[nC - 1];
  const int64_t par = cs[nC - 1] + (ntidyup == 0 ? last_w * C : 0);  // slots of full chunks
  const int64_t slots = par + last_w * ntidyup;
  out.beta = double(nnz) / double(slots);
  if (nthreads == 1) {
    out.min_beta = out.max_beta = out.beta;
    return out;
  }
  // chunk split: thread t owns chunks [start[t], start[t + 1])
  std::vector<int64_t> start(size_t(nthreads) + 1, nC), count(size_t(nthreads), 0),
      tmin(size_t(nthreads), 0);
  const int64_t per = std::max<int64_t>((par - nthreads + 1) / nthreads, 1);
  int64_t next = per, acc = 0;
  int32_t t = 0;
  start[0] = 0;
  for (int64_t c = 0; c < n_full; ++c) {
    if (acc >= next && cl[c] > 0) {
      if (t + 1 < nthreads) {
        next += count[t] * C + per;
        start[++t] = c;
      } else {
        out.clamped = true;
      }
    }
    acc += int64_t(cl[c]) * C;
    ++count[t];
  }
  // the last thread with chunks also takes the partial last chunk
  for (int32_t u = 0; u <= t; ++u) {
    const int64_t hi = u < t ? start[u + 1] : nC;
    for (int64_t c = start[u]; c < hi; ++c) tmin[u] += cmn[c];
  }
  double lo = 1.0, hi = 0.0;
  for (int32_t u = 0; u < t && nC > 1; ++u) {
    if (tmin[u] > 0) {
      const double r = double(tmin[u]) / double(cs[start[u + 1]] - cs[start[u]]);
      lo = std::min(lo, r);
      hi = std::max(hi, r);
    }
  }
  if (tmin[t] > 0) {
    const double r = double(tmin[t]) / double(slots - cs[start[t]]);
    lo = std::min(lo, r);
    hi = std::max(hi, r);
  } else {
    lo = 0.0;
  }
  out.min_beta = lo;
  out.max_beta = hi;
  return out;
}

void sell_fill(const SellHost &s, int32_t m, const int32_t *rp, const int32_t *ci,
               const void *vals, DType dt, std::vector<uint8_t> &vals_out,
               std::vector<uint8_t> &offs_out) {
  const size_t vsz = dt == DType::F64 ? 8 : 4;
  const int64_t P = std::max<int64_t>(s.padded, 1);
  vals_out.assign(size_t(P) * vsz, 0);
  offs_out.assign(size_t(P) * s.col_bytes, 0);
  const int32_t base = rp[0];
  const int64_t C = s.C;
  parallel_for(s.nC, [&](int64_t lo, int64_t hi) {
    for (int64_t c = lo; c < hi; ++c) {
      const int64_t rows = std::min<int64_t>(C, int64_t(m) - c * C);
      for (int64_t r = 0; r < rows; ++r) {
        const int64_t in = s.permd2in[c * C + r];
        const int64_t start = rp[in] - base, len = rp[in + 1] - rp[in];
        for (int64_t j = 0; j < len; ++j) {
          const int64_t at = s.cs[c] + j * C + r;
          std::memcpy(&vals_out[size_t(at) * vsz],
                      static_cast<const uint8_t *>(vals) + size_t(start + j) * vsz, vsz);
          const int64_t off = int64_t(ci[start + j] - base) - s.cmin[c];
          switch (s.col_bytes) {
            case 1: offs_out[size_t(at)] = uint8_t(off); break;
            case 2: reinterpret_cast<uint16_t *>(offs_out.data())[at] = uint16_t(off); break;
            default: reinterpret_cast<uint32_t *>(offs_out.data())[at] = uint32_t(off); break;
          }
        }
        // slots j >= len stay zero: value 0, offset 0 (padding)
      }
    }
  });
}

}  // namespace pls
