// gen.cpp — seeded synthetic matrices and vectors of the benchmark workloads
// (host-side, OpenMP).  Shared by the tests, bench.py and the selector's
// training driver; never part of the SpMV/SpMM path.
//
// Randomness: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014), one independent stream per (seed, stream
// id, row), so results do not depend on the thread count.  uniform[a,b) from a
// 53-bit fraction, N(0,1) by Box-Muller, distinct column sets by Floyd's
// algorithm (Bentley & Floyd, CACM 1987) then sorted.
//
// Two-phase matrix API: *_rows() fills row_ptr (length m+1, 0-based) and
// returns nnz; *_fill() writes col[nnz] and val[nnz].
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed, uint64_t stream, uint64_t row) {
    s = seed ^ (stream * 0xD1B54A32D192ED03ull) ^ ((row + 1) * 0x9E3779B97F4A7C15ull);
    next();
  }
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double u01() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double normal() {
    const double u1 = 1.0 - u01(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // uniform integer in [0, r)
  uint64_t below(uint64_t r) { return uint64_t((unsigned __int128)next() * r >> 64); }
};

// k distinct values from [0, w), sorted ascending (Floyd's algorithm).
void floyd(Rng &g, int64_t w, int64_t k, std::vector<int64_t> &out) {
  out.clear();
  if (k >= w) {
    for (int64_t i = 0; i < w; ++i) out.push_back(i);
    return;
  }
  size_t cap = 16;
  while (cap < size_t(2 * k)) cap <<= 1;
  std::vector<int64_t> table(cap, -1);
  auto insert = [&](int64_t v) -> bool {  // true if newly inserted
    size_t h = size_t((uint64_t(v) * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
    while (table[h] != -1) {
      if (table[h] == v) return false;
      h = (h + 1) & (cap - 1);
    }
    table[h] = v;
    return true;
  };
  for (int64_t j = w - k; j < w; ++j) {
    const int64_t t = int64_t(g.below(uint64_t(j) + 1));
    if (insert(t)) out.push_back(t);
    else {
      insert(j);
      out.push_back(j);
    }
  }
  std::sort(out.begin(), out.end());
}

template <class F> int64_t prefix_rows(int32_t m, int32_t *rp, F &&len) {
  std::vector<int64_t> l(static_cast<size_t>(m), 0);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m; ++i) l[i] = len(i);
  rp[0] = 0;
  int64_t acc = 0;
  for (int32_t i = 0; i < m; ++i) {
    acc += l[i];
    rp[i + 1] = int32_t(acc);
  }
  return acc;
}

}  // namespace

extern "C" {

void pls_gen_uniform(int64_t n, uint64_t seed, double lo, double hi, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (n + 4095) / 4096; ++b) {
    Rng g(seed, 7, uint64_t(b));
    for (int64_t i = b * 4096; i < std::min(n, (b + 1) * 4096); ++i) out[i] = lo + (hi - lo) * g.u01();
  }
}

void pls_gen_normal(int64_t n, uint64_t seed, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < (n + 4095) / 4096; ++b) {
    Rng g(seed, 8, uint64_t(b));
    for (int64_t i = b * 4096; i < std::min(n, (b + 1) * 4096); ++i) out[i] = g.normal();
  }
}

// Box stencil of radius r on an nx x ny grid (r = 1 and star = 1: 5-point
// Laplacian with diagonal 4 and off-diagonals -1; star = 0: full
// (2r+1)^2 box with diagonal (points - 1) and off-diagonals -1).
int64_t pls_gen_stencil2d(int32_t nx, int32_t ny, int32_t r, int32_t star, int32_t *rp,
                          int32_t *col, double *val) {
  const int32_t m = nx * ny;
  auto on = [&](int dy, int dx) { return star ? (dy == 0 || dx == 0) : true; };
  auto cols_of = [&](int32_t i, int32_t *c, double *v) -> int32_t {
    const int32_t y = i / nx, x = i % nx;
    int32_t cnt = 0, pts = 0;
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx)
        if (on(dy, dx)) ++pts;
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        if (!on(dy, dx)) continue;
        const int32_t yy = y + dy, xx = x + dx;
        if (yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
        if (c) {
          c[cnt] = yy * nx + xx;
          v[cnt] = (dy == 0 && dx == 0) ? double(pts - 1) : -1.0;
        }
        ++cnt;
      }
    return cnt;
  };
  if (!col) return prefix_rows(m, rp, [&](int32_t i) { return int64_t(cols_of(i, nullptr, nullptr)); });
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m; ++i) cols_of(i, col + rp[i], val + rp[i]);
  return rp[m];
}

// 3-D box stencil of radius r on nx*ny*nz (r = 1: the 27-point operator with
// diagonal 26 and neighbours -1; star = 1 gives the 7-point Laplacian).
int64_t pls_gen_stencil3d(int32_t nx, int32_t ny, int32_t nz, int32_t r, int32_t star,
                          int32_t *rp, int32_t *col, double *val) {
  const int32_t m = nx * ny * nz;
  auto on = [&](int dz, int dy, int dx) {
    return star ? ((dz == 0) + (dy == 0) + (dx == 0) >= 2) : true;
  };
  auto cols_of = [&](int32_t i, int32_t *c, double *v) -> int32_t {
    const int32_t z = i / (nx * ny), y = (i / nx) % ny, x = i % nx;
    int32_t cnt = 0, pts = 0;
    for (int dz = -r; dz <= r; ++dz)
      for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx)
          if (on(dz, dy, dx)) ++pts;
    for (int dz = -r; dz <= r; ++dz)
      for (int dy = -r; dy <= r; ++dy)
        for (int dx = -r; dx <= r; ++dx) {
          if (!on(dz, dy, dx)) continue;
          const int32_t zz = z + dz, yy = y + dy, xx = x + dx;
          if (zz < 0 || zz >= nz || yy < 0 || yy >= ny || xx < 0 || xx >= nx) continue;
          if (c) {
            c[cnt] = (zz * ny + yy) * nx + xx;
            v[cnt] = (dz == 0 && dy == 0 && dx == 0) ? double(pts - 1) : -1.0;
          }
          ++cnt;
        }
    return cnt;
  };
  if (!col) return prefix_rows(m, rp, [&](int32_t i) { return int64_t(cols_of(i, nullptr, nullptr)); });
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < m; ++i) cols_of(i, col + rp[i], val + rp[i]);
  return rp[m];
}

// Power-law rows: length round(exp(mu + sigma*z)) clipped to [lo, hi] (and
// to n); columns uniform over [0, n) without duplicates, sorted; values N(0,1).
int64_t pls_gen_powerlaw_rows(int32_t m, int32_t n, double mu, double sigma, int32_t lo,
                              int32_t hi, uint64_t seed, int32_t *rp) {
  return prefix_rows(m, rp, [&](int32_t i) {
    Rng g(seed, 1, uint64_t(i));
    double l = std::round(std::exp(mu + sigma * g.normal()));
    l = std::min<double>(std::max<double>(l, lo), hi);
    return int64_t(std::min<double>(l, n));
  });
}

void pls_gen_powerlaw_fill(int32_t m, int32_t n, const int32_t *rp, uint64_t seed, int32_t *col,
                           double *val) {
#pragma omp parallel
  {
    std::vector<int64_t> cs;
#pragma omp for schedule(dynamic, 4096)
    for (int32_t i = 0; i < m; ++i) {
      Rng g(seed, 2, uint64_t(i));
      floyd(g, n, rp[i + 1] - rp[i], cs);
      Rng gv(seed, 3, uint64_t(i));
      for (size_t j = 0; j < cs.size(); ++j) {
        col[rp[i] + j] = int32_t(cs[j]);
        val[rp[i] + j] = gv.normal();
      }
    }
  }
}

// Banded rows: nnz per row uniform in [lo, hi], columns distinct inside
// [i - bw/2, i + bw/2] clipped to [0, n); values N(0,1).
int64_t pls_gen_banded_rows(int32_t m, int32_t n, int32_t bw, int32_t lo, int32_t hi,
                            uint64_t seed, int32_t *rp) {
  return prefix_rows(m, rp, [&](int32_t i) {
    Rng g(seed, 4, uint64_t(i));
    const int64_t c0 = std::max<int64_t>(0, int64_t(i) - bw / 2);
    const int64_t c1 = std::min<int64_t>(n, int64_t(i) + bw / 2 + 1);
    const int64_t k = lo + int64_t(g.below(uint64_t(hi - lo + 1)));
    return std::min<int64_t>(k, std::max<int64_t>(0, c1 - c0));
  });
}

void pls_gen_banded_fill(int32_t m, int32_t n, int32_t bw, const int32_t *rp, uint64_t seed,
                         int32_t *col, double *val) {
#pragma omp parallel
  {
    std::vector<int64_t> cs;
#pragma omp for schedule(dynamic, 4096)
    for (int32_t i = 0; i < m; ++i) {
      const int64_t c0 = std::max<int64_t>(0, int64_t(i) - bw / 2);
      const int64_t c1 = std::min<int64_t>(n, int64_t(i) + bw / 2 + 1);
      Rng g(seed, 5, uint64_t(i));
      floyd(g, c1 - c0, rp[i + 1] - rp[i], cs);
      Rng gv(seed, 6, uint64_t(i));
      for (size_t j = 0; j < cs.size(); ++j) {
        col[rp[i] + j] = int32_t(c0 + cs[j]);
        val[rp[i] + j] = gv.normal();
      }
    }
  }
}

// Uniform random rows: nnz per row uniform in [lo, hi], columns uniform over
// [0, n) without duplicates; values N(0,1).
int64_t pls_gen_random_rows(int32_t m, int32_t n, int32_t lo, int32_t hi, uint64_t seed,
                            int32_t *rp) {
  return prefix_rows(m, rp, [&](int32_t i) {
    Rng g(seed, 9, uint64_t(i));
    return std::min<int64_t>(n, lo + int64_t(g.below(uint64_t(hi - lo + 1))));
  });
}

void pls_gen_random_fill(int32_t m, int32_t n, const int32_t *rp, uint64_t seed, int32_t *col,
                         double *val) {
  pls_gen_powerlaw_fill(m, n, rp, seed ^ 0x5a5a5a5a5a5a5a5aull, col, val);
}

// Block-structured BSR: a (bx, by, bz) grid of b x b dense blocks coupled by
// the 27-point pattern (bz = by = 1 reduces to 1-D / 2-D patterns), block
// values N(0,1), row-major blocks.  brp has bx*by*bz+1 entries.
int64_t pls_gen_block_stencil(int32_t bx, int32_t by, int32_t bz, int32_t b, uint64_t seed,
                              int32_t *brp, int32_t *bci, double *bval) {
  const int32_t mb = bx * by * bz;
  auto blocks_of = [&](int32_t i, int32_t *c) -> int32_t {
    const int32_t z = i / (bx * by), y = (i / bx) % by, x = i % bx;
    int32_t cnt = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int32_t zz = z + dz, yy = y + dy, xx = x + dx;
          if (zz < 0 || zz >= bz || yy < 0 || yy >= by || xx < 0 || xx >= bx) continue;
          if (c) c[cnt] = (zz * by + yy) * bx + xx;
          ++cnt;
        }
    return cnt;
  };
  if (!bci) return prefix_rows(mb, brp, [&](int32_t i) { return int64_t(blocks_of(i, nullptr)); });
  const int64_t bb = int64_t(b) * b;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < mb; ++i) {
    const int32_t cnt = blocks_of(i, bci + brp[i]);
    Rng g(seed, 10, uint64_t(i));
    for (int64_t q = 0; q < cnt * bb; ++q) bval[int64_t(brp[i]) * bb + q] = g.normal();
  }
  return brp[mb];
}

// Random block-structured BSR: each block row has k block columns drawn
// without duplicates from a window of w block columns around the diagonal.
int64_t pls_gen_block_random_rows(int32_t mb, int32_t nb, int32_t k, int32_t w, int32_t *brp) {
  return prefix_rows(mb, brp, [&](int32_t i) {
    const int64_t c0 = std::max<int64_t>(0, int64_t(i) - w / 2);
    const int64_t c1 = std::min<int64_t>(nb, int64_t(i) + w / 2 + 1);
    return std::min<int64_t>(k, c1 - c0);
  });
}

void pls_gen_block_random_fill(int32_t mb, int32_t nb, int32_t b, int32_t w, const int32_t *brp,
                               uint64_t seed, int32_t *bci, double *bval) {
  const int64_t bb = int64_t(b) * b;
#pragma omp parallel
  {
    std::vector<int64_t> cs;
#pragma omp for schedule(dynamic, 1024)
    for (int32_t i = 0; i < mb; ++i) {
      const int64_t c0 = std::max<int64_t>(0, int64_t(i) - w / 2);
      const int64_t c1 = std::min<int64_t>(nb, int64_t(i) + w / 2 + 1);
      Rng g(seed, 11, uint64_t(i));
      floyd(g, c1 - c0, brp[i + 1] - brp[i], cs);
      Rng gv(seed, 12, uint64_t(i));
      for (size_t j = 0; j < cs.size(); ++j) {
        bci[brp[i] + j] = int32_t(c0 + cs[j]);
        for (int64_t q = 0; q < bb; ++q) bval[(int64_t(brp[i]) + int64_t(j)) * bb + q] = gv.normal();
      }
    }
  }
}

int pls_gen_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
