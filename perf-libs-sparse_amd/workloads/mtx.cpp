// mtx.cpp — Matrix Market coordinate reader producing 0-based CSR.
//
// Same acceptance rules and error texts as the reference's benchmark reader
// (bench/pl_bench_sparse_x_dense.c:124-240); see mtx.h.  The file is read in
// one piece and parsed with a cursor (entries may span lines, as with the
// reference's fscanf); rows are bucketed with a counting sort and each row's
// entries stably sorted by column, which is a stable (row, column) sort.
#include "mtx.h"

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Fail {
  const char *msg;
};

struct Cursor {
  const char *p, *end;
  void skip_ws() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\f' ||
                       *p == '\v'))
      ++p;
  }
  bool integer(long long &v) {
    skip_ws();
    if (p >= end) return false;
    char buf[32];
    size_t k = 0;
    while (p + k < end && k < sizeof(buf) - 1 && !std::strchr(" \t\n\r\f\v", p[k])) {
      buf[k] = p[k];
      ++k;
    }
    buf[k] = 0;
    char *e = nullptr;
    errno = 0;
    v = std::strtoll(buf, &e, 10);
    if (e == buf || errno) return false;
    p += e - buf;
    return true;
  }
  bool real(double &v) {
    skip_ws();
    if (p >= end) return false;
    char buf[64];
    size_t k = 0;
    while (p + k < end && k < sizeof(buf) - 1 && !std::strchr(" \t\n\r\f\v", p[k])) {
      buf[k] = p[k];
      ++k;
    }
    buf[k] = 0;
    char *e = nullptr;
    v = std::strtod(buf, &e);
    if (e == buf) return false;
    p += e - buf;
    return true;
  }
};

// Next line [p, eol) without its terminator; false at end of file.
bool next_line(Cursor &c, std::string &line) {
  if (c.p >= c.end) return false;
  const char *e = static_cast<const char *>(std::memchr(c.p, '\n', size_t(c.end - c.p)));
  if (!e) e = c.end;
  line.assign(c.p, e);
  c.p = e < c.end ? e + 1 : e;
  return true;
}

void read_csr(const char *path, pls_mtx_csr &out) {
  std::FILE *f = std::fopen(path, "rb");
  if (!f) throw Fail{"failed to open .mtx file"};
  std::string text;
  {
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
  }
  Cursor c{text.data(), text.data() + text.size()};
  std::string line;
  if (!next_line(c, line)) throw Fail{"failed to read Matrix Market header"};
  char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, sym[64] = {0};
  if (std::sscanf(line.c_str(), "%63s %63s %63s %63s %63s", banner, object, format, field, sym) !=
          5 ||
      std::strcmp(banner, "%%MatrixMarket") || std::strcmp(object, "matrix") ||
      std::strcmp(format, "coordinate"))
    throw Fail{"expected MatrixMarket coordinate file"};
  if (!std::strcmp(field, "complex")) throw Fail{"complex Matrix Market files not supported"};
  do {
    if (!next_line(c, line)) throw Fail{"missing size line"};
  } while (!line.empty() && line[0] == '%');
  long long m = 0, n = 0, cnt = 0;
  if (std::sscanf(line.c_str(), "%lld %lld %lld", &m, &n, &cnt) != 3) throw Fail{"bad size line"};
  if (m <= 0 || n <= 0 || cnt <= 0) throw Fail{"invalid matrix dimensions"};
  if (m > INT32_MAX || n > INT32_MAX) throw Fail{"matrix too large for 32-bit indices"};

  const bool pattern = !std::strcmp(field, "pattern");
  const bool mirror = !std::strcmp(sym, "symmetric") || !std::strcmp(sym, "hermitian");
  const bool skew = !std::strcmp(sym, "skew-symmetric");
  std::vector<int32_t> er, ec;
  std::vector<double> ev;
  const size_t cap = size_t(cnt) * ((mirror || skew) ? 2 : 1);
  er.reserve(cap);
  ec.reserve(cap);
  ev.reserve(cap);
  for (long long i = 0; i < cnt; ++i) {
    long long r = 0, col = 0;
    double v = 1.0;
    if (pattern) {
      if (!c.integer(r) || !c.integer(col)) throw Fail{"bad pattern entry"};
    } else if (!c.integer(r) || !c.integer(col) || !c.real(v)) {
      throw Fail{"bad numeric entry"};
    }
    if (r < 1 || r > m || col < 1 || col > n) throw Fail{"entry index out of range"};
    er.push_back(int32_t(r - 1));
    ec.push_back(int32_t(col - 1));
    ev.push_back(v);
    if ((mirror || skew) && r != col) {
      // a mirrored entry of a non-square file can fall outside the matrix
      if (col > m || r > n) throw Fail{"entry index out of range"};
      er.push_back(int32_t(col - 1));
      ec.push_back(int32_t(r - 1));
      ev.push_back(skew ? -v : v);
    }
  }
  const size_t used = er.size();
  if (used > size_t(INT32_MAX)) throw Fail{"matrix too large for 32-bit indices"};

  out.m = int32_t(m);
  out.n = int32_t(n);
  out.nnz = int64_t(used);
  out.row_ptr = static_cast<int32_t *>(std::calloc(size_t(m) + 1, sizeof(int32_t)));
  out.col_indx = static_cast<int32_t *>(std::malloc(std::max<size_t>(used, 1) * sizeof(int32_t)));
  out.vals = static_cast<double *>(std::malloc(std::max<size_t>(used, 1) * sizeof(double)));
  if (!out.row_ptr || !out.col_indx || !out.vals) throw Fail{"CSR allocation failed"};
  for (size_t e = 0; e < used; ++e) ++out.row_ptr[er[e] + 1];
  for (long long i = 0; i < m; ++i) out.row_ptr[i + 1] += out.row_ptr[i];
  std::vector<int32_t> next(out.row_ptr, out.row_ptr + m);
  std::vector<int32_t> order(used);
  for (size_t e = 0; e < used; ++e) order[size_t(next[er[e]]++)] = int32_t(e);
  for (long long i = 0; i < m; ++i)
    std::stable_sort(order.begin() + out.row_ptr[i], order.begin() + out.row_ptr[i + 1],
                     [&](int32_t a, int32_t b) { return ec[size_t(a)] < ec[size_t(b)]; });
  for (size_t q = 0; q < used; ++q) {
    out.col_indx[q] = ec[size_t(order[q])];
    out.vals[q] = ev[size_t(order[q])];
  }
}

}  // namespace

extern "C" {

int pls_mtx_read_csr(const char *path, pls_mtx_csr *out, char *err, size_t errlen) {
  if (!out) return 1;
  *out = pls_mtx_csr{};
  try {
    read_csr(path ? path : "", *out);
  } catch (const Fail &f) {
    pls_mtx_free(out);
    if (err && errlen) std::snprintf(err, errlen, "%s", f.msg);
    return 1;
  } catch (...) {
    pls_mtx_free(out);
    if (err && errlen) std::snprintf(err, errlen, "%s", "out of memory");
    return 1;
  }
  return 0;
}

void pls_mtx_free(pls_mtx_csr *a) {
  if (!a) return;
  std::free(a->row_ptr);
  std::free(a->col_indx);
  std::free(a->vals);
  *a = pls_mtx_csr{};
}

}  // extern "C"
