// convert.cpp — host-side format conversions used when building device plans.
#include <algorithm>
#include <cstring>
#include <unordered_map>

#include "plsparse.hpp"

namespace pls {

// Expansion of a BSR handle to CSR with the reference's bsr2csr rules
// (src/compressed_sparse_rows.cpp:966-1056): rows visited block by block and
// j inside each block, explicit zeros dropped, index base kept.
void bsr_to_csr(const perflibs_spmat_top_t *A, HostCsr &out, int32_t &nnz_out) {
  const HostBsr &B = A->bsr;
  const int32_t b = B.b, base = B.rp[0];
  const bool col_major = B.layout == PERFLIBS_COL_MAJOR;
  const int64_t bb = int64_t(b) * b;
  const bool dbl = A->dt == DType::F64;
  auto val = [&](int64_t blk, int32_t i, int32_t j) -> double {
    const int64_t at = blk * bb + (col_major ? i + int64_t(b) * j : int64_t(i) * b + j);
    return dbl ? static_cast<const double *>(B.vals)[at]
               : double(static_cast<const float *>(B.vals)[at]);
  };
  out = HostCsr{};
  out.rp_own.reserve(size_t(A->m) + 1);
  out.rp_own.push_back(base);
  int32_t count = base;
  for (int32_t rb = 0; rb < B.mb; ++rb) {
    const int32_t k0 = B.rp[rb] - base, k1 = B.rp[rb + 1] - base;
    for (int32_t i = 0; i < b; ++i) {
      for (int32_t k = k0; k < k1; ++k) {
        const int32_t c0 = b * (B.ci[k] - base) + base;
        for (int32_t j = 0; j < b; ++j) {
          const double v = val(k, i, j);
          if (v == 0.0) continue;
          out.ci_own.push_back(c0 + j);
          if (dbl) out.vd_own.push_back(v);
          else out.vf_own.push_back(float(v));
          ++count;
        }
      }
      out.rp_own.push_back(count);
    }
  }
  nnz_out = count - base;
  out.point_to_own(A->dt);
}

void csr_transpose0(int32_t m, int32_t n, const int32_t *rp, const int32_t *ci,
                    const void *vals, DType dt, std::vector<int32_t> &trp,
                    std::vector<int32_t> &tci, std::vector<double> &tvd,
                    std::vector<float> &tvf) {
  const int32_t base = rp[0];
  const int64_t nnz = rp[m] - base;
  trp.assign(size_t(n) + 1, 0);
  for (int64_t j = 0; j < nnz; ++j) trp[ci[j] - base + 1]++;
  for (int32_t c = 0; c < n; ++c) trp[c + 1] += trp[c];
  std::vector<int32_t> fill(trp.begin(), trp.end() - 1);
  tci.resize(size_t(nnz));
  if (dt == DType::F64) tvd.resize(size_t(nnz));
  else tvf.resize(size_t(nnz));
  for (int32_t r = 0; r < m; ++r)
    for (int32_t j = rp[r] - base; j < rp[r + 1] - base; ++j) {
      const int32_t at = fill[ci[j] - base]++;
      tci[at] = r;
      if (dt == DType::F64) tvd[at] = static_cast<const double *>(vals)[j];
      else tvf[at] = static_cast<const float *>(vals)[j];
    }
}

// Blocks a 0-based CSR into b x b row-major blocks; block columns ascending in
// each block row, explicit zeros stored.  Duplicate (row, col) entries add up.
bool csr_to_bsr0(int32_t m, int32_t n, int32_t b, const int32_t *rp,
                 const int32_t *ci, const double *vd, const float *vf,
                 std::vector<int32_t> &brp, std::vector<int32_t> &bci,
                 std::vector<double> &bvd, std::vector<float> &bvf) {
  if (b <= 0 || m % b != 0 || n % b != 0) return false;
  const int32_t mb = m / b;
  const int64_t bb = int64_t(b) * b;
  brp.assign(size_t(mb) + 1, 0);
  bci.clear();
  bvd.clear();
  bvf.clear();
  std::vector<int32_t> cols;
  std::unordered_map<int32_t, int64_t> slot;
  for (int32_t rb = 0; rb < mb; ++rb) {
    cols.clear();
    for (int32_t i = rb * b; i < (rb + 1) * b; ++i)
      for (int32_t j = rp[i]; j < rp[i + 1]; ++j) cols.push_back(ci[j] / b);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    const int64_t first = int64_t(bci.size());
    slot.clear();
    for (size_t q = 0; q < cols.size(); ++q) {
      slot[cols[q]] = first + int64_t(q);
      bci.push_back(cols[q]);
    }
    if (vd) bvd.resize(size_t(bci.size() * bb), 0.0);
    else bvf.resize(size_t(bci.size() * bb), 0.0f);
    for (int32_t i = rb * b; i < (rb + 1) * b; ++i)
      for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
        const int64_t at = slot[ci[j] / b] * bb + int64_t(i - rb * b) * b + ci[j] % b;
        if (vd) bvd[size_t(at)] += vd[j];
        else bvf[size_t(at)] += vf[j];
      }
    brp[rb + 1] = int32_t(bci.size());
  }
  return true;
}

}  // namespace pls
