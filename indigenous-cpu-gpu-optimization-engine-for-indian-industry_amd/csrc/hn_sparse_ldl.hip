// Sparse LDL' of a symmetric matrix on the device, on the exact
// minimum-degree ordering of hn_order.cpp.
//
// Reference semantics (SparseCholesky, core/numerical/factorization.cpp):
//   * numeric factorisation (:1248-1325): B = P A P' (permute_symmetric,
//     :196-234); row i of the unit lower L is found up-looking, entry by
//     entry in column order: L(i,j) = (B(i,j) - sum_c L(j,c) L(i,c) D(c)) /
//     D(j) over the columns c < j that rows i and j share, in ascending c;
//     D(i) = B(i,i) - sum_c L(i,c) L(i,c) D(c); a pivot <= singular_tol
//     (1e-10) is replaced by singular_tol and the matrix marked singular
//     (:1309-1313).
//   * solve (:1364-1396, apply_perm :1422-1436): permute, forward
//     substitution by rows, division by the nonzero pivots, backward
//     substitution that subtracts row i's multiples from x[j] for i from the
//     last row down, permute back.
// Every sum here is formed in exactly that order with separate roundings
// (no FMA contraction), so L, D and the solution are bit-identical to the
// reference's for the same matrix.
//
// Parallel structure: a row of L (and of each triangular solve) depends only
// on the rows its entries name, all of a lower level (hn_order.cpp), so the
// rows of a level run at once, one thread per row; levels run in order
// (descending for the backward substitution, whose rows wait on higher
// rows).  Each row's own loop is the reference's sequential loop.
#include "hn_device.hpp"
#include "hn_internal.hpp"
#include "hn_sldl.hpp"

#include <algorithm>
#include <vector>

namespace hn {

namespace {

constexpr int kRowsNT = 128;

__global__ __launch_bounds__(kRowsNT) void k_sldl_factor(const int* __restrict__ rows, int cnt,
                                                         const long long* __restrict__ Lrp,
                                                         const int* __restrict__ Lci,
                                                         const long long* __restrict__ bidx,
                                                         const double* __restrict__ aval, double* L, double* D,
                                                         double tol, int* __restrict__ sing) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * kRowsNT + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  const long long b = Lrp[i], e = Lrp[i + 1];
  for (long long k = b; k < e; ++k) {
    const int j = Lci[k];
    const long long src = bidx[k];
    const double aij = src >= 0 ? aval[src] : 0.0;
    if (j < i) {
      double sum = aij;
      long long kj = Lrp[j], ki = b;
      const long long ej = Lrp[j + 1];
      while (kj < ej && ki < k) {
        const int cj = Lci[kj], ci = Lci[ki];
        if (cj == ci) {
          sum -= L[kj] * L[ki] * D[cj];
          ++kj;
          ++ki;
        } else if (cj < ci) {
          ++kj;
        } else {
          ++ki;
        }
      }
      L[k] = sum / D[j];
    } else {
      double sum = 0.0;
      for (long long kk = b; kk < k; ++kk) sum += L[kk] * L[kk] * D[Lci[kk]];
      double val = aij - sum;
      if (val <= tol) {
        val = tol;
        atomicAdd(sing, 1);
      }
      D[i] = val;
      L[k] = 1.0;
    }
  }
}

__global__ void k_sldl_perm_in(const long long* __restrict__ perm, const double* __restrict__ b,
                               double* __restrict__ xin, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) xin[r] = b[perm[r]];
}

__global__ __launch_bounds__(kRowsNT) void k_sldl_fwd(const int* __restrict__ rows, int cnt,
                                                      const long long* __restrict__ Lrp, const int* __restrict__ Lci,
                                                      const double* __restrict__ L, const double* __restrict__ D,
                                                      const double* __restrict__ xin, double* xf,
                                                      double* __restrict__ xd) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * kRowsNT + threadIdx.x;
  if (t >= cnt) return;
  const int i = rows[t];
  double sum = xin[i];
  for (long long k = Lrp[i]; k < Lrp[i + 1]; ++k) {
    const int j = Lci[k];
    if (j < i) sum -= L[k] * xf[j];
  }
  xf[i] = sum;  // the forward value (later rows read it undivided)
  const double d = D[i];
  xd[i] = (d != 0.0) ? sum / d : sum;
}

__global__ __launch_bounds__(kRowsNT) void k_sldl_bwd(const int* __restrict__ rows, int cnt,
                                                      const long long* __restrict__ Ucp, const int* __restrict__ Uri,
                                                      const long long* __restrict__ Upos,
                                                      const double* __restrict__ L, double* xd) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * kRowsNT + threadIdx.x;
  if (t >= cnt) return;
  const int j = rows[t];
  double v = xd[j];
  for (long long p = Ucp[j]; p < Ucp[j + 1]; ++p) v -= L[Upos[p]] * xd[Uri[p]];
  xd[j] = v;
}

__global__ void k_sldl_perm_out(const long long* __restrict__ perm, const double* __restrict__ xd,
                                double* __restrict__ x, int n, int accumulate) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const long long o = perm[r];
  x[o] = accumulate ? x[o] + xd[r] : xd[r];
}

__global__ void k_sldl_zero_int(int* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

template <typename T>
int upload_vec(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  HN_TRY(d.reserve(sizeof(T) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) HN_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
  return HN_OK;
}

// Symbolic arrays to the device (once per handle).
int sldl_upload(hn_sldl* h, hipStream_t s) {
  if (h->on_device) return HN_OK;
  const SldlSymbolic& S = h->S;
  const int64_t nnz = S.Lrp[S.n];
  HN_TRY(upload_vec(h->Lrp, S.Lrp, s));
  HN_TRY(upload_vec(h->Lci, S.Lci, s));
  HN_TRY(upload_vec(h->Ucp, S.Ucp, s));
  HN_TRY(upload_vec(h->Uri, S.Uri, s));
  HN_TRY(upload_vec(h->Upos, S.Upos, s));
  HN_TRY(upload_vec(h->rows, S.level_rows, s));
  HN_TRY(upload_vec(h->perm, S.perm, s));
  HN_TRY(h->L.reserve(sizeof(double) * std::max<int64_t>(nnz, 1)));
  HN_TRY(h->D.reserve(sizeof(double) * std::max<int64_t>(S.n, 1)));
  HN_TRY(h->xin.reserve(sizeof(double) * std::max<int64_t>(S.n, 1)));
  HN_TRY(h->xf.reserve(sizeof(double) * std::max<int64_t>(S.n, 1)));
  HN_TRY(h->xd.reserve(sizeof(double) * std::max<int64_t>(S.n, 1)));
  HN_TRY(h->sing.reserve(sizeof(int)));
  h->on_device = true;
  return HN_OK;
}

// bidx: for every entry (r, c) of L, the index in a's values of
// a(perm[r], perm[c]) (binary search in the row), or -1 where a has none.
int sldl_bind(hn_sldl* h, const hn_csr* a, hipStream_t s) {
  if (h->bound == a && h->bound_nnz == a->nnz) return HN_OK;
  const SldlSymbolic& S = h->S;
  HN_CHECK(a->nrows == S.n && a->ncols == S.n, HN_ERR_INVALID, "sparse LDL': matrix shape differs from the analysis");
  HN_CHECK(a->flags & HN_F64, HN_ERR_INVALID, "sparse LDL': matrix has no FP64 values");
  std::vector<int> rp(S.n + 1), ci(std::max<int64_t>(a->nnz, 1));
  HN_HIP(hipMemcpyAsync(rp.data(), a->row_ptr.p, sizeof(int) * (S.n + 1), hipMemcpyDeviceToHost, s));
  if (a->nnz) HN_HIP(hipMemcpyAsync(ci.data(), a->col.p, sizeof(int) * a->nnz, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  const int64_t nnz = S.Lrp[S.n];
  std::vector<long long> bidx(nnz, -1);
  for (int64_t r = 0; r < S.n; ++r) {
    const int64_t i = S.perm[r];
    const int* b = ci.data() + rp[i];
    const int* e = ci.data() + rp[i + 1];
    HN_CHECK(std::is_sorted(b, e), HN_ERR_INVALID, "sparse LDL': matrix rows must have ascending columns");
    for (int64_t k = S.Lrp[r]; k < S.Lrp[r + 1]; ++k) {
      const int j = (int)S.perm[S.Lci[k]];
      const int* f = std::lower_bound(b, e, j);
      if (f != e && *f == j) bidx[k] = (long long)(f - ci.data());
    }
  }
  HN_TRY(upload_vec(h->bidx, bidx, s));
  HN_HIP(hipStreamSynchronize(s));
  h->bound = a;
  h->bound_nnz = a->nnz;
  return HN_OK;
}

}  // namespace

int sldl_create(int64_t n, const int64_t* rp, const int64_t* ci, double work_limit, double max_lnnz, hn_sldl** out) {
  hn_sldl* h = new hn_sldl();
  const int st = sldl_symbolic(n, rp, ci, work_limit, max_lnnz, h->S);
  if (st != HN_OK) {
    delete h;
    return st;
  }
  *out = h;
  return HN_OK;
}

int sldl_factor(hn_sldl* h, const hn_csr* a, double singular_tol, hipStream_t s) {
  HN_CHECK(!h->S.aborted, HN_ERR_UNSUPPORTED, "sparse LDL': the ordering stopped at its work limit");
  HN_TRY(sldl_upload(h, s));
  HN_TRY(sldl_bind(h, a, s));
  const SldlSymbolic& S = h->S;
  hipLaunchKernelGGL(k_sldl_zero_int, dim3(1), dim3(64), 0, s, h->sing.as<int>());
  const int nl = (int)S.level_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int cnt = (int)(S.level_ptr[l + 1] - S.level_ptr[l]);
    hipLaunchKernelGGL(k_sldl_factor, dim3((cnt + kRowsNT - 1) / kRowsNT), dim3(kRowsNT), 0, s,
                       h->rows.as<int>() + S.level_ptr[l], cnt, h->Lrp.as<long long>(), h->Lci.as<int>(),
                       h->bidx.as<long long>(), a->val64.as<double>(), h->L.as<double>(), h->D.as<double>(),
                       singular_tol, h->sing.as<int>());
  }
  count_launch(1 + nl);
  return check_launch("k_sldl_factor");
}

int sldl_solve(const hn_sldl* h, const double* b, double* x, bool accumulate, hipStream_t s) {
  const SldlSymbolic& S = h->S;
  HN_CHECK(h->on_device && h->bound, HN_ERR_INVALID, "sparse LDL': solve before a factorisation");
  const int n = (int)S.n;
  if (n == 0) return HN_OK;
  hipLaunchKernelGGL(k_sldl_perm_in, dim3(grid_for(n, 256)), dim3(256), 0, s, h->perm.as<long long>(), b,
                     h->xin.as<double>(), n);
  const int nl = (int)S.level_ptr.size() - 1;
  for (int l = 0; l < nl; ++l) {
    const int cnt = (int)(S.level_ptr[l + 1] - S.level_ptr[l]);
    hipLaunchKernelGGL(k_sldl_fwd, dim3((cnt + kRowsNT - 1) / kRowsNT), dim3(kRowsNT), 0, s,
                       h->rows.as<int>() + S.level_ptr[l], cnt, h->Lrp.as<long long>(), h->Lci.as<int>(),
                       h->L.as<double>(), h->D.as<double>(), h->xin.as<double>(), h->xf.as<double>(),
                       h->xd.as<double>());
  }
  for (int l = nl - 1; l >= 0; --l) {
    const int cnt = (int)(S.level_ptr[l + 1] - S.level_ptr[l]);
    hipLaunchKernelGGL(k_sldl_bwd, dim3((cnt + kRowsNT - 1) / kRowsNT), dim3(kRowsNT), 0, s,
                       h->rows.as<int>() + S.level_ptr[l], cnt, h->Ucp.as<long long>(), h->Uri.as<int>(),
                       h->Upos.as<long long>(), h->L.as<double>(), h->xd.as<double>());
  }
  hipLaunchKernelGGL(k_sldl_perm_out, dim3(grid_for(n, 256)), dim3(256), 0, s, h->perm.as<long long>(),
                     h->xd.as<double>(), x, n, accumulate ? 1 : 0);
  count_launch(2 + 2 * nl);
  return check_launch("sparse LDL' solve");
}

void sldl_destroy(hn_sldl* h) { delete h; }

}  // namespace hn

using namespace hn;

extern "C" int hn_sldl_analyze(uint64_t n, const uint64_t* row_ptr, const uint64_t* col, double work_limit,
                               hn_sldl** out) {
  HN_CHECK(out, HN_ERR_INVALID, "hn_sldl_analyze: null output");
  HN_CHECK(n < (uint64_t)INT32_MAX, HN_ERR_UNSUPPORTED, "hn_sldl_analyze: dimension must be below 2^31");
  HN_CHECK(n == 0 || row_ptr, HN_ERR_INVALID, "hn_sldl_analyze: null row pointer");
  const uint64_t nnz = n ? row_ptr[n] : 0;
  HN_CHECK(nnz == 0 || col, HN_ERR_INVALID, "hn_sldl_analyze: null column indices");
  std::vector<int64_t> rp(n + 1), ci(nnz);
  for (uint64_t i = 0; i <= n; ++i) rp[i] = (int64_t)(n ? row_ptr[i] : 0);
  for (uint64_t k = 0; k < nnz; ++k) ci[k] = (int64_t)std::min<uint64_t>(col[k], (uint64_t)INT64_MAX);
  return sldl_create((int64_t)n, rp.data(), ci.data(), work_limit, 0.0, out);
}

extern "C" int hn_sldl_sizes(const hn_sldl* h, uint64_t* n, uint64_t* l_nnz, uint64_t* levels, int* aborted) {
  HN_CHECK(h && n && l_nnz && levels && aborted, HN_ERR_INVALID, "hn_sldl_sizes: null argument");
  *n = (uint64_t)h->S.n;
  *aborted = h->S.aborted ? 1 : 0;
  *l_nnz = h->S.aborted ? 0 : (uint64_t)h->S.Lrp[h->S.n];
  *levels = h->S.aborted ? 0 : (uint64_t)(h->S.level_ptr.size() - 1);
  return HN_OK;
}

extern "C" int hn_sldl_symbolic(const hn_sldl* h, uint64_t* perm, uint64_t* l_row_ptr, uint64_t* l_col,
                                uint64_t* parent, uint64_t* postorder) {
  HN_CHECK(h, HN_ERR_INVALID, "hn_sldl_symbolic: null handle");
  HN_CHECK(!h->S.aborted, HN_ERR_UNSUPPORTED, "hn_sldl_symbolic: the ordering stopped at its work limit");
  const SldlSymbolic& S = h->S;
  for (int64_t i = 0; i < S.n; ++i) {
    if (perm) perm[i] = (uint64_t)S.perm[i];
    if (parent) parent[i] = (uint64_t)S.parent[i];
    if (postorder) postorder[i] = (uint64_t)S.postorder[i];
  }
  if (l_row_ptr)
    for (int64_t i = 0; i <= S.n; ++i) l_row_ptr[i] = (uint64_t)S.Lrp[i];
  if (l_col)
    for (int64_t k = 0; k < S.Lrp[S.n]; ++k) l_col[k] = (uint64_t)S.Lci[k];
  return HN_OK;
}

extern "C" int hn_sldl_factor(hn_sldl* h, const hn_csr* a, double singular_tol, int* clamped) {
  HN_CHECK(h && a, HN_ERR_INVALID, "hn_sldl_factor: null argument");
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  h->bound = nullptr;  // map the entries of this matrix afresh
  HN_TRY(sldl_factor(h, a, singular_tol, s));
  int c = 0;
  HN_HIP(hipMemcpyAsync(&c, h->sing.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  if (clamped) *clamped = c;
  return HN_OK;
}

extern "C" int hn_sldl_factors(const hn_sldl* h, double* l_values, double* d_values) {
  HN_CHECK(h, HN_ERR_INVALID, "hn_sldl_factors: null handle");
  HN_CHECK(h->on_device && h->bound, HN_ERR_INVALID, "hn_sldl_factors: no factorisation yet");
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  const int64_t nnz = h->S.Lrp[h->S.n];
  if (l_values && nnz) HN_HIP(hipMemcpyAsync(l_values, h->L.p, sizeof(double) * nnz, hipMemcpyDeviceToHost, s));
  if (d_values && h->S.n) HN_HIP(hipMemcpyAsync(d_values, h->D.p, sizeof(double) * h->S.n, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

extern "C" int hn_sldl_solve_dev(const hn_sldl* h, const double* b, double* x, void* stream) {
  HN_CHECK(h && (h->S.n == 0 || (b && x)), HN_ERR_INVALID, "hn_sldl_solve_dev: null argument");
  HN_CHECK(b != x, HN_ERR_INVALID, "hn_sldl_solve_dev: b and x must not alias");
  HN_TRY(enter());
  return sldl_solve(h, b, x, false, pick_stream(stream));
}

extern "C" int hn_sldl_free(hn_sldl* h) {
  delete h;
  return HN_OK;
}
