// Post-solve verification on the device: the checks of
// SolutionVerifier::verify_detailed (core/validation/solution_verifier.cpp:24-88)
// over the original problem, with the reference's tolerances, order of
// summation and messages.
//
//   rows    (one thread per row)    check_primal_feasibility's row part
//                                   (:148-184) and check_complementarity
//                                   (:302-344): activity and |activity|
//                                   summed left to right, as the reference
//                                   does, so the comparisons see the same
//                                   bits.
//   columns (one thread per column) the column scale and bound checks
//                                   (:105-146), check_dual_feasibility
//                                   (:196-242: reduced cost c_j - sum A_ij y_i
//                                   over i ascending, the order the
//                                   reference's scan of the whole matrix
//                                   meets column j's entries in) and
//                                   check_integrality (:349-370).  The
//                                   column entries come from a transpose
//                                   built on the device (stable by row).
//
// Maxima use the reference's std::max, (a < b) ? b : a, so a NaN never
// enters a maximum.  The kernels are compiled without contraction of a*b+c
// into fused multiply-adds, as the reference's host build has none.
//
// The objective offset + sum c_j x_j (:70-73) is one chain of dependent
// additions in j order; no parallel sum rounds the same way.  A wavefront
// walking that chain on the device takes 13 ns per term (6.5 ms at n =
// 500,000), so a host thread walks it over the caller's c and x while the
// index conversion, the uploads and the device kernels run.  The host part
// otherwise only composes the flags and messages from the device totals.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstring>
#include <future>
#include <string>
#include <vector>

#include "hn_internal.hpp"

namespace hn {
namespace {

constexpr int kVT = 256;
constexpr unsigned kVerifyGrid = 2048;

__host__ __device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }

// Block reduction of up to four (max) doubles and one count; thread 0 holds
// the result.
template <int K>
__device__ void block_reduce(double (&v)[K], unsigned long long& cnt) {
  __shared__ double sv[K][kVT / 64];
  __shared__ unsigned long long sc[kVT / 64];
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < K; ++q) v[q] = smax(v[q], __shfl_down(v[q], off, 64));
    cnt += __shfl_down(cnt, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) sv[q][w] = v[q];
    sc[w] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int u = 1; u < kVT / 64; ++u) {
#pragma unroll
      for (int q = 0; q < K; ++q) v[q] = smax(v[q], sv[q][u]);
      cnt += sc[u];
    }
  }
}

// Row checks.  flag[i] = 1 when row i violates its constraint beyond
// ftol * (1 + |rhs| + sum |a_ij x_j|).  part[b] = (max violation, max
// complementarity), cnt[b] = flagged rows of block b.
__global__ __launch_bounds__(kVT) void k_verify_rows(const int* __restrict__ rp, const int* __restrict__ ci,
                                                     const double* __restrict__ v, const double* __restrict__ x,
                                                     const int8_t* __restrict__ sense,
                                                     const double* __restrict__ rhs, const double* __restrict__ y,
                                                     int m, double ftol, uint8_t* __restrict__ flag,
                                                     double* __restrict__ part,
                                                     unsigned long long* __restrict__ cnt_out) {
#pragma clang fp contract(off)
  double acc[2] = {0.0, 0.0};  // max primal violation, max complementarity
  unsigned long long cnt = 0;
  for (int i = blockIdx.x * kVT + threadIdx.x; i < m; i += gridDim.x * kVT) {
    double act = 0.0, aabs = 0.0;
    const int k1 = rp[i + 1];
    for (int k = rp[i]; k < k1; ++k) {
      const double t = v[k] * x[ci[k]];
      act += t;
      aabs += fabs(t);
    }
    const double r = rhs[i];
    const int sn = sense[i];
    const double scale = 1.0 + fabs(r) + aabs;
    double excess, viol = 0.0;
    if (sn == HN_LE) {
      excess = act - r - ftol * scale;
      if (excess > 0.0) viol = act - r;
    } else if (sn == HN_GE) {
      excess = r - act - ftol * scale;
      if (excess > 0.0) viol = r - act;
    } else {
      excess = fabs(act - r) - ftol * scale;
      if (excess > 0.0) viol = fabs(act - r);
    }
    uint8_t f = 0;
    if (excess > 0.0) {
      acc[0] = smax(acc[0], viol);
      f = 1;
      ++cnt;
    }
    flag[i] = f;
    if (y) {
      const double slack = sn == HN_LE ? r - act : (sn == HN_GE ? act - r : 0.0);
      double sc = 1.0;  // std::max({1.0, |rhs|, sum |a_ij x_j|})
      if (sc < fabs(r)) sc = fabs(r);
      if (sc < aabs) sc = aabs;
      const double yi = y[i];
      const double ds = smax(1.0, fabs(yi));
      acc[1] = smax(acc[1], fabs(slack) / sc * (fabs(yi) / ds));
    }
  }
  block_reduce<2>(acc, cnt);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
    cnt_out[blockIdx.x] = cnt;
  }
}

// Column checks over the transpose (tp, tr, tv).  flag[j]: bit 0 below the
// lower bound, bit 1 above the upper bound, bit 2 not finite.  dual_mode 0:
// no dual check, 1: given reduced costs, 2: c_j - sum_i a_ij y_i.
// part[b] = (max bound violation, max dual violation, max integrality
// violation), cnt[b] = bound pushes, nf[b] = first non-finite column (n if
// none).
__global__ __launch_bounds__(kVT) void k_verify_cols(
    const int* __restrict__ tp, const int* __restrict__ tr, const double* __restrict__ tv,
    const double* __restrict__ x, const double* __restrict__ lb, const double* __restrict__ ub,
    const double* __restrict__ c, const double* __restrict__ y, const double* __restrict__ rc,
    const int8_t* __restrict__ vt, int n, double ftol10, double feas, double otol, int maximize, int dual_mode,
    uint8_t* __restrict__ flag, double* __restrict__ part, unsigned long long* __restrict__ cnt_out,
    int* __restrict__ nf_out) {
#pragma clang fp contract(off)
  double acc[3] = {0.0, 0.0, 0.0};
  unsigned long long cnt = 0;
  int first_nf = n;
  for (int j = blockIdx.x * kVT + threadIdx.x; j < n; j += gridDim.x * kVT) {
    const int k0 = tp[j], k1 = tp[j + 1];
    double cs = 1.0;
    for (int k = k0; k < k1; ++k) cs = smax(cs, fabs(tv[k]));
    const double l = lb[j], u = ub[j], cj = c[j], val = x[j];
    cs = smax(cs, fabs(cj));
    if (isfinite(l)) cs = smax(cs, fabs(l));
    if (isfinite(u)) cs = smax(cs, fabs(u));
    uint8_t f = 0;
    if (!isfinite(val)) {
      f = 4;
      if (j < first_nf) first_nf = j;
    } else {
      if (val < l - ftol10 * cs) {
        acc[0] = smax(acc[0], l - val);
        f |= 1;
        ++cnt;
      }
      if (val > u + ftol10 * cs) {
        acc[0] = smax(acc[0], val - u);
        f |= 2;
        ++cnt;
      }
    }
    flag[j] = f;
    if (dual_mode) {
      double r;
      if (dual_mode == 1) {
        r = rc[j];
      } else {
        double s = 0.0;
        for (int k = k0; k < k1; ++k) s += tv[k] * y[tr[k]];
        r = cj - s;
      }
      if (maximize) r = -r;
      double vi;
      if (val <= l + feas)
        vi = (r < 0.0) ? r : 0.0;  // std::min(0.0, r)
      else if (val >= u - feas)
        vi = (0.0 < r) ? r : 0.0;  // std::max(0.0, r)
      else
        vi = fabs(r);
      vi = fabs(vi) - otol;
      if (vi > 0.0) acc[1] = smax(acc[1], vi);
    }
    if (vt && (vt[j] == HN_VAR_INTEGER || vt[j] == HN_VAR_BINARY)) acc[2] = smax(acc[2], fabs(val - round(val)));
  }
  block_reduce<3>(acc, cnt);
  // first non-finite column of the block
  __shared__ int snf[kVT / 64];
  for (int off = 32; off > 0; off >>= 1) first_nf = min(first_nf, __shfl_down(first_nf, off, 64));
  if ((threadIdx.x & 63) == 0) snf[threadIdx.x >> 6] = first_nf;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kVT / 64; ++w) first_nf = min(first_nf, snf[w]);
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
    cnt_out[blockIdx.x] = cnt;
    nf_out[blockIdx.x] = first_nf;
  }
}

// Totals: [0] max row violation, [1] max complementarity, [2] max bound
// violation, [3] max dual violation, [4] max integrality violation,
// [5] flagged rows, [6] bound pushes, [7] first non-finite column.
__global__ __launch_bounds__(kVT) void k_verify_final(const double* __restrict__ rpart,
                                                      const unsigned long long* __restrict__ rcnt, int nrb,
                                                      const double* __restrict__ cpart,
                                                      const unsigned long long* __restrict__ ccnt,
                                                      const int* __restrict__ cnf, int ncb, int n,
                                                      double* __restrict__ tot) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  unsigned long long cnt = 0;
  for (int b = threadIdx.x; b < nrb; b += kVT) {
    acc[0] = smax(acc[0], rpart[2 * b]);
    acc[1] = smax(acc[1], rpart[2 * b + 1]);
    cnt += rcnt[b];
  }
  block_reduce<4>(acc, cnt);
  if (threadIdx.x == 0) {
    tot[0] = acc[0];
    tot[1] = acc[1];
    tot[5] = (double)cnt;
  }
  __syncthreads();
  double bcc[4] = {0.0, 0.0, 0.0, 0.0};
  unsigned long long cc = 0;
  int nf = n;
  for (int b = threadIdx.x; b < ncb; b += kVT) {
    bcc[0] = smax(bcc[0], cpart[3 * b]);
    bcc[1] = smax(bcc[1], cpart[3 * b + 1]);
    bcc[2] = smax(bcc[2], cpart[3 * b + 2]);
    cc += ccnt[b];
    nf = min(nf, cnf[b]);
  }
  block_reduce<4>(bcc, cc);
  __shared__ int snf[kVT / 64];
  for (int off = 32; off > 0; off >>= 1) nf = min(nf, __shfl_down(nf, off, 64));
  if ((threadIdx.x & 63) == 0) snf[threadIdx.x >> 6] = nf;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kVT / 64; ++w) nf = min(nf, snf[w]);
    tot[2] = bcc[0];
    tot[3] = bcc[1];
    tot[4] = bcc[2];
    tot[6] = (double)cc;
    tot[7] = (double)nf;
  }
}

// Device buffers of the verifier, grown on demand and kept for the next
// call (calls are serialised by the mutex).
struct VerifyScratch {
  std::mutex mtx;
  hn_csr a, t;
  DevBuf x, sense, rhs, lb, ub, c, y, rc, vt, rflag, cflag, rpart, rcnt, cpart, ccnt, cnf, tot;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
VerifyScratch& verify_scratch() {
  static VerifyScratch* s = new VerifyScratch();
  return *s;
}

}  // namespace

void release_verify_scratch() {
  VerifyScratch& S = verify_scratch();
  std::lock_guard<std::mutex> lock(S.mtx);
  for (DevBuf* b : {&S.a.row_ptr, &S.a.col, &S.a.val64, &S.t.row_ptr, &S.t.col, &S.t.val64, &S.x, &S.sense,
                    &S.rhs, &S.lb, &S.ub, &S.c, &S.y, &S.rc, &S.vt, &S.rflag, &S.cflag, &S.rpart, &S.rcnt, &S.cpart,
                    &S.ccnt, &S.cnf, &S.tot})
    b->release();
  if (S.ev0) (void)hipEventDestroy(S.ev0);
  if (S.ev1) (void)hipEventDestroy(S.ev1);
  S.ev0 = S.ev1 = nullptr;
}

namespace {

// obj = offset + c_0 x_0 + c_1 x_1 + ... (solution_verifier.cpp:70-73),
// products and sums rounded one by one.
double host_objective(double offset, const double* c, const double* x, int64_t n) {
#pragma clang fp contract(off)
  double obj = offset;
  for (int64_t j = 0; j < n; ++j) obj += c[j] * x[j];
  return obj;
}

std::string fixed6(double v) {  // std::to_string(double) is printf("%f")
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%f", v);
  return buf;
}

}  // namespace
}  // namespace hn

using namespace hn;

extern "C" int hn_default_tolerances(hn_tolerances* t) {
  HN_CHECK(t, HN_ERR_INVALID, "hn_default_tolerances: null argument");
  t->feasibility_tol = 1e-9;
  t->optimality_tol = 1e-9;
  t->complementarity_tol = 1e-6;
  t->integrality_tol = 1e-6;
  return HN_OK;
}

extern "C" int hn_verify(const hn_lp* lp, const int8_t* var_type, const hn_solution* sol,
                         const hn_tolerances* tol_in, hn_verify_result* out) {
  HN_CHECK(lp && sol && out, HN_ERR_INVALID, "hn_verify: null argument");
  hn_tolerances tol;
  if (tol_in)
    tol = *tol_in;
  else
    hn_default_tolerances(&tol);
  uint64_t* const vlist = out->violated;
  const uint64_t vcap = out->violated_capacity;
  out->feasible = 1;
  out->optimal = 1;
  out->primal_infeasibility = out->dual_infeasibility = out->complementarity = 0.0;
  out->integrality_violation = out->objective_discrepancy = 0.0;
  out->n_violated = 0;
  out->message[0] = '\0';
  out->device_ms = 0.0;
  auto set_message = [&](const std::string& s) {
    std::snprintf(out->message, sizeof(out->message), "%s", s.c_str());
  };

  const int64_t m = lp->m, n = lp->n;
  HN_CHECK(m >= 0 && n >= 0, HN_ERR_INVALID, "hn_verify: negative dimension");
  if ((int64_t)sol->n_primal != n) {  // check_primal_feasibility (:99-103)
    out->feasible = 0;
    out->optimal = 0;
    set_message("Primal infeasible: Solution vector size mismatch");
    return HN_OK;
  }
  HN_CHECK(lp->row_ptr, HN_ERR_INVALID, "hn_verify: null row_ptr");
  const int64_t nnz = lp->row_ptr[m];
  HN_CHECK(lp->row_ptr[0] == 0 && nnz >= 0, HN_ERR_INVALID, "hn_verify: row_ptr must start at 0");
  HN_CHECK(m < INT_MAX && n < INT_MAX && nnz < INT_MAX && m + nnz < INT_MAX, HN_ERR_UNSUPPORTED,
           "hn_verify: m, n and nnz must be below 2^31");
  HN_CHECK(nnz == 0 || (lp->col && lp->val), HN_ERR_INVALID, "hn_verify: null col/val");
  HN_CHECK(m == 0 || (lp->sense && lp->rhs), HN_ERR_INVALID, "hn_verify: null sense/rhs");
  HN_CHECK(n == 0 || (lp->lb && lp->ub && lp->obj && sol->primal), HN_ERR_INVALID,
           "hn_verify: null lb/ub/obj/primal");
  const bool has_dual = (int64_t)sol->n_dual == m;
  HN_CHECK(!has_dual || m == 0 || sol->dual, HN_ERR_INVALID, "hn_verify: null dual");
  const bool use_rc = (int64_t)sol->n_reduced_costs == n;
  HN_CHECK(!use_rc || n == 0 || sol->reduced_costs, HN_ERR_INVALID, "hn_verify: null reduced_costs");
  int64_t n_int = 0;
  if (var_type)
    for (int64_t j = 0; j < n; ++j) n_int += (var_type[j] == HN_VAR_INTEGER || var_type[j] == HN_VAR_BINARY);
  // check_dual_feasibility returns clean without a dual of length m, or for
  // a MIP without reduced costs (:199-203)
  const int dual_mode = (!has_dual || (n_int > 0 && !use_rc)) ? 0 : (use_rc ? 1 : 2);

  // the objective chain runs on a host thread beside the conversion, the
  // uploads and the kernels (inline for small n, where a thread costs more)
  std::future<double> obj_f = std::async(n > (1 << 16) ? std::launch::async : std::launch::deferred, host_objective,
                                         lp->obj_offset, lp->obj, sol->primal, n);

  std::vector<int> rp32(m + 1), ci32(nnz);
  for (int64_t i = 0; i <= m; ++i) {
    if (i > 0 && lp->row_ptr[i] < lp->row_ptr[i - 1]) return fail(HN_ERR_INVALID, "hn_verify: row_ptr decreases");
    rp32[i] = (int)lp->row_ptr[i];
  }
  for (int64_t k = 0; k < nnz; ++k) {
    if (lp->col[k] < 0 || lp->col[k] >= n) return fail(HN_ERR_INVALID, "hn_verify: column index out of range");
    ci32[k] = (int)lp->col[k];
  }

  HN_TRY(enter());
  VerifyScratch& S = verify_scratch();
  std::lock_guard<std::mutex> lock(S.mtx);
  hipStream_t s = ctx().stream;
  if (!S.ev0) {
    HN_HIP(hipEventCreate(&S.ev0));
    HN_HIP(hipEventCreate(&S.ev1));
  }
  const unsigned rg = grid_for(m, kVT, kVerifyGrid), cg = grid_for(n, kVT, kVerifyGrid);
  HN_TRY(csr_alloc(&S.a, m, n, nnz, HN_F64));
  HN_TRY(csr_alloc(&S.t, n, m, nnz, HN_F64));
  const size_t md = sizeof(double) * std::max<int64_t>(m, 1), nd = sizeof(double) * std::max<int64_t>(n, 1);
  HN_TRY(S.x.reserve(nd));
  HN_TRY(S.lb.reserve(nd));
  HN_TRY(S.ub.reserve(nd));
  HN_TRY(S.c.reserve(nd));
  HN_TRY(S.sense.reserve(std::max<int64_t>(m, 1)));
  HN_TRY(S.rhs.reserve(md));
  HN_TRY(S.rflag.reserve(std::max<int64_t>(m, 1)));
  HN_TRY(S.cflag.reserve(std::max<int64_t>(n, 1)));
  HN_TRY(S.rpart.reserve(sizeof(double) * 2 * rg));
  HN_TRY(S.rcnt.reserve(sizeof(unsigned long long) * rg));
  HN_TRY(S.cpart.reserve(sizeof(double) * 3 * cg));
  HN_TRY(S.ccnt.reserve(sizeof(unsigned long long) * cg));
  HN_TRY(S.cnf.reserve(sizeof(int) * cg));
  HN_TRY(S.tot.reserve(sizeof(double) * 8));
  const bool comp_y = has_dual && m > 0;
  if (comp_y) HN_TRY(S.y.reserve(md));
  if (dual_mode == 1) HN_TRY(S.rc.reserve(nd));
  const bool has_vt = var_type && n_int > 0;
  if (has_vt) HN_TRY(S.vt.reserve(std::max<int64_t>(n, 1)));

  auto h2d = [&](void* d, const void* h, size_t b) -> int {
    if (b) HN_HIP(hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, s));
    return HN_OK;
  };
  HN_TRY(h2d(S.a.row_ptr.p, rp32.data(), sizeof(int) * (m + 1)));
  HN_TRY(h2d(S.a.col.p, ci32.data(), sizeof(int) * nnz));
  HN_TRY(h2d(S.a.val64.p, lp->val, sizeof(double) * nnz));
  HN_TRY(h2d(S.x.p, sol->primal, sizeof(double) * n));
  HN_TRY(h2d(S.lb.p, lp->lb, sizeof(double) * n));
  HN_TRY(h2d(S.ub.p, lp->ub, sizeof(double) * n));
  HN_TRY(h2d(S.c.p, lp->obj, sizeof(double) * n));
  HN_TRY(h2d(S.sense.p, lp->sense, (size_t)m));
  HN_TRY(h2d(S.rhs.p, lp->rhs, sizeof(double) * m));
  if (comp_y) HN_TRY(h2d(S.y.p, sol->dual, sizeof(double) * m));
  if (dual_mode == 1) HN_TRY(h2d(S.rc.p, sol->reduced_costs, sizeof(double) * n));
  if (has_vt) HN_TRY(h2d(S.vt.p, var_type, (size_t)n));

  HN_HIP(hipEventRecord(S.ev0, s));
  HN_TRY(transpose_into(&S.a, &S.t, s));
  hipLaunchKernelGGL(k_verify_rows, dim3(rg), dim3(kVT), 0, s, S.a.row_ptr.as<int>(), S.a.col.as<int>(),
                     S.a.val64.as<double>(), S.x.as<double>(), S.sense.as<int8_t>(), S.rhs.as<double>(),
                     comp_y ? S.y.as<double>() : (const double*)nullptr, (int)m, tol.feasibility_tol * 10.0,
                     S.rflag.as<uint8_t>(), S.rpart.as<double>(), S.rcnt.as<unsigned long long>());
  HN_TRY(check_launch("k_verify_rows"));
  hipLaunchKernelGGL(k_verify_cols, dim3(cg), dim3(kVT), 0, s, S.t.row_ptr.as<int>(), S.t.col.as<int>(),
                     S.t.val64.as<double>(), S.x.as<double>(), S.lb.as<double>(), S.ub.as<double>(),
                     S.c.as<double>(), comp_y ? S.y.as<double>() : (const double*)nullptr,
                     dual_mode == 1 ? S.rc.as<double>() : (const double*)nullptr,
                     has_vt ? S.vt.as<int8_t>() : (const int8_t*)nullptr, (int)n, tol.feasibility_tol * 10.0,
                     tol.feasibility_tol, tol.optimality_tol, lp->maximize ? 1 : 0, dual_mode,
                     S.cflag.as<uint8_t>(), S.cpart.as<double>(), S.ccnt.as<unsigned long long>(),
                     S.cnf.as<int>());
  HN_TRY(check_launch("k_verify_cols"));
  hipLaunchKernelGGL(k_verify_final, dim3(1), dim3(kVT), 0, s, S.rpart.as<double>(),
                     S.rcnt.as<unsigned long long>(), (int)rg, S.cpart.as<double>(),
                     S.ccnt.as<unsigned long long>(), S.cnf.as<int>(), (int)cg, (int)n, S.tot.as<double>());
  HN_TRY(check_launch("k_verify_final"));
  HN_HIP(hipEventRecord(S.ev1, s));
  count_launch(3);
  double tot[8];
  HN_HIP(hipMemcpyAsync(tot, S.tot.p, sizeof(tot), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  const double obj = obj_f.get();
  float ms = 0.f;
  HN_HIP(hipEventElapsedTime(&ms, S.ev0, S.ev1));
  out->device_ms = ms;

  const uint64_t rows_flagged = (uint64_t)tot[5], pushes = (uint64_t)tot[6];
  const int64_t first_nf = (int64_t)tot[7];
  std::vector<uint8_t> cf, rf;
  const bool want_list = vlist && vcap > 0;

  // check_primal_feasibility (:92-191)
  bool primal_ok = true;
  std::string primal_msg;
  if (first_nf < n) {
    // the reference returns at the first non-finite value, with the bound
    // violations of the columns before it and no maximum (:119-123)
    cf.resize(n);
    HN_HIP(hipMemcpy(cf.data(), S.cflag.p, (size_t)n, hipMemcpyDeviceToHost));
    uint64_t cnt = 0;
    for (int64_t j = 0; j < first_nf; ++j) {
      for (int bit = 0; bit < 2; ++bit)
        if (cf[j] & (1 << bit)) {
          if (want_list && cnt < vcap) vlist[cnt] = (uint64_t)j;
          ++cnt;
        }
    }
    out->n_violated = cnt;
    primal_ok = false;
    primal_msg = "Solution contains NaN or Inf values";
  } else {
    out->primal_infeasibility = smax(tot[2], tot[0]);
    const uint64_t total = pushes + rows_flagged;
    out->n_violated = total;
    if (total > 0) {
      primal_ok = false;
      primal_msg = "Max primal violation: " + fixed6(out->primal_infeasibility);
      if (want_list) {
        uint64_t cnt = 0;
        if (pushes > 0) {
          cf.resize(n);
          HN_HIP(hipMemcpy(cf.data(), S.cflag.p, (size_t)n, hipMemcpyDeviceToHost));
          for (int64_t j = 0; j < n && cnt < vcap; ++j)
            for (int bit = 0; bit < 2; ++bit)
              if ((cf[j] & (1 << bit)) && cnt < vcap) vlist[cnt++] = (uint64_t)j;
        }
        if (rows_flagged > 0 && cnt < vcap) {
          rf.resize(m);
          HN_HIP(hipMemcpy(rf.data(), S.rflag.p, (size_t)m, hipMemcpyDeviceToHost));
          for (int64_t i = 0; i < m && cnt < vcap; ++i)
            if (rf[i]) vlist[cnt++] = (uint64_t)(n + i);
        }
      }
    }
  }

  // verify_detailed (:24-88)
  if (!primal_ok) {
    out->feasible = 0;
    set_message("Primal infeasible: " + primal_msg);
  }
  out->dual_infeasibility = dual_mode ? tot[3] : 0.0;
  out->complementarity = comp_y ? tot[1] : 0.0;
  if (n_int == 0) {
    if (out->dual_infeasibility > tol.optimality_tol) {
      out->optimal = 0;
      if (out->feasible) set_message("Dual infeasible: Max dual violation: " + fixed6(out->dual_infeasibility));
    }
    if (out->complementarity > tol.complementarity_tol) {
      out->optimal = 0;
      if (out->feasible)
        set_message("Complementarity violation: Max complementarity violation: " + fixed6(out->complementarity));
    }
  }
  out->integrality_violation = has_vt ? tot[4] : 0.0;
  if (out->integrality_violation > tol.integrality_tol) {
    out->feasible = 0;
    set_message("Integrality violation: Max integrality violation: " + fixed6(out->integrality_violation));
  }
  out->objective_discrepancy = std::fabs(obj - sol->objective_value);
  if (out->objective_discrepancy > 1e-4 * (1.0 + std::fabs(obj))) {
    out->optimal = 0;
    if (out->feasible) set_message("Objective discrepancy: " + fixed6(out->objective_discrepancy));
  }
  return HN_OK;
}
