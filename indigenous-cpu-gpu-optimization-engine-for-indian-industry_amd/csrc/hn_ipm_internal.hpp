// Internal interfaces of the device IPM (normal equations and helpers).
#pragma once

#include "hn_internal.hpp"

namespace hn {

// Assembled sparse N = M D M' + r I (pattern fixed per matrix).
struct NormalEq {
  hn_csr* N = nullptr;     // pattern + FP64 values, with a SpMV tiling
  DevBuf diag_pos;         // int[m]: position of (i, i) in row i
  DevBuf part;             // reduction partials
  ~NormalEq() { delete N; }
};

// Dense factor workspace: after dense_factor, G = L^{-1} and dinv = D^{-1},
// so N^{-1} = G' D^{-1} G.
struct DenseNE {
  int m = 0;
  bool f32 = false;
  DevBuf A;              // N, then L (lower)
  DevBuf G;              // L^{-1} (lower triangular; 64 x 64 diagonal blocks stored in full)
  DevBuf dinv;           // 1 / pivot (0 for a skipped pivot)
  DevBuf u, part;        // apply workspace: D^{-1} G r (m), per-block partial sums of G' u (m per block)
  DevBuf wt, lt;         // panel W = L D and L, k-major (64 x m)
  DevBuf clamped;        // number of clamped or skipped pivots
  DevBuf diag0;  // diagonal of N (regularisation included) before elimination
};

// Skip-rule thresholds relative to the diagonal of N (HN_PIVOT_SKIP).
constexpr double kSkipRel64 = 1e-13;
constexpr double kSkipRel32 = 1e-6;
// Relative regularisation of the reference rule: r = max(reg, 1e-12 max diag N).
constexpr double kRelReg = 1e-12;

int ne_symbolic(const hn_csr* M, const hn_csr* Mt, NormalEq& ne, hipStream_t s);
// Numeric N for the current D; also writes it densely (zeros elsewhere) into
// `dense` in FP32 or FP64 when dense is not null.
// rel_reg: kRelReg under the reference's pivot rule, 0 under the skip rule.
int ne_numeric(const hn_csr* M, const hn_csr* Mt, const double* D, double reg, double rel_reg, NormalEq& ne,
               double* reg_out, DenseNE* dense, bool f32, hipStream_t s);
int dense_reserve(DenseNE& d, int m, bool f32);
// Factors the dense N (into G = L^{-1} and D^{-1}) under the pivot rule
// HN_PIVOT_CLAMP or HN_PIVOT_SKIP.
int dense_factor(DenseNE& d, int pivot_rule, hipStream_t s);
// y = N^{-1} r (accumulate: y += N^{-1} r) as G' (D^{-1} (G r)).
int dense_apply(const DenseNE& d, const double* r, double* y, bool accumulate, hipStream_t s);
// One refinement correction y += N^{-1} r that does nothing when the last
// check (*irr = ||rhs - N y||_inf, *irb = ||rhs||_inf, device scalars) has
// already met the refinement test; *sweeps counts the corrections that ran.
int dense_apply_refine(const DenseNE& d, const double* r, double* y, const double* irr, const double* irb,
                       double* sweeps, hipStream_t s);
int ne_jacobi_diag(const hn_csr* M, const double* D, double* diag, hipStream_t s);

}  // namespace hn
