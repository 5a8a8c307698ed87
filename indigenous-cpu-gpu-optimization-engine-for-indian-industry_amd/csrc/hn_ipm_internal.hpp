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

// Dense inverse workspace: after dense_factor, A holds -N^{-1}.
struct DenseNE {
  int m = 0;
  bool f32 = false;
  DevBuf A, pinv, vt, wt, clamped;
};

int ne_symbolic(const hn_csr* M, const hn_csr* Mt, NormalEq& ne, hipStream_t s);
// Numeric N for the current D; also writes it densely (zeros elsewhere) into
// `dense` in FP32 or FP64 when dense is not null.
int ne_numeric(const hn_csr* M, const hn_csr* Mt, const double* D, double reg, NormalEq& ne, double* reg_out,
               DenseNE* dense, bool f32, hipStream_t s);
int dense_reserve(DenseNE& d, int m, bool f32);
// Inverts the dense N in place (d.A becomes -N^{-1}).
int dense_factor(DenseNE& d, hipStream_t s);
int dense_apply(const DenseNE& d, const double* r, double* y, bool accumulate, hipStream_t s);
int ne_jacobi_diag(const hn_csr* M, const double* D, double* diag, hipStream_t s);

}  // namespace hn
