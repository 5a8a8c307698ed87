// Mehrotra predictor-corrector IPM with device-resident iterates.
//
// Reference: InteriorPointSolver::solve (core/lp/interior_point.cpp:402-886).
// The host keeps the reference's control flow - start point, convergence
// test, sigma rule, step rule, guards, relaxed post-loop test and status
// mapping - and reads back only the scalars it decides on.  Every vector
// operation runs on the GPU:
//   residuals  rp = M x - b, rd = M'lam + z - c        (SpMV on M and on the
//              pre-built M', then fused elementwise kernels)
//   scalings   D = x/z clamped to [0, 1e10], D rd
//   Newton     rhs = -rp - M(D rd) + M x; N dlam = rhs (normal equations,
//              hn_normal.hip) with FP64 iterative refinement
//   direction  dx = D (M'dlam + rd) - x [+ zs], dz = -(M'dlam + rd)
//   ratio test fused min-reductions of -v/d over d < 0
//   update     x += ap dx, z += ad dz, lam += ad dlam with the guards.
// Elementwise expressions are evaluated in the reference's order without
// FMA contraction, so they round like the CPU code; reductions use a fixed
// two-pass tree (deterministic, different summation order).
#include "hn_device.hpp"
#include "hn_internal.hpp"
#include "hn_ipm_internal.hpp"
#include "hn_sldl.hpp"
#include "hn_stdform.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#pragma clang fp contract(off)

namespace hn {

namespace {

enum Slot {
  S_XZ = 0,    // sum x z
  S_RPN,       // ||rp||_inf
  S_RDN,       // ||rd||_inf
  S_AP_A,      // predictor primal step
  S_AD_A,      // predictor dual step
  S_MU_AFF,    // sum (x + ap dx)(z + ad dz)
  S_AP,        // corrector primal step
  S_AD,        // corrector dual step
  S_REG,       // dynamic regularisation of N
  S_IRR,       // ||rhs - N dlam||_inf
  S_IRB,       // ||rhs||_inf
  S_CTX, S_BTY, S_DRN,  // c'x, b'y, ||M'y + s - c||_inf (final metrics)
  S_PCG_RZ, S_PCG_PQ, S_PCG_RN, S_PCG_RZ_OLD,
  S_SWEEPS,    // refinement sweeps run by device-driven iterations
  S_COUNT
};
// F_MORE: a device-driven step's solve still failed its refinement test
// after the sweeps the step holds (the host would have run more)
enum Flag { F_DIR = 0, F_UPD, F_MORE, F_COUNT };

constexpr int kNT = 256;
constexpr int kMaxParts = 512;
// nonzeros of M from which the matrix-free solver autotunes its SpMV layouts
constexpr size_t kAutotuneNnz = size_t(1) << 20;
// HN_NE_AUTO above dense_max_m: the minimum-degree ordering may merge this
// many adjacency entries (well under a second on the host) and the factor
// hold this many entries before the solve falls back to PCG ...
constexpr double kSparseAutoWork = 3e8;
constexpr double kSparseAutoLnnz = 1e8;
// ... and it is tried only when M's columns give at most this many candidate
// entries of N (the power-law LP of 200,000 rows gives 9e7)
constexpr double kSparseAutoCand = 2e7;
// ... and when the factor's critical path (sequential merge steps of one row
// per dependency level, summed over the levels) is short enough: a step is a
// dependent memory access (25fv47, 5.0e6 steps, takes 0.57 s per iteration
// with its solves), so factors with long dense rows take the PCG path
constexpr double kSparseAutoCrit = 2e5;
// singular_tol of ToleranceConfig::industrial_defaults (tolerance.hpp:70)
constexpr double kSingularTol = 1e-10;
// refinement sweeps per solve of a device-driven step once a solve of the
// LP has needed one (the refinement's own limit); until then the step holds
// none, and a solve that needs one has its step undone and run again on the
// host path
constexpr int kMaxSweeps = 5;

int parts_for(int64_t n) {
  return (int)std::min<int64_t>(kMaxParts, std::max<int64_t>(1, (n + kNT * 8 - 1) / (kNT * 8)));
}

// ---- elementwise ---------------------------------------------------------
__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ o, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = a[i] - b[i];
}

// rp = M x - b over m and rd = M' lam + z - c over n in one launch
// (interior_point.cpp:512-515)
__global__ void k_resid(const double* __restrict__ mx, const double* __restrict__ b, double* __restrict__ rp, int m,
                        const double* __restrict__ t, const double* __restrict__ z, const double* __restrict__ c,
                        double* __restrict__ rd, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) rp[i] = mx[i] - b[i];
  if (i < n) rd[i] = t[i] + z[i] - c[i];
}

__global__ void k_scaling(const double* __restrict__ x, const double* __restrict__ z, const double* __restrict__ rd,
                          double* __restrict__ D, double* __restrict__ Drd, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = x[i] / z[i];  // interior_point.cpp:532-539
  if (!(d >= 0.0) || !isfinite(d) || d > 1e10) d = 1e10;
  D[i] = d;
  Drd[i] = d * rd[i];      // :584
}

__global__ void k_rhs(const double* __restrict__ rp, const double* __restrict__ mdrd, const double* __restrict__ mx,
                      double* __restrict__ rhs, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) rhs[i] = -rp[i] - mdrd[i] + mx[i];  // :598-600
}

// dx, dz over n from t = M' dlam, and the finiteness test of dx, dz and
// dlam (over m) in one launch
__global__ void k_direction(const double* __restrict__ D, const double* __restrict__ t,
                            const double* __restrict__ rd, const double* __restrict__ x,
                            const double* __restrict__ zs, double* __restrict__ dx, double* __restrict__ dz, int n,
                            const double* __restrict__ dlam, int m, int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && !isfinite(dlam[i])) flags[F_DIR] = 1;
  if (i >= n) return;
  const double s = t[i] + rd[i];
  double a = D[i] * s - x[i];     // :610 / :668
  if (zs) a = a + zs[i];
  const double b = -s;            // :614 / :672
  dx[i] = a;
  dz[i] = b;
  if (!isfinite(a) || !isfinite(b)) flags[F_DIR] = 1;
}

__global__ void k_corrector(const double* __restrict__ dx, const double* __restrict__ dz,
                            const double* __restrict__ z, double sigma_mu, double* __restrict__ zs, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double zinv = 1.0 / z[i];  // :645-650
  zs[i] = (sigma_mu - dx[i] * dz[i]) * zinv;
}

// sigma * mu of the corrector on the device (interior_point.cpp:638-643):
// sigma = clamp((mu_aff / mu)^3, 0.1, 0.9), or 0.5 when mu_aff is not a
// finite non-negative number or mu is not positive.  The cube is formed in
// double-double (exact square, then an error-compensated product), so it is
// the correctly rounded cube except in rare near-tie cases, as the host's
// std::pow is.
__device__ __forceinline__ double cube_rn(double q) {
  const double p = q * q;
  const double e = __builtin_fma(q, q, -p);
  const double r = p * q;
  const double e2 = __builtin_fma(p, q, -r);
  return r + (e * q + e2);
}

__device__ __forceinline__ double sigma_mu_dev(const double* __restrict__ S, int n) {
  const double mu = S[S_XZ] / (double)n;
  const double mu_aff = S[S_MU_AFF] / (double)n;
  double sigma;
  if (!isfinite(mu_aff) || mu_aff < 0.0 || !(mu > 0.0)) {
    sigma = 0.5;
  } else {
    const double c = cube_rn(mu_aff / fmax(mu, 1e-300));
    sigma = (c < 0.1) ? 0.1 : ((0.9 < c) ? 0.9 : c);
  }
  return sigma * mu;
}

// k_corrector with sigma * mu formed from S (each thread forms the same value)
__global__ void k_corrector_s(const double* __restrict__ dx, const double* __restrict__ dz,
                              const double* __restrict__ z, const double* __restrict__ S, double* __restrict__ zs,
                              int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sigma_mu = sigma_mu_dev(S, n);
  const double zinv = 1.0 / z[i];
  zs[i] = (sigma_mu - dx[i] * dz[i]) * zinv;
}

// The update is skipped when the step's direction was not finite (the
// host finds the flag at its next read of the flags and stops there, as
// the reference stops before updating, :617-620, :675-678).
__global__ void k_update(double* __restrict__ x, double* __restrict__ z, const double* __restrict__ dx,
                         const double* __restrict__ dz, const double* __restrict__ S, int n,
                         double* __restrict__ lam, const double* __restrict__ dlam, int m,
                         int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (flags[F_DIR]) return;
  if (i < m) lam[i] = lam[i] + S[S_AD] * dlam[i];  // :694-696
  if (i >= n) return;
  const double xi = x[i] + S[S_AP] * dx[i];  // :690-693
  const double zi = z[i] + S[S_AD] * dz[i];
  x[i] = xi;
  z[i] = zi;
  if (xi < 0.0 || zi < 0.0 || !isfinite(xi) || !isfinite(zi)) flags[F_UPD] = 1;  // :698-708
}

__global__ void k_clear_flag(int* __restrict__ flags, int f) {
  if (threadIdx.x == 0) flags[f] = 0;
}

// A device-driven step keeps the point it started from (x, z, lam), so that
// a step whose solve needs more refinement sweeps than it holds can be
// undone and run again on the host path.
__global__ void k_keep_point(const double* __restrict__ x, const double* __restrict__ z,
                             const double* __restrict__ lam, double* __restrict__ xk, double* __restrict__ zk,
                             double* __restrict__ lk, int n, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    xk[i] = x[i];
    zk[i] = z[i];
  }
  if (i < m) lk[i] = lam[i];
}

__global__ void k_undo_step(double* __restrict__ x, double* __restrict__ z, double* __restrict__ lam,
                            const double* __restrict__ xk, const double* __restrict__ zk,
                            const double* __restrict__ lk, int n, int m, int* __restrict__ flags,
                            double* __restrict__ S, double sweeps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] = xk[i];
    z[i] = zk[i];
  }
  if (i < m) lam[i] = lk[i];
  if (i == 0) {
    for (int f = 0; f < F_COUNT; ++f) flags[f] = 0;
    S[S_SWEEPS] = sweeps;
  }
}

__global__ void k_fill(double* __restrict__ p, double v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- reductions (fixed two-pass trees) -------------------------------------
// sum x z over n; max |rp| over m; max |rd| over n
__global__ __launch_bounds__(kNT) void k_red_start(const double* __restrict__ x, const double* __restrict__ z,
                                                   const double* __restrict__ rp, const double* __restrict__ rd,
                                                   int n, int m, double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  double s = 0.0, a = 0.0, b = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
    s += x[i] * z[i];
    b = nan_ignoring_max(b, fabs(rd[i]));
  }
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) a = nan_ignoring_max(a, fabs(rp[i]));
  s = block_sum<double, kNT>(s, red);
  a = block_max<double, kNT>(a, red);
  b = block_max<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[kMaxParts + blockIdx.x] = a;
    part[2 * kMaxParts + blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kNT) void k_fin_start(const double* __restrict__ part, int np, double* __restrict__ S) {
  __shared__ double red[kNT / kWave];
  double s = 0.0, a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    s += part[i];
    a = nan_ignoring_max(a, part[kMaxParts + i]);
    b = nan_ignoring_max(b, part[2 * kMaxParts + i]);
  }
  s = block_sum<double, kNT>(s, red);
  a = block_max<double, kNT>(a, red);
  b = block_max<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    S[S_XZ] = s;
    S[S_RPN] = a;
    S[S_RDN] = b;
  }
}

// ratio test of (x, dx) and (z, dz): min over d < 0 of -v/d
__global__ __launch_bounds__(kNT) void k_red_step(const double* __restrict__ x, const double* __restrict__ dx,
                                                  const double* __restrict__ z, const double* __restrict__ dz, int n,
                                                  double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  double a = 1.0, b = 1.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
    const double d1 = dx[i], d2 = dz[i];
    if (d1 < 0.0) {
      const double q = -x[i] / d1;
      a = (q < a) ? q : a;
    }
    if (d2 < 0.0) {
      const double q = -z[i] / d2;
      b = (q < b) ? q : b;
    }
  }
  a = block_min<double, kNT>(a, red);
  b = block_min<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxParts + blockIdx.x] = b;
  }
}

__global__ __launch_bounds__(kNT) void k_fin_step(const double* __restrict__ part, int np, double* __restrict__ S,
                                                  int slot_p, int slot_d) {
  __shared__ double red[kNT / kWave];
  double a = 1.0, b = 1.0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    a = (part[i] < a) ? part[i] : a;
    b = (part[kMaxParts + i] < b) ? part[kMaxParts + i] : b;
  }
  a = block_min<double, kNT>(a, red);
  b = block_min<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    const double sa = 0.99 * a, sb = 0.99 * b;  // interior_point.cpp:141
    S[slot_p] = (sa < 1.0) ? sa : 1.0;
    S[slot_d] = (sb < 1.0) ? sb : 1.0;
  }
}

// sum (x + ap dx)(z + ad dz) with the predictor steps (:632-636)
__global__ __launch_bounds__(kNT) void k_red_muaff(const double* __restrict__ x, const double* __restrict__ dx,
                                                   const double* __restrict__ z, const double* __restrict__ dz, int n,
                                                   const double* __restrict__ S, double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  const double ap = S[S_AP_A], ad = S[S_AD_A];
  double s = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) s += (x[i] + ap * dx[i]) * (z[i] + ad * dz[i]);
  s = block_sum<double, kNT>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kNT) void k_fin_sum(const double* __restrict__ part, int np, double* __restrict__ S,
                                                 int slot) {
  __shared__ double red[kNT / kWave];
  double s = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) s += part[i];
  s = block_sum<double, kNT>(s, red);
  if (threadIdx.x == 0) S[slot] = s;
}

// refinement residual: r = rhs - Nd, ||r||_inf and ||rhs||_inf
// (irr, irb given: a refinement check that is skipped once the solution has
// met the test, like the correction before it; see ne_refined)
__device__ __forceinline__ bool refined_dev(const double* irr, const double* irb) {
  if (!irr) return false;
  const double nb = *irb;
  return nb == 0.0 || nb != nb || *irr / nb <= 1e-10;
}

__global__ __launch_bounds__(kNT) void k_red_resid(const double* __restrict__ rhs, const double* __restrict__ nd,
                                                   double* __restrict__ r, int m, double* __restrict__ part,
                                                   const double* irr, const double* irb) {
  __shared__ double red[kNT / kWave];
  if (refined_dev(irr, irb)) return;
  double a = 0.0, b = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) {
    const double ri = rhs[i] - nd[i];
    r[i] = ri;
    a = nan_ignoring_max(a, fabs(ri));
    b = nan_ignoring_max(b, fabs(rhs[i]));
  }
  a = block_max<double, kNT>(a, red);
  b = block_max<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a;
    part[kMaxParts + blockIdx.x] = b;
  }
}

// guarded: skipped once the previous check has passed; more (optional): set
// when this check fails the refinement test
__global__ __launch_bounds__(kNT) void k_fin_max2(const double* __restrict__ part, int np, double* __restrict__ S,
                                                  int slot_a, int slot_b, int guarded, int* __restrict__ more) {
  __shared__ double red[kNT / kWave];
  if (guarded && refined_dev(S + slot_a, S + slot_b)) return;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    a = nan_ignoring_max(a, part[i]);
    b = nan_ignoring_max(b, part[kMaxParts + i]);
  }
  a = block_max<double, kNT>(a, red);
  b = block_max<double, kNT>(b, red);
  if (threadIdx.x == 0) {
    S[slot_a] = a;
    S[slot_b] = b;
    if (more && !(b == 0.0 || b != b || a / b <= 1e-10)) *more = 1;
  }
}

// final metrics: c'x, b'y (over different lengths) and ||M'y + s - c||_inf
__global__ __launch_bounds__(kNT) void k_red_final(const double* __restrict__ c, const double* __restrict__ x,
                                                   const double* __restrict__ b, const double* __restrict__ y,
                                                   const double* __restrict__ rd, int n, int m,
                                                   double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  double s1 = 0.0, s2 = 0.0, a = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < n; i += gridDim.x * kNT) {
    s1 += c[i] * x[i];
    a = nan_ignoring_max(a, fabs(rd[i]));
  }
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) s2 += b[i] * y[i];
  s1 = block_sum<double, kNT>(s1, red);
  s2 = block_sum<double, kNT>(s2, red);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1;
    part[kMaxParts + blockIdx.x] = s2;
    part[2 * kMaxParts + blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(kNT) void k_fin_final(const double* __restrict__ part, int np, double* __restrict__ S) {
  __shared__ double red[kNT / kWave];
  double s1 = 0.0, s2 = 0.0, a = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    s1 += part[i];
    s2 += part[kMaxParts + i];
    a = nan_ignoring_max(a, part[2 * kMaxParts + i]);
  }
  s1 = block_sum<double, kNT>(s1, red);
  s2 = block_sum<double, kNT>(s2, red);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) {
    S[S_CTX] = s1;
    S[S_BTY] = s2;
    S[S_DRN] = a;
  }
}

// ---- PCG pieces (Jacobi, matrix-free N = M D M' + r I) -----------------------
// Vectors of type T (FP64, or FP32 for the bulk solves of mixed precision);
// scalars and reductions in FP64.
template <typename T>
__global__ void k_mul(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ o, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = a[i] * b[i];
}

template <typename Ti, typename To>
__global__ void k_cvt(const Ti* __restrict__ a, To* __restrict__ o, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = (To)a[i];
}

// y += (double) e
__global__ void k_add_f32(const float* __restrict__ e, double* __restrict__ y, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = y[i] + (double)e[i];
}

// q += r p; partial p'q
template <typename T>
__global__ __launch_bounds__(kNT) void k_pcg_pq(T* __restrict__ q, const T* __restrict__ p,
                                                const double* __restrict__ S, int m, double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  const T reg = (T)S[S_REG];
  double s = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) {
    const T qi = q[i] + reg * p[i];
    q[i] = qi;
    s += (double)(p[i] * qi);
  }
  s = block_sum<double, kNT>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// x += a p; r -= a q; z = r / diag; partials r'z and max |r|
template <typename T>
__global__ __launch_bounds__(kNT) void k_pcg_step(T* __restrict__ xk, T* __restrict__ r, T* __restrict__ zk,
                                                  const T* __restrict__ p, const T* __restrict__ q,
                                                  const T* __restrict__ dinv, const double* __restrict__ S, int m,
                                                  double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  const T alpha = (T)(S[S_PCG_RZ] / S[S_PCG_PQ]);
  double s = 0.0, a = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) {
    xk[i] = xk[i] + alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = ri * dinv[i];
    zk[i] = zi;
    s += (double)(ri * zi);
    a = nan_ignoring_max(a, fabs((double)ri));
  }
  s = block_sum<double, kNT>(s, red);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[kMaxParts + blockIdx.x] = a;
  }
}

__global__ __launch_bounds__(kNT) void k_pcg_fin(const double* __restrict__ part, int np, double* __restrict__ S) {
  __shared__ double red[kNT / kWave];
  double s = 0.0, a = 0.0;
  for (int i = threadIdx.x; i < np; i += kNT) {
    s += part[i];
    a = nan_ignoring_max(a, part[kMaxParts + i]);
  }
  s = block_sum<double, kNT>(s, red);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) {
    S[S_PCG_RZ_OLD] = S[S_PCG_RZ];
    S[S_PCG_RZ] = s;
    S[S_PCG_RN] = a;
  }
}

template <typename T>
__global__ void k_pcg_dir(T* __restrict__ p, const T* __restrict__ zk, const double* __restrict__ S, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const T beta = (T)(S[S_PCG_RZ] / S[S_PCG_RZ_OLD]);
  p[i] = zk[i] + beta * p[i];
}

// r = rhs (x0 = 0), z = r/diag, p = z, dinv = 1/(diag + reg); partials r'z, max|r|
template <typename T>
__global__ __launch_bounds__(kNT) void k_pcg_init(const T* __restrict__ rhs, const double* __restrict__ diag,
                                                  const double* __restrict__ S, T* __restrict__ xk,
                                                  T* __restrict__ r, T* __restrict__ zk, T* __restrict__ p,
                                                  T* __restrict__ dinv, int m, double* __restrict__ part) {
  __shared__ double red[kNT / kWave];
  const double reg = S[S_REG];
  double s = 0.0, a = 0.0;
  for (int i = blockIdx.x * kNT + threadIdx.x; i < m; i += gridDim.x * kNT) {
    const T di = (T)(1.0 / (diag[i] + reg));
    dinv[i] = di;
    xk[i] = T(0);
    const T ri = rhs[i];
    r[i] = ri;
    const T zi = ri * di;
    zk[i] = zi;
    p[i] = zi;
    s += (double)(ri * zi);
    a = nan_ignoring_max(a, fabs((double)ri));
  }
  s = block_sum<double, kNT>(s, red);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s;
    part[kMaxParts + blockIdx.x] = a;
  }
}

// r = max(reg, 1e-12 max diag) for the matrix-free operator
__global__ __launch_bounds__(kNT) void k_pcg_reg(const double* __restrict__ diag, int m, double reg,
                                                 double rel_reg, double* __restrict__ S) {
  __shared__ double red[kNT / kWave];
  double a = 0.0;
  for (int i = threadIdx.x; i < m; i += kNT) a = nan_ignoring_max(a, diag[i]);
  a = block_max<double, kNT>(a, red);
  if (threadIdx.x == 0) S[S_REG] = fmax(reg, rel_reg * a);
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double s() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

// ---------------------------------------------------------------------------
struct Solver {
  const hn_ipm_opts& o;
  hn_ipm_result& res;
  hipStream_t s;
  StandardForm f;
  int m = 0, n = 0;
  hn_csr* M = nullptr;
  hn_csr* Mt = nullptr;
  // device vectors
  DevBuf b, c, x, z, lam, rp, rd, D, Drd, tn, tm, mdrd, rhs, rhs2, dlam, dx, dz, zs, mz, ir_r, ir_nd, ir_e;
  DevBuf pcg_r, pcg_z, pcg_p, pcg_q, pcg_dinv, pcg_diag, pcg_tn;
  // FP32 copies for the bulk PCG of mixed precision
  DevBuf f_r, f_z, f_p, f_q, f_x, f_dinv, f_tn, f_D, f_rhs;
  DevBuf S, flags, part;
  DevBuf keep_x, keep_z, keep_lam;  // the point a device-driven step started from
  double* hS = nullptr;  // pinned mirror of S
  int* hF = nullptr;
  NormalEq ne;
  DenseNE dense;
  bool use_dense = true;
  bool use_sparse = false;   // sparse LDL' on the exact minimum-degree ordering
  hn_sldl* sl = nullptr;
  bool fp64_solve = true;  // current normal-equation precision (dense)
  double reg_now = 0.0;
  bool dense_ready = false;

  Solver(const hn_ipm_opts& oo, hn_ipm_result& rr, hipStream_t ss) : o(oo), res(rr), s(ss) {}
  ~Solver() {
    sldl_destroy(sl);
    delete M;
    delete Mt;
    if (hS) (void)hipHostFree(hS);
    if (hF) (void)hipHostFree(hF);
  }

  dim3 g(int64_t cnt) const { return dim3(grid_for(cnt, 256)); }

  int pull() {  // scalars and flags to the host
    HN_HIP(hipMemcpyAsync(hS, S.p, sizeof(double) * S_COUNT, hipMemcpyDeviceToHost, s));
    HN_HIP(hipMemcpyAsync(hF, flags.p, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, s));
    HN_HIP(hipStreamSynchronize(s));
    return HN_OK;
  }

  int setup() {
    m = (int)f.m;
    n = (int)f.n;
    HN_CHECK(f.m < INT_MAX / 2 && f.n < INT_MAX / 2, HN_ERR_UNSUPPORTED, "hn_ipm_solve: LP too large");
    std::vector<uint64_t> rp64(f.rp.begin(), f.rp.end()), ci64(f.ci.begin(), f.ci.end());
    const int flags32 = HN_F64 | (o.precision == HN_PREC_MIXED ? HN_F32 : 0);
    HN_TRY(hn_csr_upload(f.v.data(), ci64.data(), rp64.data(), f.m, f.n, f.v.size(), flags32, &M));
    HN_TRY(hn_csr_transpose(M, &Mt));
    const size_t bm = sizeof(double) * std::max(m, 1), bn = sizeof(double) * std::max(n, 1);
    for (DevBuf* d : {&b, &lam, &rp, &tm, &mdrd, &rhs, &rhs2, &dlam, &mz, &ir_r, &ir_nd, &ir_e}) HN_TRY(d->reserve(bm));
    for (DevBuf* d : {&c, &x, &z, &rd, &D, &Drd, &tn, &dx, &dz, &zs}) HN_TRY(d->reserve(bn));
    HN_TRY(S.reserve(sizeof(double) * S_COUNT));
    HN_TRY(flags.reserve(sizeof(int) * F_COUNT));
    HN_TRY(part.reserve(sizeof(double) * 3 * kMaxParts));
    HN_HIP(hipHostMalloc((void**)&hS, sizeof(double) * S_COUNT, hipHostMallocDefault));
    HN_HIP(hipHostMalloc((void**)&hF, sizeof(int) * F_COUNT, hipHostMallocDefault));
    HN_HIP(hipMemsetAsync(S.p, 0, sizeof(double) * S_COUNT, s));
    HN_HIP(hipMemsetAsync(flags.p, 0, sizeof(int) * F_COUNT, s));
    HN_HIP(hipMemcpyAsync(b.p, f.rhs.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(c.p, f.c.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    use_dense = o.solver == HN_NE_DENSE || (o.solver == HN_NE_AUTO && f.m <= o.dense_max_m);
    // candidate entries of N (sum over columns of M of their squared counts)
    // bound its pattern and the ordering's first pass
    double cand = 0.0;
    if (!use_dense && o.solver == HN_NE_AUTO) {
      std::vector<int64_t> cc(n, 0);
      for (const int64_t j : f.ci) ++cc[j];
      for (const int64_t k : cc) cand += (double)k * (double)k;
    }
    if (!use_dense && (o.solver == HN_NE_SPARSE ||
                       (o.solver == HN_NE_AUTO && o.pivot_rule == HN_PIVOT_CLAMP && o.precision == HN_PREC_F64 &&
                        cand <= kSparseAutoCand))) {
      // the reference's factorisation: symbolic analysis once per solve
      // (interior_point.cpp:549-576; D > 0 on every iterate keeps N's
      // pattern), numeric LDL' per iteration.  Under AUTO the ordering gets a
      // work budget and a factor that fills beyond it falls back to PCG.
      HN_CHECK(o.solver != HN_NE_SPARSE || (o.pivot_rule == HN_PIVOT_CLAMP && o.precision == HN_PREC_F64),
               HN_ERR_UNSUPPORTED, "hn_ipm_solve: the sparse LDL' solver takes the reference pivot rule in FP64");
      HN_TRY(ne_symbolic(M, Mt, ne, s));
      std::vector<int> nrp(m + 1), nci((size_t)std::max<int64_t>(ne.N->nnz, 1));
      HN_HIP(hipMemcpyAsync(nrp.data(), ne.N->row_ptr.p, sizeof(int) * (m + 1), hipMemcpyDeviceToHost, s));
      if (ne.N->nnz)
        HN_HIP(hipMemcpyAsync(nci.data(), ne.N->col.p, sizeof(int) * ne.N->nnz, hipMemcpyDeviceToHost, s));
      HN_HIP(hipStreamSynchronize(s));
      std::vector<int64_t> rp64n(nrp.begin(), nrp.end()), ci64n(nci.begin(), nci.begin() + ne.N->nnz);
      const bool auto_pick = o.solver == HN_NE_AUTO;
      HN_TRY(sldl_create(m, rp64n.data(), ci64n.data(), auto_pick ? kSparseAutoWork : 0.0,
                         auto_pick ? kSparseAutoLnnz : 0.0, &sl));
      if (!sl->S.aborted && auto_pick && sl->S.crit_steps > kSparseAutoCrit) sl->S.aborted = true;
      if (!sl->S.aborted) {
        use_sparse = true;
        res.sparse_l_nnz = sl->S.Lrp[m];
      } else {
        sldl_destroy(sl);
        sl = nullptr;
        HN_CHECK(auto_pick, HN_ERR_UNSUPPORTED, "hn_ipm_solve: sparse LDL' ordering stopped");
      }
    }
    if (use_dense) {
      HN_TRY(ne_symbolic(M, Mt, ne, s));
    } else if (!use_sparse) {
      // the matrix-free path spends its time in products with M and M':
      // on large matrices both get the SpMV layout that is fastest on them
      // (a few ms of trial products, once per solve)
      if (f.v.size() >= kAutotuneNnz) {
        HN_TRY(hn_csr_autotune(M, 3, nullptr));
        HN_TRY(hn_csr_autotune(Mt, 3, nullptr));
      }
      for (DevBuf* d : {&pcg_r, &pcg_z, &pcg_p, &pcg_q, &pcg_dinv, &pcg_diag}) HN_TRY(d->reserve(bm));
      HN_TRY(pcg_tn.reserve(bn));
      if (o.precision == HN_PREC_MIXED) {
        const size_t fm = sizeof(float) * std::max(m, 1), fn = sizeof(float) * std::max(n, 1);
        for (DevBuf* d : {&f_r, &f_z, &f_p, &f_q, &f_x, &f_dinv, &f_rhs}) HN_TRY(d->reserve(fm));
        for (DevBuf* d : {&f_tn, &f_D}) HN_TRY(d->reserve(fn));
      }
    }
    fp64_solve = o.precision != HN_PREC_MIXED;
    HN_HIP(hipStreamSynchronize(s));
    return HN_OK;
  }

  int spmv(const hn_csr* A, const double* xin, double* yout) { return spmv_launch(A, HN_F64, xin, yout, s); }

  double rel_reg() const { return o.pivot_rule == HN_PIVOT_SKIP ? 0.0 : kRelReg; }

  // Factor (dense) or prepare (PCG) the normal equations for the current D.
  int prepare_ne() {
    if (use_sparse) {
      HN_TRY(ne_numeric(M, Mt, D.as<double>(), o.regularization, rel_reg(), ne, S.as<double>() + S_REG, nullptr,
                        false, s));
      HN_TRY(sldl_factor(sl, ne.N, kSingularTol, s));
      dense_ready = true;
    } else if (use_dense) {
      HN_TRY(ne_numeric(M, Mt, D.as<double>(), o.regularization, rel_reg(), ne, S.as<double>() + S_REG, &dense,
                        !fp64_solve, s));
      HN_TRY(dense_factor(dense, o.pivot_rule, s));
      dense_ready = true;
    } else {
      HN_TRY(ne_jacobi_diag(M, D.as<double>(), pcg_diag.as<double>(), s));
      hipLaunchKernelGGL(k_pcg_reg, dim3(1), dim3(kNT), 0, s, pcg_diag.as<double>(), m, o.regularization,
                         rel_reg(), S.as<double>());
      if (!fp64_solve) hipLaunchKernelGGL((k_cvt<double, float>), g(n), dim3(256), 0, s, D.as<double>(), f_D.as<float>(), n);
      HN_TRY(check_launch("k_pcg_reg"));
    }
    return HN_OK;
  }

  // out (+)= N^{-1} r by the dense inverse or the sparse factor
  int ne_apply(const double* r, double* out, bool accumulate, hipStream_t st) {
    if (use_sparse) return sldl_solve(sl, r, out, accumulate, st);
    return dense_apply(dense, r, out, accumulate, st);
  }

  // N d = v by the matrix-free operator (FP64).
  int apply_N(const double* v, double* out) {
    if (use_dense || use_sparse) return spmv(ne.N, v, out);
    HN_TRY(spmv(Mt, v, pcg_tn.as<double>()));
    hipLaunchKernelGGL((k_mul<double>), g(n), dim3(256), 0, s, pcg_tn.as<double>(), D.as<double>(),
                       pcg_tn.as<double>(), n);
    HN_TRY(spmv(M, pcg_tn.as<double>(), out));
    return HN_OK;
  }

  // q = M (D (M' p)) in FP32 (the regularisation is added by k_pcg_pq)
  int apply_N32(const float* p, float* q) {
    HN_TRY(spmv_launch(Mt, HN_F32, p, f_tn.as<float>(), s));
    hipLaunchKernelGGL((k_mul<float>), g(n), dim3(256), 0, s, f_tn.as<float>(), f_D.as<float>(), f_tn.as<float>(),
                       n);
    return spmv_launch(M, HN_F32, f_tn.as<float>(), q, s);
  }

  // Dense: dlam = N^{-1} rhs then refinement against N (refinement.cpp:7-36):
  // up to 5 corrections, stop when ||rhs - N dlam||_inf <= 1e-10 ||rhs||_inf.
  // solve_dense_begin forms the first solution and its residual check on
  // the device; dense_first_ok reads the check after a pull;
  // solve_dense_rest runs the corrections when it failed.
  int solve_dense_begin(const double* r_in, double* out) {
    HN_TRY(ne_apply(r_in, out, false, s));
    ++res.ne_solves;
    const int np = parts_for(m);
    HN_TRY(apply_N(out, ir_nd.as<double>()));
    hipLaunchKernelGGL(k_red_resid, dim3(np), dim3(kNT), 0, s, r_in, ir_nd.as<double>(), ir_r.as<double>(), m,
                       part.as<double>(), nullptr, nullptr);
    hipLaunchKernelGGL(k_fin_max2, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>(), S_IRR, S_IRB, 0, nullptr);
    return check_launch("dense solve");
  }
  bool dense_first_ok() const {
    const double nb = hS[S_IRB];
    return nb == 0.0 || nb != nb || hS[S_IRR] / nb <= 1e-10;
  }
  int solve_dense_rest(const double* r_in, double* out, bool* converged) {
    const int np = parts_for(m);
    *converged = false;
    for (int it = 1; it <= 5; ++it) {
      HN_TRY(ne_apply(ir_r.as<double>(), out, true, s));
      ++res.refine_sweeps;
      HN_TRY(apply_N(out, ir_nd.as<double>()));
      hipLaunchKernelGGL(k_red_resid, dim3(np), dim3(kNT), 0, s, r_in, ir_nd.as<double>(), ir_r.as<double>(), m,
                         part.as<double>(), nullptr, nullptr);
      hipLaunchKernelGGL(k_fin_max2, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>(), S_IRR, S_IRB, 0, nullptr);
      HN_TRY(pull());
      if (dense_first_ok()) {
        *converged = true;
        return HN_OK;
      }
    }
    return HN_OK;
  }
  int solve_dense(const double* r_in, double* out, bool* converged) {
    HN_TRY(solve_dense_begin(r_in, out));
    HN_TRY(pull());
    if (dense_first_ok()) {
      *converged = true;
      return HN_OK;
    }
    return solve_dense_rest(r_in, out, converged);
  }

  // The dense solve with its refinement on the device: the first solution
  // and its check, then `sweeps` correction sweeps enqueued ahead, each
  // doing nothing once a check has passed (ne_refined).  The same kernels in
  // the same order as solve_dense, with no host round trip, so a whole
  // iteration can be captured in a graph.  When the last check still fails
  // and sweeps < kMaxSweeps (the host would run another sweep), F_MORE is set.
  int solve_dense_dev(const double* r_in, double* out, int sweeps) {
    const int np = parts_for(m);
    double* Sd = S.as<double>();
    int* more = flags.as<int>() + F_MORE;
    HN_TRY(dense_apply(dense, r_in, out, false, s));
    HN_TRY(apply_N(out, ir_nd.as<double>()));
    hipLaunchKernelGGL(k_red_resid, dim3(np), dim3(kNT), 0, s, r_in, ir_nd.as<double>(), ir_r.as<double>(), m,
                       part.as<double>(), nullptr, nullptr);
    hipLaunchKernelGGL(k_fin_max2, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, Sd, S_IRR, S_IRB, 0,
                       sweeps == 0 ? more : nullptr);
    for (int it = 1; it <= sweeps; ++it) {
      HN_TRY(dense_apply_refine(dense, ir_r.as<double>(), out, Sd + S_IRR, Sd + S_IRB, Sd + S_SWEEPS, s));
      HN_TRY(apply_N(out, ir_nd.as<double>()));
      hipLaunchKernelGGL(k_red_resid, dim3(np), dim3(kNT), 0, s, r_in, ir_nd.as<double>(), ir_r.as<double>(), m,
                         part.as<double>(), Sd + S_IRR, Sd + S_IRB);
      hipLaunchKernelGGL(k_fin_max2, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, Sd, S_IRR, S_IRB, 1,
                         (it == sweeps && it < kMaxSweeps) ? more : nullptr);
    }
    count_launch(2 + 4 * sweeps);
    return check_launch("dense solve (device refinement)");
  }

  // PCG on the matrix-free N with a Jacobi preconditioner; the relative
  // tolerance follows mu (tighter as the iterates approach the optimum).
  int solve_pcg(const double* r_in, double* out, double mu) {
    const int np = parts_for(m);
    hipLaunchKernelGGL((k_pcg_init<double>), dim3(np), dim3(kNT), 0, s, r_in, pcg_diag.as<double>(), S.as<double>(),
                       out, pcg_r.as<double>(), pcg_z.as<double>(), pcg_p.as<double>(), pcg_dinv.as<double>(), m,
                       part.as<double>());
    hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>());
    HN_TRY(pull());
    const double r0 = hS[S_PCG_RN];
    const double rel = std::max(1e-11, std::min(1e-3, 0.1 * mu));
    if (r0 == 0.0) return HN_OK;
    ++res.ne_solves;
    for (int it = 0; it < o.pcg_max_iter; ++it) {
      HN_TRY(apply_N(pcg_p.as<double>(), pcg_q.as<double>()));
      hipLaunchKernelGGL((k_pcg_pq<double>), dim3(np), dim3(kNT), 0, s, pcg_q.as<double>(), pcg_p.as<double>(),
                         S.as<double>(), m, part.as<double>());
      hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>(), (int)S_PCG_PQ);
      hipLaunchKernelGGL((k_pcg_step<double>), dim3(np), dim3(kNT), 0, s, out, pcg_r.as<double>(), pcg_z.as<double>(),
                         pcg_p.as<double>(), pcg_q.as<double>(), pcg_dinv.as<double>(), S.as<double>(), m,
                         part.as<double>());
      hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>());
      hipLaunchKernelGGL((k_pcg_dir<double>), g(m), dim3(256), 0, s, pcg_p.as<double>(), pcg_z.as<double>(),
                         S.as<double>(), m);
      ++res.cg_iterations;
      if ((it & 7) == 7) {
        HN_TRY(pull());
        if (!(hS[S_PCG_RN] > rel * r0)) break;
      }
    }
    return check_launch("pcg");
  }

  // Mixed precision: FP32 Jacobi-PCG on the FP32 operator for the bulk of
  // the work, FP64 refinement around it - the residual r_in - N out is
  // recomputed in FP64 and solved for again, until it meets the tolerance
  // of the FP64 PCG (or 8 sweeps).
  int solve_pcg_mixed(const double* r_in, double* out, double mu) {
    const int np = parts_for(m);
    const double rel = std::max(1e-11, std::min(1e-3, 0.1 * mu));
    hipLaunchKernelGGL(k_fill, g(m), dim3(256), 0, s, out, 0.0, m);
    const double* res_src = r_in;
    double r0 = -1.0;
    ++res.ne_solves;
    for (int sweep = 0; sweep < 8; ++sweep) {
      if (sweep > 0) {
        HN_TRY(apply_N(out, ir_nd.as<double>()));
        hipLaunchKernelGGL(k_red_resid, dim3(np), dim3(kNT), 0, s, r_in, ir_nd.as<double>(), ir_r.as<double>(), m,
                           part.as<double>(), nullptr, nullptr);
        hipLaunchKernelGGL(k_fin_max2, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>(), S_IRR, S_IRB, 0, nullptr);
        HN_TRY(pull());
        r0 = hS[S_IRB];
        if (!(hS[S_IRR] > rel * r0)) break;
        res_src = ir_r.as<double>();
        ++res.refine_sweeps;
      }
      hipLaunchKernelGGL((k_cvt<double, float>), g(m), dim3(256), 0, s, res_src, f_rhs.as<float>(), m);
      hipLaunchKernelGGL((k_pcg_init<float>), dim3(np), dim3(kNT), 0, s, f_rhs.as<float>(), pcg_diag.as<double>(),
                         S.as<double>(), f_x.as<float>(), f_r.as<float>(), f_z.as<float>(), f_p.as<float>(),
                         f_dinv.as<float>(), m, part.as<double>());
      hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>());
      HN_TRY(pull());
      const double f0 = hS[S_PCG_RN];
      if (f0 == 0.0) break;
      // FP32 can resolve about 1e-6 of the residual it starts from
      const double frel = std::max(1e-6, sweep == 0 ? rel : 1e-6);
      for (int it = 0; it < o.pcg_max_iter; ++it) {
        HN_TRY(apply_N32(f_p.as<float>(), f_q.as<float>()));
        hipLaunchKernelGGL((k_pcg_pq<float>), dim3(np), dim3(kNT), 0, s, f_q.as<float>(), f_p.as<float>(),
                           S.as<double>(), m, part.as<double>());
        hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>(),
                           (int)S_PCG_PQ);
        hipLaunchKernelGGL((k_pcg_step<float>), dim3(np), dim3(kNT), 0, s, f_x.as<float>(), f_r.as<float>(),
                           f_z.as<float>(), f_p.as<float>(), f_q.as<float>(), f_dinv.as<float>(), S.as<double>(), m,
                           part.as<double>());
        hipLaunchKernelGGL(k_pcg_fin, dim3(1), dim3(kNT), 0, s, part.as<double>(), np, S.as<double>());
        hipLaunchKernelGGL((k_pcg_dir<float>), g(m), dim3(256), 0, s, f_p.as<float>(), f_z.as<float>(),
                           S.as<double>(), m);
        ++res.cg_iterations;
        if ((it & 7) == 7) {
          HN_TRY(pull());
          if (!(hS[S_PCG_RN] > frel * f0)) break;
        }
      }
      hipLaunchKernelGGL(k_add_f32, g(m), dim3(256), 0, s, f_x.as<float>(), out, m);
    }
    return check_launch("pcg mixed");
  }

  // FP32 inverse could not be refined: switch to FP64 for good and solve again
  int escalate(const double* r_in, double* out, bool* escalated) {
    fp64_solve = true;
    *escalated = true;
    HN_TRY(ne_numeric(M, Mt, D.as<double>(), o.regularization, rel_reg(), ne, S.as<double>() + S_REG, &dense, false,
                      s));
    HN_TRY(dense_factor(dense, o.pivot_rule, s));
    bool ok = false;
    return solve_dense(r_in, out, &ok);
  }

  int solve(const double* r_in, double* out, double mu, bool* escalated) {
    *escalated = false;
    if (!use_dense && !use_sparse) return fp64_solve ? solve_pcg(r_in, out, mu) : solve_pcg_mixed(r_in, out, mu);
    bool ok = false;
    HN_TRY(solve_dense(r_in, out, &ok));
    if (!ok && !fp64_solve) HN_TRY(escalate(r_in, out, escalated));
    return HN_OK;
  }

  // The dense solve with its first check left to the caller's next pull
  // (solve_begin, then the caller's kernels and pull, then solve_settle):
  // the same kernels as solve() in the same order.  *redo: the first
  // solution failed its check, the solve went on here, and the caller's
  // kernels that used it must run again.
  int solve_begin(const double* r_in, double* out) { return solve_dense_begin(r_in, out); }
  int solve_settle(const double* r_in, double* out, bool* escalated, bool* redo) {
    *escalated = false;
    *redo = false;
    if (dense_first_ok()) return HN_OK;
    *redo = true;
    bool ok = false;
    HN_TRY(solve_dense_rest(r_in, out, &ok));
    if (!ok && !fp64_solve) HN_TRY(escalate(r_in, out, escalated));
    return HN_OK;
  }
};

}  // namespace

int ipm_solve_impl(const hn_lp& lp, const hn_ipm_opts& o, hn_ipm_result& res) {
  Timer clock;
  res.status = HN_STATUS_UNKNOWN;
  res.iterations = 0;
  res.objective = 0.0;
  res.sf_m = res.sf_n = 0;
  res.rp_norm = res.rd_norm = res.mu = 0.0;
  res.rel_primal = res.rel_dual = res.rel_gap = 0.0;
  res.scaled_rel_primal = res.scaled_rel_dual = res.scaled_rel_gap = 0.0;
  res.ne_solves = res.refine_sweeps = res.cg_iterations = 0;
  res.fp64_from_iter = -1;
  res.warm_started = 0;
  res.projected = 0;
  res.sparse_l_nnz = 0;
  res.device_steps = res.host_redos = 0;
  auto finish = [&]() { res.solve_ms = 1e3 * clock.s(); };

  if (lp.n == 0) {  // interior_point.cpp:406-410
    res.status = HN_STATUS_OPTIMAL;
    res.objective = lp.obj_offset;
    finish();
    return HN_OK;
  }
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  Solver sv(o, res, s);
  HN_TRY(build_standard_form(lp, sv.f));
  StandardForm& f = sv.f;
  res.sf_m = f.m;
  res.sf_n = f.n;
  std::vector<double> primal(lp.n, 0.0);
  auto copy_primal = [&]() {
    if (res.primal) std::memcpy(res.primal, primal.data(), sizeof(double) * lp.n);
  };
  if (f.n == 0) {  // :416-421
    res.status = HN_STATUS_OPTIMAL;
    res.objective = f.offset;
    to_original_primal(f, nullptr, primal.data());
    copy_primal();
    finish();
    return HN_OK;
  }
  if (f.m == 0) {  // :423-441
    bool unbounded = false;
    for (int64_t j = 0; j < f.n; ++j)
      if (f.c[j] < -1e-9) {
        unbounded = true;
        break;
      }
    res.status = unbounded ? HN_STATUS_UNBOUNDED : HN_STATUS_OPTIMAL;
    std::vector<double> y(f.n, 0.0);
    to_original_primal(f, y.data(), primal.data());
    double obj = lp.obj_offset;
    for (int64_t j = 0; j < lp.n; ++j) obj += lp.obj[j] * primal[j];
    res.objective = obj;
    copy_primal();
    finish();
    return HN_OK;
  }
  HN_TRY(sv.setup());
  const int m = sv.m, n = sv.n;
  double* S = sv.S.as<double>();
  int* F = sv.flags.as<int>();
  double* hS = sv.hS;

  // start point (:443-453) or the caller's warm start
  auto cold_start = [&]() -> int {
    std::vector<double> x0(n), z0(n), l0(m, 0.0);
    double b_norm = 1.0, c_norm = 1.0;
    for (double v : f.rhs) b_norm = std::max(b_norm, std::abs(v));
    for (double v : f.c) c_norm = std::max(c_norm, std::abs(v));
    for (int j = 0; j < n; ++j) {
      x0[j] = b_norm + 1.0;
      z0[j] = std::max(c_norm + 1.0, std::abs(f.c[j]));
    }
    HN_HIP(hipMemcpyAsync(sv.x.p, x0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(sv.z.p, z0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(sv.lam.p, l0.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    HN_HIP(hipStreamSynchronize(s));
    return HN_OK;
  };
  {
    std::vector<double> x0(n), z0(n), l0(m, 0.0);
    // the caller's warm start (options first, then the LP's own), if it
    // passes the acceptance test; otherwise the standard start point
    const double* wx = o.x0 ? o.x0 : lp.x0;
    const double* wy = o.x0 ? o.y0 : lp.y0;
    const double* ws = o.x0 ? o.s0 : lp.s0;
    if (wx && wy && ws && warm_start_acceptable(f, wx, wy, ws)) {
      std::memcpy(x0.data(), wx, sizeof(double) * n);
      std::memcpy(z0.data(), ws, sizeof(double) * n);
      std::memcpy(l0.data(), wy, sizeof(double) * m);
      res.warm_started = 1;
    } else {
      double b_norm = 1.0, c_norm = 1.0;
      for (double v : f.rhs) b_norm = std::max(b_norm, std::abs(v));
      for (double v : f.c) c_norm = std::max(c_norm, std::abs(v));
      for (int j = 0; j < n; ++j) {
        x0[j] = b_norm + 1.0;
        z0[j] = std::max(c_norm + 1.0, std::abs(f.c[j]));
      }
    }
    HN_HIP(hipMemcpyAsync(sv.x.p, x0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(sv.z.p, z0.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(sv.lam.p, l0.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
    HN_HIP(hipStreamSynchronize(s));
  }

  const dim3 B(256);
  const int np_n = parts_for(std::max(n, m));
  bool done = false, time_up = false, interrupted = false, numerical_error = false, stopped = false;
  bool warm_live = res.warm_started != 0;
  int stuck = 0;
  auto deadline = [&]() { return o.time_limit_seconds > 0.0 && clock.s() >= o.time_limit_seconds; };
  // residuals, mu and their norms on the device; host copies in hS
  const dim3 g_nm = sv.g(std::max(n, m));
  auto residual_kernels = [&]() -> int {  // :508-522
    HN_TRY(sv.spmv(sv.M, sv.x.as<double>(), sv.tm.as<double>()));
    HN_TRY(sv.spmv(sv.Mt, sv.lam.as<double>(), sv.tn.as<double>()));
    hipLaunchKernelGGL(k_resid, g_nm, B, 0, s, sv.tm.as<double>(), sv.b.as<double>(), sv.rp.as<double>(), m,
                       sv.tn.as<double>(), sv.z.as<double>(), sv.c.as<double>(), sv.rd.as<double>(), n);
    hipLaunchKernelGGL(k_red_start, dim3(np_n), dim3(kNT), 0, s, sv.x.as<double>(), sv.z.as<double>(),
                       sv.rp.as<double>(), sv.rd.as<double>(), n, m, sv.part.as<double>());
    hipLaunchKernelGGL(k_fin_start, dim3(1), dim3(kNT), 0, s, sv.part.as<double>(), np_n, S);
    count_launch(3);
    return check_launch("residuals");
  };
  auto residuals = [&]() -> int {
    HN_TRY(residual_kernels());
    return sv.pull();
  };
  // direction (:606-620 / :664-678) from dlam, and its ratio test (:629-631)
  auto direction = [&](const double* zs, int slot_p, int slot_d) -> int {
    HN_TRY(sv.spmv(sv.Mt, sv.dlam.as<double>(), sv.tn.as<double>()));
    hipLaunchKernelGGL(k_direction, g_nm, B, 0, s, sv.D.as<double>(), sv.tn.as<double>(), sv.rd.as<double>(),
                       sv.x.as<double>(), zs, sv.dx.as<double>(), sv.dz.as<double>(), n, sv.dlam.as<double>(), m, F);
    hipLaunchKernelGGL(k_red_step, dim3(np_n), dim3(kNT), 0, s, sv.x.as<double>(), sv.dx.as<double>(),
                       sv.z.as<double>(), sv.dz.as<double>(), n, sv.part.as<double>());
    hipLaunchKernelGGL(k_fin_step, dim3(1), dim3(kNT), 0, s, sv.part.as<double>(), np_n, S, slot_p, slot_d);
    count_launch(3);
    return check_launch("direction");
  };
  auto update = [&]() -> int {  // :687-708; skipped on the device if the direction was not finite
    hipLaunchKernelGGL(k_update, g_nm, B, 0, s, sv.x.as<double>(), sv.z.as<double>(), sv.dx.as<double>(),
                       sv.dz.as<double>(), S, n, sv.lam.as<double>(), sv.dlam.as<double>(), m, F);
    count_launch(1);
    return check_launch("update");
  };

  // ---- device-driven iterations (dense normal equations in FP64) ----------
  // Everything between two reads of the residuals runs on the device: the
  // scaling, N and its factor, the predictor with its ratio test and mu_aff,
  // sigma (formed by k_corrector_s), the corrector, both solves with their
  // refinement (solve_dense_dev), the update, and the next point's
  // residuals.  A predictor direction that is not finite sets the flag the
  // update checks, so the step leaves the point unchanged and the host finds
  // the flag with the residuals, as it finds the corrector's.  The step
  // holds no refinement sweep until a solve of this LP has needed one (the
  // FP64 factor rarely leaves one to do); a solve that needs more sweeps
  // than the step holds sets F_MORE, and the host then puts back the point
  // the step started from, runs the iteration on the host path, and has
  // later steps hold all kMaxSweeps (guarded) sweeps.  So the iterates are
  // those of the host path in every case.  With HN_LOOP_GRAPH the first such iteration runs directly (it
  // sizes every workspace), the next is captured in a graph and later ones
  // replay it: an iteration costs one graph launch and one read of the
  // scalars.  The time limit is checked between iterations only (the host
  // path also checks it after factoring N).
  const bool use_graph = o.loop == HN_LOOP_GRAPH;
  struct StepGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    void reset() {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (g) (void)hipGraphDestroy(g);
      exec = nullptr;
      g = nullptr;
    }
    ~StepGraph() { reset(); }
  } graph;
  bool stepped = false;
  int dev_sweeps = 0;
  uint64_t step_launches = 0;
  auto device_step_body = [&]() -> int {
    hipLaunchKernelGGL(k_keep_point, g_nm, B, 0, s, sv.x.as<double>(), sv.z.as<double>(), sv.lam.as<double>(),
                       sv.keep_x.as<double>(), sv.keep_z.as<double>(), sv.keep_lam.as<double>(), n, m);
    hipLaunchKernelGGL(k_scaling, sv.g(n), B, 0, s, sv.x.as<double>(), sv.z.as<double>(), sv.rd.as<double>(),
                       sv.D.as<double>(), sv.Drd.as<double>(), n);
    count_launch(2);
    HN_TRY(sv.prepare_ne());
    // predictor (:584-643)
    HN_TRY(sv.spmv(sv.M, sv.Drd.as<double>(), sv.mdrd.as<double>()));
    hipLaunchKernelGGL(k_rhs, sv.g(m), B, 0, s, sv.rp.as<double>(), sv.mdrd.as<double>(), sv.tm.as<double>(),
                       sv.rhs.as<double>(), m);
    HN_TRY(sv.solve_dense_dev(sv.rhs.as<double>(), sv.dlam.as<double>(), dev_sweeps));
    HN_TRY(direction(nullptr, S_AP_A, S_AD_A));
    hipLaunchKernelGGL(k_red_muaff, dim3(np_n), dim3(kNT), 0, s, sv.x.as<double>(), sv.dx.as<double>(),
                       sv.z.as<double>(), sv.dz.as<double>(), n, S, sv.part.as<double>());
    hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(kNT), 0, s, sv.part.as<double>(), np_n, S, (int)S_MU_AFF);
    // sigma and corrector (:638-685)
    hipLaunchKernelGGL(k_corrector_s, sv.g(n), B, 0, s, sv.dx.as<double>(), sv.dz.as<double>(), sv.z.as<double>(),
                       (const double*)S, sv.zs.as<double>(), n);
    HN_TRY(sv.spmv(sv.M, sv.zs.as<double>(), sv.mz.as<double>()));
    hipLaunchKernelGGL(k_sub, sv.g(m), B, 0, s, sv.rhs.as<double>(), sv.mz.as<double>(), sv.rhs2.as<double>(), m);
    count_launch(5);
    HN_TRY(sv.solve_dense_dev(sv.rhs2.as<double>(), sv.dlam.as<double>(), dev_sweeps));
    HN_TRY(direction(sv.zs.as<double>(), S_AP, S_AD));
    HN_TRY(update());
    HN_TRY(check_launch("device step"));
    return residual_kernels();
  };
  auto device_step = [&]() -> int {
    if (graph.exec) {
      HN_HIP(hipGraphLaunch(graph.exec, s));
      count_launch(step_launches);
      return HN_OK;
    }
    if (!stepped || !use_graph) {
      if (!stepped) {
        HN_TRY(sv.keep_x.reserve(sizeof(double) * std::max(n, 1)));
        HN_TRY(sv.keep_z.reserve(sizeof(double) * std::max(n, 1)));
        HN_TRY(sv.keep_lam.reserve(sizeof(double) * std::max(m, 1)));
      }
      stepped = true;
      return device_step_body();
    }
    const uint64_t l0 = ctx().launches.load();
    HN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = device_step_body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (g) graph.g = g;
    HN_TRY(st);
    HN_HIP(e);
    HN_HIP(hipGraphInstantiate(&graph.exec, graph.g, nullptr, nullptr, 0));
    step_launches = ctx().launches.load() - l0;
    HN_HIP(hipGraphLaunch(graph.exec, s));
    return HN_OK;
  };

  int64_t iter = 0;
  double mu = 0.0;
  // The flags of a step's corrector direction and update are read with the
  // next residuals (one host round trip fewer per iteration): a failed step
  // ends the solve there, with the iteration count and state the reference
  // has when it stops inside that step.
  auto step_failed = [&]() { return sv.hF[F_DIR] != 0 || sv.hF[F_UPD] != 0; };
  double prev_mu = 0.0, prev_rp = 0.0, prev_rd = 0.0;
  bool fresh = false;  // hS already holds this point's residuals (read after a device step)
  for (iter = 0; iter < o.max_iterations; ++iter) {
    if (!fresh) HN_TRY(residuals());
    fresh = false;
    if (iter > 0 && o.verbose)
      std::fprintf(stderr, "IPM it=%lld mu=%g rp=%g rd=%g ap=%g ad=%g%s\n", (long long)(iter - 1), prev_mu, prev_rp,
                   prev_rd, hS[S_AP], hS[S_AD], sv.fp64_solve ? "" : " [fp32 NE]");
    if (step_failed()) {
      numerical_error = true;
      res.iterations = iter - 1;
      break;
    }
    if (o.interrupt_callback && o.interrupt_callback(o.interrupt_user)) {  // :490-493
      interrupted = true;
      break;
    }
    if (deadline()) {  // :494-501
      time_up = true;
      break;
    }
    mu = hS[S_XZ] / (double)n;
    prev_mu = mu;
    prev_rp = hS[S_RPN];
    prev_rd = hS[S_RDN];
    if (hS[S_RPN] < o.convergence_tol && hS[S_RDN] < o.convergence_tol && mu < o.complementarity_tol) {
      bool gap_ok = true;
      if (o.gap_tol > 0.0) {  // optional relative-gap bar on the caller's LP (not in the reference)
        std::vector<double> gx(n), gz(n), gl(m);
        HN_HIP(hipMemcpyAsync(gx.data(), sv.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HN_HIP(hipMemcpyAsync(gz.data(), sv.z.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HN_HIP(hipMemcpyAsync(gl.data(), sv.lam.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
        HN_HIP(hipStreamSynchronize(s));
        OriginalKkt kk;
        original_kkt(lp, f, gx.data(), gl.data(), gz.data(), kk);
        gap_ok = kk.rel_gap <= o.gap_tol;
      }
      if (gap_ok) {
        done = true;  // :525-530
        break;
      }
    }
    if (o.mu_target > 0.0 && mu <= o.mu_target) {  // early stop (not in the reference)
      stopped = true;
      break;
    }
    // a warm start that lets mu fall below its tolerance while the residuals
    // stay above theirs for 3 iterations has stalled: restart once from the
    // standard point
    if (warm_live) {
      stuck = (mu < o.complementarity_tol &&
               (hS[S_RPN] >= o.convergence_tol || hS[S_RDN] >= o.convergence_tol)) ? stuck + 1 : 0;
      if (stuck >= 3) {
        warm_live = false;
        res.warm_started = 2;
        HN_TRY(cold_start());
        HN_TRY(residuals());
        mu = hS[S_XZ] / (double)n;
      }
    }
    if (!sv.fp64_solve && mu < o.mu_switch) {
      sv.fp64_solve = true;
      if (res.fp64_from_iter < 0) res.fp64_from_iter = iter;
    }
    if (o.loop != HN_LOOP_HOST && sv.use_dense && sv.fp64_solve) {
      const double sweeps_before = hS[S_SWEEPS];
      HN_TRY(device_step());
      HN_TRY(sv.pull());
      ++res.device_steps;
      if (!sv.hF[F_MORE]) {
        res.ne_solves += 2;
        res.iterations = iter + 1;
        fresh = true;
        continue;
      }
      // a solve needed more refinement sweeps than the step holds: back to
      // the point this iteration started from, and the host path below
      ++res.host_redos;
      dev_sweeps = kMaxSweeps;
      graph.reset();  // the stream is idle: pull() synchronised it
      hipLaunchKernelGGL(k_undo_step, g_nm, B, 0, s, sv.x.as<double>(), sv.z.as<double>(), sv.lam.as<double>(),
                         sv.keep_x.as<double>(), sv.keep_z.as<double>(), sv.keep_lam.as<double>(), n, m, F, S,
                         sweeps_before);
      HN_TRY(check_launch("undo step"));
      HN_TRY(residuals());
    }
    hipLaunchKernelGGL(k_scaling, sv.g(n), B, 0, s, sv.x.as<double>(), sv.z.as<double>(), sv.rd.as<double>(),
                       sv.D.as<double>(), sv.Drd.as<double>(), n);
    HN_TRY(sv.prepare_ne());
    if (deadline()) {
      time_up = true;
      break;
    }
    // predictor (:584-615)
    HN_TRY(sv.spmv(sv.M, sv.Drd.as<double>(), sv.mdrd.as<double>()));
    hipLaunchKernelGGL(k_rhs, sv.g(m), B, 0, s, sv.rp.as<double>(), sv.mdrd.as<double>(), sv.tm.as<double>(),
                       sv.rhs.as<double>(), m);
    bool esc = false;
    // the dense solve's first check is read with the predictor's scalars
    // (one host round trip fewer); if it failed, the solve is finished and
    // the predictor's kernels run again.  Only for the FP64 factor: an FP32
    // one almost never passes its first check, so the predictor would run
    // twice.
    const bool deferred = (sv.use_dense || sv.use_sparse) && sv.fp64_solve;
    if (deferred)
      HN_TRY(sv.solve_begin(sv.rhs.as<double>(), sv.dlam.as<double>()));
    else
      HN_TRY(sv.solve(sv.rhs.as<double>(), sv.dlam.as<double>(), mu, &esc));
    auto predictor = [&]() -> int {
      // direction, ratio test and mu_aff (:606-643)
      HN_TRY(direction(nullptr, S_AP_A, S_AD_A));
      hipLaunchKernelGGL(k_red_muaff, dim3(np_n), dim3(kNT), 0, s, sv.x.as<double>(), sv.dx.as<double>(),
                         sv.z.as<double>(), sv.dz.as<double>(), n, S, sv.part.as<double>());
      hipLaunchKernelGGL(k_fin_sum, dim3(1), dim3(kNT), 0, s, sv.part.as<double>(), np_n, S, (int)S_MU_AFF);
      HN_TRY(check_launch("predictor"));
      return sv.pull();
    };
    HN_TRY(predictor());
    if (deferred) {
      bool redo = false;
      HN_TRY(sv.solve_settle(sv.rhs.as<double>(), sv.dlam.as<double>(), &esc, &redo));
      if (redo) {  // the direction flag of the unrefined solution does not count
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, F, (int)F_DIR);
        HN_TRY(predictor());
      }
    }
    if (esc && res.fp64_from_iter < 0) res.fp64_from_iter = iter;
    if (sv.hF[F_DIR]) {
      numerical_error = true;
      break;
    }
    const double mu_aff = hS[S_MU_AFF] / (double)n;
    double sigma;
    if (!std::isfinite(mu_aff) || mu_aff < 0.0 || !(mu > 0.0))
      sigma = 0.5;
    else
      sigma = std::clamp(std::pow(mu_aff / std::max(mu, 1e-300), 3.0), 0.1, 0.9);
    // corrector (:645-685)
    hipLaunchKernelGGL(k_corrector, sv.g(n), B, 0, s, sv.dx.as<double>(), sv.dz.as<double>(), sv.z.as<double>(),
                       sigma * mu, sv.zs.as<double>(), n);
    HN_TRY(sv.spmv(sv.M, sv.zs.as<double>(), sv.mz.as<double>()));
    hipLaunchKernelGGL(k_sub, sv.g(m), B, 0, s, sv.rhs.as<double>(), sv.mz.as<double>(), sv.rhs2.as<double>(), m);
    HN_TRY(sv.solve(sv.rhs2.as<double>(), sv.dlam.as<double>(), mu, &esc));
    if (esc && res.fp64_from_iter < 0) res.fp64_from_iter = iter;
    HN_TRY(direction(sv.zs.as<double>(), S_AP, S_AD));
    HN_TRY(update());
    res.iterations = iter + 1;
  }

  res.refine_sweeps += (int64_t)hS[S_SWEEPS];
  if (!done && !interrupted && !stopped) {  // :728-753
    if (!fresh) HN_TRY(residuals());
    if (iter == o.max_iterations) {  // the last step's flags
      if (iter > 0 && o.verbose)
        std::fprintf(stderr, "IPM it=%lld mu=%g rp=%g rd=%g ap=%g ad=%g%s\n", (long long)(iter - 1), prev_mu,
                     prev_rp, prev_rd, hS[S_AP], hS[S_AD], sv.fp64_solve ? "" : " [fp32 NE]");
      if (step_failed()) {
        numerical_error = true;
        res.iterations = iter - 1;
      }
    }
    mu = hS[S_XZ] / (double)n;
    if (hS[S_RPN] < 1e-4 && hS[S_RDN] < 1e-4 && mu < 1e-4) done = true;
  } else if (interrupted || stopped) {
    HN_TRY(residuals());
    mu = hS[S_XZ] / (double)n;
  }
  std::vector<double> hx(n), hz(n), hl(m);
  HN_HIP(hipMemcpyAsync(hx.data(), sv.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HN_HIP(hipMemcpyAsync(hz.data(), sv.z.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  HN_HIP(hipMemcpyAsync(hl.data(), sv.lam.p, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  // final metrics on the scaled form
  hipLaunchKernelGGL(k_red_final, dim3(np_n), dim3(kNT), 0, s, sv.c.as<double>(), sv.x.as<double>(),
                     sv.b.as<double>(), sv.lam.as<double>(), sv.rd.as<double>(), n, m, sv.part.as<double>());
  hipLaunchKernelGGL(k_fin_final, dim3(1), dim3(kNT), 0, s, sv.part.as<double>(), np_n, S);
  HN_TRY(sv.pull());
  const double rpn = hS[S_RPN];
  to_original_primal(f, hx.data(), primal.data());  // :757
  double obj = 0.0;                                  // :759-769
  bool obj_ok = std::isfinite(lp.obj_offset);
  for (int64_t j = 0; j < lp.n && obj_ok; ++j) {
    const double term = lp.obj[j] * primal[j];
    if (!std::isfinite(term)) obj_ok = false;
    obj += term;
  }
  if (obj_ok) obj += lp.obj_offset;
  if (!obj_ok) numerical_error = true;
  res.objective = obj_ok ? obj : 0.0;
  const bool feasible_incumbent = rpn <= 1e-9;       // :773-780
  res.status = numerical_error ? HN_STATUS_NUMERICAL_ERROR
               : interrupted   ? HN_STATUS_INTERRUPTED
               : done          ? HN_STATUS_OPTIMAL
               : stopped       ? HN_STATUS_ITER_LIMIT
               : feasible_incumbent ? HN_STATUS_SUBOPTIMAL
               : time_up       ? HN_STATUS_TIME_LIMIT
                               : HN_STATUS_ITER_LIMIT;
  res.rp_norm = rpn;
  res.rd_norm = hS[S_RDN];
  res.mu = mu;
  scaled_kkt(f, hx.data(), hl.data(), rpn, hS[S_DRN], res);
  OriginalKkt kk;
  original_kkt(lp, f, hx.data(), hl.data(), hz.data(), kk);
  res.rel_primal = kk.rel_primal;
  res.rel_dual = kk.rel_dual;
  res.rel_gap = kk.rel_gap;
  if (res.x && res.cap_n >= n) std::memcpy(res.x, hx.data(), sizeof(double) * n);
  if (res.s && res.cap_n >= n) std::memcpy(res.s, hz.data(), sizeof(double) * n);
  if (res.y && res.cap_m >= m) std::memcpy(res.y, hl.data(), sizeof(double) * m);
  copy_primal();
  finish();
  return HN_OK;
}

}  // namespace hn

extern "C" int hn_ipm_default_opts(hn_ipm_opts* o) {
  if (!o) return HN_ERR_INVALID;
  std::memset(o, 0, sizeof(*o));
  o->max_iterations = 100;
  o->time_limit_seconds = 0.0;
  o->convergence_tol = 1e-8;
  o->complementarity_tol = 1e-8;
  o->regularization = 1e-8;
  o->solver = HN_NE_AUTO;
  o->precision = HN_PREC_F64;
  o->mu_switch = 1e-4;
  o->dense_max_m = 40000;
  o->pcg_max_iter = 2000;
  o->pivot_rule = HN_PIVOT_CLAMP;
  return HN_OK;
}

extern "C" int hn_ipm_solve(const hn_lp* lp, const hn_ipm_opts* opts, hn_ipm_result* res) {
  HN_CHECK(lp && res, HN_ERR_INVALID, "hn_ipm_solve: null argument");
  hn_ipm_opts o;
  if (opts)
    o = *opts;
  else
    hn_ipm_default_opts(&o);
  HN_CHECK(o.precision == HN_PREC_F64 || o.precision == HN_PREC_MIXED, HN_ERR_INVALID, "hn_ipm_solve: bad precision");
  HN_CHECK(o.solver >= HN_NE_AUTO && o.solver <= HN_NE_SPARSE, HN_ERR_INVALID, "hn_ipm_solve: bad solver");
  HN_CHECK(o.max_iterations >= 0, HN_ERR_INVALID, "hn_ipm_solve: negative max_iterations");
  HN_CHECK(o.pivot_rule == HN_PIVOT_CLAMP || o.pivot_rule == HN_PIVOT_SKIP, HN_ERR_INVALID,
           "hn_ipm_solve: bad pivot_rule");
  HN_CHECK(o.loop >= HN_LOOP_GRAPH && o.loop <= HN_LOOP_HOST, HN_ERR_INVALID, "hn_ipm_solve: bad loop");
  return hn::ipm_solve_impl(*lp, o, *res);
}

// The ratio test as a device operation, on the solver's own kernels.
extern "C" int hn_ratio_test_dev(const double* x, const double* dx, const double* z, const double* dz, uint64_t n,
                                 double* out, void* stream) {
  HN_CHECK(out && (n == 0 || (x && dx && z && dz)), HN_ERR_INVALID, "hn_ratio_test_dev: null argument");
  HN_CHECK(n < (uint64_t)INT_MAX, HN_ERR_UNSUPPORTED, "hn_ratio_test_dev: n >= 2^31");
  HN_TRY(hn::enter());
  static std::mutex mtx;
  static hn::DevBuf* part = new hn::DevBuf();  // partial minima, kept until the process ends
  std::lock_guard<std::mutex> lk(mtx);
  HN_TRY(part->reserve(sizeof(double) * 2 * hn::kMaxParts));
  hipStream_t s = hn::pick_stream(stream);
  const int np = hn::parts_for((int64_t)n);
  hipLaunchKernelGGL(hn::k_red_step, dim3(np), dim3(hn::kNT), 0, s, x, dx, z, dz, (int)n, part->as<double>());
  hipLaunchKernelGGL(hn::k_fin_step, dim3(1), dim3(hn::kNT), 0, s, (const double*)part->as<double>(), np, out, 0, 1);
  hn::count_launch(2);
  return hn::check_launch("ratio test");
}
