// Normal equations N = M D M' + r I of the IPM Newton step, on the device.
//
// Reference: build_normal_equations (core/lp/interior_point.cpp:144-200)
// forms every product m_ij d_j m_kj of the columns with D_j > 0, merges equal
// (i,k) with from_triplets, and adds r = max(reg, 1e-12 max diag) to the
// diagonal; SparseCholesky factors it (factorization.cpp:1248-1325, pivots
// <= 1e-10 clamped to 1e-10) and iterative_refinement (refinement.cpp:7-36)
// polishes each solve.
//
// Here:
//   * symbolic  - the pattern (union of column cliques plus the diagonal) is
//     built once per matrix: candidates gathered per row through the
//     pre-built transpose, segment-sorted, de-duplicated.  Integer work,
//     bit-exact with the reference's pattern;
//   * numeric   - one thread per row accumulates its row in column order
//     (deterministic, no atomics), then r is added on the diagonal;
//   * dense     - for moderate m the matrix is scattered into a dense
//     row-major buffer and inverted in place by the blocked symmetric sweep
//     operator (Gauss-Jordan on SPD blocks): each step inverts an NB x NB
//     pivot block in LDS (its scalar pivots are the LDL' pivots in natural
//     order, clamped like the reference) and applies a rank-NB update to the
//     whole matrix with LDS-staged 64x64 tiles.  Solves are then one dense
//     GEMV; FP64 refinement against the sparse N corrects them;
//   * PCG       - matrix-free products M (D (M' p)) + r p and the Jacobi
//     diagonal sum_j m_ij^2 D_j + r, for LPs too large to densify.
#include "hn_device.hpp"
#include "hn_internal.hpp"
#include "hn_ipm_internal.hpp"

#include <algorithm>
#include <climits>

namespace hn {

// ---------------------------------------------------------------------------
// symbolic pattern
// ---------------------------------------------------------------------------
__global__ void k_ne_cand_count(const int* __restrict__ rp, const int* __restrict__ ci,
                                const int* __restrict__ trp, int m, int* __restrict__ cnt,
                                unsigned long long* __restrict__ total) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  long long s = 1;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int j = ci[k];
    s += trp[j + 1] - trp[j];
  }
  cnt[i] = (int)min(s, (long long)INT_MAX);
  atomicAdd(total, (unsigned long long)s);
}

__global__ void k_ne_cand_fill(const int* __restrict__ rp, const int* __restrict__ ci,
                               const int* __restrict__ trp, const int* __restrict__ tci, int m,
                               const int* __restrict__ cptr, int* __restrict__ cand) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int w = cptr[i];
  cand[w++] = i;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int j = ci[k];
    for (int q = trp[j]; q < trp[j + 1]; ++q) cand[w++] = tci[q];
  }
}

__global__ void k_ne_unique_count(const int* __restrict__ cptr, const int* __restrict__ cand, int m,
                                  int* __restrict__ ucnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int u = 0, prev = -1;
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int v = cand[k];
    u += (v != prev);
    prev = v;
  }
  ucnt[i] = u;
}

__global__ void k_ne_unique_fill(const int* __restrict__ cptr, const int* __restrict__ cand, int m,
                                 const int* __restrict__ nrp, int* __restrict__ ncol,
                                 int* __restrict__ diag_pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  int w = nrp[i], prev = -1;
  for (int k = cptr[i]; k < cptr[i + 1]; ++k) {
    const int v = cand[k];
    if (v != prev) {
      if (v == i) diag_pos[i] = w;
      ncol[w++] = v;
      prev = v;
    }
  }
}

int ne_symbolic(const hn_csr* M, const hn_csr* Mt, NormalEq& ne, hipStream_t s) {
  const int m = (int)M->nrows;
  DevBuf cnt, cptr, cand, ucnt, scan_tmp, total;
  SortScratch sc;
  HN_TRY(cnt.reserve(sizeof(int) * std::max(m, 1)));
  HN_TRY(cptr.reserve(sizeof(int) * (m + 1)));
  HN_TRY(total.reserve(sizeof(unsigned long long)));
  HN_HIP(hipMemsetAsync(total.p, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_ne_cand_count, dim3(grid_for(m, 256)), dim3(256), 0, s, M->row_ptr.as<int>(),
                     M->col.as<int>(), Mt->row_ptr.as<int>(), m, cnt.as<int>(), total.as<unsigned long long>());
  HN_TRY(check_launch("k_ne_cand_count"));
  unsigned long long tot = 0;
  HN_HIP(hipMemcpyAsync(&tot, total.p, sizeof(tot), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  HN_CHECK(tot < (unsigned long long)INT_MAX, HN_ERR_UNSUPPORTED,
           "normal equations: more than 2^31 candidate entries; use the PCG solver");
  HN_TRY(exclusive_scan(cnt.as<int>(), m, cptr.as<int>(), scan_tmp, s));
  HN_TRY(cand.reserve(sizeof(int) * std::max<unsigned long long>(tot, 1)));
  hipLaunchKernelGGL(k_ne_cand_fill, dim3(grid_for(m, 256)), dim3(256), 0, s, M->row_ptr.as<int>(),
                     M->col.as<int>(), Mt->row_ptr.as<int>(), Mt->col.as<int>(), m, cptr.as<int>(), cand.as<int>());
  HN_TRY(check_launch("k_ne_cand_fill"));
  HN_TRY(segmented_sort(cptr.as<int>(), m, cand.as<int>(), (int64_t)tot, sc, s));
  HN_TRY(ucnt.reserve(sizeof(int) * std::max(m, 1)));
  hipLaunchKernelGGL(k_ne_unique_count, dim3(grid_for(m, 256)), dim3(256), 0, s, cptr.as<int>(), cand.as<int>(), m,
                     ucnt.as<int>());
  HN_TRY(check_launch("k_ne_unique_count"));
  hn_csr* N = new hn_csr();
  ne.N = N;
  N->nrows = m;
  N->ncols = m;
  N->flags = HN_F64;
  HN_TRY(N->row_ptr.reserve(sizeof(int) * (m + 1)));
  HN_TRY(exclusive_scan(ucnt.as<int>(), m, N->row_ptr.as<int>(), scan_tmp, s));
  int nnz = 0;
  HN_HIP(hipMemcpyAsync(&nnz, N->row_ptr.as<int>() + m, sizeof(int), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  N->nnz = nnz;
  HN_TRY(N->col.reserve(sizeof(int) * std::max(nnz, 1)));
  HN_TRY(N->val64.reserve(sizeof(double) * std::max(nnz, 1)));
  HN_TRY(ne.diag_pos.reserve(sizeof(int) * std::max(m, 1)));
  hipLaunchKernelGGL(k_ne_unique_fill, dim3(grid_for(m, 256)), dim3(256), 0, s, cptr.as<int>(), cand.as<int>(), m,
                     N->row_ptr.as<int>(), N->col.as<int>(), ne.diag_pos.as<int>());
  HN_TRY(check_launch("k_ne_unique_fill"));
  HN_TRY(build_tiling(N, 8, s));
  HN_HIP(hipStreamSynchronize(s));
  count_launch(6);
  return HN_OK;
}

// ---------------------------------------------------------------------------
// numeric values
// ---------------------------------------------------------------------------
__global__ void k_ne_numeric(const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                             const int* __restrict__ trp, const int* __restrict__ tci, const double* __restrict__ tv,
                             const double* __restrict__ D, int m, const int* __restrict__ nrp,
                             const int* __restrict__ ncol, double* __restrict__ nval) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int b = nrp[i], e = nrp[i + 1];
  for (int k = b; k < e; ++k) nval[k] = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const int j = ci[k];
    const double d = D[j];
    if (d <= 0.0) continue;  // interior_point.cpp:178
    const double a = v[k];
    for (int q = trp[j]; q < trp[j + 1]; ++q) {
      const int r = tci[q];
      int lo = b, hi = e;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ncol[mid] < r) lo = mid + 1; else hi = mid;
      }
      nval[lo] += (a * tv[q]) * d;
    }
  }
}

// Per-block maximum of the diagonal, then r = max(reg, 1e-12 * max) added.
__global__ __launch_bounds__(256) void k_ne_diag_max(const double* __restrict__ nval,
                                                     const int* __restrict__ diag_pos, int m,
                                                     double* __restrict__ part) {
  __shared__ double red[4];
  double mx = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256)
    mx = nan_ignoring_max(mx, nval[diag_pos[i]]);
  mx = block_max<double, 256>(mx, red);
  if (threadIdx.x == 0) part[blockIdx.x] = mx;
}

__global__ __launch_bounds__(256) void k_ne_add_reg(const double* __restrict__ part, int np, double reg,
                                                    double* __restrict__ nval, const int* __restrict__ diag_pos,
                                                    int m, double* __restrict__ reg_out) {
  // every block recomputes the (tiny) final maximum so no extra launch is needed
  __shared__ double red[4];
  double mx = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) mx = nan_ignoring_max(mx, part[i]);
  mx = block_max<double, 256>(mx, red);
  const double r = fmax(reg, 1e-12 * mx);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) nval[diag_pos[i]] += r;
  if (blockIdx.x == 0 && threadIdx.x == 0 && reg_out) *reg_out = r;
}

int ne_numeric(const hn_csr* M, const hn_csr* Mt, const double* D, double reg, NormalEq& ne, double* reg_out,
               hipStream_t s) {
  const int m = (int)M->nrows;
  hn_csr* N = ne.N;
  hipLaunchKernelGGL(k_ne_numeric, dim3(grid_for(m, 128)), dim3(128), 0, s, M->row_ptr.as<int>(), M->col.as<int>(),
                     M->val64.as<double>(), Mt->row_ptr.as<int>(), Mt->col.as<int>(), Mt->val64.as<double>(), D, m,
                     N->row_ptr.as<int>(), N->col.as<int>(), N->val64.as<double>());
  HN_TRY(check_launch("k_ne_numeric"));
  const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (m + 255) / 256));
  HN_TRY(ne.part.reserve(sizeof(double) * 256));
  hipLaunchKernelGGL(k_ne_diag_max, dim3(nb), dim3(256), 0, s, N->val64.as<double>(), ne.diag_pos.as<int>(), m,
                     ne.part.as<double>());
  hipLaunchKernelGGL(k_ne_add_reg, dim3(nb), dim3(256), 0, s, ne.part.as<double>(), nb, reg, N->val64.as<double>(),
                     ne.diag_pos.as<int>(), m, reg_out);
  count_launch(3);
  return check_launch("k_ne_add_reg");
}

// ---------------------------------------------------------------------------
// dense inverse by blocked symmetric sweeps
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_dense_fill(T* __restrict__ A, long long total) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x)
    A[i] = T(0);
}

template <typename T>
__global__ void k_dense_scatter(const int* __restrict__ nrp, const int* __restrict__ ncol,
                                const double* __restrict__ nval, int m, T* __restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  for (int k = nrp[i]; k < nrp[i + 1]; ++k) A[(long long)i * m + ncol[k]] = (T)nval[k];
}

constexpr int kGjNB = 64;
constexpr int kGjTile = 64;

// Inverts the pivot block A[k0:k0+b, k0:k0+b] by b scalar sweeps in LDS.
// Pivots <= tol are replaced by tol (the reference's LDL' clamp); the count
// of clamped pivots is added to *clamped.  Writes Pinv (NB x NB, row-major).
template <typename T>
__global__ __launch_bounds__(256) void k_gj_pivot(const T* __restrict__ A, int m, int k0, int b,
                                                  T* __restrict__ Pinv, double tol, int* __restrict__ clamped) {
  constexpr int NB = kGjNB;
  __shared__ T P[NB][NB + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += 256) {
    const int i = e / NB, j = e % NB;
    P[i][j] = (i < b && j < b) ? A[(long long)(k0 + i) * m + (k0 + j)] : T(i == j ? 1 : 0);
  }
  __syncthreads();
  constexpr int PER = NB * NB / 256;
  T nv[PER];
  for (int k = 0; k < b; ++k) {
    T piv = P[k][k];
    if (piv <= T(tol)) {
      piv = T(tol);
      if (tid == 0) atomicAdd(clamped, 1);
    }
    const T rinv = T(1) / piv;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * 256;
      const int i = e / NB, j = e % NB;
      T val;
      if (i == k && j == k)
        val = -rinv;
      else if (i == k)
        val = P[k][j] * rinv;
      else if (j == k)
        val = P[i][k] * rinv;
      else
        val = P[i][j] - P[i][k] * (P[k][j] * rinv);
      nv[u] = val;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * 256;
      P[e / NB][e % NB] = nv[u];
    }
    __syncthreads();
  }
  for (int e = tid; e < NB * NB; e += 256) Pinv[e] = -P[e / NB][e % NB];
}

// Vt[k][i] = A[i][k0+k]; Wt[c][i] = sum_k Vt[k][i] Pinv[k][c]  (k-major, m columns)
template <typename T>
__global__ __launch_bounds__(256) void k_gj_panel(const T* __restrict__ A, int m, int k0, int b,
                                                  const T* __restrict__ Pinv, T* __restrict__ Vt,
                                                  T* __restrict__ Wt) {
  constexpr int NB = kGjNB;
  __shared__ T Ps[NB][NB];
  for (int e = threadIdx.x; e < NB * NB; e += 256) Ps[e / NB][e % NB] = Pinv[e];
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  T v[NB];
  const T* row = A + (long long)i * m + k0;
#pragma unroll
  for (int k = 0; k < NB; ++k) v[k] = (k < b) ? row[k] : T(0);
#pragma unroll
  for (int k = 0; k < NB; ++k)
    if (k < b) Vt[(long long)k * m + i] = v[k];
  for (int c = 0; c < b; ++c) {
    T acc = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) acc += v[k] * Ps[k][c];
    Wt[(long long)c * m + i] = acc;
  }
}

// A -= W V' outside the pivot rows/columns; pivot rows/columns get W, the
// pivot block gets -Pinv.  64 x 64 output tile per workgroup, 4 x 4 per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_update(T* __restrict__ A, int m, int k0, int b,
                                                   const T* __restrict__ Vt, const T* __restrict__ Wt,
                                                   const T* __restrict__ Pinv) {
  constexpr int NB = kGjNB, TL = kGjTile;
  __shared__ T Ws[NB][TL];
  __shared__ T Vs[NB][TL];
  const int i0 = blockIdx.y * TL, j0 = blockIdx.x * TL;
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * TL; e += 256) {
    const int k = e / TL, r = e % TL;
    const bool okk = k < b;
    Ws[k][r] = (okk && i0 + r < m) ? Wt[(long long)k * m + i0 + r] : T(0);
    Vs[k][r] = (okk && j0 + r < m) ? Vt[(long long)k * m + j0 + r] : T(0);
  }
  __syncthreads();
  const int tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0;
  for (int k = 0; k < b; ++k) {
    T wv[4], vv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) wv[a] = Ws[k][ty + 16 * a];
#pragma unroll
    for (int c = 0; c < 4; ++c) vv[c] = Vs[k][tx + 16 * c];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][c] += wv[a] * vv[c];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= m) continue;
    const bool ik = (i >= k0 && i < k0 + b);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + tx + 16 * c;
      if (j >= m) continue;
      const bool jk = (j >= k0 && j < k0 + b);
      T* dst = A + (long long)i * m + j;
      if (!ik && !jk)
        *dst -= acc[a][c];
      else if (!ik)
        *dst = Wt[(long long)(j - k0) * m + i];
      else if (!jk)
        *dst = Wt[(long long)(i - k0) * m + j];
      else
        *dst = -Pinv[(i - k0) * NB + (j - k0)];
    }
  }
}

template <typename T>
static int dense_invert_typed(DenseNE& d, int m, double tol, hipStream_t s) {
  T* A = d.A.as<T>();
  T* Pinv = d.pinv.as<T>();
  T* Vt = d.vt.as<T>();
  T* Wt = d.wt.as<T>();
  const dim3 g2((m + kGjTile - 1) / kGjTile, (m + kGjTile - 1) / kGjTile);
  for (int k0 = 0; k0 < m; k0 += kGjNB) {
    const int b = std::min(kGjNB, m - k0);
    hipLaunchKernelGGL((k_gj_pivot<T>), dim3(1), dim3(256), 0, s, A, m, k0, b, Pinv, tol, d.clamped.as<int>());
    hipLaunchKernelGGL((k_gj_panel<T>), dim3(grid_for(m, 256)), dim3(256), 0, s, A, m, k0, b, Pinv, Vt, Wt);
    hipLaunchKernelGGL((k_gj_update<T>), g2, dim3(256), 0, s, A, m, k0, b, Vt, Wt, Pinv);
  }
  count_launch(3 * ((m + kGjNB - 1) / kGjNB));
  return check_launch("dense invert");
}

int dense_reserve(DenseNE& d, int m, bool f32) {
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  HN_TRY(d.A.reserve(es * (size_t)m * m));
  HN_TRY(d.pinv.reserve(es * kGjNB * kGjNB));
  HN_TRY(d.vt.reserve(es * (size_t)kGjNB * m));
  HN_TRY(d.wt.reserve(es * (size_t)kGjNB * m));
  HN_TRY(d.clamped.reserve(sizeof(int)));
  d.m = m;
  return HN_OK;
}

// Dense inverse of the assembled N in the chosen precision: afterwards
// d.A holds -N^{-1}.
int dense_factor(const NormalEq& ne, DenseNE& d, bool f32, hipStream_t s) {
  const int m = (int)ne.N->nrows;
  HN_TRY(dense_reserve(d, m, f32));
  d.f32 = f32;
  const long long tot = (long long)m * m;
  const double tol = 1e-10;  // singular_tol of ToleranceConfig::industrial_defaults
  HN_HIP(hipMemsetAsync(d.clamped.p, 0, sizeof(int), s));
  if (f32) {
    hipLaunchKernelGGL((k_dense_fill<float>), dim3(grid_for(tot, 256, 16384)), dim3(256), 0, s, d.A.as<float>(), tot);
    hipLaunchKernelGGL((k_dense_scatter<float>), dim3(grid_for(m, 256)), dim3(256), 0, s, ne.N->row_ptr.as<int>(),
                       ne.N->col.as<int>(), ne.N->val64.as<double>(), m, d.A.as<float>());
    count_launch(2);
    return dense_invert_typed<float>(d, m, tol, s);
  }
  hipLaunchKernelGGL((k_dense_fill<double>), dim3(grid_for(tot, 256, 16384)), dim3(256), 0, s, d.A.as<double>(), tot);
  hipLaunchKernelGGL((k_dense_scatter<double>), dim3(grid_for(m, 256)), dim3(256), 0, s, ne.N->row_ptr.as<int>(),
                     ne.N->col.as<int>(), ne.N->val64.as<double>(), m, d.A.as<double>());
  count_launch(2);
  return dense_invert_typed<double>(d, m, tol, s);
}

// y = -A r  (A holds -N^{-1}); one wave per row, FP64 accumulation.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_gemv(const T* __restrict__ A, int m, const double* __restrict__ r,
                                                    double* __restrict__ y, int accumulate) {
  const int row = (blockIdx.x * 256 + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= m) return;
  const T* a = A + (long long)row * m;
  double s = 0.0;
  for (int j = lane; j < m; j += kWave) s += (double)a[j] * r[j];
  s = wave_sum(s);
  if (lane == 0) y[row] = accumulate ? y[row] - s : -s;
}

int dense_apply(const DenseNE& d, const double* r, double* y, bool accumulate, hipStream_t s) {
  const int m = d.m;
  const long long threads = (long long)m * kWave;
  const dim3 g((unsigned)((threads + 255) / 256));
  if (d.f32)
    hipLaunchKernelGGL((k_dense_gemv<float>), g, dim3(256), 0, s, d.A.as<float>(), m, r, y, (int)accumulate);
  else
    hipLaunchKernelGGL((k_dense_gemv<double>), g, dim3(256), 0, s, d.A.as<double>(), m, r, y, (int)accumulate);
  count_launch();
  return check_launch("k_dense_gemv");
}

// ---------------------------------------------------------------------------
// matrix-free operator pieces for PCG
// ---------------------------------------------------------------------------
// diag_i = sum_j m_ij^2 D_j (columns with D_j > 0), the diagonal of M D M'.
__global__ void k_ne_jacobi_diag(const int* __restrict__ rp, const double* __restrict__ v,
                                 const int* __restrict__ ci, const double* __restrict__ D, int m,
                                 double* __restrict__ diag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; ++k) {
    const double d = D[ci[k]];
    if (d > 0.0) s += (v[k] * v[k]) * d;
  }
  diag[i] = s;
}

int ne_jacobi_diag(const hn_csr* M, const double* D, double* diag, hipStream_t s) {
  const int m = (int)M->nrows;
  hipLaunchKernelGGL(k_ne_jacobi_diag, dim3(grid_for(m, 256)), dim3(256), 0, s, M->row_ptr.as<int>(),
                     M->val64.as<double>(), M->col.as<int>(), D, m, diag);
  count_launch();
  return check_launch("k_ne_jacobi_diag");
}

}  // namespace hn
