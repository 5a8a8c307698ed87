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
// One wave per row i.  Lane l owns pattern entries b+l, b+l+64, ... of the
// row; every product m_ij d_j m_rj (columns j of row i in order, rows r of
// column j in order) is broadcast across the wave and added by the lane that
// owns column r.  Each entry is therefore summed in the same order as a
// sequential loop, without atomics.  The row is also written into the dense
// matrix when one is given (entries outside the pattern must be zero).
template <typename T>
__global__ __launch_bounds__(256) void k_ne_numeric(
    const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
    const int* __restrict__ trp, const int* __restrict__ tci, const double* __restrict__ tv,
    const double* __restrict__ D, int m, const int* __restrict__ nrp, const int* __restrict__ ncol,
    double* __restrict__ nval, T* __restrict__ dense) {
  constexpr int CH = 8;  // rows of up to 512 pattern entries in registers; longer rows loop
  const int i = (blockIdx.x * 256 + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= m) return;
  const int b = nrp[i], e = nrp[i + 1];
  for (int base = b; base < e; base += CH * kWave) {
    int mycol[CH];
    double acc[CH];
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const int p = base + t * kWave + lane;
      mycol[t] = (p < e) ? ncol[p] : -1;
      acc[t] = 0.0;
    }
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = ci[k];
      const double d = D[j];
      if (d <= 0.0) continue;  // interior_point.cpp:178
      const double a = v[k];
      const int q0 = trp[j], q1 = trp[j + 1];
      for (int qb = q0; qb < q1; qb += kWave) {
        int r = -1;
        double val = 0.0;
        if (qb + lane < q1) {
          r = tci[qb + lane];
          val = (a * tv[qb + lane]) * d;
        }
        const int cnt = min(kWave, q1 - qb);
        for (int l = 0; l < cnt; ++l) {
          const int rr = __shfl(r, l, kWave);
          const double vv = __shfl(val, l, kWave);
#pragma unroll
          for (int t = 0; t < CH; ++t)
            if (mycol[t] == rr) acc[t] += vv;
        }
      }
    }
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      const int p = base + t * kWave + lane;
      if (p < e) {
        nval[p] = acc[t];
        if (dense) dense[(long long)i * m + mycol[t]] = (T)acc[t];
      }
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
                                                    int m, double* __restrict__ reg_out, double* __restrict__ d64,
                                                    float* __restrict__ d32) {
  // every block recomputes the (tiny) final maximum so no extra launch is needed
  __shared__ double red[4];
  double mx = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) mx = nan_ignoring_max(mx, part[i]);
  mx = block_max<double, 256>(mx, red);
  const double r = fmax(reg, 1e-12 * mx);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
    const double dv = nval[diag_pos[i]] + r;
    nval[diag_pos[i]] = dv;
    if (d64) d64[(long long)i * m + i] = dv;
    if (d32) d32[(long long)i * m + i] = (float)dv;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && reg_out) *reg_out = r;
}

template <typename T>
__global__ void k_dense_fill(T* __restrict__ A, long long total) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x)
    A[i] = T(0);
}

int ne_numeric(const hn_csr* M, const hn_csr* Mt, const double* D, double reg, NormalEq& ne, double* reg_out,
               DenseNE* dense, bool f32, hipStream_t s) {
  const int m = (int)M->nrows;
  hn_csr* N = ne.N;
  const dim3 gw((unsigned)(((long long)m * kWave + 255) / 256));
  double* d64 = nullptr;
  float* d32 = nullptr;
  if (dense) {
    HN_TRY(dense_reserve(*dense, m, f32));
    dense->f32 = f32;
    const long long tot = (long long)m * m;
    if (f32) {
      d32 = dense->A.as<float>();
      hipLaunchKernelGGL((k_dense_fill<float>), dim3(grid_for(tot, 256, 16384)), dim3(256), 0, s, d32, tot);
    } else {
      d64 = dense->A.as<double>();
      hipLaunchKernelGGL((k_dense_fill<double>), dim3(grid_for(tot, 256, 16384)), dim3(256), 0, s, d64, tot);
    }
  }
  if (f32 && dense)
    hipLaunchKernelGGL((k_ne_numeric<float>), gw, dim3(256), 0, s, M->row_ptr.as<int>(), M->col.as<int>(),
                       M->val64.as<double>(), Mt->row_ptr.as<int>(), Mt->col.as<int>(), Mt->val64.as<double>(), D, m,
                       N->row_ptr.as<int>(), N->col.as<int>(), N->val64.as<double>(), d32);
  else
    hipLaunchKernelGGL((k_ne_numeric<double>), gw, dim3(256), 0, s, M->row_ptr.as<int>(), M->col.as<int>(),
                       M->val64.as<double>(), Mt->row_ptr.as<int>(), Mt->col.as<int>(), Mt->val64.as<double>(), D, m,
                       N->row_ptr.as<int>(), N->col.as<int>(), N->val64.as<double>(), d64);
  HN_TRY(check_launch("k_ne_numeric"));
  const int nb = (int)std::min<int64_t>(256, std::max<int64_t>(1, (m + 255) / 256));
  HN_TRY(ne.part.reserve(sizeof(double) * 256));
  hipLaunchKernelGGL(k_ne_diag_max, dim3(nb), dim3(256), 0, s, N->val64.as<double>(), ne.diag_pos.as<int>(), m,
                     ne.part.as<double>());
  hipLaunchKernelGGL(k_ne_add_reg, dim3(nb), dim3(256), 0, s, ne.part.as<double>(), nb, reg, N->val64.as<double>(),
                     ne.diag_pos.as<int>(), m, reg_out, d64, d32);
  count_launch(dense ? 4 : 3);
  return check_launch("k_ne_add_reg");
}

// ---------------------------------------------------------------------------
// dense inverse by blocked symmetric sweeps
// ---------------------------------------------------------------------------
constexpr int kGjNB = 64;
constexpr int kGjTile = 64;

// Symmetric sweep inverse of a 64 x 64 block held in registers by a
// 256-thread workgroup: thread (ti, tj) = (tid / 16, tid % 16) owns rows
// 4ti..4ti+3 and columns 4tj..4tj+3.  Step k publishes row k and column k in
// double-buffered LDS (one barrier per step), then every thread updates its
// 16 entries:  a_kk <- -1/p,  a_kj <- a_kj/p,  a_ik <- a_ik/p,
// a_ij <- a_ij - a_ik a_kj / p.  Pivots p <= tol (and NaN-free non-positive
// ones) are replaced by tol, the reference's LDL' clamp; steps k >= b leave
// the block untouched.  Returns the number of clamped pivots (thread 0).
template <typename T>
__device__ int sweep_block64(T (&P)[4][4], int b, double tol, T (*rk)[64], T (*ck)[64]) {
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  int nclamp = 0;
  for (int k = 0; k < b; ++k) {
    const int par = k & 1;
    const int kk = k & 3;
    if (ti == (k >> 2)) {
#pragma unroll
      for (int c = 0; c < 4; ++c) rk[par][4 * tj + c] = P[kk][c];
    }
    if (tj == (k >> 2)) {
#pragma unroll
      for (int a = 0; a < 4; ++a) ck[par][4 * ti + a] = P[a][kk];
    }
    __syncthreads();
    T piv = rk[par][k];
    if (piv <= T(tol)) {
      piv = T(tol);
      ++nclamp;
    }
    const T rinv = T(1) / piv;
    T rkj[4], cik[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) rkj[c] = rk[par][4 * tj + c] * rinv;
#pragma unroll
    for (int a = 0; a < 4; ++a) cik[a] = ck[par][4 * ti + a];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = 4 * ti + a;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = 4 * tj + c;
        T v;
        if (i == k && j == k)
          v = -rinv;
        else if (i == k)
          v = rkj[c];
        else if (j == k)
          v = cik[a] * rinv;
        else
          v = P[a][c] - cik[a] * rkj[c];
        P[a][c] = v;
      }
    }
  }
  return nclamp;
}

// Inverts the pivot block A[k0:k0+b, k0:k0+b] (identity beyond b) and writes
// Pinv (64 x 64, row-major); clamped pivots are added to *clamped.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_pivot(const T* __restrict__ A, int m, int k0, int b,
                                                  T* __restrict__ Pinv, double tol, int* __restrict__ clamped) {
  __shared__ T rk[2][64];
  __shared__ T ck[2][64];
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  T P[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * ti + a, j = 4 * tj + c;
      P[a][c] = (i < b && j < b) ? A[(long long)(k0 + i) * m + k0 + j] : T(i == j ? 1 : 0);
    }
  const int nclamp = sweep_block64<T>(P, b, tol, rk, ck);
  if (tid == 0 && nclamp) atomicAdd(clamped, nclamp);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) Pinv[(4 * ti + a) * 64 + 4 * tj + c] = -P[a][c];
}

// Vt[k][i] = A[i][k0+k]; Wt[c][i] = sum_k Vt[k][i] Pinv[k][c]  (k-major, m
// columns).  64 rows per workgroup: the rows' pivot columns and Pinv are
// staged in LDS, each thread produces 16 outputs.
template <typename T>
__global__ __launch_bounds__(256) void k_gj_panel(const T* __restrict__ A, int m, int k0, int b,
                                                  const T* __restrict__ Pinv, T* __restrict__ Vt,
                                                  T* __restrict__ Wt) {
  constexpr int NB = kGjNB, TL = 64;
  __shared__ T Ps[NB][NB + 1];
  __shared__ T Vs[TL][NB + 1];
  const int i0 = blockIdx.x * TL;
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += 256) Ps[e / NB][e % NB] = Pinv[e];
  for (int e = tid; e < TL * NB; e += 256) {
    const int r = e / NB, k = e % NB;
    const int i = i0 + r;
    const T val = (i < m && k < b) ? A[(long long)i * m + k0 + k] : T(0);
    Vs[r][k] = val;
  }
  __syncthreads();
  for (int e = tid; e < TL * NB; e += 256) {
    const int k = e / TL, r = e % TL;
    if (i0 + r < m && k < b) Vt[(long long)k * m + i0 + r] = Vs[r][k];
  }
  // outputs: column c = tid / 4 (0..63), rows r = (tid % 4) + 4 t
  const int c = tid >> 2;
#pragma unroll 4
  for (int t = 0; t < TL / 4; ++t) {
    const int r = (tid & 3) + 4 * t;
    T acc = 0;
#pragma unroll
    for (int k = 0; k < NB; ++k) acc += Vs[r][k] * Ps[k][c];
    if (i0 + r < m && c < b) Wt[(long long)c * m + i0 + r] = acc;
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
    hipLaunchKernelGGL((k_gj_panel<T>), dim3((m + 63) / 64), dim3(256), 0, s, A, m, k0, b, Pinv, Vt, Wt);
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

// In-place inverse of the dense N written by ne_numeric: afterwards d.A
// holds -N^{-1}.
int dense_factor(DenseNE& d, hipStream_t s) {
  const double tol = 1e-10;  // singular_tol of ToleranceConfig::industrial_defaults
  HN_HIP(hipMemsetAsync(d.clamped.p, 0, sizeof(int), s));
  return d.f32 ? dense_invert_typed<float>(d, d.m, tol, s) : dense_invert_typed<double>(d, d.m, tol, s);
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
