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
//     row-major buffer and factored N = L D L' by 64-wide blocks (pivots in
//     natural order, clamped like the reference or skipped), forming
//     G = L^{-1} alongside.  A solve applies G, D^{-1} and G' (two
//     triangular matrix-vector passes, the bytes of one dense GEMV); FP64
//     refinement against the sparse N corrects it;
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

// Per-block maximum of the diagonal, then r = max(reg, rel_reg * max) added
// (rel_reg = 1e-12 under the reference's rule, 0 under the skip rule).
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
                                                    double rel_reg, double* __restrict__ nval,
                                                    const int* __restrict__ diag_pos, int m,
                                                    double* __restrict__ reg_out, double* __restrict__ d64,
                                                    float* __restrict__ d32, double* __restrict__ diag0) {
  // every block recomputes the (tiny) final maximum so no extra launch is needed
  __shared__ double red[4];
  double mx = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) mx = nan_ignoring_max(mx, part[i]);
  mx = block_max<double, 256>(mx, red);
  const double r = fmax(reg, rel_reg * mx);
  for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
    const double dv = nval[diag_pos[i]] + r;
    nval[diag_pos[i]] = dv;
    if (diag0) diag0[i] = dv;
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

int ne_numeric(const hn_csr* M, const hn_csr* Mt, const double* D, double reg, double rel_reg, NormalEq& ne,
               double* reg_out, DenseNE* dense, bool f32, hipStream_t s) {
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
  hipLaunchKernelGGL(k_ne_add_reg, dim3(nb), dim3(256), 0, s, ne.part.as<double>(), nb, reg, rel_reg,
                     N->val64.as<double>(), ne.diag_pos.as<int>(), m, reg_out, d64, d32,
                     dense ? dense->diag0.as<double>() : nullptr);
  count_launch(dense ? 4 : 3);
  return check_launch("k_ne_add_reg");
}

// ---------------------------------------------------------------------------
// dense N^{-1} through a blocked LDL' factorisation
// ---------------------------------------------------------------------------
// N = L D L' (right-looking, 64-wide blocks, natural pivot order - the pivots
// are the reference's LDL' pivots and are clamped or skipped like them), and
// G = L^{-1} (block rows from the inverted diagonal blocks), so that
// N^{-1} = G' D^{-1} G.  L^{-1} formed from the triangular factor is as
// accurate as triangular solves; a Gauss-Jordan sweep of N itself loses
// several more digits once N's condition passes ~1e10, which the skip rule
// reaches.
constexpr int kNB = 64;

// acc[a][c] += sum_k As[k][ty + 16 a] * Bs[k][tx + 16 c], k < kw.
template <typename T>
__device__ __forceinline__ void tile_acc(T (&acc)[4][4], const T (*As)[kNB], const T (*Bs)[kNB], int kw) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  for (int k = 0; k < kw; ++k) {
    T av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = As[k][ty + 16 * a];
#pragma unroll
    for (int c = 0; c < 4; ++c) bv[c] = Bs[k][tx + 16 * c];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[a][c] += av[a] * bv[c];
  }
}

template <typename T>
__device__ __forceinline__ void tile_zero(T (&acc)[4][4]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = T(0);
}

// Value of v in lane `lane` (a wave-uniform lane index), through readlane.
__device__ __forceinline__ double read_lane(double v, int lane) {
  const long long w = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)w, lane);
  const int hi = __builtin_amdgcn_readlane((int)(w >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ float read_lane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// (1) LDL' of the diagonal block A[k0:k0+b, k0:k0+b] by four waves (a
// 64 x 64 block, rows past b the identity).  Writes the strictly lower part
// of L, dinv[k0:k0+b] (1/pivot; 0 for a skipped pivot) and X = L_KK^{-1} as
// the full 64 x 64 diagonal block of G = L^{-1} (zeros above).  Pivots
// p <= tol are clamped to tol (skip_rel == 0) or, with skip_rel > 0,
// p <= skip_rel * diag0 is skipped.
//
// Right-looking: step k forms d_k, the multipliers l_j = A(k, j) / d_k of
// column k (the trailing block stays symmetric, so lane j's own entry of
// row k gives its multiplier) and updates a_ij -= l_i a_kj below row k.
// Wave w keeps rows 16w..16w+15 of the block in registers (lane j: column
// j).  The 16 steps of block kb change only rows >= 16kb, and wave kb's rows
// only through its own steps, so wave kb runs them alone: the next pivot
// and its reciprocal first (from values known before the step's
// reciprocal: lane k+1's multiplier and diagonal entry), then its other
// rows.  It then publishes the 16 multiplier vectors and pivot rows in LDS,
// and the waves below apply the 16 steps to their rows, each row in step
// order.  Every entry receives the same multiply-adds in the same order as
// in a one-step-at-a-time sweep, so the factor does not depend on the
// split.  The pivot steps form one dependent chain (pivot, reciprocal,
// multiplier, next pivot); the four-wave split keeps 48 rows' updates off
// it.
template <typename T>
__global__ __launch_bounds__(256) void k_ldl_diag(T* __restrict__ A, int m, int k0, int b, T* __restrict__ dinv,
                                                  T* __restrict__ G, double tol, const double* __restrict__ diag0,
                                                  double skip_rel, int* __restrict__ clamped) {
  __shared__ T S[kNB][kNB + 1];
  __shared__ T Xs[kNB][kNB + 1];
  __shared__ T Ts[3 * 256];
  __shared__ T Lk[2][16][kNB];  // multipliers of a block's 16 steps, by block parity
  __shared__ T Ak[2][16][kNB];  // the 16 pivot rows A(k, j), by block parity
  __shared__ T Dk[kNB];     // 1 / pivot of each step
  const int tid = threadIdx.x;
  const int j = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  {
    T v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int ii = min(16 * w + r, b - 1);
      const int jj = min(j, ii);
      v[r] = A[(long long)(k0 + ii) * m + k0 + jj];
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 16 * w + r;
      if (j <= i) {
        const T x = (i < b) ? v[r] : T(i == j ? 1 : 0);
        S[i][j] = x;
        S[j][i] = x;
      }
    }
  }
  __syncthreads();
  T c[16];  // c[r] = A(16 w + r, j)
#pragma unroll
  for (int r = 0; r < 16; ++r) c[r] = S[16 * w + r][j];
  const double d0j = (skip_rel > 0.0 && j < b) ? diag0[k0 + j] : 1.0;
  int ncl = 0;
  // 1 / pivot under the pivot rule, from the pivot's value (wave-uniform)
  auto pivot_inverse_of = [&](T piv, int k, bool live) -> T {
    const bool skipped = skip_rel > 0.0 && !((double)piv > skip_rel * read_lane(d0j, k));
    const bool clamped = !(skip_rel > 0.0) && piv <= T(tol);
    const T r = T(1) / (clamped ? T(tol) : piv);
    ncl += (live && (skipped || clamped)) ? 1 : 0;
    return skipped ? T(0) : r;
  };
  auto pivot_inverse = [&](T cv, int k, bool live) -> T { return pivot_inverse_of(read_lane(cv, k), k, live); };

  T rinv = T(0);
  if (w == 0) rinv = pivot_inverse(c[0], 0, true);
  __syncthreads();  // every row of S is in registers before S takes L
  for (int kb = 0; kb < 4; ++kb) {
    if (16 * kb >= b) break;
    const int buf = kb & 1;
    if (w == kb) {
#pragma unroll
      for (int kk = 0; kk < 16; ++kk) {
        const int k = 16 * kb + kk;
        if (k >= b) continue;  // (not break: the loop must unroll fully)
        const T akj = c[kk];  // A(k, j)
        Dk[k] = rinv;  // every lane stores the same value
        const T l = (j > k) ? akj * rinv : T(0);
        Lk[buf][kk][j] = l;
        Ak[buf][kk][j] = akj;
        const T a = (j > k) ? akj : T(0);
        if (kk < 15) {
          // the next pivot from values known before this step's reciprocal:
          // row k+1's multiplier A(k, k+1) / d_k and its diagonal entry
          const T ak1 = read_lane(akj, k + 1);
          const T lk1 = ak1 * rinv;  // lane k+1's l
          const T piv = read_lane(c[kk + 1], k + 1) - lk1 * ak1;
          c[kk + 1] = c[kk + 1] - lk1 * a;
          rinv = pivot_inverse_of(piv, k + 1, k + 1 < b);
        }
#pragma unroll
        for (int r = kk + 2; r < 16; ++r) c[r] = c[r] - Lk[buf][kk][16 * kb + r] * a;
      }
    }
    __syncthreads();
    const int ke = min(16, b - 16 * kb);
    // columns 16kb.. of L into S (row j of L), by all waves
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int kk = 4 * w + q;
      if (kk < ke) S[j][16 * kb + kk] = Lk[buf][kk][j];
    }
    if (w > kb) {
      // the block's steps on this wave's rows, each row in step order (a
      // full block unrolled, row 0 first: for wave kb+1 it is the next
      // pivot row)
      if (ke == 16) {
        T a[16];
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) a[kk] = (j > 16 * kb + kk) ? Ak[buf][kk][j] : T(0);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          T x = c[r];
#pragma unroll
          for (int kk = 0; kk < 16; ++kk) x = x - Lk[buf][kk][16 * w + r] * a[kk];
          c[r] = x;
        }
      } else {
        for (int kk = 0; kk < ke; ++kk) {
          const T a = (j > 16 * kb + kk) ? Ak[buf][kk][j] : T(0);
#pragma unroll
          for (int r = 0; r < 16; ++r) c[r] = c[r] - Lk[buf][kk][16 * w + r] * a;
        }
      }
      if (w == kb + 1) rinv = pivot_inverse(c[0], 16 * w, 16 * w < b);
    }
  }
  __syncthreads();
  if (tid < b) dinv[k0 + tid] = Dk[tid];
  for (int i = 1 + w; i < b; i += 4)  // L below the diagonal, row by row (coalesced)
    if (j < i) A[(long long)(k0 + i) * m + k0 + j] = S[i][j];
  if (j == 0 && ncl) atomicAdd(clamped, ncl);
  // X = L_KK^{-1} by 16 x 16 blocks: the four diagonal blocks by forward
  // substitution (a column per lane of wave 0), then block rows I = 1..3:
  // X_IJ = -X_II sum_{J<=K<I} L_IK X_KJ, the entries spread over the four
  // waves.  Only the strictly lower part of S (= L) is read.
  if (w == 0) {
    const int o = 16 * (j >> 4), cc = j & 15;
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T sx = (i == cc) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) sx -= S[o + i][o + k] * x[k];
      x[i] = sx;
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) Xs[o + i][o + cc] = x[i];
  }
  __syncthreads();
  const int r = (tid >> 4) & 15, cc = tid & 15;
#pragma unroll
  for (int I = 1; I < 4; ++I) {
#pragma unroll
    for (int J = 0; J < I; ++J) {
      T sx = T(0);
#pragma unroll
      for (int K = J; K < I; ++K)
#pragma unroll
        for (int k = 0; k < 16; ++k) sx += S[16 * I + r][16 * K + k] * Xs[16 * K + k][16 * J + cc];
      Ts[(J << 8) + tid] = sx;
    }
    __syncthreads();
#pragma unroll
    for (int J = 0; J < I; ++J) {
      T sx = T(0);
#pragma unroll
      for (int k = 0; k < 16; ++k) sx += Xs[16 * I + r][16 * I + k] * Ts[(J << 8) + (k << 4) + cc];
      Xs[16 * I + r][16 * J + cc] = -sx;
    }
    __syncthreads();
  }
  for (int i = w; i < b; i += 4)
    if (j < b) G[(long long)(k0 + i) * m + k0 + j] = (j <= i) ? Xs[i][j] : T(0);
}

// (2) panel, 64 rows r >= k0 + b per workgroup: W = A[r, K] X' (so that
// W L_KK' = A[r, K]) and L[r, K] = W dinv; W and L are also written k-major
// for the trailing update.  Workgroups past the panel finish block row K of
// G = L^{-1}: G[K, J] = X_K R[K, J] for J < K, where R[K, J] holds the
// accumulated -sum L[K, K'] G[K', J] of the earlier steps (G[K, K] = X_K
// already).  Both are 64 x 64 x 64 products, LDS-staged, 4 x 4 per thread.
template <typename T>
__global__ __launch_bounds__(256) void k_ldl_panel(T* __restrict__ A, int m, int k0, int b, int npanel,
                                                   const T* __restrict__ dinv, T* __restrict__ G,
                                                   T* __restrict__ Wt, T* __restrict__ Lt) {
  __shared__ T As[kNB][kNB];
  __shared__ T Bs[kNB][kNB];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
  tile_zero(acc);
  if ((int)blockIdx.x >= npanel) {
    const int j0 = ((int)blockIdx.x - npanel) * kNB;  // a full block left of K
    for (int e = tid; e < kNB * kNB; e += 256) {
      const int r = e / kNB, k = e % kNB;
      As[k][r] = (r < b && k < b) ? G[(long long)(k0 + r) * m + k0 + k] : T(0);  // X[r][k]
      Bs[r][k] = r < b ? G[(long long)(k0 + r) * m + j0 + k] : T(0);              // R[r][c], c = k
    }
    __syncthreads();
    tile_acc(acc, As, Bs, b);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int r = ty + 16 * a;
      if (r >= b) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) G[(long long)(k0 + r) * m + j0 + tx + 16 * q] = acc[a][q];
    }
    return;
  }
  const int i0 = k0 + b + blockIdx.x * kNB;
  for (int e = tid; e < kNB * kNB; e += 256) {
    const int r = e / kNB, k = e % kNB;
    As[k][r] = (i0 + r < m && k < b) ? A[(long long)(i0 + r) * m + k0 + k] : T(0);  // V[r][k]
    Bs[k][r] = (r < b && k < b) ? G[(long long)(k0 + r) * m + k0 + k] : T(0);       // X[c][k], c = r
  }
  __syncthreads();
  tile_acc(acc, As, Bs, b);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= m) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = tx + 16 * q;
      if (c >= b) continue;
      const T w = acc[a][q];
      const T l = w * dinv[k0 + c];
      A[(long long)i * m + k0 + c] = l;
      Wt[(long long)c * m + i] = w;
      Lt[(long long)c * m + i] = l;
    }
  }
}

// (3) trailing update.  Grid x: K + 1 G-columns then nt A-columns; grid y:
// the nt block rows below K.  A tiles (J >= I excluded):
// A[I, J] -= W[I, K] L[J, K]'.  G tiles (J <= K): R[I, J] = -L[I, K] X_K for
// J = K (first touch), R[I, J] -= L[I, K] G[K, J] for J < K.
template <typename T>
__global__ __launch_bounds__(256) void k_ldl_update(T* __restrict__ A, T* __restrict__ G, int m, int k0, int b,
                                                    int K, const T* __restrict__ Wt, const T* __restrict__ Lt) {
  __shared__ T Ws[kNB][kNB];
  __shared__ T Ls[kNB][kNB];
  const int r0 = k0 + b;
  const int tI = blockIdx.y;
  const int i0 = r0 + kNB * tI;
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
  tile_zero(acc);
  if ((int)blockIdx.x <= K) {  // G tile (I, J)
    const int J = blockIdx.x;
    const int j0 = J * kNB;
    for (int e = tid; e < kNB * kNB; e += 256) {
      const int k = e / kNB, q = e % kNB;
      Ws[k][q] = (k < b && i0 + q < m) ? Lt[(long long)k * m + i0 + q] : T(0);
      Ls[k][q] = k < b ? G[(long long)(k0 + k) * m + j0 + q] : T(0);
    }
    __syncthreads();
    tile_acc(acc, Ws, Ls, b);
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = i0 + ty + 16 * a;
      if (i >= m) continue;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        T* dst = G + (long long)i * m + j0 + tx + 16 * c;
        *dst = (J == K) ? -acc[a][c] : *dst - acc[a][c];
      }
    }
    return;
  }
  const int tJ = (int)blockIdx.x - (K + 1);
  if (tJ > tI) return;
  const int j0 = r0 + kNB * tJ;
  for (int e = tid; e < kNB * kNB; e += 256) {
    const int k = e / kNB, q = e % kNB;
    Ws[k][q] = (k < b && i0 + q < m) ? Wt[(long long)k * m + i0 + q] : T(0);
    Ls[k][q] = (k < b && j0 + q < m) ? Lt[(long long)k * m + j0 + q] : T(0);
  }
  __syncthreads();
  tile_acc(acc, Ws, Ls, b);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int i = i0 + ty + 16 * a;
    if (i >= m) continue;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + tx + 16 * c;
      if (j <= i && j < m) A[(long long)i * m + j] -= acc[a][c];
    }
  }
}

__global__ void k_zero_int(int* p) {
  if (threadIdx.x == 0) *p = 0;
}

template <typename T>
static int dense_invert_typed(DenseNE& d, int m, double tol, double skip_rel, hipStream_t s) {
  T* A = d.A.as<T>();
  T* G = d.G.as<T>();
  T* dinv = d.dinv.as<T>();
  T* Wt = d.wt.as<T>();
  T* Lt = d.lt.as<T>();
  const int nb = (m + kNB - 1) / kNB;
  for (int K = 0; K < nb; ++K) {
    const int k0 = K * kNB;
    const int b = std::min(kNB, m - k0);
    hipLaunchKernelGGL((k_ldl_diag<T>), dim3(1), dim3(256), 0, s, A, m, k0, b, dinv, G, tol, d.diag0.as<double>(),
                       skip_rel, d.clamped.as<int>());
    const int rest = m - k0 - b;
    const int nt = (rest + kNB - 1) / kNB;
    if (nt + K > 0)
      hipLaunchKernelGGL((k_ldl_panel<T>), dim3(nt + K), dim3(256), 0, s, A, m, k0, b, nt, dinv, G, Wt, Lt);
    if (nt > 0)
      hipLaunchKernelGGL((k_ldl_update<T>), dim3(K + 1 + nt, nt), dim3(256), 0, s, A, G, m, k0, b, K, Wt, Lt);
  }
  count_launch(3 * nb);
  return check_launch("dense factor");
}

int dense_reserve(DenseNE& d, int m, bool f32) {
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  HN_TRY(d.A.reserve(es * (size_t)m * m));
  HN_TRY(d.G.reserve(es * (size_t)m * m));
  HN_TRY(d.dinv.reserve(es * (size_t)m));
  HN_TRY(d.wt.reserve(es * (size_t)kNB * m));
  HN_TRY(d.lt.reserve(es * (size_t)kNB * m));
  HN_TRY(d.clamped.reserve(sizeof(int)));
  HN_TRY(d.diag0.reserve(sizeof(double) * (size_t)m));
  HN_TRY(d.u.reserve(sizeof(double) * (size_t)m));
  HN_TRY(d.part.reserve(sizeof(double) * (size_t)m * ((m + kNB - 1) / kNB)));
  d.m = m;
  return HN_OK;
}

// Factors the dense N written by ne_numeric: afterwards d.G holds L^{-1}
// and d.dinv the inverse pivots.
int dense_factor(DenseNE& d, int pivot_rule, hipStream_t s) {
  const double tol = 1e-10;  // singular_tol of ToleranceConfig::industrial_defaults
  // a kernel, not a memset: this factorisation is captured in graphs
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, d.clamped.as<int>());
  const bool skip = pivot_rule == HN_PIVOT_SKIP;
  return d.f32 ? dense_invert_typed<float>(d, d.m, tol, skip ? kSkipRel32 : 0.0, s)
               : dense_invert_typed<double>(d, d.m, tol, skip ? kSkipRel64 : 0.0, s);
}

// Applying N^{-1} = G' D^{-1} G: the same bytes as one GEMV with the
// explicit inverse (each half of G once), and a pivot of D that was clamped
// to 1e-10 scales its own component only, as in the reference's triangular
// solves (an explicit inverse would spread its 1e10 through every entry).
//
// Refinement guard: with irr/irb (||rhs - N x||_inf and ||rhs||_inf of the
// last check) given, a correction sweep does nothing once the solution has
// met the refinement test (refinement.cpp:7-36: rel <= 1e-10, or a zero or
// NaN right-hand side), so a fixed number of sweeps can be enqueued ahead.
__device__ __forceinline__ bool ne_refined(const double* irr, const double* irb) {
  if (!irr) return false;
  const double nb = *irb;
  return nb == 0.0 || nb != nb || *irr / nb <= 1e-10;
}

// (1) u = D^{-1} (G r): a wave per row k, over the columns up to the end of
// k's 64-row block (G is lower block-triangular; the upper part of each
// diagonal block is zero).
template <typename T>
__global__ __launch_bounds__(256) void k_gapply_lower(const T* __restrict__ G, const T* __restrict__ dinv, int m,
                                                      const double* __restrict__ r, double* __restrict__ u,
                                                      const double* irr, const double* irb) {
  if (ne_refined(irr, irb)) return;
  const int row = (blockIdx.x * 256 + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= m) return;
  const int len = min(m, (row / kNB + 1) * kNB);
  const T* g = G + (long long)row * m;
  double s = 0.0;
  for (int j = lane; j < len; j += kWave) s += (double)g[j] * r[j];
  s = wave_sum(s);
  if (lane == 0) u[row] = (double)dinv[row] * s;
}

// (2) part[K][i] = sum over the rows k of block K of G[k][i] u[k], for the
// columns i below the block's end (the rest of those rows is zero).
template <typename T>
__global__ __launch_bounds__(256) void k_gapply_upper_part(const T* __restrict__ G, int m,
                                                           const double* __restrict__ u, double* __restrict__ part,
                                                           const double* irr, const double* irb) {
  if (ne_refined(irr, irb)) return;
  const int K = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int kb = K * kNB, ke = min(m, kb + kNB);
  if (i >= ke) return;
  double s = 0.0;
  for (int k = kb; k < ke; ++k) s += (double)G[(long long)k * m + i] * u[k];
  part[(long long)K * m + i] = s;
}

// (3) y_i (+)= sum over the blocks K from i's own up of part[K][i], in block
// order (a fixed order: the result is deterministic).
__global__ void k_gapply_sum(const double* __restrict__ part, int m, int nb, double* __restrict__ y,
                             int accumulate, const double* irr, const double* irb, double* sweeps) {
  if (ne_refined(irr, irb)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (sweeps && i == 0) *sweeps += 1.0;  // a correction sweep that ran
  if (i >= m) return;
  double s = 0.0;
  for (int K = i / kNB; K < nb; ++K) s += part[(long long)K * m + i];
  y[i] = accumulate ? y[i] + s : s;
}

static int dense_apply_impl(const DenseNE& d, const double* r, double* y, bool accumulate, const double* irr,
                            const double* irb, double* sweeps, hipStream_t s) {
  const int m = d.m;
  const int nb = (m + kNB - 1) / kNB;
  const long long threads = (long long)m * kWave;
  const dim3 g((unsigned)((threads + 255) / 256));
  const dim3 gp((unsigned)((m + 255) / 256), (unsigned)nb);
  if (d.f32) {
    hipLaunchKernelGGL((k_gapply_lower<float>), g, dim3(256), 0, s, d.G.as<float>(), d.dinv.as<float>(), m, r,
                       d.u.as<double>(), irr, irb);
    hipLaunchKernelGGL((k_gapply_upper_part<float>), gp, dim3(256), 0, s, d.G.as<float>(), m, d.u.as<double>(),
                       d.part.as<double>(), irr, irb);
  } else {
    hipLaunchKernelGGL((k_gapply_lower<double>), g, dim3(256), 0, s, d.G.as<double>(), d.dinv.as<double>(), m, r,
                       d.u.as<double>(), irr, irb);
    hipLaunchKernelGGL((k_gapply_upper_part<double>), gp, dim3(256), 0, s, d.G.as<double>(), m, d.u.as<double>(),
                       d.part.as<double>(), irr, irb);
  }
  hipLaunchKernelGGL(k_gapply_sum, dim3(grid_for(m, 256)), dim3(256), 0, s, d.part.as<double>(), m, nb, y,
                     (int)accumulate, irr, irb, sweeps);
  count_launch(3);
  return check_launch("dense apply");
}

int dense_apply(const DenseNE& d, const double* r, double* y, bool accumulate, hipStream_t s) {
  return dense_apply_impl(d, r, y, accumulate, nullptr, nullptr, nullptr, s);
}

int dense_apply_refine(const DenseNE& d, const double* r, double* y, const double* irr, const double* irb,
                       double* sweeps, hipStream_t s) {
  return dense_apply_impl(d, r, y, true, irr, irb, sweeps, s);
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

extern "C" int hn_csr_adat_pattern(const hn_csr* a, hn_csr** out) {
  HN_CHECK(a && out, HN_ERR_INVALID, "hn_csr_adat_pattern: null argument");
  *out = nullptr;
  HN_TRY(hn::enter());
  hipStream_t s = hn::ctx().stream;
  hn_csr t;
  HN_TRY(hn::csr_alloc(&t, a->ncols, a->nrows, a->nnz, a->flags));
  HN_TRY(hn::transpose_into(a, &t, s));
  hn::NormalEq ne;
  HN_TRY(hn::ne_symbolic(a, &t, ne, s));
  HN_HIP(hipMemsetAsync(ne.N->val64.p, 0, sizeof(double) * std::max<int64_t>(ne.N->nnz, 1), s));
  HN_HIP(hipStreamSynchronize(s));
  *out = ne.N;
  ne.N = nullptr;
  return HN_OK;
}
