// Batched interior-point solves: thousands of small LPs in one persistent
// kernel, one workgroup per LP.
//
// Each workgroup takes the next LP from a device work queue (largest first,
// so the long solves start early), runs the whole Mehrotra loop of
// InteriorPointSolver::solve (core/lp/interior_point.cpp:443-753) on it and
// takes another, until the queue is empty; every workgroup reaches the
// queue's end, so the grid always drains.  Inside a workgroup:
//   * SpMV on M and on M' (pre-built transpose) one row per thread, summed
//     left to right - the order of the reference's CPU loops;
//   * N = M D M' + r I assembled densely one row per thread, then factored
//     in place by a right-looking blocked LDL' (64-wide blocks: the diagonal
//     block in registers with one barrier per column, the panel one row per
//     thread, the trailing update by LDS-staged 64x64 tiles), with the
//     reference's pivot clamp (factorization.cpp:1309-1313) or, under
//     HN_PIVOT_SKIP, Wright's skipping of tiny pivots;
//   * solves by blocked forward/backward substitution and FP64 iterative
//     refinement against the matrix-free N (refinement.cpp:7-36);
//   * in mixed precision the factor is FP32 and the refinement FP64; an LP
//     whose FP32 factor cannot be refined to 1e-10 in 5 sweeps, or whose mu
//     falls below mu_switch, continues with an FP64 factor.
// Everything else (start point, sigma, ratio test, guards, relaxed post-loop
// test, status) follows the reference exactly as the single-LP driver does.
#include "hn_device.hpp"
#include "hn_internal.hpp"
#include "hn_stdform.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

#pragma clang fp contract(off)

namespace hn {

namespace {

constexpr int NT = 256;
constexpr int BK = 64;  // LDL' block width
constexpr int kLdw = 2 * BK;  // row stride of W: the two blocks of a trailing update

struct Inst {
  int m, n;
  long long mp, mz;   // M: row_ptr offset (m+1 ints), col/val offset
  long long tp, tz;   // M': row_ptr offset (n+1 ints), col/val offset
  long long vm, vn;   // device-only m- and n-vector scratch blocks
  long long om, on;   // b and lambda (m each), c (n) and [x, z] (2n) in the host-visible arrays
  long long dn, w;    // dense N (m*m elements), panel workspace (m*64 elements)
  int warm;           // 1: x, z, lam hold a warm start accepted on the host; 2: one predicted on the
                      // device, tested by the kernel before use
  int mixed;          // 1: FP32 factor with FP64 refinement until mu < mu_switch; 0: FP64
  double off;         // objective constant of the standard form: c'x of the caller's LP = c_s'x_s + off
};

// Per-LP vectors.  B, LAM, C, X and Z cross the host link and live in their
// own arrays (b and c uploaded, lam, x and z uploaded for warm starts and
// downloaded); the rest is device-only scratch.
enum VM { VM_B = 0, VM_LAM, VM_RP, VM_TM, VM_MDRD, VM_RHS, VM_RHS2, VM_DLAM, VM_MZ, VM_R, VM_ND, VM_DG, VM_COUNT };
enum VN { VN_C = 0, VN_X, VN_Z, VN_RD, VN_D, VN_T, VN_DX, VN_DZ, VN_ZS, VN_COUNT };
constexpr int kVmIo = 2, kVnIo = 3;  // leading slots held in the host-visible arrays

constexpr int kStallIters = 3;

// Phases of a solve, timed per workgroup when Params::timing is set (the
// verbose >= 2 diagnostic): wall-clock ticks of the workgroup's first thread,
// taken at phase boundaries.
enum Phase {
  PH_START = 0,  // start point, warm-start acceptance
  PH_RESID,      // residuals, mu and the stop tests
  PH_PROJECT,    // optimal-face projection attempts, their factorisations included
  PH_RHS,        // D, M (D rd) and the right-hand side
  PH_ASSEMBLE,   // dense N: zeroing and the column products
  PH_FACTOR,     // blocked LDL': the rest (block bookkeeping)
  PH_DIAG,       // blocked LDL': 64 x 64 diagonal blocks
  PH_PANEL,      // blocked LDL': inverse of L11 and the panel W, L21
  PH_UPDATE,     // blocked LDL': trailing updates on the matrix cores
  PH_TRSV,       // triangular solves with the factor
  PH_REFINE,     // refinement products N v and residuals
  PH_STEP,       // directions, ratio tests, sigma, corrector right-hand side, update
  kPhases
};

struct Out {
  int status, iterations, fp64_from, ne_solves, refine, restarted, accepted, projected;
  double rp, rd, mu;
  unsigned long long ph[kPhases];  // phase ticks (Params::timing), else 0
};

// Phase clock: ph points at kPhases accumulators in LDS followed by the last
// timestamp; only thread 0 reads the clock and writes them.
struct Clock {
  unsigned long long* ph;
  int on;
  __device__ __forceinline__ void mark(int p) const {
    if (on && threadIdx.x == 0) {
      const unsigned long long now = (unsigned long long)wall_clock64();
      ph[p] += now - ph[kPhases];
      ph[kPhases] = now;
    }
  }
};

struct Params {
  int max_iterations;
  double conv_tol, comp_tol, reg, mu_switch;
  int skip;  // HN_PIVOT_SKIP: r = reg only, tiny pivots skipped
  double mu_target;  // > 0: stop at the first iterate with mu <= mu_target
  double gap_tol;    // > 0: the stop test also needs |c_s'x_s - b_s'lam| <= gap_tol (1 + |c'x|), c'x the
                     // caller's objective (the gap of the caller's LP when no column has a bound shift)
  double endgame_mu; // > 0: optimal-face projection once mu <= endgame_mu (then after each halving of
                     // mu or every 4 iterations)
  int timing;        // 1: per-phase wall-clock ticks into Out::ph
};

// A skipped pivot: its entry of D becomes "infinite", so its column of L and
// its component of every solve vanish.
template <typename T>
__device__ __forceinline__ T skipped_pivot();
template <>
__device__ __forceinline__ double skipped_pivot<double>() { return 1e128; }
template <>
__device__ __forceinline__ float skipped_pivot<float>() { return 1e30f; }

// ---- workgroup helpers -----------------------------------------------------
// A workgroup runs one wave per SIMD, so a loop that waits for each of its
// loads before the next is bound by memory latency.  The helpers below keep
// several independent loads in flight per thread while every row and every
// thread-local sum still accumulates in its original order (results are
// bit-identical to the one-at-a-time loops).

// y = A x, one row per thread, summed left to right.
__device__ __forceinline__ void wg_spmv(const int* __restrict__ rp, const int* __restrict__ ci,
                                        const double* __restrict__ v, const double* __restrict__ x,
                                        double* __restrict__ y, int rows) {
  for (int i = threadIdx.x; i < rows; i += NT) {
    double s = 0.0;
    int k = rp[i];
    const int e = rp[i + 1];
    for (; k + 4 <= e; k += 4) {  // four independent loads in flight, summed in order
      const int c0 = ci[k], c1 = ci[k + 1], c2 = ci[k + 2], c3 = ci[k + 3];
      const double v0 = v[k], v1 = v[k + 1], v2 = v[k + 2], v3 = v[k + 3];
      const double x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
      s += v0 * x0;
      s += v1 * x1;
      s += v2 * x2;
      s += v3 * x3;
    }
    for (; k < e; ++k) s += v[k] * x[ci[k]];
    y[i] = s;
  }
}

// for j = tid, tid + NT, ... < n: fetch(j, u) loads element j's operands
// into slot u, apply(j, u) computes and stores.  The loads of kEachU
// consecutive elements of the thread are issued before any of their
// stores; elements are applied in the plain loop's order.
constexpr int kEachU = 4;
template <class Fetch, class Apply>
__device__ __forceinline__ void wg_each(int n, Fetch&& fetch, Apply&& apply) {
  for (int j0 = threadIdx.x; j0 < n; j0 += kEachU * NT) {
#pragma unroll
    for (int u = 0; u < kEachU; ++u)
      if (j0 + u * NT < n) fetch(j0 + u * NT, u);
#pragma unroll
    for (int u = 0; u < kEachU; ++u)
      if (j0 + u * NT < n) apply(j0 + u * NT, u);
  }
}

__device__ __forceinline__ double wg_sum(double v, double* red) { return block_sum<double, NT>(v, red); }
__device__ __forceinline__ double wg_max(double v, double* red) { return block_max<double, NT>(v, red); }
__device__ __forceinline__ double wg_min(double v, double* red) { return block_min<double, NT>(v, red); }

// min over d < 0 of -v/d, then the reference's min(1, 0.99 alpha)
__device__ double wg_step(const double* __restrict__ v, const double* __restrict__ d, int n, double* red) {
  double a = 1.0;
  double vv[kEachU], dd[kEachU];
  wg_each(
      n,
      [&](int i, int u) {
        vv[u] = v[i];
        dd[u] = d[i];
      },
      [&](int, int u) {
        if (dd[u] < 0.0) {
          const double q = -vv[u] / dd[u];
          a = (q < a) ? q : a;
        }
      });
  a = wg_min(a, red);
  const double s = 0.99 * a;
  return (s < 1.0) ? s : 1.0;
}

// ---- dense LDL' in place (row-major) -------------------------------------------
// The factor keeps L (unit, strictly below the diagonal) and D (on it) in
// place of N's lower triangle and a copy of L' above the diagonal, where
// N's upper half (never read after the assembly) was: the forward solve
// then reads L's block columns as rows of L', contiguous across the threads
// that handle consecutive rows, as the backward solve reads L itself.
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
// v_mfma_f64_16x16x4 / v_mfma_f32_16x16x4f32: lane l holds row l%16 of A
// and column l%16 of B at k = l/16.
__device__ __forceinline__ d4_t mfma_16x16x4(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4_t mfma_16x16x4(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
// Tile row and column of output element q of block a (a: row half a >> 1,
// column half a & 1 of the wave's 32 x 32 quarter at (rq, cq)): FP64 lane
// l, element q is row 4q + l/16, column l%16 of a block; FP32 row
// 4(l/16) + q (both read back from the device).
template <typename T>
__device__ __forceinline__ void mfma_entry(int ln, int rq, int cq, int a, int q, int& r, int& c) {
  constexpr bool kF64 = std::is_same<T, double>::value;
  r = rq + 16 * (a >> 1) + (kF64 ? 4 * q + (ln >> 4) : 4 * (ln >> 4) + q);
  c = cq + 16 * (a & 1) + (ln & 15);
}

// LDS of the factorisation (FP64: 67 KB; gfx950 has 160 KB per workgroup).
// L11 (factor and inverse phases) and the update tiles share storage.
template <typename T>
struct SharedLDL {
  struct {
    T rk[2][BK];
    T D[BK];
    double dgb[BK];  // dg of the block (skip rule)
    T L11[BK][BK + 1];
  } f;
  union {
    struct {
      T X[BK][65];   // L11^{-1}
      T Ts[64][65];  // a 64-row tile of A21, then of W
    } p;
    struct {
      T Wr[kLdw][65];  // W of the current tile row, column-major (k, row)
      T Ls[BK][65];    // L slice of the current step, (k, row)
    } u;
  };
};

// The blocked LDL' runs its three parts per 64-column block as separate
// functions, so that each gets the registers of its own working set (values
// live in one part would otherwise hold registers through the others, and
// the loads of a part wait one at a time).
//
// (1) The diagonal block: returns the number of clamped (or skipped)
// pivots.  skip_rel == 0: the reference's clamp of pivots <= tol;
// skip_rel > 0: pivots <= skip_rel times dg (the diagonal of N before
// elimination, staged in LDS for the block) are skipped.  Leaves L11 (unit)
// and D in LDS.
template <typename T>
__device__ __noinline__ int ldl_diag(T* __restrict__ A, int m, int k0, int b, SharedLDL<T>& sh, double tol,
                                     const double* __restrict__ dg, double skip_rel) {
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  int nclamp = 0;
  if (skip_rel > 0.0 && tid < b) sh.f.dgb[tid] = dg[k0 + tid];
  // symmetric copy of the block in registers
  T P[4][4];
  if (b == BK) {  // a full block: no per-entry test around the reads
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = 4 * ti + a, j = 4 * tj + c;
        const int hi = max(i, j), lo = min(i, j);
        P[a][c] = A[(long long)(k0 + hi) * m + k0 + lo];
      }
  } else {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int i = 4 * ti + a, j = 4 * tj + c;
        const int hi = max(i, j), lo = min(i, j);
        P[a][c] = (i < b && j < b) ? A[(long long)(k0 + hi) * m + k0 + lo] : T(i == j ? 1 : 0);
      }
  }
  // four steps per pass, so the pivot column's register index is a
  // constant (a variable index puts P in scratch memory)
  for (int kb = 0; kb < b; kb += 4)
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    const int k = kb + kk, par = kk & 1;
    if (k >= b) break;
    if (tj == (kb >> 2)) {
#pragma unroll
      for (int a = 0; a < 4; ++a) sh.f.rk[par][4 * ti + a] = P[a][kk];
    }
    __syncthreads();
    T piv = sh.f.rk[par][k];
    T rinv;
    if (skip_rel > 0.0) {
      if (!((double)piv > skip_rel * sh.f.dgb[k])) {
        piv = skipped_pivot<T>();
        rinv = T(0);
        ++nclamp;
      } else {
        rinv = T(1) / piv;
      }
    } else {
      if (piv <= T(tol)) {
        piv = T(tol);
        ++nclamp;
      }
      rinv = T(1) / piv;
    }
    T ci[4], cj[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) ci[a] = sh.f.rk[par][4 * ti + a];
#pragma unroll
    for (int c = 0; c < 4; ++c) cj[c] = sh.f.rk[par][4 * tj + c];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int i = 4 * ti + a;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int j = 4 * tj + c;
        if (i == k && j == k)
          P[a][c] = piv;
        else if (j == k && i > k)
          P[a][c] = ci[a] * rinv;
        else if (i == k && j > k)
          P[a][c] = cj[c] * rinv;
        else if (i > k && j > k)
          P[a][c] = P[a][c] - ci[a] * (cj[c] * rinv);
      }
    }
  }
  // write back L (unit, below the diagonal), D, and above the diagonal
  // L' (the elimination leaves L[j][i] at (i, j) for j > i); keep L11, D
  // in LDS
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 4 * ti + a, j = 4 * tj + c;
      if (i < b && j < b) {
        A[(long long)(k0 + i) * m + k0 + j] = P[a][c];
        sh.f.L11[i][j] = (j < i) ? P[a][c] : T(0);
        if (i == j) sh.f.D[i] = P[a][c];
      }
    }
  __syncthreads();
  return nclamp;
}

// (2) The panel, for the rows below the block (k0 + b < m).  X = L11^{-1}
// (unit lower), one column per thread of the first wave, held in registers;
// then W = A21 L11^{-T} = A21 X' and L21 = W / D for 64 rows at a time: the
// rows are staged through LDS by coalesced loads, thread (r = tid % 64,
// g = tid / 64) forms 16 columns of row r in column order, and W, L21 and
// L21' leave through LDS by coalesced stores.
template <typename T>
__device__ __noinline__ void ldl_panel(T* __restrict__ A, int m, int k0, int b, T* __restrict__ W, int wc,
                                       SharedLDL<T>& sh) {
  const int tid = threadIdx.x;
  typedef T t4 __attribute__((ext_vector_type(4)));
  T(*X)[65] = sh.p.X;
  T(*Ts)[65] = sh.p.Ts;
  const int ln = tid & (kWave - 1), wq = tid / kWave;
  const int rq = 32 * (wq >> 1), cq = 32 * (wq & 1);
  if (tid < 64) {
    const int j = tid;
    T x[BK];
#pragma unroll
    for (int i = 0; i < BK; ++i) {
      // x[k] = 0 for k < j: those terms add exact zeros
      T acc = 0;
#pragma unroll
      for (int k = 0; k < i; ++k) acc += sh.f.L11[i][k] * x[k];
      x[i] = (i > j) ? -acc : T(i == j ? 1 : 0);
    }
#pragma unroll
    for (int i = 0; i < BK; ++i) X[i][j] = x[i];
  }
  for (int i0 = k0 + b; i0 < m; i0 += 64) {
    const int rows = min(64, m - i0);
    if (b == BK && rows == 64) {  // no per-entry test around the loads
#pragma unroll
      for (int u = 0; u < 64 * BK / NT; ++u) {
        const int e = tid + NT * u, rr = e / BK, cc = e % BK;
        Ts[rr][cc] = A[(long long)(i0 + rr) * m + k0 + cc];
      }
    } else {
#pragma unroll
      for (int u = 0; u < 64 * BK / NT; ++u) {
        const int e = tid + NT * u, rr = e / BK, cc = e % BK;
        Ts[rr][cc] = (rr < rows && cc < b) ? A[(long long)(i0 + rr) * m + k0 + cc] : T(0);
      }
    }
    __syncthreads();
    // W tile (r, c) = sum_k Ts[r][k] X[c][k] on the matrix cores, in the
    // operand and output layouts of the trailing update below: each wave
    // forms a 32 x 32 quarter as four 16 x 16 blocks, k in steps of 4
    t4 dd[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) dd[a] = t4{0, 0, 0, 0};
#pragma unroll
    for (int c0 = 0; c0 < BK; c0 += 4) {
      const int kk = c0 + (ln >> 4);
      const T a0 = Ts[rq + (ln & 15)][kk];
      const T a1 = Ts[rq + 16 + (ln & 15)][kk];
      const T b0 = X[cq + (ln & 15)][kk];
      const T b1 = X[cq + 16 + (ln & 15)][kk];
      dd[0] = mfma_16x16x4(a0, b0, dd[0]);
      dd[1] = mfma_16x16x4(a0, b1, dd[1]);
      dd[2] = mfma_16x16x4(a1, b0, dd[2]);
      dd[3] = mfma_16x16x4(a1, b1, dd[3]);
    }
    __syncthreads();  // every read of the staged rows precedes the writes below
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        int rr, cc;
        mfma_entry<T>(ln, rq, cq, a, q, rr, cc);
        Ts[rr][cc] = dd[a][q];
      }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 64 * BK / NT; ++u) {
      const int e = tid + NT * u, rr = e / BK, cc = e % BK;
      if (rr < rows && cc < b) {
        const T w = Ts[rr][cc];
        W[(long long)(i0 + rr) * kLdw + wc + cc] = w;
        A[(long long)(i0 + rr) * m + k0 + cc] = w / sh.f.D[cc];
      }
    }
#pragma unroll
    for (int u = 0; u < 64 * BK / NT; ++u) {
      const int e = tid + NT * u, cc = e / 64, rr = e % 64;
      if (rr < rows && cc < b) A[(long long)(k0 + cc) * m + i0 + rr] = Ts[rr][cc] / sh.f.D[cc];  // L' above the diagonal
    }
    __syncthreads();  // the next rows are staged into Ts
  }
  __syncthreads();
}

// (3) A trailing update of the lower triangle: A_ij -= sum_c W_ic L_jc over
// kw <= 128 columns c (L_jc = A[j][kc + c], W_ic = W[i][c]), for the rows
// i >= r0 and the first ntc tile columns from r0, by 64 x 64 tiles
// (I, J <= I), row by row, in steps of 64 columns c.  Folding two blocks'
// updates into one pass over the trailing matrix halves its reads and
// writes per multiply-add.  A tile row's W stays in LDS for all its tiles.
// While a step runs on the matrix cores, the next step's L slice (columns
// < r0 of A, which the update does not write), at a new tile the entries of
// A it updates and at a new tile row its W rows are loaded into registers;
// they are stored to LDS rows padded to 65 (a store's lanes of one row fall
// in different banks) when that step starts.
template <typename T>
__device__ __noinline__ void ldl_update(T* __restrict__ A, int m, int kc, int kw, int r0, int ntc,
                                        const T* __restrict__ W, SharedLDL<T>& sh) {
  typedef T t4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x;
  const int nt = (m - r0 + 63) / 64;
  const int nch = (kw + BK - 1) / BK;  // 64-column steps per tile
  // tiles: tile row tI has min(tI + 1, ntc) of them
  const int ntiles = ntc >= nt ? nt * (nt + 1) / 2 : ntc * (ntc + 1) / 2 + (nt - ntc) * ntc;
  const int nsteps = ntiles * nch;
  constexpr int kSlots = 64 * BK / NT;  // a thread's (row, column) slots of a 64 x 64 slice
  const int ln = tid & (kWave - 1), wq = tid / kWave;
  const int rq = 32 * (wq >> 1), cq = 32 * (wq & 1);
  T pl[kSlots], pw[kSlots], pa[16];
  // Slices that lie wholly inside the matrix (all but the last tile row and
  // column) load without per-entry tests: written with them, every load
  // becomes a branch of its own.  Entries above the diagonal of a diagonal
  // tile are loaded (they are inside the matrix) but not written back.
  auto fetch_l = [&](int tJ, int ch) {  // L rows of tile column tJ, step ch
    const int j0 = r0 + 64 * tJ, cw = min(BK, kw - BK * ch), c0 = kc + BK * ch;
    if (cw == BK && j0 + 64 <= m) {
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int e = tid + NT * u;
        pl[u] = A[(long long)(j0 + e / BK) * m + c0 + e % BK];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int e = tid + NT * u;
        const int r = e / BK, c = e % BK;
        pl[u] = (j0 + r < m && c < cw) ? A[(long long)(j0 + r) * m + c0 + c] : T(0);
      }
    }
  };
  auto fetch_w = [&](int tI, int ch) {  // W rows of tile row tI, step ch
    const int i0 = r0 + 64 * tI, cw = min(BK, kw - BK * ch), c0 = BK * ch;
    if (cw == BK && i0 + 64 <= m) {
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int e = tid + NT * u;
        pw[u] = W[(long long)(i0 + e / BK) * kLdw + c0 + e % BK];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int e = tid + NT * u;
        const int r = e / BK, c = e % BK;
        pw[u] = (i0 + r < m && c < cw) ? W[(long long)(i0 + r) * kLdw + c0 + c] : T(0);
      }
    }
  };
  auto fetch_a = [&](int tI, int tJ) {  // the thread's 16 entries of tile (tI, tJ)
    const int i0 = r0 + 64 * tI, j0 = r0 + 64 * tJ;
    if (i0 + 64 <= m) {  // j0 <= i0
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int r, c;
          mfma_entry<T>(ln, rq, cq, a, q, r, c);
          pa[4 * a + q] = A[(long long)(i0 + r) * m + j0 + c];
        }
    } else {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          int r, c;
          mfma_entry<T>(ln, rq, cq, a, q, r, c);
          const int i = i0 + r, j = j0 + c;
          pa[4 * a + q] = (i < m && j < m && j <= i) ? A[(long long)i * m + j] : T(0);
        }
    }
  };
  int tI = 0, tJ = 0, ch = 0;
  if (nsteps > 0) {
    fetch_w(0, 0);
    fetch_l(0, 0);
    fetch_a(0, 0);
  }
  T aold[16];
  t4 dd[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) dd[a] = t4{0, 0, 0, 0};
  for (int st = 0; st < nsteps; ++st) {
    if (ch == 0) {
#pragma unroll
      for (int u = 0; u < 16; ++u) aold[u] = pa[u];
    }
    if (tJ == 0) {  // a new tile row: its W into LDS
#pragma unroll
      for (int u = 0; u < kSlots; ++u) {
        const int e = tid + NT * u;
        sh.u.Wr[BK * ch + e % BK][e / BK] = pw[u];
      }
    }
#pragma unroll
    for (int u = 0; u < kSlots; ++u) {
      const int e = tid + NT * u;
      sh.u.Ls[e % BK][e / BK] = pl[u];
    }
    __syncthreads();
    int nI = tI, nJ = tJ, nc = ch + 1;  // the next step
    if (nc == nch) {
      nc = 0;
      if (++nJ > nI || nJ >= ntc) {
        nJ = 0;
        ++nI;
      }
    }
    if (st + 1 < nsteps) {
      fetch_l(nJ, nc);
      if (nJ == 0) fetch_w(nI, nc);
      if (nc == 0) fetch_a(nI, nJ);
    }
#pragma unroll
    for (int c0 = 0; c0 < BK; c0 += 4) {
      const int kk = c0 + (ln >> 4);
      const T a0 = sh.u.Wr[BK * ch + kk][rq + (ln & 15)];
      const T a1 = sh.u.Wr[BK * ch + kk][rq + 16 + (ln & 15)];
      const T b0 = sh.u.Ls[kk][cq + (ln & 15)];
      const T b1 = sh.u.Ls[kk][cq + 16 + (ln & 15)];
      dd[0] = mfma_16x16x4(a0, b0, dd[0]);
      dd[1] = mfma_16x16x4(a0, b1, dd[1]);
      dd[2] = mfma_16x16x4(a1, b0, dd[2]);
      dd[3] = mfma_16x16x4(a1, b1, dd[3]);
    }
    __syncthreads();  // every read of Ls and Wr precedes the next step's stores
    if (ch == nch - 1) {  // the tile's last step: subtract it from A
      const int i0 = r0 + 64 * tI, j0 = r0 + 64 * tJ;
      if (i0 + 64 <= m && tJ < tI) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            int r, c;
            mfma_entry<T>(ln, rq, cq, a, q, r, c);
            A[(long long)(i0 + r) * m + j0 + c] = aold[4 * a + q] - dd[a][q];
          }
      } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            int r, c;
            mfma_entry<T>(ln, rq, cq, a, q, r, c);
            const int i = i0 + r, j = j0 + c;
            if (i < m && j < m && j <= i) A[(long long)i * m + j] = aold[4 * a + q] - dd[a][q];
          }
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) dd[a] = t4{0, 0, 0, 0};
    }
    tI = nI;
    tJ = nJ;
    ch = nc;
  }
  __syncthreads();  // the next block reads what the update wrote
}

// Blocked LDL' by pairs of 64-column blocks a = [K, K + 64) and b = [K + 64,
// K + 128): factor a, update block column b by a (one tile column), factor
// b, then update the trailing matrix from K + 128 by a and b together.
template <typename T>
__device__ __noinline__ int wg_ldlt(T* __restrict__ A, int m, T* __restrict__ W, SharedLDL<T>& sh, double tol,
                                    const double* __restrict__ dg, double skip_rel, Clock ck) {
  int nclamp = 0;
  for (int K = 0; K < m; K += 2 * BK) {
    const int ba = min(BK, m - K);
    nclamp += ldl_diag<T>(A, m, K, ba, sh, tol, dg, skip_rel);
    ck.mark(PH_DIAG);
    if (K + ba >= m) break;
    ldl_panel<T>(A, m, K, ba, W, 0, sh);
    ck.mark(PH_PANEL);
    const int kb = K + BK, bb = min(BK, m - kb);
    ldl_update<T>(A, m, K, ba, kb, 1, W, sh);
    ck.mark(PH_UPDATE);
    nclamp += ldl_diag<T>(A, m, kb, bb, sh, tol, dg, skip_rel);
    ck.mark(PH_DIAG);
    if (kb + bb >= m) break;
    ldl_panel<T>(A, m, kb, bb, W, BK, sh);
    ck.mark(PH_PANEL);
    const int r0 = kb + bb;
    ldl_update<T>(A, m, K, ba + bb, r0, (m - r0 + 63) / 64, W, sh);
    ck.mark(PH_UPDATE);
  }
  return nclamp;
}

// x = N^{-1} r with the LDL' factor in A (in-place on x, which holds r).
// Every row segment is loaded in full before it is used, so each thread keeps
// up to 64 loads in flight; the current block of x is shared through LDS.
// x[i] -= sum_c A[k0 + c][i] xb[c] over the block's b columns, for the rows
// i = i0, i0 + NT, ... below i1 (one per thread; the block's entries
// A[k0 + c][i] of consecutive rows are contiguous), summed in column order.
// A full block (b = 64, every block but the last) takes a path without the
// per-column test: written with it, the compiler turns each of the 64 reads
// into a branch of its own and spills their addresses.  xb is re-read from
// LDS for every row: kept in registers across the rows, its 64 values push
// the row's own 64 out to scratch memory.
template <typename T>
__device__ __forceinline__ void solve_rows(const T* __restrict__ A, int m, int k0, int b, int i0, int i1,
                                           double* __restrict__ x, const double* xb) {
  if (b == BK) {
    for (int i = i0; i < i1; i += NT) {
      asm volatile("" ::: "memory");
      const T* col = A + (long long)k0 * m + i;
      T lv[BK];
#pragma unroll
      for (int c = 0; c < BK; ++c) lv[c] = col[(long long)c * m];
      double s = x[i];
#pragma unroll
      for (int c = 0; c < BK; ++c) s -= (double)lv[c] * xb[c];
      x[i] = s;
    }
  } else {
    for (int i = i0; i < i1; i += NT) {
      asm volatile("" ::: "memory");
      double s = x[i];
      for (int c = 0; c < b; ++c) s -= (double)A[(long long)(k0 + c) * m + i] * xb[c];
      x[i] = s;
    }
  }
}

template <typename T>
__device__ __noinline__ void wg_ldlt_solve(const T* __restrict__ A, int m, double* __restrict__ x) {
  __shared__ double xb[BK];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // forward: L y = r
  for (int k0 = 0; k0 < m; k0 += BK) {
    const int b = min(BK, m - k0);
    if (tid < 64) {
      T lr[BK];
#pragma unroll
      for (int c = 0; c < BK; ++c) lr[c] = (lane < b && c < lane) ? A[(long long)(k0 + c) * m + k0 + lane] : T(0);
      double y = (lane < b) ? x[k0 + lane] : 0.0;
#pragma unroll
      for (int c = 0; c < BK; ++c) {
        const double yc = __shfl(y, c, 64);
        if (lane > c) y -= (double)lr[c] * yc;
      }
      if (lane < b) {
        x[k0 + lane] = y;
        xb[lane] = y;
      }
    }
    __syncthreads();
    solve_rows<T>(A, m, k0, b, k0 + b + tid, m, x, xb);  // L[i][k0 + c] read from L' above the diagonal
    __syncthreads();
  }
  for (int i = tid; i < m; i += NT) {
    const double d = (double)A[(long long)i * m + i];
    if (d != 0.0) x[i] = x[i] / d;
  }
  __syncthreads();
  // backward: L' x = y
  const int nb = (m + BK - 1) / BK;
  for (int bi = nb - 1; bi >= 0; --bi) {
    const int k0 = bi * BK;
    const int b = min(BK, m - k0);
    if (tid < 64) {
      T lc[BK];
#pragma unroll
      for (int c = 0; c < BK; ++c) lc[c] = (c < b && lane < c) ? A[(long long)(k0 + c) * m + k0 + lane] : T(0);
      double y = (lane < b) ? x[k0 + lane] : 0.0;
#pragma unroll
      for (int c = BK - 1; c >= 0; --c) {
        const double xc = __shfl(y, c, 64);
        if (lane < c) y -= (double)lc[c] * xc;
      }
      if (lane < b) {
        x[k0 + lane] = y;
        xb[lane] = y;
      }
    }
    __syncthreads();
    solve_rows<T>(A, m, k0, b, tid, k0, x, xb);
    __syncthreads();
  }
}

// ---- the per-LP Mehrotra loop ------------------------------------------------
struct Ctx {
  const int *mrp, *mci, *trp, *tci;
  const double *mv, *tv;
  double *vm[VM_COUNT], *vn[VN_COUNT];
  int m, n;
};

// nd = N v = M (D (M' v)) + reg v  (matrix-free, FP64)
__device__ __noinline__ void wg_apply_N(const Ctx& c, const double* __restrict__ v, double* __restrict__ nd, double reg) {
  double* t = c.vn[VN_T];
  wg_spmv(c.trp, c.tci, c.tv, v, t, c.n);
  __syncthreads();
  const double* D = c.vn[VN_D];
  double dq[kEachU], tq[kEachU];
  wg_each(
      c.n,
      [&](int j, int u) {
        dq[u] = D[j];
        tq[u] = t[j];
      },
      [&](int j, int u) { t[j] = dq[u] * tq[u]; });
  __syncthreads();
  wg_spmv(c.mrp, c.mci, c.mv, t, nd, c.m);
  __syncthreads();
  for (int i = threadIdx.x; i < c.m; i += NT) nd[i] = nd[i] + reg * v[i];
  __syncthreads();
}

// dlam = N^{-1} rhs with refinement; returns true when refined to 1e-10.
template <typename T>
__device__ bool wg_solve(const Ctx& c, const T* A, const double* __restrict__ rhs, double* __restrict__ dlam,
                         double reg, double* red, int* refine, Clock ck) {
  for (int i = threadIdx.x; i < c.m; i += NT) dlam[i] = rhs[i];
  __syncthreads();
  wg_ldlt_solve<T>(A, c.m, dlam);
  ck.mark(PH_TRSV);
  double* r = c.vm[VM_R];
  double* nd = c.vm[VM_ND];
  double nb = 0.0;
  for (int i = threadIdx.x; i < c.m; i += NT) nb = nan_ignoring_max(nb, fabs(rhs[i]));
  nb = wg_max(nb, red);
  if (nb == 0.0 || nb != nb) return true;
  for (int it = 0; it <= 5; ++it) {
    wg_apply_N(c, dlam, nd, reg);
    double a = 0.0;
    for (int i = threadIdx.x; i < c.m; i += NT) {
      const double ri = rhs[i] - nd[i];
      r[i] = ri;
      a = nan_ignoring_max(a, fabs(ri));
    }
    a = wg_max(a, red);
    ck.mark(PH_REFINE);
    if (a / nb <= 1e-10) return true;
    if (it == 5) break;
    wg_ldlt_solve<T>(A, c.m, r);
    for (int i = threadIdx.x; i < c.m; i += NT) dlam[i] = dlam[i] + r[i];
    __syncthreads();
    ck.mark(PH_TRSV);
    if (threadIdx.x == 0) ++*refine;
  }
  return false;
}

// Dense N (row-major, full rows written) + max-diag regularisation; returns r.
template <typename T>
__device__ __noinline__ double wg_assemble(const Ctx& c, T* __restrict__ A, double reg_min, double rel_reg,
                                           double* red) {
  const int m = c.m;
  const long long tot = (long long)m * m;
  for (long long e = threadIdx.x; e < tot; e += NT) A[e] = T(0);
  __syncthreads();
  const double* D = c.vn[VN_D];
  double mxd = 0.0;
  // Row i of N, one thread per row: for each column j of row i (in order)
  // with d_j > 0 and each entry (rr, j) of that column (in order),
  // N[i][rr] += (a_ij a_rrj) d_j.  Every such column holds row i itself,
  // so the diagonal is summed in registers, in the same order and type as
  // its read-modify-writes would have, and stored once; only the other
  // rows' entries are updated in memory.  The columns' indices, scales and
  // first two entries are loaded kAsmG columns at a time ahead of the
  // updates, which stay in order.
  constexpr int kAsmG = 4;
  for (int i = threadIdx.x; i < m; i += NT) {
    double diag = 0.0;
    T diag_t = T(0);
    T* row = A + (long long)i * m;
    const int kb = c.mrp[i], ke = c.mrp[i + 1];
    for (int k0 = kb; k0 < ke; k0 += kAsmG) {
      int jq[kAsmG], q0[kAsmG], q1[kAsmG], r0[kAsmG], r1[kAsmG];
      double aq[kAsmG], dq[kAsmG], t0[kAsmG], t1[kAsmG];
#pragma unroll
      for (int u = 0; u < kAsmG; ++u)
        if (k0 + u < ke) {
          jq[u] = c.mci[k0 + u];
          aq[u] = c.mv[k0 + u];
        }
#pragma unroll
      for (int u = 0; u < kAsmG; ++u)
        if (k0 + u < ke) {
          dq[u] = D[jq[u]];
          q0[u] = c.trp[jq[u]];
          q1[u] = c.trp[jq[u] + 1];
        }
#pragma unroll
      for (int u = 0; u < kAsmG; ++u)
        if (k0 + u < ke) {
          if (q0[u] < q1[u]) {
            r0[u] = c.tci[q0[u]];
            t0[u] = c.tv[q0[u]];
          }
          if (q0[u] + 1 < q1[u]) {
            r1[u] = c.tci[q0[u] + 1];
            t1[u] = c.tv[q0[u] + 1];
          }
        }
#pragma unroll
      for (int u = 0; u < kAsmG; ++u) {
        if (k0 + u >= ke || dq[u] <= 0.0) continue;
        const double a = aq[u], d = dq[u];
        for (int q = q0[u]; q < q1[u]; ++q) {
          const int rr = (q == q0[u]) ? r0[u] : (q == q0[u] + 1) ? r1[u] : c.tci[q];
          const double tq = (q == q0[u]) ? t0[u] : (q == q0[u] + 1) ? t1[u] : c.tv[q];
          const double val = (a * tq) * d;
          if (rr == i) {
            diag += val;
            diag_t = diag_t + (T)val;
          } else {
            row[rr] = row[rr] + (T)val;
          }
        }
      }
    }
    row[i] = diag_t;
    mxd = nan_ignoring_max(mxd, diag);
    c.vm[VM_DG][i] = diag;
  }
  mxd = wg_max(mxd, red);
  const double r = fmax(reg_min, rel_reg * mxd);
  for (int i = threadIdx.x; i < m; i += NT) {
    A[(long long)i * m + i] = A[(long long)i * m + i] + (T)r;
    c.vm[VM_DG][i] = c.vm[VM_DG][i] + r;
  }
  __syncthreads();
  return r;
}

// Assembles and factors N; *reg_out receives its regularisation.
template <typename T>
__device__ void wg_factor(const Ctx& c, T* A, T* W, SharedLDL<T>& sh, double* red, double reg_min, bool skip,
                          double* reg_out, Clock ck) {
  *reg_out = wg_assemble<T>(c, A, reg_min, skip ? 0.0 : 1e-12, red);
  ck.mark(PH_ASSEMBLE);
  wg_ldlt<T>(A, c.m, W, sh, 1e-10, c.vm[VM_DG], skip ? (sizeof(T) == 4 ? 1e-6 : 1e-13) : 0.0, ck);
  __syncthreads();
  ck.mark(PH_FACTOR);
}

__global__ __launch_bounds__(NT) void k_batch_ipm(const Inst* __restrict__ inst, const int* __restrict__ order,
                                                  int count, int* __restrict__ queue, const int* __restrict__ mrp,
                                                  const int* __restrict__ mci, const double* __restrict__ mv,
                                                  const int* __restrict__ trp, const int* __restrict__ tci,
                                                  const double* __restrict__ tv, double* __restrict__ vecm,
                                                  double* __restrict__ vecn, double* __restrict__ bio,
                                                  double* __restrict__ lio, double* __restrict__ cio,
                                                  double* __restrict__ xzio, double* __restrict__ dense,
                                                  double* __restrict__ work, Params prm, Out* __restrict__ out) {
  __shared__ double red[NT / kWave];
  __shared__ int s_next;
  __shared__ SharedLDL<double> sh64;
  __shared__ unsigned long long s_ph[kPhases + 1];
  SharedLDL<float>& sh32 = *reinterpret_cast<SharedLDL<float>*>(&sh64);
  const int tid = threadIdx.x;
  const Clock ck{s_ph, prm.timing}, ck_off{s_ph, 0};
  for (;;) {
    if (tid == 0) s_next = atomicAdd(queue, 1);
    __syncthreads();
    const int q = s_next;
    __syncthreads();
    if (q >= count) return;
    const int id = order[q];
    const Inst I = inst[id];
    if (prm.timing && tid == 0) {
      for (int p = 0; p < kPhases; ++p) s_ph[p] = 0;
      s_ph[kPhases] = (unsigned long long)wall_clock64();
    }
    Ctx c;
    c.m = I.m;
    c.n = I.n;
    c.mrp = mrp + I.mp;
    c.mci = mci + I.mz;
    c.mv = mv + I.mz;
    c.trp = trp + I.tp;
    c.tci = tci + I.tz;
    c.tv = tv + I.tz;
    c.vm[VM_B] = bio + I.om;
    c.vm[VM_LAM] = lio + I.om;
    for (int k = kVmIo; k < VM_COUNT; ++k) c.vm[k] = vecm + I.vm + (long long)(k - kVmIo) * I.m;
    c.vn[VN_C] = cio + I.on;
    c.vn[VN_X] = xzio + 2 * I.on;
    c.vn[VN_Z] = xzio + 2 * I.on + I.n;
    for (int k = kVnIo; k < VN_COUNT; ++k) c.vn[k] = vecn + I.vn + (long long)(k - kVnIo) * I.n;
    const int m = c.m, n = c.n;
    double* A64 = dense + I.dn;
    float* A32 = reinterpret_cast<float*>(A64);
    double* W64 = work + I.w;
    float* W32 = reinterpret_cast<float*>(W64);
    double* b = c.vm[VM_B];
    double* lam = c.vm[VM_LAM];
    double* rp = c.vm[VM_RP];
    double* tm = c.vm[VM_TM];
    double* cc = c.vn[VN_C];
    double* x = c.vn[VN_X];
    double* z = c.vn[VN_Z];
    double* rd = c.vn[VN_RD];
    double* D = c.vn[VN_D];
    double* t = c.vn[VN_T];
    double* dx = c.vn[VN_DX];
    double* dz = c.vn[VN_DZ];
    double* zs = c.vn[VN_ZS];
    double* dlam = c.vm[VM_DLAM];
    double* rhs = c.vm[VM_RHS];
    double* rhs2 = c.vm[VM_RHS2];
    double* mdrd = c.vm[VM_MDRD];
    double* mz = c.vm[VM_MZ];
    int fp64_from = -1, ne_solves = 0, refine = 0;
    bool fp64 = !I.mixed;
    // start point (interior_point.cpp:443-453), unless the host accepted a
    // warm start and packed it into x, z, lam
    auto cold_start = [&]() {
      double bn = 1.0, cn = 1.0;
      for (int i = tid; i < m; i += NT) bn = nan_ignoring_max(bn, fabs(b[i]));
      for (int j = tid; j < n; j += NT) cn = nan_ignoring_max(cn, fabs(cc[j]));
      bn = wg_max(bn, red);
      cn = wg_max(cn, red);
      for (int j = tid; j < n; j += NT) {
        x[j] = bn + 1.0;
        z[j] = fmax(cn + 1.0, fabs(cc[j]));
      }
      for (int i = tid; i < m; i += NT) lam[i] = 0.0;
      __syncthreads();
    };
    // A predicted start passes the host's test (warm_start_acceptable) here:
    // finite and strictly interior, mu below the cold start's, every x_j s_j
    // at least 1e-3 mu, and a primal residual no larger than the cold start's.
    auto warm_acceptable = [&]() -> bool {
      double bn = 1.0, cn = 1.0;
      for (int i = tid; i < m; i += NT) bn = nan_ignoring_max(bn, fabs(b[i]));
      for (int j = tid; j < n; j += NT) cn = nan_ignoring_max(cn, fabs(cc[j]));
      bn = wg_max(bn, red);
      cn = wg_max(cn, red);
      double bad = 0.0, mc = 0.0, m0 = 0.0, pmin = HUGE_VAL;
      for (int j = tid; j < n; j += NT) {
        if (!(x[j] > 0.0) || !(z[j] > 0.0) || !isfinite(x[j]) || !isfinite(z[j])) bad = 1.0;
        mc += (bn + 1.0) * fmax(cn + 1.0, fabs(cc[j]));
        const double pj = x[j] * z[j];
        m0 += pj;
        pmin = fmin(pmin, pj);
      }
      for (int i = tid; i < m; i += NT)
        if (!isfinite(lam[i])) bad = 1.0;
      bad = wg_max(bad, red);
      mc = wg_sum(mc, red) / (double)n;
      m0 = wg_sum(m0, red) / (double)n;
      pmin = wg_min(pmin, red);
      if (bad > 0.0 || !(m0 < mc) || !(pmin >= 1e-3 * m0)) return false;
      wg_spmv(c.mrp, c.mci, c.mv, x, tm, m);
      __syncthreads();
      double rw = 0.0, rc = 0.0;
      for (int i = tid; i < m; i += NT) {
        double sc = 0.0;
        for (int k = c.mrp[i]; k < c.mrp[i + 1]; ++k) sc += c.mv[k] * (bn + 1.0);
        rw = fmax(rw, fabs(tm[i] - b[i]));
        rc = fmax(rc, fabs(sc - b[i]));
      }
      rw = wg_max(rw, red);
      rc = wg_max(rc, red);
      return rw <= rc;
    };
    bool accepted = I.warm == 1;
    if (I.warm == 2) accepted = warm_acceptable();
    if (!accepted) cold_start();
    __syncthreads();
    ck.mark(PH_START);
    bool done = false, numerr = false, stopped = false;
    bool warm_live = accepted, restarted = false;
    int stuck = 0;
    int iterations = 0;
    double mu = 0.0, rpn = 0.0, rdn = 0.0;
    auto residuals = [&]() {
      wg_spmv(c.mrp, c.mci, c.mv, x, tm, m);
      wg_spmv(c.trp, c.tci, c.tv, lam, t, n);
      __syncthreads();
      double a = 0.0, bb = 0.0, s = 0.0;
      for (int i = tid; i < m; i += NT) {
        const double v = tm[i] - b[i];
        rp[i] = v;
        a = nan_ignoring_max(a, fabs(v));
      }
      {
        double tq[kEachU], zq[kEachU], cq[kEachU], xq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              tq[u] = t[j];
              zq[u] = z[j];
              cq[u] = cc[j];
              xq[u] = x[j];
            },
            [&](int j, int u) {
              const double v = tq[u] + zq[u] - cq[u];
              rd[j] = v;
              bb = nan_ignoring_max(bb, fabs(v));
              s += xq[u] * zq[u];
            });
      }
      rpn = wg_max(a, red);
      rdn = wg_max(bb, red);
      mu = wg_sum(s, red) / (double)n;
    };
    // Optimal-face projection (not in the reference; opt-in by endgame_mu).
    // Primal: over B = {j : x_j >= theta z_j} and with D_B = x_B (0 off B),
    // N_B = M D_B M' + r I:
    //   x' = x_B + D_B M' w,   N_B w = b - M x_B        (x'_N = 0, M x' = b)
    // tried with theta = 1, then 0.1, then 0.01 until x' is primal feasible
    // to the tolerance: on a degenerate LP a basic column can still have
    // x_j < z_j when the iterates stall (a pair that has not separated), and
    // without it M_B cannot reach b.  The supports are nested, so a theta
    // that adds no column is skipped.
    // Dual: a weighted least squares for z'_B = 0,
    //   y' = lam + d,   (M W M' + r I) d = M W (c - M'lam)
    // first with the IPM's own scaling W = x/z over {j : x_j >= z_j}, in
    // which the strongly basic columns (x/z ~ 1e5 and up) outweigh a weakly
    // basic one that would otherwise leave its share of the misfit as
    // negative reduced costs on the others; W is kept unclamped (up to
    // 1e200), since capping it at the IPM's 1e10 flattens the strongly basic
    // columns and brings those negative reduced costs back.  If z' is not
    // dual feasible, W = x over the same support, then W = x/z over
    // {x_j >= 100 z_j}.
    //   z' = c - M'y'
    // x' and z' are clamped at 0 (each by at most conv_tol) and the point is
    // accepted when it meets the reference's stop test (||M x' - b||,
    // ||rd|| = ||max(-z', 0)|| and mu below their tolerances) and, with
    // gap_tol, the relative gap; the solve then ends there.  Otherwise the
    // iterates are left unchanged.  N_B is singular when M_B has fewer than
    // m independent columns; both right-hand sides lie in its range, and
    // the reference's regularisation (clamped pivots, r >= 1e-12 max diag)
    // keeps the solutions' null-space parts at rounding size - a skipped
    // pivot would instead fix them at an arbitrary value.
    auto project = [&]() -> bool {
      double reg = 0.0, bad = 0.0, xneg = 0.0, rp2 = 0.0, prev_nb = -1.0;
      bool primal_ok = false;
      for (int ti = 0; ti < 3 && !primal_ok; ++ti) {
        const double th = (ti == 0) ? 1.0 : (ti == 1 ? 0.1 : 0.01);
        double nb = 0.0;
        for (int j = tid; j < n; j += NT) {
          const bool in = x[j] >= th * z[j];
          D[j] = in ? x[j] : 0.0;
          t[j] = in ? x[j] : 0.0;
          nb += in ? 1.0 : 0.0;
        }
        nb = wg_sum(nb, red);
        if (nb == prev_nb) continue;
        prev_nb = nb;
        wg_factor<double>(c, A64, W64, sh64, red, prm.reg, false, &reg, ck_off);
        wg_spmv(c.mrp, c.mci, c.mv, t, tm, m);
        __syncthreads();
        for (int i = tid; i < m; i += NT) rhs[i] = b[i] - tm[i];
        __syncthreads();
        (void)wg_solve<double>(c, A64, rhs, dlam, reg, red, &refine, ck_off);
        wg_spmv(c.trp, c.tci, c.tv, dlam, t, n);
        __syncthreads();
        // x'_B below zero by no more than the tolerance (degenerate basic
        // components) is clamped to 0; the residual is that of the clamped
        // point
        double b1 = 0.0, xn = 0.0;
        for (int j = tid; j < n; j += NT) {
          const double v = (D[j] > 0.0) ? x[j] + D[j] * t[j] : 0.0;
          dx[j] = fmax(v, 0.0);
          xn = nan_ignoring_max(xn, -v);
          if (!isfinite(v)) b1 = 1.0;
        }
        __syncthreads();
        wg_spmv(c.mrp, c.mci, c.mv, dx, tm, m);
        __syncthreads();
        double r2 = 0.0;
        for (int i = tid; i < m; i += NT) r2 = nan_ignoring_max(r2, fabs(tm[i] - b[i]));
        bad = wg_max(b1, red);
        xneg = wg_max(xn, red);
        rp2 = wg_max(r2, red);
        primal_ok = bad == 0.0 && xneg < prm.conv_tol && rp2 < prm.conv_tol;
      }
      ++ne_solves;
      if (!primal_ok) return false;
      for (int di = 0; di < 3; ++di) {
        // weights x/z over {x >= z}, then x over {x >= z}, then x/z over
        // {x >= 100 z}: the last drops columns whose pair has not
        // separated, for when {x >= z} holds more columns than M has rank
        // and z'_B = 0 cannot hold on all of them
        const double th = (di == 2) ? 100.0 : 1.0;
        for (int j = tid; j < n; j += NT) {
          double w = 0.0;
          if (x[j] >= th * z[j]) {
            w = (di == 1) ? x[j] : x[j] / z[j];
            if (!(w <= 1e200)) w = 1e200;  // also catches a NaN quotient
          }
          D[j] = w;
        }
        __syncthreads();
        wg_factor<double>(c, A64, W64, sh64, red, prm.reg, false, &reg, ck_off);
        wg_spmv(c.trp, c.tci, c.tv, lam, t, n);
        __syncthreads();
        for (int j = tid; j < n; j += NT) zs[j] = D[j] * (cc[j] - t[j]);
        __syncthreads();
        wg_spmv(c.mrp, c.mci, c.mv, zs, rhs2, m);
        __syncthreads();
        (void)wg_solve<double>(c, A64, rhs2, mdrd, reg, red, &refine, ck_off);
        for (int i = tid; i < m; i += NT) mdrd[i] = lam[i] + mdrd[i];  // y'
        __syncthreads();
        wg_spmv(c.trp, c.tci, c.tv, mdrd, t, n);
        __syncthreads();
        double zneg = 0.0, cx = 0.0, by = 0.0, xz = 0.0, b2 = 0.0;
        for (int j = tid; j < n; j += NT) {
          const double v = cc[j] - t[j];
          dz[j] = v;
          if (!isfinite(v)) b2 = 1.0;
          zneg = nan_ignoring_max(zneg, -v);
          cx += cc[j] * dx[j];
          xz += dx[j] * fmax(v, 0.0);
        }
        for (int i = tid; i < m; i += NT) {
          by += b[i] * mdrd[i];
          if (!isfinite(mdrd[i])) b2 = 1.0;
        }
        b2 = wg_max(b2, red);
        zneg = wg_max(zneg, red);
        cx = wg_sum(cx, red);
        by = wg_sum(by, red);
        xz = wg_sum(xz, red) / (double)n;
        // the reference's stop test on the projected point (rd = max(-z', 0))
        bool ok = b2 == 0.0 && zneg < prm.conv_tol && xz < prm.comp_tol;
        if (ok && prm.gap_tol > 0.0) ok = fabs(cx - by) <= prm.gap_tol * (1.0 + fabs(cx + I.off));
        if (ok) {
          for (int j = tid; j < n; j += NT) {
            x[j] = dx[j];
            z[j] = fmax(dz[j], 0.0);
          }
          for (int i = tid; i < m; i += NT) lam[i] = mdrd[i];
          __syncthreads();
          return true;
        }
      }
      return false;
    };
    double endgame_next = prm.endgame_mu;
    int projected = 0, endgame_iter = 0;
    for (int iter = 0; iter < prm.max_iterations; ++iter) {
      residuals();
      ck.mark(PH_RESID);
      if (rpn < prm.conv_tol && rdn < prm.conv_tol && mu < prm.comp_tol) {
        bool gap_ok = true;
        if (prm.gap_tol > 0.0) {  // optional relative-gap bar (not in the reference)
          double cx = 0.0, by = 0.0;
          for (int j = tid; j < n; j += NT) cx += cc[j] * x[j];
          for (int i = tid; i < m; i += NT) by += b[i] * lam[i];
          cx = wg_sum(cx, red);
          by = wg_sum(by, red);
          gap_ok = fabs(cx - by) <= prm.gap_tol * (1.0 + fabs(cx + I.off));
        }
        ck.mark(PH_RESID);
        if (gap_ok) {
          done = true;
          break;
        }
      }
      if (prm.mu_target > 0.0 && mu <= prm.mu_target) {
        stopped = true;
        break;
      }
      if (endgame_next > 0.0 && rpn < 1e-4 && rdn < 1e-4 &&
          (mu <= endgame_next || (mu <= prm.endgame_mu && iter >= endgame_iter + 4))) {
        if (project()) {  // passed the stop test: the solve ends at the projected point
          residuals();
          ck.mark(PH_PROJECT);
          projected = 1;
          done = true;
          break;
        }
        endgame_next = 0.5 * mu;  // next try once mu has halved, or after 4 iterations
        endgame_iter = iter;
        residuals();  // the attempt overwrote M x (tm) and M'lam (t), which the step reads
        ck.mark(PH_PROJECT);
      }
      // a warm start that lets mu fall below its tolerance while the
      // residuals stay above theirs for kStallIters iterations has stalled:
      // restart once from the standard point
      if (warm_live) {
        stuck = (mu < prm.comp_tol && (rpn >= prm.conv_tol || rdn >= prm.conv_tol)) ? stuck + 1 : 0;
        if (stuck >= kStallIters) {
          warm_live = false;
          restarted = true;
          cold_start();
          residuals();
        }
      }
      if (!fp64 && mu < prm.mu_switch) {
        fp64 = true;
        if (fp64_from < 0) fp64_from = iter;
      }
      {
        double xq[kEachU], zq[kEachU], rq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              xq[u] = x[j];
              zq[u] = z[j];
              rq[u] = rd[j];
            },
            [&](int j, int u) {
              double d = xq[u] / zq[u];
              if (!(d >= 0.0) || !isfinite(d) || d > 1e10) d = 1e10;
              D[j] = d;
              t[j] = d * rq[u];
            });
      }
      __syncthreads();
      wg_spmv(c.mrp, c.mci, c.mv, t, mdrd, m);
      __syncthreads();
      for (int i = tid; i < m; i += NT) rhs[i] = -rp[i] - mdrd[i] + tm[i];
      ck.mark(PH_RHS);
      double reg = 0.0;
      bool ok;
      if (fp64) {
        wg_factor<double>(c, A64, W64, sh64, red, prm.reg, prm.skip != 0, &reg, ck);
        ok = wg_solve<double>(c, A64, rhs, dlam, reg, red, &refine, ck);
      } else {
        wg_factor<float>(c, A32, W32, sh32, red, prm.reg, prm.skip != 0, &reg, ck);
        ok = wg_solve<float>(c, A32, rhs, dlam, reg, red, &refine, ck);
        if (!ok) {  // FP32 factor not refinable: FP64 from here on
          fp64 = true;
          if (fp64_from < 0) fp64_from = iter;
          wg_factor<double>(c, A64, W64, sh64, red, prm.reg, prm.skip != 0, &reg, ck);
          ok = wg_solve<double>(c, A64, rhs, dlam, reg, red, &refine, ck);
        }
      }
      ++ne_solves;
      // predictor direction (:607-627)
      wg_spmv(c.trp, c.tci, c.tv, dlam, t, n);
      __syncthreads();
      int bad = 0;
      {
        double tq[kEachU], rq[kEachU], dq[kEachU], xq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              tq[u] = t[j];
              rq[u] = rd[j];
              dq[u] = D[j];
              xq[u] = x[j];
            },
            [&](int j, int u) {
              const double s = tq[u] + rq[u];
              const double a = dq[u] * s - xq[u];
              dx[j] = a;
              dz[j] = -s;
              if (!isfinite(a) || !isfinite(s)) bad = 1;
            });
      }
      for (int i = tid; i < m; i += NT)
        if (!isfinite(dlam[i])) bad = 1;
      if (wg_max((double)bad, red) > 0.0) {
        numerr = true;
        break;
      }
      const double ap_a = wg_step(x, dx, n, red);
      const double ad_a = wg_step(z, dz, n, red);
      double s = 0.0;
      {
        double xq[kEachU], dxq[kEachU], zq[kEachU], dzq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              xq[u] = x[j];
              dxq[u] = dx[j];
              zq[u] = z[j];
              dzq[u] = dz[j];
            },
            [&](int, int u) { s += (xq[u] + ap_a * dxq[u]) * (zq[u] + ad_a * dzq[u]); });
      }
      const double mu_aff = wg_sum(s, red) / (double)n;
      double sigma;
      if (!isfinite(mu_aff) || mu_aff < 0.0 || !(mu > 0.0)) {
        sigma = 0.5;
      } else {
        sigma = pow(mu_aff / fmax(mu, 1e-300), 3.0);
        sigma = sigma < 0.1 ? 0.1 : (sigma > 0.9 ? 0.9 : sigma);
      }
      const double smu = sigma * mu;
      {
        double zq[kEachU], dxq[kEachU], dzq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              zq[u] = z[j];
              dxq[u] = dx[j];
              dzq[u] = dz[j];
            },
            [&](int j, int u) {
              const double zinv = 1.0 / zq[u];
              zs[j] = (smu - dxq[u] * dzq[u]) * zinv;
            });
      }
      __syncthreads();
      wg_spmv(c.mrp, c.mci, c.mv, zs, mz, m);
      __syncthreads();
      for (int i = tid; i < m; i += NT) rhs2[i] = rhs[i] - mz[i];
      __syncthreads();
      ck.mark(PH_STEP);
      if (fp64)
        ok = wg_solve<double>(c, A64, rhs2, dlam, reg, red, &refine, ck);
      else
        ok = wg_solve<float>(c, A32, rhs2, dlam, reg, red, &refine, ck);
      ++ne_solves;
      // corrector direction (:666-685)
      wg_spmv(c.trp, c.tci, c.tv, dlam, t, n);
      __syncthreads();
      bad = 0;
      {
        double tq[kEachU], rq[kEachU], dq[kEachU], xq[kEachU], sq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              tq[u] = t[j];
              rq[u] = rd[j];
              dq[u] = D[j];
              xq[u] = x[j];
              sq[u] = zs[j];
            },
            [&](int j, int u) {
              const double s2 = tq[u] + rq[u];
              const double a = dq[u] * s2 - xq[u] + sq[u];
              dx[j] = a;
              dz[j] = -s2;
              if (!isfinite(a) || !isfinite(s2)) bad = 1;
            });
      }
      for (int i = tid; i < m; i += NT)
        if (!isfinite(dlam[i])) bad = 1;
      if (wg_max((double)bad, red) > 0.0) {
        numerr = true;
        break;
      }
      const double ap = wg_step(x, dx, n, red);
      const double ad = wg_step(z, dz, n, red);
      bad = 0;
      {
        double xq[kEachU], dxq[kEachU], zq[kEachU], dzq[kEachU];
        wg_each(
            n,
            [&](int j, int u) {
              xq[u] = x[j];
              dxq[u] = dx[j];
              zq[u] = z[j];
              dzq[u] = dz[j];
            },
            [&](int j, int u) {
              const double xi = xq[u] + ap * dxq[u];
              const double zi = zq[u] + ad * dzq[u];
              x[j] = xi;
              z[j] = zi;
              if (xi < 0.0 || zi < 0.0 || !isfinite(xi) || !isfinite(zi)) bad = 1;
            });
      }
      for (int i = tid; i < m; i += NT) lam[i] = lam[i] + ad * dlam[i];
      if (wg_max((double)bad, red) > 0.0) {
        numerr = true;
        break;
      }
      iterations = iter + 1;
      ck.mark(PH_STEP);
    }
    if (!done && !stopped) {  // :728-753
      residuals();
      if (rpn < 1e-4 && rdn < 1e-4 && mu < 1e-4) done = true;
      ck.mark(PH_RESID);
    }
    if (tid == 0) {
      Out o;
      o.status = numerr ? HN_STATUS_NUMERICAL_ERROR
                 : done ? HN_STATUS_OPTIMAL
                 : stopped ? HN_STATUS_ITER_LIMIT
                 : (rpn <= 1e-9) ? HN_STATUS_SUBOPTIMAL
                                 : HN_STATUS_ITER_LIMIT;
      o.iterations = iterations;
      o.fp64_from = fp64_from;
      o.ne_solves = ne_solves;
      o.refine = refine;
      o.restarted = restarted ? 1 : 0;
      o.accepted = accepted ? 1 : 0;
      o.projected = projected;
      o.rp = rpn;
      o.rd = rdn;
      o.mu = mu;
      for (int p = 0; p < kPhases; ++p) o.ph[p] = prm.timing ? s_ph[p] : 0ULL;
      out[id] = o;
    }
    __syncthreads();
  }
}

// Counting-sort transpose on the host (entries of each column in row order,
// as build_normal_equations scatters them, interior_point.cpp:153-173).
// CSR of M' by a counting scatter in row order (stable), written into
// caller-provided slices: trp (n+1), tci and tv (nnz).
void host_transpose(const StandardForm& f, int* trp, int* tci, double* tv) {
  std::fill(trp, trp + f.n + 1, 0);
  for (int64_t k = 0; k < (int64_t)f.ci.size(); ++k) trp[f.ci[k] + 1]++;
  for (int64_t j = 0; j < f.n; ++j) trp[j + 1] += trp[j];
  std::vector<int> cur(trp, trp + f.n);
  for (int64_t i = 0; i < f.m; ++i)
    for (int64_t k = f.rp[i]; k < f.rp[i + 1]; ++k) {
      const int p = cur[f.ci[k]]++;
      tci[p] = (int)i;
      tv[p] = f.v[k];
    }
}

// Page-locked host staging that only grows (hipHostMalloc; copies from it
// run at the host link's DMA rate).
struct PinnedBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t want) {
    if (want <= bytes) return HN_OK;
    if (p) hipHostFree(p);
    p = nullptr;
    bytes = 0;
    HN_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
    bytes = want;
    return HN_OK;
  }
  ~PinnedBuf() {
    if (p) hipHostFree(p);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Runs f(k) for k in [0, count) on up to 16 host threads.
template <typename F>
void parallel_for(uint32_t count, F&& f) {
  const unsigned nth = std::max(1u, std::min({16u, std::thread::hardware_concurrency(), count}));
  if (nth <= 1) {
    for (uint32_t k = 0; k < count; ++k) f(k);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned w = 0; w < nth; ++w)
    th.emplace_back([&, w]() {
      for (uint32_t k = w; k < count; k += nth) f(k);
    });
  for (auto& t : th) t.join();
}

struct BatchWorkspace {
  std::mutex mtx;
  DevBuf inst, order, queue, mrp, mci, mv, trp, tci, tv, vecm, vecn, bio, lio, cio, xzio, dense, work, out;
  PinnedBuf h_mrp, h_mci, h_mv, h_trp, h_tci, h_tv, h_bio, h_lio, h_cio, h_xzio;
  hipEvent_t ev[2] = {nullptr, nullptr};  // brackets the kernel for the verbose timing line
};
BatchWorkspace& bws() {
  static BatchWorkspace* w = new BatchWorkspace();
  return *w;
}

template <typename V>
int upload(DevBuf& d, const std::vector<V>& h, hipStream_t s) {
  HN_TRY(d.reserve(sizeof(V) * std::max<size_t>(h.size(), 1)));
  if (!h.empty()) HN_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
  return HN_OK;
}

// Uploads n elements of pinned staging into d (grown as needed).
template <typename V>
int upload_pinned(DevBuf& d, const PinnedBuf& h, long long n, hipStream_t s) {
  HN_TRY(d.reserve(sizeof(V) * std::max<long long>(n, 1)));
  if (n > 0) HN_HIP(hipMemcpyAsync(d.p, h.p, sizeof(V) * n, hipMemcpyHostToDevice, s));
  return HN_OK;
}

}  // namespace

namespace {

// The autotuned table of the device the library runs on (hn_batch_autotune).
struct BatchTuning {
  std::mutex mtx;
  bool valid = false;
  hn_batch_tuning t{};
};
BatchTuning& batch_tuning() {
  static BatchTuning* b = new BatchTuning();  // reset by hn_finalize
  return *b;
}
constexpr int kTuneEdges[HN_TUNE_BUCKETS] = {128, 256, 512, 1024, 2048, 20000};
int tune_bucket(int64_t m) {
  int b = 0;
  while (b + 1 < HN_TUNE_BUCKETS && m > kTuneEdges[b]) ++b;
  return b;
}

// 1 when LP k (standard-form m rows) runs with the FP32 factor.
int instance_mixed(int precision, int64_t m, const hn_batch_tuning* table) {
  if (precision == HN_PREC_MIXED) return 1;
  if (precision != HN_PREC_AUTO || !table) return 0;
  return table->precision[tune_bucket(m)] == HN_PREC_MIXED ? 1 : 0;
}

}  // namespace

void reset_batch_tuning() {
  BatchTuning& bt = batch_tuning();
  std::lock_guard<std::mutex> lk(bt.mtx);
  bt.valid = false;
}


int ipm_solve_batch_impl(const hn_lp* lps, uint32_t count, const hn_ipm_opts& o, hn_ipm_result* res,
                         hn_gnn* net = nullptr, hn_gnn_graph* graph = nullptr) {
  const auto t0 = std::chrono::steady_clock::now();
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  if (count == 0) return HN_OK;
  // host preparation, in parallel over LPs
  std::vector<StandardForm> f(count);
  std::vector<int> st(count, HN_OK);
  std::vector<int> warm_ok(count, 0), predict(count, 0);
  std::vector<std::string> err(count);
  parallel_for(count, [&](uint32_t k) {
    st[k] = build_standard_form(lps[k], f[k]);
    if (st[k] != HN_OK) {
      err[k] = get_error();
      return;
    }
    if (lps[k].x0 && lps[k].y0 && lps[k].s0)
      warm_ok[k] = warm_start_acceptable(f[k], lps[k].x0, lps[k].y0, lps[k].s0) ? 1 : 0;
    else if (net && gnn_pattern_matches(graph, f[k].m, f[k].n, f[k].rp, f[k].ci))
      predict[k] = 1;  // the network predicts its start on the device
  });
  const auto t_prep = std::chrono::steady_clock::now();
  for (uint32_t k = 0; k < count; ++k) {
    if (st[k] != HN_OK) return fail(st[k], "batch LP " + std::to_string(k) + ": " + err[k]);
    if (f[k].m == 0 || f[k].n == 0 || lps[k].n == 0)
      return fail(HN_ERR_UNSUPPORTED, "batch LP " + std::to_string(k) + " has an empty standard form");
    if (f[k].m > 20000) return fail(HN_ERR_UNSUPPORTED, "batch LPs are limited to 20000 standard-form rows");
  }
  // the autotuned precision per m bucket, when asked for
  hn_batch_tuning table{};
  bool have_table = false;
  if (o.precision == HN_PREC_AUTO) {
    BatchTuning& bt = batch_tuning();
    std::lock_guard<std::mutex> lk(bt.mtx);
    have_table = bt.valid;
    table = bt.t;
  }
  // pack: offsets first, then every LP's slices filled in parallel straight
  // into page-locked staging
  std::vector<Inst> inst(count);
  long long n_mp = 0, n_mz = 0, n_tp = 0, n_vm = 0, n_vn = 0, n_om = 0, n_on = 0, dn = 0, wk = 0;
  bool any_warm = false;
  for (uint32_t k = 0; k < count; ++k) {
    const StandardForm& F = f[k];
    Inst& I = inst[k];
    I.m = (int)F.m;
    I.n = (int)F.n;
    I.mp = n_mp;
    n_mp += F.m + 1;
    I.mz = n_mz;
    I.tz = n_mz;
    n_mz += (long long)F.ci.size();
    I.tp = n_tp;
    n_tp += F.n + 1;
    I.vm = n_vm;
    n_vm += (long long)(VM_COUNT - kVmIo) * F.m;
    I.vn = n_vn;
    n_vn += (long long)(VN_COUNT - kVnIo) * F.n;
    I.om = n_om;
    n_om += F.m;
    I.on = n_on;
    n_on += F.n;
    I.dn = dn;
    dn += (long long)F.m * F.m;
    I.w = wk;
    wk += (long long)F.m * kLdw;
    I.warm = warm_ok[k] ? 1 : (predict[k] ? 2 : 0);
    I.mixed = instance_mixed(o.precision, F.m, have_table ? &table : nullptr);
    I.off = F.offset;
    any_warm = any_warm || warm_ok[k];
  }
  // every instance's slices lie inside the arenas sized above (checked on
  // the host before anything is launched)
  for (uint32_t k = 0; k < count; ++k) {
    const Inst& I = inst[k];
    const long long m = f[k].m, n = f[k].n, nz = (long long)f[k].ci.size();
    const bool inside = I.mp >= 0 && I.mp + m + 1 <= n_mp && I.mz >= 0 && I.mz + nz <= n_mz && I.tp >= 0 &&
                        I.tp + n + 1 <= n_tp && I.tz == I.mz && I.vm >= 0 &&
                        I.vm + (long long)(VM_COUNT - kVmIo) * m <= n_vm && I.vn >= 0 &&
                        I.vn + (long long)(VN_COUNT - kVnIo) * n <= n_vn && I.om >= 0 && I.om + m <= n_om &&
                        I.on >= 0 && I.on + n <= n_on && I.dn >= 0 && I.dn + m * m <= dn && I.w >= 0 &&
                        I.w + m * kLdw <= wk;
    HN_CHECK(inside, HN_ERR_INVALID, "batch LP " + std::to_string(k) + ": a slice lies outside its arena");
  }
  BatchWorkspace& w = bws();
  std::lock_guard<std::mutex> lk(w.mtx);
  HN_TRY(w.h_mrp.reserve(sizeof(int) * n_mp));
  HN_TRY(w.h_mci.reserve(sizeof(int) * std::max(n_mz, 1LL)));
  HN_TRY(w.h_mv.reserve(sizeof(double) * std::max(n_mz, 1LL)));
  HN_TRY(w.h_trp.reserve(sizeof(int) * n_tp));
  HN_TRY(w.h_tci.reserve(sizeof(int) * std::max(n_mz, 1LL)));
  HN_TRY(w.h_tv.reserve(sizeof(double) * std::max(n_mz, 1LL)));
  HN_TRY(w.h_bio.reserve(sizeof(double) * n_om));
  HN_TRY(w.h_lio.reserve(sizeof(double) * n_om));
  HN_TRY(w.h_cio.reserve(sizeof(double) * n_on));
  HN_TRY(w.h_xzio.reserve(sizeof(double) * 2 * n_on));
  {
    int* mrp = w.h_mrp.as<int>();
    int* mci = w.h_mci.as<int>();
    double* mv = w.h_mv.as<double>();
    int* trp = w.h_trp.as<int>();
    int* tci = w.h_tci.as<int>();
    double* tv = w.h_tv.as<double>();
    double* bio = w.h_bio.as<double>();
    double* lio = w.h_lio.as<double>();
    double* cio = w.h_cio.as<double>();
    double* xzio = w.h_xzio.as<double>();
    parallel_for(count, [&](uint32_t k) {
      const StandardForm& F = f[k];
      const Inst& I = inst[k];
      for (int64_t i = 0; i <= F.m; ++i) mrp[I.mp + i] = (int)F.rp[i];
      for (size_t q = 0; q < F.ci.size(); ++q) mci[I.mz + q] = (int)F.ci[q];
      std::memcpy(mv + I.mz, F.v.data(), sizeof(double) * F.v.size());
      host_transpose(F, trp + I.tp, tci + I.tz, tv + I.tz);
      std::memcpy(bio + I.om, F.rhs.data(), sizeof(double) * F.m);
      std::memcpy(cio + I.on, F.c.data(), sizeof(double) * F.n);
      if (I.warm == 1) {  // the caller's start (predicted starts are written on the device)
        std::memcpy(xzio + 2 * I.on, lps[k].x0, sizeof(double) * F.n);
        std::memcpy(xzio + 2 * I.on + F.n, lps[k].s0, sizeof(double) * F.n);
        std::memcpy(lio + I.om, lps[k].y0, sizeof(double) * F.m);
      }
    });
  }
  const auto t_pack = std::chrono::steady_clock::now();
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].m > inst[b].m; });
  HN_TRY(upload(w.inst, inst, s));
  HN_TRY(upload(w.order, order, s));
  HN_TRY(upload_pinned<int>(w.mrp, w.h_mrp, n_mp, s));
  HN_TRY(upload_pinned<int>(w.mci, w.h_mci, n_mz, s));
  HN_TRY(upload_pinned<double>(w.mv, w.h_mv, n_mz, s));
  HN_TRY(upload_pinned<int>(w.trp, w.h_trp, n_tp, s));
  HN_TRY(upload_pinned<int>(w.tci, w.h_tci, n_mz, s));
  HN_TRY(upload_pinned<double>(w.tv, w.h_tv, n_mz, s));
  HN_TRY(upload_pinned<double>(w.bio, w.h_bio, n_om, s));
  HN_TRY(upload_pinned<double>(w.cio, w.h_cio, n_on, s));
  HN_TRY(w.lio.reserve(sizeof(double) * n_om));
  HN_TRY(w.xzio.reserve(sizeof(double) * 2 * n_on));
  if (any_warm) {  // cold LPs' entries are overwritten by the kernel's start point
    HN_HIP(hipMemcpyAsync(w.lio.p, w.h_lio.p, sizeof(double) * n_om, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(w.xzio.p, w.h_xzio.p, sizeof(double) * 2 * n_on, hipMemcpyHostToDevice, s));
  }
  {  // learned starts, predicted in place from the packed standard forms
    std::vector<GnnJob> jobs;
    for (uint32_t k = 0; k < count; ++k) {
      if (inst[k].warm != 2) continue;
      const Inst& I = inst[k];
      double* xz = w.xzio.as<double>() + 2 * I.on;
      jobs.push_back({w.mv.as<double>() + I.mz, w.bio.as<double>() + I.om, w.cio.as<double>() + I.on, xz,
                      w.lio.as<double>() + I.om, xz + I.n});
    }
    HN_TRY(gnn_warm_start_many(net, graph, jobs, s));
  }
  HN_TRY(w.vecm.reserve(sizeof(double) * std::max(n_vm, 1LL)));
  HN_TRY(w.vecn.reserve(sizeof(double) * std::max(n_vn, 1LL)));
  HN_HIP(hipMemsetAsync(w.vecm.p, 0, sizeof(double) * n_vm, s));
  HN_HIP(hipMemsetAsync(w.vecn.p, 0, sizeof(double) * n_vn, s));
  HN_TRY(w.dense.reserve(sizeof(double) * std::max<long long>(dn, 1)));
  HN_TRY(w.work.reserve(sizeof(double) * std::max<long long>(wk, 1)));
  HN_TRY(w.out.reserve(sizeof(Out) * count));
  HN_TRY(w.queue.reserve(sizeof(int)));
  HN_HIP(hipMemsetAsync(w.queue.p, 0, sizeof(int), s));
  for (hipEvent_t& e : w.ev)
    if (!e) HN_HIP(hipEventCreate(&e));
  Params prm;
  prm.max_iterations = o.max_iterations;
  prm.conv_tol = o.convergence_tol;
  prm.comp_tol = o.complementarity_tol;
  prm.reg = o.regularization;
  prm.mu_switch = o.mu_switch;
  prm.skip = o.pivot_rule == HN_PIVOT_SKIP;
  prm.mu_target = o.mu_target;
  prm.gap_tol = o.gap_tol;
  prm.endgame_mu = o.endgame_mu;
  prm.timing = o.verbose >= 2 ? 1 : 0;
  const int grid = (int)std::min<uint32_t>(count, 2u * (uint32_t)std::max(1, ctx().info.compute_units));
  HN_HIP(hipEventRecord(w.ev[0], s));
  hipLaunchKernelGGL(k_batch_ipm, dim3(grid), dim3(NT), 0, s, w.inst.as<Inst>(), w.order.as<int>(), (int)count,
                     w.queue.as<int>(), w.mrp.as<int>(), w.mci.as<int>(), w.mv.as<double>(), w.trp.as<int>(),
                     w.tci.as<int>(), w.tv.as<double>(), w.vecm.as<double>(), w.vecn.as<double>(),
                     w.bio.as<double>(), w.lio.as<double>(), w.cio.as<double>(), w.xzio.as<double>(),
                     w.dense.as<double>(), w.work.as<double>(), prm, w.out.as<Out>());
  HN_TRY(check_launch("k_batch_ipm"));
  count_launch();
  HN_HIP(hipEventRecord(w.ev[1], s));
  std::vector<Out> out(count);
  HN_HIP(hipMemcpyAsync(out.data(), w.out.p, sizeof(Out) * count, hipMemcpyDeviceToHost, s));
  HN_HIP(hipMemcpyAsync(w.h_lio.p, w.lio.p, sizeof(double) * n_om, hipMemcpyDeviceToHost, s));
  HN_HIP(hipMemcpyAsync(w.h_xzio.p, w.xzio.p, sizeof(double) * 2 * n_on, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  const auto t_dev = std::chrono::steady_clock::now();
  const double ms = 1e3 * std::chrono::duration<double>(t_dev - t0).count();
  // results (objective as interior_point.cpp:757-780)
  const double* lio = w.h_lio.as<double>();
  const double* xzio = w.h_xzio.as<double>();
  parallel_for(count, [&](uint32_t k) {
    const StandardForm& F = f[k];
    const Inst& I = inst[k];
    hn_ipm_result& r = res[k];
    const double* x = xzio + 2 * I.on;
    const double* z = x + I.n;
    const double* lam = lio + I.om;
    std::vector<double> primal(lps[k].n);
    to_original_primal(F, x, primal.data());
    double obj = 0.0;
    bool ok = std::isfinite(lps[k].obj_offset);
    for (int64_t j = 0; j < lps[k].n && ok; ++j) {
      const double term = lps[k].obj[j] * primal[j];
      if (!std::isfinite(term)) ok = false;
      obj += term;
    }
    if (ok) obj += lps[k].obj_offset;
    r.status = ok ? out[k].status : HN_STATUS_NUMERICAL_ERROR;
    r.iterations = out[k].iterations;
    r.objective = ok ? obj : 0.0;
    r.solve_ms = ms;
    r.sf_m = F.m;
    r.sf_n = F.n;
    r.rp_norm = out[k].rp;
    r.rd_norm = out[k].rd;
    r.mu = out[k].mu;
    scaled_kkt(F, x, lam, out[k].rp, out[k].rd, r);
    OriginalKkt kk;
    original_kkt(lps[k], F, x, lam, z, kk);
    r.rel_primal = kk.rel_primal;
    r.rel_dual = kk.rel_dual;
    r.rel_gap = kk.rel_gap;
    r.ne_solves = out[k].ne_solves;
    r.refine_sweeps = out[k].refine;
    r.cg_iterations = 0;
    r.fp64_from_iter = out[k].fp64_from;
    r.warm_started = out[k].accepted ? (out[k].restarted ? 2 : 1) : 0;
    r.projected = out[k].projected;
    r.sparse_l_nnz = 0;
    r.device_steps = r.host_redos = 0;
    if (r.primal) std::memcpy(r.primal, primal.data(), sizeof(double) * lps[k].n);
    if (r.x && r.cap_n >= F.n) std::memcpy(r.x, x, sizeof(double) * F.n);
    if (r.s && r.cap_n >= F.n) std::memcpy(r.s, z, sizeof(double) * F.n);
    if (r.y && r.cap_m >= F.m) std::memcpy(r.y, lam, sizeof(double) * F.m);
  });
  if (o.verbose) {
    float kms = 0.f;
    HN_HIP(hipEventElapsedTime(&kms, w.ev[0], w.ev[1]));
    auto span = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
      return 1e3 * std::chrono::duration<double>(b - a).count();
    };
    std::fprintf(stderr,
                 "[hn batch] %u LPs: standard forms %.1f ms, packing %.1f ms, upload+kernel+download %.1f ms "
                 "(kernel %.1f ms), results %.1f ms\n",
                 count, span(t0, t_prep), span(t_prep, t_pack), span(t_pack, t_dev), (double)kms,
                 span(t_dev, std::chrono::steady_clock::now()));
    if (prm.timing) {  // workgroup time per phase, summed over the LPs
      int khz = 0;
      HN_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx().device));
      static const char* names[kPhases] = {"start", "residuals", "projection", "rhs",  "assembly", "factor",
                                           "diag",  "panel",     "update",     "trsv", "refine",   "step"};
      double tot[kPhases] = {0.0}, all = 0.0;
      long long its = 0;
      for (uint32_t k = 0; k < count; ++k) {
        its += out[k].iterations;
        for (int p = 0; p < kPhases; ++p) tot[p] += (double)out[k].ph[p];
      }
      for (int p = 0; p < kPhases; ++p) all += tot[p];
      std::fprintf(stderr, "[hn batch] workgroup time %.1f ms over %lld iterations (clock %d kHz):", all / khz, its,
                   khz);
      for (int p = 0; p < kPhases; ++p)
        std::fprintf(stderr, " %s %.1f%%", names[p], all > 0.0 ? 100.0 * tot[p] / all : 0.0);
      std::fprintf(stderr, "\n");
    }
  }
  return HN_OK;
}

}  // namespace hn

extern "C" int hn_ipm_solve_batch(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, hn_ipm_result* results) {
  HN_CHECK(count == 0 || (lps && results), HN_ERR_INVALID, "hn_ipm_solve_batch: null argument");
  hn_ipm_opts o;
  if (opts)
    o = *opts;
  else
    hn_ipm_default_opts(&o);
  HN_CHECK(o.precision == HN_PREC_F64 || o.precision == HN_PREC_MIXED || o.precision == HN_PREC_AUTO,
           HN_ERR_INVALID, "hn_ipm_solve_batch: bad precision");
  HN_CHECK(o.pivot_rule == HN_PIVOT_CLAMP || o.pivot_rule == HN_PIVOT_SKIP, HN_ERR_INVALID,
           "hn_ipm_solve_batch: bad pivot_rule");
  return hn::ipm_solve_batch_impl(lps, count, o, results);
}

extern "C" int hn_ipm_solve_batch_gnn(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, hn_gnn* net,
                                      hn_gnn_graph* graph, hn_ipm_result* results) {
  HN_CHECK(count == 0 || (lps && results), HN_ERR_INVALID, "hn_ipm_solve_batch_gnn: null argument");
  HN_CHECK(net && graph, HN_ERR_INVALID, "hn_ipm_solve_batch_gnn: null network or graph");
  hn_ipm_opts o;
  if (opts)
    o = *opts;
  else
    hn_ipm_default_opts(&o);
  HN_CHECK(o.precision == HN_PREC_F64 || o.precision == HN_PREC_MIXED || o.precision == HN_PREC_AUTO,
           HN_ERR_INVALID, "hn_ipm_solve_batch_gnn: bad precision");
  HN_CHECK(o.pivot_rule == HN_PIVOT_CLAMP || o.pivot_rule == HN_PIVOT_SKIP, HN_ERR_INVALID,
           "hn_ipm_solve_batch_gnn: bad pivot_rule");
  return hn::ipm_solve_batch_impl(lps, count, o, results, net, graph);
}

extern "C" int hn_batch_autotune(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, uint32_t per_bucket,
                                 hn_batch_tuning* out) {
  using namespace hn;
  HN_CHECK(count > 0 && lps, HN_ERR_INVALID, "hn_batch_autotune: empty sample");
  HN_CHECK(per_bucket > 0, HN_ERR_INVALID, "hn_batch_autotune: per_bucket must be positive");
  hn_ipm_opts o;
  if (opts)
    o = *opts;
  else
    hn_ipm_default_opts(&o);
  HN_CHECK(o.pivot_rule == HN_PIVOT_CLAMP || o.pivot_rule == HN_PIVOT_SKIP, HN_ERR_INVALID,
           "hn_batch_autotune: bad pivot_rule");
  o.verbose = 0;
  // bucket the sample by standard-form rows
  std::vector<std::vector<hn_lp>> bucket(HN_TUNE_BUCKETS);
  for (uint32_t k = 0; k < count; ++k) {
    StandardForm f;
    HN_TRY(build_standard_form(lps[k], f));
    std::vector<hn_lp>& b = bucket[tune_bucket(f.m)];
    if (b.size() < per_bucket) b.push_back(lps[k]);
  }
  std::vector<hn_ipm_result> res(count);
  std::vector<std::vector<double>> primal(count);
  auto solve = [&](const std::vector<hn_lp>& set, int precision, double* ms) -> int {
    for (size_t k = 0; k < set.size(); ++k) {
      primal[k].assign(std::max<int64_t>(set[k].n, 1), 0.0);
      std::memset(&res[k], 0, sizeof(hn_ipm_result));
      res[k].primal = primal[k].data();
    }
    hn_ipm_opts oo = o;
    oo.precision = precision;
    const auto t0 = std::chrono::steady_clock::now();
    HN_TRY(ipm_solve_batch_impl(set.data(), (uint32_t)set.size(), oo, res.data()));
    *ms = 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return HN_OK;
  };
  {  // warm-up: grows the arenas and staging to the whole sample's size
    std::vector<hn_lp> all;
    for (const auto& b : bucket) all.insert(all.end(), b.begin(), b.end());
    double ms = 0.0;
    HN_TRY(solve(all, HN_PREC_F64, &ms));
  }
  hn_batch_tuning t{};
  for (int b = 0; b < HN_TUNE_BUCKETS; ++b) {
    t.m_hi[b] = kTuneEdges[b];
    t.lps[b] = (int)bucket[b].size();
    t.precision[b] = HN_PREC_F64;
    if (bucket[b].empty()) continue;
    HN_TRY(solve(bucket[b], HN_PREC_F64, &t.ms_f64[b]));
    HN_TRY(solve(bucket[b], HN_PREC_MIXED, &t.ms_mixed[b]));
    if (t.ms_mixed[b] < t.ms_f64[b]) t.precision[b] = HN_PREC_MIXED;
  }
  {
    BatchTuning& bt = batch_tuning();
    std::lock_guard<std::mutex> lk(bt.mtx);
    bt.t = t;
    bt.valid = true;
  }
  if (out) *out = t;
  return HN_OK;
}
