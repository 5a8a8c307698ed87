// CSR sparse matrix-vector products for gfx950.
//
// Reference semantics: y[i] = sum_k val[k] * x[col[k]] over row i
// (CPUBackend::spmv, core/execution/gpu_backend.cpp:89-103; matvec,
// core/lp/interior_point.cpp:52-59).  Empty rows give 0.
//
// Design (merge-path tiles): the sequence of row ends and nonzeros is split
// into tiles of 256*IPT items, so every workgroup does the same amount of
// work whatever the row-length skew.  Per tile:
//   1. the tile's nonzeros are streamed with coalesced non-temporal loads
//      (val/col are touched once; keeping them out of L2 leaves L2 to the
//      randomly gathered x) and their products staged in LDS, together with
//      the tile's row ends;
//   2. each thread walks IPT items of the merge path sequentially, writing
//      rows it owns completely;
//   3. the partial sums of rows that straddle threads are combined with a
//      wave64 segmented scan (shuffles) plus a 4-entry cross-wave pass;
//   4. a row that straddles tiles leaves a (row, partial) carry; a tiny
//      fixup kernel adds each run of carries to y in tile order.
// Everything is deterministic for a fixed tiling.
#include "hn_device.hpp"
#include "hn_internal.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

namespace hn {

// ---- tiling -------------------------------------------------------------
__global__ void k_tile_coords(const int* __restrict__ rowptr, int nrows, int nnz, int items_per_tile,
                              int ntiles, int2* __restrict__ coord) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > ntiles) return;
  const long long total = (long long)nrows + nnz;
  long long dl = (long long)t * items_per_tile;
  if (dl > total) dl = total;
  const int d = (int)dl;
  // Smallest i with row_end[i] > d - i - 1, row_end[i] = rowptr[i + 1].
  int lo = max(d - nnz, 0), hi = min(d, nrows);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (rowptr[mid + 1] <= d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  coord[t] = make_int2(lo, d - lo);
}

int build_tiling(hn_csr* a, int ipt, hipStream_t s) {
  HN_CHECK(ipt == 4 || ipt == 8 || ipt == 16, HN_ERR_INVALID, "items_per_thread must be 4, 8 or 16");
  Tiling& tl = a->tiling;
  const int64_t items = a->nrows + a->nnz;
  const int64_t per_tile = (int64_t)kSpmvBlock * ipt;
  const int64_t ntiles = std::max<int64_t>(1, (items + per_tile - 1) / per_tile);
  HN_CHECK(ntiles < INT_MAX / 2, HN_ERR_UNSUPPORTED, "too many SpMV tiles");
  HN_TRY(tl.coord.reserve(sizeof(int2) * (ntiles + 1)));
  HN_TRY(tl.carry_row.reserve(sizeof(int) * ntiles));
  HN_TRY(tl.carry_val64.reserve(sizeof(double) * ntiles));
  HN_TRY(tl.carry_val32.reserve(sizeof(float) * ntiles));
  const int nt = (int)ntiles;
  hipLaunchKernelGGL(k_tile_coords, dim3(grid_for(nt + 1, 256)), dim3(256), 0, s, a->row_ptr.as<int>(),
                     (int)a->nrows, (int)a->nnz, (int)per_tile, nt, tl.coord.as<int2>());
  HN_TRY(check_launch("k_tile_coords"));
  tl.ipt = ipt;
  tl.ntiles = nt;
  return HN_OK;
}

// ---- SpMV main kernel ----------------------------------------------------
template <typename V, int IPT>
__global__ __launch_bounds__(kSpmvBlock) void k_spmv_merge(
    const int* __restrict__ rowptr, const int* __restrict__ col, const V* __restrict__ val,
    const V* __restrict__ x, V* __restrict__ y, const int2* __restrict__ coord,
    int* __restrict__ carry_row, V* __restrict__ carry_val, int nrows) {
  constexpr int NT = kSpmvBlock;
  constexpr int T = NT * IPT;
  constexpr int NW = NT / kWave;
  __shared__ int s_end[T + 1];
  __shared__ V s_prod[T];
  __shared__ int s_wkey[NW], s_wfirst[NW];
  __shared__ V s_wval[NW];

  const int tile = blockIdx.x;
  const int tid = threadIdx.x;
  const int2 c0 = coord[tile];
  const int2 c1 = coord[tile + 1];
  const int r0 = c0.x, e0 = c0.y;
  const int nr = c1.x - r0;  // row ends inside the tile
  const int ne = c1.y - e0;  // nonzeros inside the tile

#pragma unroll
  for (int u = 0; u <= IPT; ++u) {
    const int i = tid + u * NT;
    if (i <= nr) {
      const int r = r0 + i;
      s_end[i] = (r < nrows) ? rowptr[r + 1] : INT_MAX;
    }
  }
#pragma unroll
  for (int u = 0; u < IPT; ++u) {
    const int k = tid + u * NT;
    if (k < ne) {
      const int e = e0 + k;
      const int c = __builtin_nontemporal_load(col + e);
      const V v = __builtin_nontemporal_load(val + e);
      s_prod[k] = v * x[c];
    }
  }
  __syncthreads();

  // Locate this thread's start on the tile's merge path.
  const int items = nr + ne;
  const int d = min(tid * IPT, items);
  int lo = max(d - ne, 0), hi = min(d, nr);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (s_end[mid] <= e0 + d - mid - 1)
      lo = mid + 1;
    else
      hi = mid;
  }
  int i = lo, j = d - lo;
  const int dend = min(d + IPT, items);
  V acc = 0;
  int first_i = -1;
  V first_v = 0;
#pragma unroll
  for (int u = 0; u < IPT; ++u) {
    if (d + u < dend) {
      if (e0 + j < s_end[i]) {
        acc += s_prod[j];
        ++j;
      } else {
        if (first_i < 0) {
          first_i = i;
          first_v = acc;
        } else {
          y[r0 + i] = acc;
        }
        acc = 0;
        ++i;
      }
    }
  }

  // Block-wide inclusive segmented scan of the per-thread partial rows.
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const int key = i;
  V v = wave_segscan(key, acc);
  if (lane == kWave - 1) {
    s_wkey[w] = key;
    s_wval[w] = v;
  }
  if (lane == 0) s_wfirst[w] = key;
  __syncthreads();
  int kin = -1;
  V cin = 0;
  for (int u = 0; u < w; ++u) {
    const int ku = s_wkey[u];
    V vu = s_wval[u];
    if (s_wfirst[u] == ku && kin == ku) vu += cin;
    kin = ku;
    cin = vu;
  }
  if (key == kin) v += cin;

  // Value of the previous thread's (scanned) partial row.
  int pkey = __shfl_up(key, 1, kWave);
  V pv = __shfl_up(v, 1, kWave);
  if (lane == 0) {
    pkey = kin;
    pv = cin;
  }
  if (first_i >= 0) {
    const V carry_in = (tid > 0 && pkey == first_i) ? pv : V(0);
    y[r0 + first_i] = first_v + carry_in;
  }
  if (tid == NT - 1) {
    const int r = r0 + key;
    carry_row[tile] = (r < nrows) ? r : -1;
    carry_val[tile] = v;
  }
}

template <typename V>
__global__ void k_spmv_fixup(const int* __restrict__ carry_row, const V* __restrict__ carry_val,
                             int ntiles, V* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int r = carry_row[t];
  if (r < 0) return;
  if (t > 0 && carry_row[t - 1] == r) return;  // not the first tile of the run
  V s = carry_val[t];
  for (int u = t + 1; u < ntiles && carry_row[u] == r; ++u) s += carry_val[u];
  y[r] += s;
}

template <typename V>
__global__ void k_fill(V* __restrict__ p, V v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename V, int IPT>
static int spmv_run(const hn_csr* a, const V* val, const V* x, V* y, hipStream_t s) {
  const Tiling& tl = a->tiling;
  V* cval = std::is_same<V, double>::value ? (V*)tl.carry_val64.p : (V*)tl.carry_val32.p;
  hipLaunchKernelGGL((k_spmv_merge<V, IPT>), dim3(tl.ntiles), dim3(kSpmvBlock), 0, s,
                     a->row_ptr.as<int>(), a->col.as<int>(), val, x, y, tl.coord.as<int2>(),
                     tl.carry_row.as<int>(), cval, (int)a->nrows);
  HN_TRY(check_launch("k_spmv_merge"));
  hipLaunchKernelGGL((k_spmv_fixup<V>), dim3(grid_for(tl.ntiles, 256)), dim3(256), 0, s,
                     tl.carry_row.as<int>(), (const V*)cval, tl.ntiles, y);
  HN_TRY(check_launch("k_spmv_fixup"));
  count_launch(2);
  return HN_OK;
}

template <typename V>
static int spmv_typed(const hn_csr* a, const V* val, const V* x, V* y, hipStream_t s) {
  if (a->nrows == 0) return HN_OK;
  if (a->nnz == 0) {
    hipLaunchKernelGGL((k_fill<V>), dim3(grid_for(a->nrows, 256, 4096)), dim3(256), 0, s, y, V(0),
                       (long long)a->nrows);
    return check_launch("k_fill");
  }
  if (a->tiling.chunk_ipt > 0) return spmv_chunk_launch<V>(a, x, y, s);
  if (a->tiling.slab_log2 > 0) return spmv_slab_launch<V>(a, x, y, s);
  switch (a->tiling.ipt) {
    case 4: return spmv_run<V, 4>(a, val, x, y, s);
    case 8: return spmv_run<V, 8>(a, val, x, y, s);
    case 16: return spmv_run<V, 16>(a, val, x, y, s);
    default: return fail(HN_ERR_INVALID, "matrix has no SpMV tiling");
  }
}

int spmv_launch(const hn_csr* a, int prec, const void* x, void* y, hipStream_t s) {
  if (prec == HN_F64) {
    HN_CHECK(a->flags & HN_F64, HN_ERR_INVALID, "matrix has no FP64 values");
    return spmv_typed<double>(a, a->val64.as<double>(), (const double*)x, (double*)y, s);
  }
  if (prec == HN_F32) {
    HN_CHECK(a->flags & HN_F32, HN_ERR_INVALID, "matrix has no FP32 values");
    return spmv_typed<float>(a, a->val32.as<float>(), (const float*)x, (float*)y, s);
  }
  return fail(HN_ERR_INVALID, "precision must be HN_F64 or HN_F32");
}

int fill_launch(double* p, double v, int64_t n, hipStream_t s) {
  if (n <= 0) return HN_OK;
  hipLaunchKernelGGL((k_fill<double>), dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, v, (long long)n);
  return check_launch("k_fill");
}

// ---- SpMM: C (+)= A * B, B row-major (ncols x k) -----------------------------
// One wave per row; lanes cover 64 consecutive columns of B/C, so every
// nonzero reads one contiguous strip of a B row.  Each entry of C is summed
// over the row's nonzeros in order.  A batch of products over the same
// pattern runs as gridDim.y products: product y reads val, B and writes C
// at y * istride bytes past the given pointers.
template <typename V>
__device__ __forceinline__ const V* shift(const V* p, long long bytes) {
  return reinterpret_cast<const V*>(reinterpret_cast<const char*>(p) + bytes);
}
template <typename V>
__device__ __forceinline__ V* shift(V* p, long long bytes) {
  return reinterpret_cast<V*>(reinterpret_cast<char*>(p) + bytes);
}

template <typename V>
__global__ __launch_bounds__(256) void k_spmm_rows(const int* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const V* __restrict__ val_,
                                                   const V* __restrict__ B_, V* __restrict__ C_,
                                                   int nrows, int k, int accumulate, long long istride) {
  const long long off = (long long)blockIdx.y * istride;
  const V* __restrict__ val = shift(val_, off);
  const V* __restrict__ B = shift(B_, off);
  V* __restrict__ C = shift(C_, off);
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (wave >= nrows) return;
  const int beg = rowptr[wave], end = rowptr[wave + 1];
  for (int c0 = 0; c0 < k; c0 += kWave) {
    const int c = c0 + lane;
    V acc = 0;
    if (c < k) {
      for (int e = beg; e < end; ++e) acc += val[e] * B[(long long)col[e] * k + c];
      V* o = C + (long long)wave * k + c;
      *o = accumulate ? *o + acc : acc;
    }
  }
}

// The same product for long rows: one workgroup per row; wave w takes the
// nonzeros w, w + 4, w + 8, ... into two running sums (w + 8t and
// w + 8t + 4), and the four waves' partial rows are added in a fixed order,
// so the result is deterministic.
template <typename V>
__global__ __launch_bounds__(256) void k_spmm_rows_long(const int* __restrict__ rowptr,
                                                        const int* __restrict__ col,
                                                        const V* __restrict__ val_,
                                                        const V* __restrict__ B_, V* __restrict__ C_, int k,
                                                        int accumulate, long long istride) {
  const long long off = (long long)blockIdx.y * istride;
  const V* __restrict__ val = shift(val_, off);
  const V* __restrict__ B = shift(B_, off);
  V* __restrict__ C = shift(C_, off);
  __shared__ V part[4][kWave];
  const int r = blockIdx.x;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int k0 = rowptr[r], k1 = rowptr[r + 1];
  for (int c0 = 0; c0 < k; c0 += kWave) {
    const int c = c0 + lane;
    V a0 = 0, a1 = 0;
    if (c < k) {
      int e = k0 + w;
      for (; e + 4 < k1; e += 8) {
        const int j0 = col[e], j1 = col[e + 4];
        a0 += val[e] * B[(long long)j0 * k + c];
        a1 += val[e + 4] * B[(long long)j1 * k + c];
      }
      for (; e < k1; e += 4) a0 += val[e] * B[(long long)col[e] * k + c];
    }
    part[w][lane] = a0 + a1;
    __syncthreads();
    if (w == 0 && c < k) {
      const V acc = ((part[0][lane] + part[1][lane]) + part[2][lane]) + part[3][lane];
      V* o = C + (long long)r * k + c;
      *o = accumulate ? *o + acc : acc;
    }
    __syncthreads();
  }
}

// Rows of exactly 64 columns (the GNN's hidden width), 16-byte lanes: a
// row of B or C is 256 B (float) or 512 B (double), so kW64 = 16 (float)
// or 32 (double) lanes cover one row with one vector load each.
//
// Short rows: each group of kW64 lanes owns a row (4 or 2 rows per wave),
// every output element summed over the row's nonzeros in order - the
// arithmetic of k_spmm_rows, with a quarter (or half) of its waves and
// each group's next two gathers issued together.
template <typename V>
struct Vec4Of;
template <>
struct Vec4Of<float> {
  typedef float4 T;
  static constexpr int kPer = 4;
};
template <>
struct Vec4Of<double> {
  typedef double2 T;
  static constexpr int kPer = 2;
};

template <typename V>
__device__ __forceinline__ void axpy_vec(float4& a, float w, const float4& b) {
  a.x += w * b.x;
  a.y += w * b.y;
  a.z += w * b.z;
  a.w += w * b.w;
}
template <typename V>
__device__ __forceinline__ void axpy_vec(double2& a, double w, const double2& b) {
  a.x += w * b.x;
  a.y += w * b.y;
}
__device__ __forceinline__ float4 add_vec(const float4& a, const float4& b) {
  return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}
__device__ __forceinline__ double2 add_vec(const double2& a, const double2& b) {
  return make_double2(a.x + b.x, a.y + b.y);
}

template <typename V>
__global__ __launch_bounds__(256) void k_spmm_rows64(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                     const V* __restrict__ val_, const V* __restrict__ B_,
                                                     V* __restrict__ C_, int nrows, int accumulate,
                                                     long long istride) {
  typedef typename Vec4Of<V>::T VT;
  constexpr int kW64 = 64 / Vec4Of<V>::kPer;  // lanes per row
  const long long off = (long long)blockIdx.y * istride;
  const V* __restrict__ val = shift(val_, off);
  const VT* __restrict__ B = reinterpret_cast<const VT*>(shift(B_, off));
  VT* __restrict__ C = reinterpret_cast<VT*>(shift(C_, off));
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long r = t / kW64;
  const int li = (int)(t % kW64);
  if (r >= nrows) return;
  const int beg = rowptr[r], end = rowptr[r + 1];
  VT acc;
  acc = VT{};
  int e = beg;
  for (; e + 1 < end; e += 2) {
    const int j0 = col[e], j1 = col[e + 1];
    const V w0 = val[e], w1 = val[e + 1];
    const VT b0 = B[(long long)j0 * kW64 + li];
    const VT b1 = B[(long long)j1 * kW64 + li];
    axpy_vec<V>(acc, w0, b0);
    axpy_vec<V>(acc, w1, b1);
  }
  if (e < end) axpy_vec<V>(acc, val[e], B[(long long)col[e] * kW64 + li]);
  VT* o = C + r * kW64 + li;
  *o = accumulate ? add_vec(*o, acc) : acc;
}

// Long rows: one wave per row (4 rows per workgroup).  Lane group g of the
// wave takes the nonzeros g, g + 4, g + 8, ... into two running sums
// (g + 8t and g + 8t + 4) and the groups' partial rows are added in the
// order ((p0 + p1) + p2) + p3 - the split and order of k_spmm_rows_long's
// four waves, so the result is the same bit for bit, with 16-byte gathers
// and eight of them in flight per wave.
template <typename V>
__global__ __launch_bounds__(256) void k_spmm_rows64_long(const int* __restrict__ rowptr,
                                                          const int* __restrict__ col,
                                                          const V* __restrict__ val_, const V* __restrict__ B_,
                                                          V* __restrict__ C_, int nrows, int accumulate,
                                                          long long istride) {
  typedef typename Vec4Of<V>::T VT;
  constexpr int kW64 = 64 / Vec4Of<V>::kPer;  // lanes per row slice
  constexpr int kG = kWave / kW64;            // lane groups per wave (4 float, 2 double)
  static_assert(kG == 4 || kG == 2, "lane groups per wave");
  const long long off = (long long)blockIdx.y * istride;
  const V* __restrict__ val = shift(val_, off);
  const VT* __restrict__ B = reinterpret_cast<const VT*>(shift(B_, off));
  VT* __restrict__ C = reinterpret_cast<VT*>(shift(C_, off));
  __shared__ VT part[4][4][kW64];  // [wave][split][lane]
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int g = lane / kW64, li = lane % kW64;
  const long long r = (long long)blockIdx.x * 4 + w;
  const bool live = r < nrows;
  const int k0 = live ? rowptr[r] : 0, k1 = live ? rowptr[r + 1] : 0;
  // with 2 groups (double) each group plays two of the four splits
#pragma unroll
  for (int h = 0; h < 4 / kG; ++h) {
    const int sp = g + h * kG;  // the split this pass plays: nonzeros sp, sp + 4, ...
    VT a0, a1;
    a0 = VT{};
    a1 = VT{};
    int e = k0 + sp;
    for (; e + 4 < k1; e += 8) {
      const int j0 = col[e], j1 = col[e + 4];
      const V w0 = val[e], w1 = val[e + 4];
      const VT b0 = B[(long long)j0 * kW64 + li];
      const VT b1 = B[(long long)j1 * kW64 + li];
      axpy_vec<V>(a0, w0, b0);
      axpy_vec<V>(a1, w1, b1);
    }
    for (; e < k1; e += 4) axpy_vec<V>(a0, val[e], B[(long long)col[e] * kW64 + li]);
    part[w][sp][li] = add_vec(a0, a1);
  }
  __syncthreads();
  if (live && g == 0) {
    const VT acc = add_vec(add_vec(add_vec(part[w][0][li], part[w][1][li]), part[w][2][li]), part[w][3][li]);
    VT* o = C + r * kW64 + li;
    *o = accumulate ? add_vec(*o, acc) : acc;
  }
}

template <typename V>
int spmm_rows_launch_batch(const int* rowptr, const int* col, const V* val, int64_t nrows, int64_t nnz, const V* B,
                           V* C, int k, bool accumulate, int batch, long long istride, hipStream_t s) {
  if (nrows <= 0 || k <= 0 || batch <= 0) return HN_OK;
  HN_CHECK(nrows * (int64_t)kWave < (int64_t)INT_MAX * 4, HN_ERR_UNSUPPORTED, "spmm: too many rows");
  HN_CHECK(batch <= 65535, HN_ERR_UNSUPPORTED, "spmm: too many products in one batch");
  const bool vec_ok = k == 64 && reinterpret_cast<uintptr_t>(B) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(C) % 16 == 0 && istride % 16 == 0;
  if (vec_ok) {
    if (nnz >= 32 * nrows) {
      hipLaunchKernelGGL((k_spmm_rows64_long<V>), dim3((unsigned)((nrows + 3) / 4), (unsigned)batch), dim3(256), 0,
                         s, rowptr, col, val, B, C, (int)nrows, accumulate ? 1 : 0, istride);
    } else {
      const int64_t threads = nrows * (int64_t)(64 / Vec4Of<V>::kPer);
      hipLaunchKernelGGL((k_spmm_rows64<V>), dim3((unsigned)((threads + 255) / 256), (unsigned)batch), dim3(256), 0,
                         s, rowptr, col, val, B, C, (int)nrows, accumulate ? 1 : 0, istride);
    }
    count_launch();
    return check_launch("spmm");
  }
  if (nnz >= 32 * nrows) {
    HN_CHECK(nrows < INT_MAX, HN_ERR_UNSUPPORTED, "spmm: too many rows");
    hipLaunchKernelGGL((k_spmm_rows_long<V>), dim3((unsigned)nrows, (unsigned)batch), dim3(256), 0, s, rowptr, col,
                       val, B, C, k, accumulate ? 1 : 0, istride);
  } else {
    const int64_t threads = nrows * (int64_t)kWave;
    hipLaunchKernelGGL((k_spmm_rows<V>), dim3((unsigned)((threads + 255) / 256), (unsigned)batch), dim3(256), 0, s,
                       rowptr, col, val, B, C, (int)nrows, k, accumulate ? 1 : 0, istride);
  }
  count_launch();
  return check_launch("spmm");
}

template <typename V>
int spmm_rows_launch(const int* rowptr, const int* col, const V* val, int64_t nrows, int64_t nnz, const V* B, V* C,
                     int k, bool accumulate, hipStream_t s) {
  return spmm_rows_launch_batch<V>(rowptr, col, val, nrows, nnz, B, C, k, accumulate, 1, 0, s);
}

template int spmm_rows_launch<double>(const int*, const int*, const double*, int64_t, int64_t, const double*,
                                      double*, int, bool, hipStream_t);
template int spmm_rows_launch<float>(const int*, const int*, const float*, int64_t, int64_t, const float*, float*,
                                     int, bool, hipStream_t);
template int spmm_rows_launch_batch<float>(const int*, const int*, const float*, int64_t, int64_t, const float*,
                                           float*, int, bool, int, long long, hipStream_t);

}  // namespace hn

using namespace hn;

namespace {

template <typename V>
__global__ void k_narrow_copy(const double* __restrict__ in, V* __restrict__ out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (V)in[i];
}

}  // namespace

namespace hn {

int csr_alloc(hn_csr* a, int64_t nrows, int64_t ncols, int64_t nnz, int flags) {
  a->nrows = nrows;
  a->ncols = ncols;
  a->nnz = nnz;
  a->flags = flags;
  HN_TRY(a->row_ptr.reserve(sizeof(int) * (nrows + 1)));
  HN_TRY(a->col.reserve(sizeof(int) * std::max<int64_t>(nnz, 1)));
  if (flags & HN_F64) HN_TRY(a->val64.reserve(sizeof(double) * std::max<int64_t>(nnz, 1)));
  if (flags & HN_F32) HN_TRY(a->val32.reserve(sizeof(float) * std::max<int64_t>(nnz, 1)));
  return HN_OK;
}

}  // namespace hn

extern "C" int hn_csr_upload(const double* values, const uint64_t* col_indices,
                             const uint64_t* row_ptr, uint64_t nrows, uint64_t ncols,
                             uint64_t nnz, int flags, hn_csr** out) {
  HN_CHECK(out, HN_ERR_INVALID, "hn_csr_upload: null out pointer");
  *out = nullptr;
  HN_CHECK(row_ptr, HN_ERR_INVALID, "hn_csr_upload: null row_ptr");
  HN_CHECK(nnz == 0 || (values && col_indices), HN_ERR_INVALID, "hn_csr_upload: null values/col_indices");
  HN_CHECK((flags & (HN_F64 | HN_F32)) != 0 && (flags & ~(HN_F64 | HN_F32)) == 0, HN_ERR_INVALID,
           "hn_csr_upload: flags must be HN_F64 and/or HN_F32");
  HN_CHECK(nrows < (uint64_t)INT_MAX && ncols < (uint64_t)INT_MAX && nnz < (uint64_t)INT_MAX,
           HN_ERR_UNSUPPORTED, "hn_csr_upload: nrows, ncols and nnz must be below 2^31");
  HN_CHECK(nrows + nnz < (uint64_t)INT_MAX, HN_ERR_UNSUPPORTED, "hn_csr_upload: nrows + nnz must be below 2^31");
  HN_CHECK(row_ptr[0] == 0 && row_ptr[nrows] == nnz, HN_ERR_INVALID,
           "hn_csr_upload: row_ptr must start at 0 and end at nnz");
  std::vector<int> rp32(nrows + 1), ci32(nnz);
  for (uint64_t i = 0; i <= nrows; ++i) {
    if (i > 0 && row_ptr[i] < row_ptr[i - 1])
      return fail(HN_ERR_INVALID, "hn_csr_upload: row_ptr is not non-decreasing");
    rp32[i] = (int)row_ptr[i];
  }
  for (uint64_t k = 0; k < nnz; ++k) {
    if (col_indices[k] >= ncols) return fail(HN_ERR_INVALID, "hn_csr_upload: column index out of range");
    ci32[k] = (int)col_indices[k];
  }
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  hn_csr* a = new hn_csr();
  int st = csr_alloc(a, (int64_t)nrows, (int64_t)ncols, (int64_t)nnz, flags);
  if (st == HN_OK) {
    hipError_t e = hipMemcpyAsync(a->row_ptr.p, rp32.data(), sizeof(int) * (nrows + 1), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nnz) e = hipMemcpyAsync(a->col.p, ci32.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && nnz) {
      DevBuf staging;
      double* d64 = nullptr;
      if (flags & HN_F64) {
        d64 = a->val64.as<double>();
      } else {
        st = staging.reserve(sizeof(double) * nnz);
        d64 = staging.as<double>();
      }
      if (st == HN_OK) {
        e = hipMemcpyAsync(d64, values, sizeof(double) * nnz, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && (flags & HN_F32)) {
          hipLaunchKernelGGL((k_narrow_copy<float>), dim3(grid_for(nnz, 256, 8192)), dim3(256), 0, s, d64,
                             a->val32.as<float>(), (long long)nnz);
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(s);
      }
    }
    if (st == HN_OK && e != hipSuccess) st = hip_fail(e, "hn_csr_upload copy", __FILE__, __LINE__);
  }
  if (st == HN_OK) st = build_tiling(a, 8, s);
  if (st == HN_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = hip_fail(e, "hn_csr_upload sync", __FILE__, __LINE__);
  }
  if (st != HN_OK) {
    delete a;
    return st;
  }
  *out = a;
  return HN_OK;
}

extern "C" int hn_csr_free(hn_csr* a) {
  if (!a) return HN_OK;
  if (ctx().initialized) {
    (void)hipSetDevice(ctx().device);
    (void)hipStreamSynchronize(ctx().stream);
  }
  delete a;
  return HN_OK;
}

extern "C" int hn_csr_dims(const hn_csr* a, uint64_t* nrows, uint64_t* ncols, uint64_t* nnz) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_dims: null matrix");
  if (nrows) *nrows = (uint64_t)a->nrows;
  if (ncols) *ncols = (uint64_t)a->ncols;
  if (nnz) *nnz = (uint64_t)a->nnz;
  return HN_OK;
}

extern "C" int hn_csr_download(const hn_csr* a, double* values, uint64_t* col_indices, uint64_t* row_ptr) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_download: null matrix");
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  std::vector<int> rp(a->nrows + 1), ci(a->nnz);
  HN_HIP(hipMemcpyAsync(rp.data(), a->row_ptr.p, sizeof(int) * (a->nrows + 1), hipMemcpyDeviceToHost, s));
  if (a->nnz) HN_HIP(hipMemcpyAsync(ci.data(), a->col.p, sizeof(int) * a->nnz, hipMemcpyDeviceToHost, s));
  if (values && a->nnz) {
    if (a->flags & HN_F64) {
      HN_HIP(hipMemcpyAsync(values, a->val64.p, sizeof(double) * a->nnz, hipMemcpyDeviceToHost, s));
    } else {
      std::vector<float> v32(a->nnz);
      HN_HIP(hipMemcpyAsync(v32.data(), a->val32.p, sizeof(float) * a->nnz, hipMemcpyDeviceToHost, s));
      HN_HIP(hipStreamSynchronize(s));
      for (int64_t k = 0; k < a->nnz; ++k) values[k] = v32[k];
    }
  }
  HN_HIP(hipStreamSynchronize(s));
  if (row_ptr)
    for (int64_t i = 0; i <= a->nrows; ++i) row_ptr[i] = (uint64_t)rp[i];
  if (col_indices)
    for (int64_t k = 0; k < a->nnz; ++k) col_indices[k] = (uint64_t)ci[k];
  return HN_OK;
}

extern "C" int hn_csr_set_tiling(hn_csr* a, int items_per_thread) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_set_tiling: null matrix");
  HN_TRY(enter());
  HN_TRY(build_tiling(a, items_per_thread, ctx().stream));
  a->tiling.slab_log2 = 0;
  a->tiling.chunk_ipt = 0;
  HN_HIP(hipStreamSynchronize(ctx().stream));
  return HN_OK;
}

extern "C" int hn_csr_set_slab(hn_csr* a, int slab_log2) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_set_slab: null matrix");
  HN_TRY(enter());
  HN_TRY(build_slab(a, slab_log2, ctx().stream));
  a->tiling.chunk_ipt = 0;
  return HN_OK;
}

extern "C" int hn_csr_set_chunks(hn_csr* a, int items_per_thread, int span_log2, int threads) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_set_chunks: null matrix");
  HN_TRY(enter());
  return build_chunks(a, items_per_thread, span_log2, threads, ctx().stream);
}

extern "C" int hn_csr_get_chunks(const hn_csr* a, int* items_per_thread, int* span_log2, int* threads) {
  HN_CHECK(a && items_per_thread && span_log2 && threads, HN_ERR_INVALID, "hn_csr_get_chunks: null argument");
  const bool on = a->tiling.chunk_ipt > 0;
  *items_per_thread = a->tiling.chunk_ipt;
  *span_log2 = on ? a->tiling.chunk_span : 0;
  *threads = on ? a->tiling.chunk_threads : 0;
  return HN_OK;
}

extern "C" int hn_csr_get_slab(const hn_csr* a, int* slab_log2) {
  HN_CHECK(a && slab_log2, HN_ERR_INVALID, "hn_csr_get_slab: null argument");
  *slab_log2 = a->tiling.slab_log2;
  return HN_OK;
}

extern "C" int hn_csr_get_tiling(const hn_csr* a, int* items_per_thread) {
  HN_CHECK(a && items_per_thread, HN_ERR_INVALID, "hn_csr_get_tiling: null argument");
  *items_per_thread = a->tiling.ipt;
  return HN_OK;
}

extern "C" int hn_csr_autotune(hn_csr* a, int iters, double* best_us) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_csr_autotune: null matrix");
  HN_CHECK(iters > 0, HN_ERR_INVALID, "hn_csr_autotune: iters must be positive");
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  const int prec = (a->flags & HN_F64) ? HN_F64 : HN_F32;
  const size_t vb = prec == HN_F64 ? sizeof(double) : sizeof(float);
  DevBuf x, y;
  HN_TRY(x.reserve(vb * std::max<int64_t>(a->ncols, 1)));
  HN_TRY(y.reserve(vb * std::max<int64_t>(a->nrows, 1)));
  HN_HIP(hipMemsetAsync(x.p, 0, vb * std::max<int64_t>(a->ncols, 1), s));
  hipEvent_t e0, e1;
  HN_HIP(hipEventCreate(&e0));
  HN_HIP(hipEventCreate(&e1));
  // candidates: merge path with 4, 8, 16 items per thread; the slab layout
  // with 2^14 .. 2^19 columns per slab and the chunked stream (256-, 512-
  // or 1024-thread workgroups, 4 or 8 items per thread, chunks spanning
  // 2^17 .. 2^19 columns), both only when x spans several slabs
  struct Cand {
    int ipt, slab, chunk, span, threads;
  };
  std::vector<Cand> cand = {{4, 0, 0, 0, 0}, {8, 0, 0, 0, 0}, {16, 0, 0, 0, 0}};
  const bool layouts = a->nnz > 0 && a->nrows > 0;
  for (int lg = 14; lg <= 19; ++lg)
    if (layouts && (a->ncols >> lg) >= 1) cand.push_back({8, lg, 0, 0, 0});
  for (int nt : {256, 512, 1024})
    for (int ci : {4, 8})
      for (int sp = 17; sp <= 19; ++sp)
        if (layouts && (a->ncols >> 14) >= 1) cand.push_back({8, 0, ci, sp, nt});
  Cand best{8, 0, 0, 0, 0};
  float best_ms = 1e30f;
  int st = HN_OK;
  for (size_t c = 0; c < cand.size() && st == HN_OK; ++c) {
    st = build_tiling(a, cand[c].ipt, s);
    if (st == HN_OK) st = build_slab(a, cand[c].slab, s);
    if (st == HN_OK) st = build_chunks(a, cand[c].chunk, cand[c].span, cand[c].threads, s);
    if (st == HN_OK) st = spmv_launch(a, prec, x.p, y.p, s);  // warm-up
    if (st != HN_OK) break;
    (void)hipEventRecord(e0, s);
    for (int it = 0; it < iters && st == HN_OK; ++it) st = spmv_launch(a, prec, x.p, y.p, s);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms < best_ms) {
      best_ms = ms;
      best = cand[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HN_TRY(st);
  HN_TRY(build_tiling(a, best.ipt, s));
  HN_TRY(build_slab(a, best.slab, s));
  HN_TRY(build_chunks(a, best.chunk, best.span, best.threads, s));
  HN_HIP(hipStreamSynchronize(s));
  if (best_us) *best_us = 1000.0 * best_ms / iters;
  return HN_OK;
}

extern "C" int hn_spmv_dev(const hn_csr* a, int prec, const void* x, void* y, void* stream) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_spmv_dev: null matrix");
  HN_CHECK((x || a->ncols == 0 || a->nnz == 0) && (y || a->nrows == 0), HN_ERR_INVALID,
           "hn_spmv_dev: null vector");
  HN_TRY(enter());
  return spmv_launch(a, prec, x, y, pick_stream(stream));
}

extern "C" int hn_spmv(const hn_csr* a, const double* x, double* y) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_spmv: null matrix");
  HN_CHECK(a->flags & HN_F64, HN_ERR_INVALID, "hn_spmv: matrix has no FP64 values");
  HN_CHECK((x || a->ncols == 0) && (y || a->nrows == 0), HN_ERR_INVALID, "hn_spmv: null vector");
  HN_TRY(enter());
  HostOpScratch& sc = host_scratch();
  std::lock_guard<std::mutex> lk(sc.mtx);
  hipStream_t s = ctx().stream;
  HN_TRY(sc.a.reserve(sizeof(double) * std::max<int64_t>(a->ncols, 1)));
  HN_TRY(sc.b.reserve(sizeof(double) * std::max<int64_t>(a->nrows, 1)));
  if (a->ncols) HN_HIP(hipMemcpyAsync(sc.a.p, x, sizeof(double) * a->ncols, hipMemcpyHostToDevice, s));
  HN_TRY(spmv_launch(a, HN_F64, sc.a.p, sc.b.p, s));
  if (a->nrows) HN_HIP(hipMemcpyAsync(y, sc.b.p, sizeof(double) * a->nrows, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

namespace hn {
static int spmm_launch(const hn_csr* a, const double* B, double* C, int64_t k, hipStream_t s) {
  HN_CHECK(a->flags & HN_F64, HN_ERR_INVALID, "spmm: matrix has no FP64 values");
  HN_CHECK(k < INT_MAX, HN_ERR_UNSUPPORTED, "spmm: shape too large");
  return spmm_rows_launch<double>(a->row_ptr.as<int>(), a->col.as<int>(), a->val64.as<double>(), a->nrows, a->nnz,
                                  B, C, (int)k, false, s);
}
}  // namespace hn

extern "C" int hn_spmm_dev(const hn_csr* a, const double* B, double* C, uint64_t k, void* stream) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_spmm_dev: null matrix");
  HN_TRY(enter());
  return spmm_launch(a, B, C, (int64_t)k, pick_stream(stream));
}

extern "C" int hn_spmm(const hn_csr* a, const double* B, double* C, uint64_t k) {
  HN_CHECK(a, HN_ERR_INVALID, "hn_spmm: null matrix");
  if (a->nrows == 0 || k == 0) return HN_OK;
  HN_CHECK(C && (B || a->ncols == 0), HN_ERR_INVALID, "hn_spmm: null buffer");
  HN_TRY(enter());
  HostOpScratch& sc = host_scratch();
  std::lock_guard<std::mutex> lk(sc.mtx);
  hipStream_t s = ctx().stream;
  const size_t bb = sizeof(double) * std::max<int64_t>(a->ncols * (int64_t)k, 1);
  const size_t cb = sizeof(double) * a->nrows * k;
  HN_TRY(sc.a.reserve(bb));
  HN_TRY(sc.b.reserve(cb));
  if (a->ncols) HN_HIP(hipMemcpyAsync(sc.a.p, B, sizeof(double) * a->ncols * k, hipMemcpyHostToDevice, s));
  HN_TRY(spmm_launch(a, sc.a.as<double>(), sc.b.as<double>(), (int64_t)k, s));
  HN_HIP(hipMemcpyAsync(C, sc.b.p, cb, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}
