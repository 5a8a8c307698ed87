// CSR transpose on the device by a stable counting sort.
//
// Reference layout: build_normal_equations scatters row i's entries into
// column buckets in increasing i (core/lp/interior_point.cpp:153-173), so the
// entries of each transposed row are ordered by source row, and by source
// position within a row.  The same order is produced here, bit for bit:
//   1. column histogram (integer atomics: the counts do not depend on order);
//   2. exclusive scan -> transposed row pointer;
//   3. atomic scatter of source positions k into their column bucket;
//   4. each bucket sorted by k (keys are unique, so the result is the stable
//      order): short buckets by one thread, buckets up to 4096 by an LDS
//      bitonic sort, longer ones by LDS-sorted runs merged by rank;
//   5. gather of source row (binary search of row_ptr) and value.
#include "hn_device.hpp"
#include "hn_internal.hpp"

#include <algorithm>
#include <climits>

namespace hn {

__global__ void k_hist(const int* __restrict__ col, long long nnz, int* __restrict__ cnt) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nnz;
       k += (long long)gridDim.x * blockDim.x)
    atomicAdd(&cnt[col[k]], 1);
}

// ---- device-wide exclusive scan of int counts (n entries -> n+1) ----------
constexpr int kScanNT = 256;
constexpr int kScanIPT = 8;
constexpr int kScanChunk = kScanNT * kScanIPT;

__device__ __forceinline__ int block_exclusive_scan_int(int v, int* red, int* total) {
  // Inclusive wave scan.
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  int inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int o = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += o;
  }
  if (lane == kWave - 1) red[w] = inc;
  __syncthreads();
  int wprefix = 0, tot = 0;
#pragma unroll
  for (int u = 0; u < kScanNT / kWave; ++u) {
    if (u < w) wprefix += red[u];
    tot += red[u];
  }
  __syncthreads();
  *total = tot;
  return wprefix + inc - v;
}

__global__ __launch_bounds__(kScanNT) void k_scan_partial(const int* __restrict__ in, long long n,
                                                          long long* __restrict__ sums) {
  __shared__ int red[kScanNT / kWave];
  const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanIPT;
  int s = 0;
#pragma unroll
  for (int u = 0; u < kScanIPT; ++u)
    if (base + u < n) s += in[base + u];
  int tot;
  (void)block_exclusive_scan_int(s, red, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of the block sums in place.
__global__ __launch_bounds__(kScanNT) void k_scan_sums(long long* __restrict__ sums, int nb) {
  __shared__ long long red[kScanNT / kWave];
  long long carry = 0;
  for (int base = 0; base < nb; base += kScanNT) {
    const int i = base + threadIdx.x;
    long long v = (i < nb) ? sums[i] : 0;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    long long inc = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      long long o = __shfl_up(inc, off, kWave);
      if (lane >= off) inc += o;
    }
    if (lane == kWave - 1) red[w] = inc;
    __syncthreads();
    long long wp = 0, tot = 0;
#pragma unroll
    for (int u = 0; u < kScanNT / kWave; ++u) {
      if (u < w) wp += red[u];
      tot += red[u];
    }
    if (i < nb) sums[i] = carry + wp + inc - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[nb] = carry;
}

__global__ __launch_bounds__(kScanNT) void k_scan_final(const int* __restrict__ in, long long n,
                                                        const long long* __restrict__ sums,
                                                        int* __restrict__ out) {
  __shared__ int red[kScanNT / kWave];
  const long long base = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanIPT;
  int vals[kScanIPT];
  int s = 0;
#pragma unroll
  for (int u = 0; u < kScanIPT; ++u) {
    vals[u] = (base + u < n) ? in[base + u] : 0;
    s += vals[u];
  }
  int tot;
  int ex = block_exclusive_scan_int(s, red, &tot);
  long long run = sums[blockIdx.x] + ex;
#pragma unroll
  for (int u = 0; u < kScanIPT; ++u) {
    if (base + u < n) out[base + u] = (int)run;
    run += vals[u];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (int)sums[gridDim.x];
}

// out[0..n] = exclusive scan of in[0..n-1] with out[n] = total (must fit int).
int exclusive_scan(const int* in, int64_t n, int* out, DevBuf& tmp, hipStream_t s) {
  const int64_t nb = std::max<int64_t>(1, (n + kScanChunk - 1) / kScanChunk);
  HN_TRY(tmp.reserve(sizeof(long long) * (nb + 1)));
  long long* sums = tmp.as<long long>();
  hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)nb), dim3(kScanNT), 0, s, in, (long long)n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanNT), 0, s, sums, (int)nb);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kScanNT), 0, s, in, (long long)n, sums, out);
  return check_launch("exclusive_scan");
}

// ---- bucket sort helpers --------------------------------------------------
__global__ void k_scatter_pos(const int* __restrict__ col, long long nnz, int* __restrict__ cursor,
                              int* __restrict__ perm) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < nnz;
       k += (long long)gridDim.x * blockDim.x) {
    const int p = atomicAdd(&cursor[col[k]], 1);
    perm[p] = (int)k;
  }
}

constexpr int kShortSeg = 16;
constexpr int kBlockSeg = 4096;

// Sorts short segments in place; appends longer segments to two work lists.
__global__ void k_sort_short(const int* __restrict__ ptr, int nseg, int* __restrict__ perm,
                             int* __restrict__ mid_list, int* __restrict__ long_list,
                             int* __restrict__ counters) {
  const int sgm = blockIdx.x * blockDim.x + threadIdx.x;
  if (sgm >= nseg) return;
  const int b = ptr[sgm], e = ptr[sgm + 1], len = e - b;
  if (len <= 1) return;
  if (len <= kShortSeg) {
    int v[kShortSeg];
    for (int u = 0; u < len; ++u) v[u] = perm[b + u];
    for (int u = 1; u < len; ++u) {
      const int key = v[u];
      int w = u - 1;
      while (w >= 0 && v[w] > key) {
        v[w + 1] = v[w];
        --w;
      }
      v[w + 1] = key;
    }
    for (int u = 0; u < len; ++u) perm[b + u] = v[u];
  } else if (len <= kBlockSeg) {
    mid_list[atomicAdd(&counters[0], 1)] = sgm;
  } else {
    long_list[atomicAdd(&counters[1], 1)] = sgm;
  }
}

__device__ void lds_bitonic_sort(int* s, int npow2) {
  for (int k = 2; k <= npow2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < npow2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const int a = s[i], b = s[ixj];
          if ((a > b) == up) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(256) void k_sort_mid(const int* __restrict__ ptr, const int* __restrict__ list,
                                                  int* __restrict__ perm) {
  __shared__ int s[kBlockSeg];
  const int sgm = list[blockIdx.x];
  const int b = ptr[sgm], len = ptr[sgm + 1] - b;
  int np = 1;
  while (np < len) np <<= 1;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s[i] = (i < len) ? perm[b + i] : INT_MAX;
  __syncthreads();
  lds_bitonic_sort(s, np);
  for (int i = threadIdx.x; i < len; i += blockDim.x) perm[b + i] = s[i];
}

// One block per long segment: sort 4096-chunks in LDS, then merge runs by
// rank: an element's output slot is its index plus the number of partner-run
// keys ordered before it (strictly smaller for run A, not larger for run B,
// so equal keys keep run order).
__global__ __launch_bounds__(256) void k_sort_long(const int* __restrict__ ptr, const int* __restrict__ list,
                                                   int* __restrict__ perm, int* __restrict__ tmp) {
  __shared__ int s[kBlockSeg];
  const int sgm = list[blockIdx.x];
  const int b = ptr[sgm], len = ptr[sgm + 1] - b;
  for (int c0 = 0; c0 < len; c0 += kBlockSeg) {
    const int cl = min(kBlockSeg, len - c0);
    for (int i = threadIdx.x; i < kBlockSeg; i += blockDim.x) s[i] = (i < cl) ? perm[b + c0 + i] : INT_MAX;
    __syncthreads();
    lds_bitonic_sort(s, kBlockSeg);
    for (int i = threadIdx.x; i < cl; i += blockDim.x) perm[b + c0 + i] = s[i];
    __syncthreads();
  }
  int* src = perm + b;
  int* dst = tmp + b;
  for (int w = kBlockSeg; w < len; w <<= 1) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
      const int run = i / (2 * w);
      const int a0 = run * 2 * w;
      const int a1 = min(a0 + w, len);
      const int b1 = min(a0 + 2 * w, len);
      const int key = src[i];
      int pos;
      if (i < a1) {  // element of run A: count keys < key in run B
        int lo = a1, hi = b1;
        while (lo < hi) {
          const int m = (lo + hi) >> 1;
          if (src[m] < key) lo = m + 1; else hi = m;
        }
        pos = a0 + (i - a0) + (lo - a1);
      } else {  // element of run B: count keys <= key in run A (A first on ties)
        int lo = a0, hi = a1;
        while (lo < hi) {
          const int m = (lo + hi) >> 1;
          if (src[m] <= key) lo = m + 1; else hi = m;
        }
        pos = a0 + (i - a1) + (lo - a0);
      }
      dst[pos] = key;
    }
    __syncthreads();
    int* t = src;
    src = dst;
    dst = t;
  }
  if (src != perm + b) {
    for (int i = threadIdx.x; i < len; i += blockDim.x) perm[b + i] = src[i];
  }
}


int segmented_sort(const int* ptr, int nseg, int* keys, int64_t total, SortScratch& sc, hipStream_t s) {
  if (nseg <= 0 || total <= 0) return HN_OK;
  HN_TRY(sc.lists.reserve(sizeof(int) * 2 * (size_t)nseg));
  HN_TRY(sc.counters.reserve(sizeof(int) * 2));
  HN_HIP(hipMemsetAsync(sc.counters.p, 0, sizeof(int) * 2, s));
  int* mid_list = sc.lists.as<int>();
  int* long_list = sc.lists.as<int>() + nseg;
  hipLaunchKernelGGL(k_sort_short, dim3(grid_for(nseg, 256)), dim3(256), 0, s, ptr, nseg, keys, mid_list, long_list,
                     sc.counters.as<int>());
  HN_TRY(check_launch("k_sort_short"));
  int hc[2] = {0, 0};
  HN_HIP(hipMemcpyAsync(hc, sc.counters.p, sizeof(hc), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  if (hc[0] > 0) {
    hipLaunchKernelGGL(k_sort_mid, dim3(hc[0]), dim3(256), 0, s, ptr, mid_list, keys);
    HN_TRY(check_launch("k_sort_mid"));
  }
  if (hc[1] > 0) {
    HN_TRY(sc.tmp.reserve(sizeof(int) * total));
    hipLaunchKernelGGL(k_sort_long, dim3(hc[1]), dim3(256), 0, s, ptr, long_list, keys, sc.tmp.as<int>());
    HN_TRY(check_launch("k_sort_long"));
  }
  return HN_OK;
}

template <bool F64, bool F32>
__global__ void k_gather_transpose(const int* __restrict__ rowptr, int nrows, const int* __restrict__ perm,
                                   long long nnz, const double* __restrict__ v64, const float* __restrict__ v32,
                                   int* __restrict__ out_col, double* __restrict__ o64, float* __restrict__ o32) {
  for (long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x; p < nnz;
       p += (long long)gridDim.x * blockDim.x) {
    const int k = perm[p];
    // Source row: last r with rowptr[r] <= k.
    int lo = 0, hi = nrows;
    while (lo < hi) {
      const int m = (lo + hi + 1) >> 1;
      if (rowptr[m] <= k) lo = m; else hi = m - 1;
    }
    // The last row starting at or before k is the one holding k (empty rows
    // before it share its start offset and are skipped by "last").
    out_col[p] = lo;
    if (F64) o64[p] = v64[k];
    if (F32) o32[p] = v32[k];
  }
}

}  // namespace hn

using namespace hn;

namespace hn {

int transpose_into(const hn_csr* a, hn_csr* t, hipStream_t s) {
  DevBuf cnt, perm, scan_tmp;
  SortScratch scratch;
  const int64_t nnz = a->nnz, nc = a->ncols;
  HN_TRY(cnt.reserve(sizeof(int) * std::max<int64_t>(nc, 1)));
  HN_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int) * std::max<int64_t>(nc, 1), s));
  if (nnz > 0) {
    hipLaunchKernelGGL(k_hist, dim3(grid_for(nnz, 256, 16384)), dim3(256), 0, s, a->col.as<int>(),
                       (long long)nnz, cnt.as<int>());
    HN_TRY(check_launch("k_hist"));
  }
  HN_TRY(exclusive_scan(cnt.as<int>(), nc, t->row_ptr.as<int>(), scan_tmp, s));
  if (nnz == 0) return HN_OK;
  // cursor = row pointer copy (reuse cnt)
  HN_HIP(hipMemcpyAsync(cnt.p, t->row_ptr.p, sizeof(int) * nc, hipMemcpyDeviceToDevice, s));
  HN_TRY(perm.reserve(sizeof(int) * nnz));
  hipLaunchKernelGGL(k_scatter_pos, dim3(grid_for(nnz, 256, 16384)), dim3(256), 0, s, a->col.as<int>(),
                     (long long)nnz, cnt.as<int>(), perm.as<int>());
  HN_TRY(check_launch("k_scatter_pos"));
  // segmented_sort synchronises the stream, so the local buffers outlive their use
  HN_TRY(segmented_sort(t->row_ptr.as<int>(), (int)nc, perm.as<int>(), nnz, scratch, s));
  const bool f64 = a->flags & HN_F64, f32 = a->flags & HN_F32;
  const dim3 g(grid_for(nnz, 256, 16384)), bdim(256);
  if (f64 && f32)
    hipLaunchKernelGGL((k_gather_transpose<true, true>), g, bdim, 0, s, a->row_ptr.as<int>(), (int)a->nrows,
                       perm.as<int>(), (long long)nnz, a->val64.as<double>(), a->val32.as<float>(),
                       t->col.as<int>(), t->val64.as<double>(), t->val32.as<float>());
  else if (f64)
    hipLaunchKernelGGL((k_gather_transpose<true, false>), g, bdim, 0, s, a->row_ptr.as<int>(), (int)a->nrows,
                       perm.as<int>(), (long long)nnz, a->val64.as<double>(), (const float*)nullptr,
                       t->col.as<int>(), t->val64.as<double>(), (float*)nullptr);
  else
    hipLaunchKernelGGL((k_gather_transpose<false, true>), g, bdim, 0, s, a->row_ptr.as<int>(), (int)a->nrows,
                       perm.as<int>(), (long long)nnz, (const double*)nullptr, a->val32.as<float>(),
                       t->col.as<int>(), (double*)nullptr, t->val32.as<float>());
  HN_TRY(check_launch("k_gather_transpose"));
  // perm is read by the gather: wait before the local buffers are released
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
}

}  // namespace hn

extern "C" int hn_csr_transpose(const hn_csr* a, hn_csr** out) {
  HN_CHECK(a && out, HN_ERR_INVALID, "hn_csr_transpose: null argument");
  *out = nullptr;
  HN_TRY(enter());
  hipStream_t s = ctx().stream;
  hn_csr* t = new hn_csr();
  int st = csr_alloc(t, a->ncols, a->nrows, a->nnz, a->flags);
  if (st == HN_OK) st = transpose_into(a, t, s);
  if (st == HN_OK) st = build_tiling(t, 8, s);
  if (st == HN_OK) {
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) st = hip_fail(e, "hn_csr_transpose sync", __FILE__, __LINE__);
  }
  if (st != HN_OK) {
    (void)hipStreamSynchronize(s);
    delete t;
    return st;
  }
  count_launch(6);
  *out = t;
  return HN_OK;
}
