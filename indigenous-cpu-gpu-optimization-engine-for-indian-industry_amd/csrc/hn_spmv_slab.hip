// Slab layout for CSR SpMV on matrices whose x does not fit in L2.
//
// Reference semantics as the merge-path kernel: y[i] = sum_k val[k] *
// x[col[k]] over row i (CPUBackend::spmv, core/execution/gpu_backend.cpp:
// 89-103).  Empty rows give 0.
//
// Why: with columns spread over millions of entries of x, nearly every
// gather x[col[k]] of a row-ordered product misses the 4 MB L2 of its XCD
// and pulls a whole 128-byte line for 8 useful bytes; that, not the matrix
// stream, bounds the product.  The slab layout orders the work so that, at
// any moment, the workgroups of an XCD gather from one narrow slab of x
// that stays L2-resident.
//
// Layout (built once on the device from the CSR):
//   * rows are cut into blocks of at most kSlabRows consecutive rows with a
//     balanced number of nonzeros (one workgroup per block);
//   * columns are cut into K slabs of 2^slab_log2 columns;
//   * inside a block the nonzeros are stored slab by slab, and inside a
//     slab in CSR order (row, then column), each as its value and one
//     32-bit word (row within the block << 21 | column within the slab);
//   * seg[b*K + s] is where block b's slab s starts (seg[nb*K] = nnz).
// Product: the workgroup of block b keeps the block's partial sums in LDS
// and walks the slabs in order.  Per chunk of up to 256*IPT nonzeros of a
// slab it stages the products and row indices in LDS (entry k of the chunk
// at k = u*256 + thread, so neighbouring lanes touch neighbouring banks);
// the thread that holds the first entry of each row's run sums the run in
// order and adds it to the row's LDS accumulator.  Each row has one run per
// chunk, so every accumulator is updated by one thread at a time, in a
// fixed order: the result is deterministic.  Blocks are sized so that all
// of them are resident at once; all walk the slabs in the same order, so
// the workgroups of an XCD gather from the same L2-resident slab.  What
// remains is the gather itself: one L2 access per nonzero.
#include "hn_device.hpp"
#include "hn_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

namespace hn {

namespace {

constexpr int kNT = 256;
constexpr int kSlabIpt = 4;  // nonzeros per thread and chunk
constexpr int kRowsPerThread = kSlabRows / kNT;
static_assert(kSlabRows % kNT == 0, "block rows must split evenly over the threads");
constexpr int kSlabColBits = 21;  // column within a slab; the row within the block takes the upper 11 bits
static_assert(kSlabRows <= (1 << (32 - kSlabColBits)), "row within a block must fit the packed word");

// Exclusive prefix of v over the workgroup (thread order); *total receives
// the sum.  sh holds kNT/64 ints.
__device__ int block_exclusive_scan(int v, int* sh, int* total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  int inc = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const int t = __shfl_up(inc, off, kWave);
    if (lane >= off) inc += t;
  }
  if (lane == kWave - 1) sh[w] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int u = 0; u < kNT / kWave; ++u) {
    if (u < w) base += sh[u];
    tot += sh[u];
  }
  __syncthreads();  // sh is reused by the next call
  *total = tot;
  return base + inc - v;
}

// One workgroup per block.  Thread t owns rows [t*8, t*8+8) of the block
// and keeps a cursor into each; per slab it counts each row's run (columns
// are sorted within a row, so a run is the next stretch of the row), scans
// the counts in row order and copies the runs to their places.
__global__ __launch_bounds__(kNT) void k_slab_build(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const double* __restrict__ v64, const float* __restrict__ v32,
                                                    const int* __restrict__ rb, int K, int slab_log2,
                                                    int* __restrict__ seg, unsigned* __restrict__ pk,
                                                    double* __restrict__ sv64, float* __restrict__ sv32) {
  __shared__ int sh[kNT / kWave];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int r0 = rb[b], r1 = rb[b + 1];
  int cur[kRowsPerThread], end[kRowsPerThread];
#pragma unroll
  for (int q = 0; q < kRowsPerThread; ++q) {
    const int i = r0 + tid * kRowsPerThread + q;
    cur[q] = (i < r1) ? rowptr[i] : 0;
    end[q] = (i < r1) ? rowptr[i + 1] : 0;
  }
  int off = rowptr[r0];
  for (int s = 0; s < K; ++s) {
    const long long base = (long long)s << slab_log2;
    const long long lim = base + (1LL << slab_log2);
    int cnt[kRowsPerThread];
    int mine = 0;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
      int c = 0;
      while (cur[q] + c < end[q] && (long long)col[cur[q] + c] < lim) ++c;
      cnt[q] = c;
      mine += c;
    }
    int total;
    int dst = off + block_exclusive_scan(mine, sh, &total);
    if (tid == 0) seg[(long long)b * K + s] = off;
#pragma unroll
    for (int q = 0; q < kRowsPerThread; ++q) {
      const unsigned row_bits = (unsigned)(tid * kRowsPerThread + q) << kSlabColBits;
      for (int e = 0; e < cnt[q]; ++e, ++dst) {
        const int src = cur[q] + e;
        pk[dst] = row_bits | (unsigned)(col[src] - base);
        if (sv64) sv64[dst] = v64[src];
        if (sv32) sv32[dst] = v32[src];
      }
      cur[q] += cnt[q];
    }
    off += total;
  }
}

template <typename V, int IPT>
__global__ __launch_bounds__(kNT) void k_spmv_slab(const int* __restrict__ rb, const int* __restrict__ seg,
                                                   const unsigned* __restrict__ pk, const V* __restrict__ sval,
                                                   const V* __restrict__ x, V* __restrict__ y, int K, int slab_log2) {
  constexpr int CH = kNT * IPT;
  constexpr unsigned kColMask = (1u << kSlabColBits) - 1u;
  __shared__ V ys[kSlabRows];
  __shared__ V prod[CH];
  __shared__ unsigned short lr[CH];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int r0 = rb[b];
  const int R = rb[b + 1] - r0;
  for (int i = tid; i < R; i += kNT) ys[i] = V(0);
  const int* sg = seg + (long long)b * K;
  for (int s = 0; s < K; ++s) {
    const int beg = sg[s], end = sg[s + 1];
    const V* xs = x + ((long long)s << slab_log2);
    for (int c0 = beg; c0 < end; c0 += CH) {
      const int n = min(CH, end - c0);
#pragma unroll
      for (int u = 0; u < IPT; ++u) {
        const int k = u * kNT + tid;
        if (k < n) {
          const unsigned w = __builtin_nontemporal_load(pk + c0 + k);
          const V v = __builtin_nontemporal_load(sval + c0 + k);
          lr[k] = (unsigned short)(w >> kSlabColBits);
          prod[k] = v * xs[w & kColMask];
        }
      }
      __syncthreads();  // also orders the zeroing of ys before the first adds
#pragma unroll
      for (int u = 0; u < IPT; ++u) {
        const int q = u * kNT + tid;
        if (q < n && (q == 0 || lr[q] != lr[q - 1])) {
          const int key = lr[q];
          V acc = prod[q];
          for (int q2 = q + 1; q2 < n && lr[q2] == key; ++q2) acc += prod[q2];
          ys[key] += acc;
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();  // blocks whose slabs are all empty still zeroed ys
  for (int i = tid; i < R; i += kNT) y[r0 + i] = ys[i];
}

}  // namespace

int build_slab(hn_csr* a, int slab_log2, hipStream_t s) {
  Tiling& tl = a->tiling;
  if (slab_log2 == 0) {
    tl.slab_log2 = 0;
    return HN_OK;
  }
  HN_CHECK(slab_log2 >= 4 && slab_log2 <= kSlabColBits, HN_ERR_INVALID, "slab_log2 must be 0 or in [4, 21]");
  if (a->nrows == 0 || a->nnz == 0 || a->ncols == 0) {  // nothing to gather: the merge path handles these
    tl.slab_log2 = 0;
    return HN_OK;
  }
  std::vector<int> rp(a->nrows + 1);
  HN_HIP(hipMemcpyAsync(rp.data(), a->row_ptr.p, sizeof(int) * (a->nrows + 1), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  // Blocks of at most kSlabRows rows and about nnz / (0.97 * resident
  // workgroups) nonzeros, so that every block is resident at once and all
  // walk the slabs together (a second round would start at slab 0 while the
  // first is at the end, and the two would evict each other's slabs).
  int per_cu = 0;
  if (a->flags & HN_F64)
    HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv_slab<double, kSlabIpt>, kNT, 0));
  else
    HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv_slab<float, kSlabIpt>, kNT, 0));
  const double slots = 0.97 * std::max(1, per_cu) * std::max(1, ctx().info.compute_units);
  const long long target = std::max<long long>(1024, (long long)std::ceil((double)a->nnz / slots));
  std::vector<int> rb{0};
  for (int64_t r = 0; r < a->nrows;) {
    const int64_t start = r;
    long long acc = 0;
    while (r < a->nrows && r - start < kSlabRows) {
      const long long len = rp[r + 1] - rp[r];
      if (r > start && acc + len > target) break;
      acc += len;
      ++r;
    }
    rb.push_back((int)r);
  }
  const long long nb = (long long)rb.size() - 1;
  const long long K = (a->ncols + (1LL << slab_log2) - 1) >> slab_log2;
  HN_CHECK(nb * K < INT_MAX && nb < INT_MAX, HN_ERR_UNSUPPORTED, "slab layout: too many blocks x slabs");
  HN_TRY(tl.slab_rb.reserve(sizeof(int) * (nb + 1)));
  HN_TRY(tl.slab_seg.reserve(sizeof(int) * (nb * K + 1)));
  HN_TRY(tl.slab_pk.reserve(sizeof(unsigned) * a->nnz));
  if (a->flags & HN_F64) HN_TRY(tl.slab_val64.reserve(sizeof(double) * a->nnz));
  if (a->flags & HN_F32) HN_TRY(tl.slab_val32.reserve(sizeof(float) * a->nnz));
  HN_HIP(hipMemcpyAsync(tl.slab_rb.p, rb.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s));
  const int last = (int)a->nnz;
  HN_HIP(hipMemcpyAsync(tl.slab_seg.as<int>() + nb * K, &last, sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_slab_build, dim3((unsigned)nb), dim3(kNT), 0, s, a->row_ptr.as<int>(), a->col.as<int>(),
                     (a->flags & HN_F64) ? a->val64.as<double>() : nullptr,
                     (a->flags & HN_F32) ? a->val32.as<float>() : nullptr, tl.slab_rb.as<int>(), (int)K, slab_log2,
                     tl.slab_seg.as<int>(), tl.slab_pk.as<unsigned>(),
                     (a->flags & HN_F64) ? tl.slab_val64.as<double>() : nullptr,
                     (a->flags & HN_F32) ? tl.slab_val32.as<float>() : nullptr);
  HN_TRY(check_launch("k_slab_build"));
  HN_HIP(hipStreamSynchronize(s));  // rb and last live on this stack frame
  tl.slab_log2 = slab_log2;
  tl.slab_nb = (int)nb;
  tl.slab_K = (int)K;
  return HN_OK;
}

template <typename V>
int spmv_slab_launch(const hn_csr* a, const V* x, V* y, hipStream_t s) {
  const Tiling& tl = a->tiling;
  const V* sval = std::is_same<V, double>::value ? (const V*)tl.slab_val64.p : (const V*)tl.slab_val32.p;
  hipLaunchKernelGGL((k_spmv_slab<V, kSlabIpt>), dim3(tl.slab_nb), dim3(kNT), 0, s, tl.slab_rb.as<int>(),
                     tl.slab_seg.as<int>(), tl.slab_pk.as<unsigned>(), sval, x, y, tl.slab_K, tl.slab_log2);
  HN_TRY(check_launch("k_spmv_slab"));
  count_launch();
  return HN_OK;
}

template int spmv_slab_launch<double>(const hn_csr*, const double*, double*, hipStream_t);
template int spmv_slab_launch<float>(const hn_csr*, const float*, float*, hipStream_t);

}  // namespace hn
