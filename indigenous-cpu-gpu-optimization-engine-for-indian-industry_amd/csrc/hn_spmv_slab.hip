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
// chunked stream: the packed word holds the row within the block in its
// upper bits and the column relative to the chunk's first column in the low
// cb bits; cb = 19 (8,192 rows) unless the blocks must be taller to be
// resident at once, down to kChunkMinColBits (rows up to kChunkMaxRows)
constexpr int kChunkColBits = 19;
constexpr int kChunkMinColBits = 17;
constexpr int kChunkMaxRows = 80 * kNT;  // 20,480 rows: k_slab_build<80>
static_assert(kChunkMaxRows <= (1 << (32 - kChunkMinColBits)), "rows per block must fit the packed word");
constexpr int kChunkMaxItems = 8192;  // nonzeros per chunk (13-bit position in the pack sort key)

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

// One workgroup per block.  Thread t owns rows [t*RPT, t*RPT+RPT) of the
// block and keeps a cursor into each; per slab it counts each row's run
// (columns are sorted within a row, so a run is the next stretch of the
// row), scans the counts in row order and copies the runs to their places as
// (row within the block << shift | column within the slab).
template <int RPT>
__global__ __launch_bounds__(kNT) void k_slab_build(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const double* __restrict__ v64, const float* __restrict__ v32,
                                                    const int* __restrict__ rb, int K, int slab_log2, int shift,
                                                    int* __restrict__ seg, unsigned* __restrict__ pk,
                                                    double* __restrict__ sv64, float* __restrict__ sv32) {
  __shared__ int sh[kNT / kWave];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int r0 = rb[b], r1 = rb[b + 1];
  int cur[RPT], end[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int i = r0 + tid * RPT + q;
    cur[q] = (i < r1) ? rowptr[i] : 0;
    end[q] = (i < r1) ? rowptr[i + 1] : 0;
  }
  int off = rowptr[r0];
  for (int s = 0; s < K; ++s) {
    const long long base = (long long)s << slab_log2;
    const long long lim = base + (1LL << slab_log2);
    int cnt[RPT];
    int mine = 0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      int c = 0;
      while (cur[q] + c < end[q] && (long long)col[cur[q] + c] < lim) ++c;
      cnt[q] = c;
      mine += c;
    }
    int total;
    int dst = off + block_exclusive_scan(mine, sh, &total);
    if (tid == 0) seg[(long long)b * K + s] = off;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const unsigned row_bits = (unsigned)(tid * RPT + q) << shift;
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


// ---- chunked stream ---------------------------------------------------------
// The per-slab walk above pays one dependent round trip (segment bounds, the
// matrix stream, the gather, two barriers) per slab for a few hundred
// nonzeros.  The chunked stream keeps the column-slab order of each block
// but cuts the block's stream into chunks of up to 256*IPT nonzeros that run
// across slab boundaries: each thread owns IPT consecutive nonzeros (vector
// loads), gathers all of them at once, and the chunk needs one barrier.
// Inside a chunk the nonzeros are sorted by row, so a row forms one run per
// chunk; runs are summed in registers (within a thread), by a wave64
// segmented scan (across threads) and a 4-entry cross-wave pass, always in
// the same order: deterministic.  The packed word holds the row within the
// block and the column relative to the chunk's first column, so a chunk may
// span at most 2^21 (or the chosen 2^span) columns.

// One thread per block: cuts [seg[b*K], seg[b*K+K]) into chunks of at most
// CH nonzeros whose slabs lie in [s, s + span_slabs), s the slab of the
// chunk's first nonzero.  Block b writes its chunks to rec[cs[b]...] as
// (stream offset, count, first column of slab s, block) and their number to
// cnt[b]; cs[] reserves enough slots (see build_chunks).
__global__ void k_chunk_cut(const int* __restrict__ seg, int nb, int K, int slab_log2, int span_slabs, int CH,
                            const int* __restrict__ cs, int4* __restrict__ rec, int* __restrict__ cnt) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int* sg = seg + (long long)b * K;
  int o = sg[0];
  const int e = sg[K];
  int s = 0, n = 0;
  int4* out = rec + cs[b];
  const int cap = cs[b + 1] - cs[b];
  while (o < e) {
    if (n == cap) {  // cannot happen with the host's slot count; reported, never written past
      cnt[b] = -1;
      return;
    }
    while (sg[s + 1] <= o) ++s;  // seg[s] <= o < seg[s+1]; ends at s <= K-1 since o < seg[K]
    const int lim = sg[min(K, s + span_slabs)];
    const int end = min(o + CH, lim);
    out[n++] = make_int4(o, end - o, s << slab_log2, b);
    o = end;
  }
  cnt[b] = n;
}

// One workgroup per chunk: sorts the chunk's nonzeros by (row, stream
// position) in LDS (bitonic) and writes them at the chunk's padded offset
// as (row << cb | column - first column) and value.  src = (stream offset,
// count, first column, block); dst_off = padded output offset.  The stream
// words are (row << cb | column within the slab).
__global__ __launch_bounds__(1024) void k_chunk_pack(const int4* __restrict__ src, const int* __restrict__ dst_off,
                                                     const int* __restrict__ seg, int K, int slab_log2, int cb,
                                                     const unsigned* __restrict__ spk,
                                                     const double* __restrict__ sv64, const float* __restrict__ sv32,
                                                     unsigned* __restrict__ pk, double* __restrict__ v64,
                                                     float* __restrict__ v32) {
  __shared__ unsigned key[kChunkMaxItems];
  const int c = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
  const int4 r = src[c];
  const int off = r.x, n = r.y, base = r.z, b = r.w;
  const int out = dst_off[c];
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = tid; i < P; i += nt)
    key[i] = (i < n) ? (((spk[off + i] >> cb) << 13) | (unsigned)i) : 0xFFFFFFFFu;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += nt) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned ki = key[i], kl = key[l];
          if ((ki > kl) == ((i & k) == 0)) {
            key[i] = kl;
            key[l] = ki;
          }
        }
      }
      __syncthreads();
    }
  }
  const int* sg = seg + (long long)b * K;
  const unsigned in_slab = (1u << slab_log2) - 1u;
  for (int i = tid; i < n; i += nt) {
    const unsigned kk = key[i];
    const int p = off + (int)(kk & 8191u);
    // slab of stream position p: the largest s with seg[s] <= p
    int lo = 0, hi = K - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (sg[mid] <= p)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int colg = (lo << slab_log2) | (int)(spk[p] & in_slab);
    pk[out + i] = ((kk >> 13) << cb) | (unsigned)(colg - base);
    if (v64) v64[out + i] = sv64[p];
    if (v32) v32[out + i] = sv32[p];
  }
}

typedef unsigned hn_u4 __attribute__((ext_vector_type(4)));
typedef double hn_d2 __attribute__((ext_vector_type(2)));
typedef float hn_f4 __attribute__((ext_vector_type(4)));

template <int IPT>
__device__ __forceinline__ void load_words(const unsigned* p, unsigned (&w)[IPT]) {
#pragma unroll
  for (int q = 0; q < IPT / 4; ++q) {
    const hn_u4 t = __builtin_nontemporal_load(reinterpret_cast<const hn_u4*>(p) + q);
    w[4 * q + 0] = t.x;
    w[4 * q + 1] = t.y;
    w[4 * q + 2] = t.z;
    w[4 * q + 3] = t.w;
  }
}

template <int IPT>
__device__ __forceinline__ void load_vals(const double* p, double (&v)[IPT]) {
#pragma unroll
  for (int q = 0; q < IPT / 2; ++q) {
    const hn_d2 t = __builtin_nontemporal_load(reinterpret_cast<const hn_d2*>(p) + q);
    v[2 * q + 0] = t.x;
    v[2 * q + 1] = t.y;
  }
}

template <int IPT>
__device__ __forceinline__ void load_vals(const float* p, float (&v)[IPT]) {
#pragma unroll
  for (int q = 0; q < IPT / 4; ++q) {
    const hn_f4 t = __builtin_nontemporal_load(reinterpret_cast<const hn_f4*>(p) + q);
    v[4 * q + 0] = t.x;
    v[4 * q + 1] = t.y;
    v[4 * q + 2] = t.z;
    v[4 * q + 3] = t.w;
  }
}

// Wave-wide data movement on the DPP lanes of the VALU (no LDS round trip):
// row_shr:k within each row of 16 lanes, row_bcast:15/31 across rows,
// wave_shr:1 / wave_shl:1 across the wave.  Lanes without a source read 0.
constexpr int kDppRowShr = 0x110;
constexpr int kDppWaveShl1 = 0x130;
constexpr int kDppWaveShr1 = 0x138;
constexpr int kDppRowBcast15 = 0x142;
constexpr int kDppRowBcast31 = 0x143;

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, true);
}

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = dpp<CTRL, ROWS>((int)(b & 0xffffffffll)), hi = dpp<CTRL, ROWS>((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(float, dpp<CTRL, ROWS>(__builtin_bit_cast(int, v)));
}

__device__ __forceinline__ double readlane(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float readlane(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// One step of a segmented scan: (f, v) absorbs the pair (fp, vp) that comes
// before it; a set flag marks the start of a segment.
template <typename V>
__device__ __forceinline__ void seg_absorb(int& f, V& v, int fp, V vp) {
  if (!f) v = vp + v;
  f |= fp;
}

template <int CTRL, int ROWS, typename V>
__device__ __forceinline__ void seg_level(int& f, V& v) {
  const int fp = dpp<CTRL, ROWS>(f);
  const V vp = dpp<CTRL, ROWS>(v);
  seg_absorb(f, v, fp, vp);
}

// Inclusive segmented scan within each row of 16 lanes.
template <typename V>
__device__ __forceinline__ void seg_scan_rows(int& f, V& v) {
  seg_level<kDppRowShr + 1, 0xf>(f, v);
  seg_level<kDppRowShr + 2, 0xf>(f, v);
  seg_level<kDppRowShr + 4, 0xf>(f, v);
  seg_level<kDppRowShr + 8, 0xf>(f, v);
}

// Inclusive segmented scan over the wave: rows, then each row absorbs the
// row before it (lane 15 into row 1, lane 47 into row 3; then lane 31 into
// rows 2 and 3).  A fixed order, so the sums are deterministic.
template <typename V>
__device__ __forceinline__ void seg_scan_wave(int& f, V& v) {
  seg_scan_rows(f, v);
  seg_level<kDppRowBcast15, 0xa>(f, v);
  seg_level<kDppRowBcast31, 0xc>(f, v);
}

// A register copy the compiler cannot fold away: the copy of chunk c+1's
// stream into the current registers happens here, after its loads have
// arrived, so the registers that the next loads write are free again (a
// folded copy would sit at the loop's end, behind those next loads, and
// wait for them).
__device__ __forceinline__ unsigned vmov(unsigned v) {
  unsigned o;
  asm volatile("v_mov_b32 %0, %1" : "=v"(o) : "v"(v));
  return o;
}

__device__ __forceinline__ float vmov(float v) {
  return __builtin_bit_cast(float, vmov(__builtin_bit_cast(unsigned, v)));
}

__device__ __forceinline__ double vmov(double v) {
  double o;
  asm volatile("v_mov_b64 %0, %1" : "=v"(o) : "v"(v));
  return o;
}

// pk and val are not __restrict__: the compiler would then treat their loads
// as invariant and move them past the fence that keeps the next chunk's
// loads ahead of the chunk's reduction.
//
// Pipeline per chunk c: wait for chunk c's gathers and chunk c+1's stream
// (all issued one step earlier); form chunk c's products and rows; copy
// chunk c+1's stream, which has arrived, into the current registers; issue
// chunk c+1's gathers and chunk c+2's stream; then reduce chunk c (scans,
// the barrier, LDS sums) while those loads are in flight, so the next
// chunk's gathers overlap the barrier.
template <typename V, int IPT, int NT>
__global__ __launch_bounds__(NT) void k_spmv_chunk(const int* __restrict__ rb, const int* __restrict__ bc,
                                                   const int4* __restrict__ rec, const unsigned* pk, const V* val,
                                                   const V* __restrict__ x, V* __restrict__ y, int cb) {
  constexpr int NW = NT / kWave;
  static_assert(NW <= 16, "the cross-wave pass scans one row of 16 lanes");
  const unsigned col_mask = (1u << cb) - 1u;
  extern __shared__ __align__(16) unsigned char hn_chunk_lds[];
  V* ys = reinterpret_cast<V*>(hn_chunk_lds);  // the block's row sums
  // per wave and chunk parity: inclusive tail value at lane 63, whether a
  // run starts in the wave, the first row of lane 0 and the last of lane 63
  __shared__ V wv[2][NW];
  __shared__ int wf[2][NW], wfirst[2][NW], wlast[2][NW];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  const int r0 = rb[b];
  const int R = rb[b + 1] - r0;
  for (int i = tid; i < R; i += NT) ys[i] = V(0);
  const int c_beg = bc[b], c_end = bc[b + 1];
  const int k0 = tid * IPT;
  // Every load is issued unconditionally, so the compiler can wait for the
  // ones it needs alone: a chunk index past the end reads the block's last
  // chunk again (n = 0; its lines are still in L2), a thread past the end of
  // a chunk reads the chunk's first words, and words past the end inside a
  // thread are zero padding, so every gather stays inside the chunk's
  // columns; such values are never summed.
  unsigned wdC[IPT], wdN[IPT];  // current chunk (arrived), next chunk (in flight)
  V vC[IPT], vN[IPT], xv[IPT];
  int nC = 0, baseC = 0, nN = 0, baseN = 0;
  auto load_next = [&](int c) {
    const bool ok = c < c_end;
    const int4 r = rec[ok ? c : c_end - 1];
    nN = ok ? r.y : 0;
    baseN = r.z;
    const int kk = (k0 < r.y) ? k0 : 0;
    load_words<IPT>(pk + r.x + kk, wdN);
    load_vals<IPT>(val + r.x + kk, vN);
  };
  auto advance = [&]() {  // next -> current, then the current chunk's gathers
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
      wdC[u] = vmov(wdN[u]);
      vC[u] = vmov(vN[u]);
    }
    nC = nN;
    baseC = baseN;
    const V* xb = x + baseC;
#pragma unroll
    for (int u = 0; u < IPT; ++u) xv[u] = xb[wdC[u] & col_mask];
  };
  if (c_beg < c_end) {
    load_next(c_beg);
    advance();
    load_next(c_beg + 1);
  }
  for (int c = c_beg; c < c_end; ++c) {
    const int par = (c - c_beg) & 1;
    const int nv = min(IPT, max(0, nC - k0));  // valid nonzeros of this thread (a prefix)
    int row[IPT];
    V p[IPT];
#pragma unroll
    for (int u = 0; u < IPT; ++u) {
      row[u] = (u < nv) ? (int)(wdC[u] >> cb) : -1;
      p[u] = vC[u] * xv[u];
    }
    __builtin_amdgcn_sched_barrier(0);
    advance();
    load_next(c + 2);
    // keep these loads here, ahead of the reduction (an empty asm with a
    // memory clobber stops the optimiser from sinking them; it reads no
    // register, so it waits for nothing)
    asm volatile("" ::: "memory");
    __builtin_amdgcn_sched_barrier(0);
    V sum[IPT];  // sum[u]: sum of the thread's nonzeros of row[u] up to u
#pragma unroll
    for (int u = 0; u < IPT; ++u) sum[u] = (u > 0 && row[u] == row[u - 1]) ? sum[u - 1] + p[u] : p[u];
    int first = -1, last = -1;
    V tail = V(0);
    if (nv > 0) first = row[0];
#pragma unroll
    for (int u = 0; u < IPT; ++u)
      if (u == nv - 1) {
        last = row[u];
        tail = sum[u];
      }
    const bool single = nv > 0 && first == last;
    // does this thread's first run continue the previous thread's last run?
    const int prev_last = dpp<kDppWaveShr1>(last);
    bool cont = lane > 0 && nv > 0 && prev_last == first;
    // segmented scan of the tail runs; lane 0 of a wave assumes it continues
    // the previous wave when single (its carry is added below or not)
    int f = (nv == 0 || !single) ? 1 : ((lane == 0) ? 0 : (cont ? 0 : 1));
    V sv = tail;
    seg_scan_wave(f, sv);
    if (lane == kWave - 1) {
      wv[par][w] = sv;
      wf[par][w] = f;
      wlast[par][w] = last;
    }
    if (lane == 0) wfirst[par][w] = first;
    __syncthreads();
    // tail value accumulated through the end of wave w-1: a segmented scan
    // over the waves' records, lane q holding wave q's (a segment starts at
    // wave q unless its first row continues wave q-1's last)
    int cf = 0;
    V cv = V(0);
    if (lane < NW) {
      const bool cq = lane > 0 && wfirst[par][lane] >= 0 && wlast[par][lane - 1] == wfirst[par][lane];
      cf = (wf[par][lane] || !cq) ? 1 : 0;
      cv = wv[par][lane];
    }
    seg_scan_rows(cf, cv);
    const V carry = (w > 0) ? readlane(cv, w - 1) : V(0);
    if (lane == 0) cont = w > 0 && nv > 0 && wlast[par][w - 1] == first;
    const bool wave_cont = w > 0 && wfirst[par][w] >= 0 && wlast[par][w - 1] == wfirst[par][w];
    const V S = f ? sv : ((wave_cont ? carry : V(0)) + sv);  // inclusive block-wide tail value
    V S_prev = dpp<kDppWaveShr1>(S);
    if (lane == 0) S_prev = carry;
    int next_first = dpp<kDppWaveShl1>(first);
    if (lane == kWave - 1) next_first = (w + 1 < NW) ? wfirst[par][w + 1] : -1;
    if (nv > 0) {
      bool first_run = true;
#pragma unroll
      for (int u = 0; u + 1 < IPT; ++u) {
        if (u + 1 < nv && row[u + 1] != row[u]) {
          V r = sum[u];
          if (first_run && cont) r = S_prev + r;
          first_run = false;
          ys[row[u]] += r;
        }
      }
      if (next_first != last) ys[last] += S;
    }
  }
  __syncthreads();  // blocks without chunks still zeroed ys
  for (int i = tid; i < R; i += NT) y[r0 + i] = ys[i];
}
}  // namespace

namespace {

// Row blocks of at most kSlabRows rows and about nnz / (0.97 * resident
// workgroups) nonzeros, so that every block is resident at once and all walk
// the columns together (a second round would start at the first columns
// while the first is at the end, and the two would evict each other's part
// of x from L2).
std::vector<int> row_blocks(const std::vector<int>& rp, int64_t nrows, int64_t nnz, int per_cu, int max_rows) {
  const double slots = 0.97 * std::max(1, per_cu) * std::max(1, ctx().info.compute_units);
  const long long target = std::max<long long>(1024, (long long)std::ceil((double)nnz / slots));
  std::vector<int> rb{0};
  for (int64_t r = 0; r < nrows;) {
    const int64_t start = r;
    long long acc = 0;
    while (r < nrows && r - start < max_rows) {
      const long long len = rp[r + 1] - rp[r];
      if (r > start && acc + len > target) break;
      acc += len;
      ++r;
    }
    rb.push_back((int)r);
  }
  return rb;
}

// Orders the nonzeros of every row block by column slab (then row, then
// column) into pk/sv64/sv32 as (row within the block << shift | column
// within the slab), with seg[b*K + s] the start of block b's slab s and
// seg[nb*K] = nnz.  drb receives rb.  Blocks hold at most 2048 rows
// (rows_per_thread 8), 8192 (32) or 20,480 (80).
int slab_stream(const hn_csr* a, const std::vector<int>& rb, int slab_log2, long long K, int shift,
                int rows_per_thread, DevBuf& drb, DevBuf& seg, DevBuf& pk, DevBuf& sv64, DevBuf& sv32,
                hipStream_t s) {
  const long long nb = (long long)rb.size() - 1;
  HN_TRY(drb.reserve(sizeof(int) * (nb + 1)));
  HN_TRY(seg.reserve(sizeof(int) * (nb * K + 1)));
  HN_TRY(pk.reserve(sizeof(unsigned) * a->nnz));
  if (a->flags & HN_F64) HN_TRY(sv64.reserve(sizeof(double) * a->nnz));
  if (a->flags & HN_F32) HN_TRY(sv32.reserve(sizeof(float) * a->nnz));
  HN_HIP(hipMemcpyAsync(drb.p, rb.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s));
  const int last = (int)a->nnz;
  HN_HIP(hipMemcpyAsync(seg.as<int>() + nb * K, &last, sizeof(int), hipMemcpyHostToDevice, s));
  const double* v64 = (a->flags & HN_F64) ? a->val64.as<double>() : nullptr;
  const float* v32 = (a->flags & HN_F32) ? a->val32.as<float>() : nullptr;
  double* o64 = (a->flags & HN_F64) ? sv64.as<double>() : nullptr;
  float* o32 = (a->flags & HN_F32) ? sv32.as<float>() : nullptr;
#define HN_SLAB_BUILD(RPT)                                                                                      \
  hipLaunchKernelGGL(k_slab_build<RPT>, dim3((unsigned)nb), dim3(kNT), 0, s, a->row_ptr.as<int>(), a->col.as<int>(), \
                     v64, v32, drb.as<int>(), (int)K, slab_log2, shift, seg.as<int>(), pk.as<unsigned>(), o64, o32)
  if (rows_per_thread == 8)
    HN_SLAB_BUILD(8);
  else if (rows_per_thread == 32)
    HN_SLAB_BUILD(32);
  else
    HN_SLAB_BUILD(80);
#undef HN_SLAB_BUILD
  HN_TRY(check_launch("k_slab_build"));
  HN_HIP(hipStreamSynchronize(s));  // rb and last live in the caller's frame
  return HN_OK;
}

int download_row_ptr(const hn_csr* a, std::vector<int>& rp, hipStream_t s) {
  rp.resize(a->nrows + 1);
  HN_HIP(hipMemcpyAsync(rp.data(), a->row_ptr.p, sizeof(int) * (a->nrows + 1), hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  return HN_OK;
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
  std::vector<int> rp;
  HN_TRY(download_row_ptr(a, rp, s));
  int per_cu = 0;
  if (a->flags & HN_F64)
    HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv_slab<double, kSlabIpt>, kNT, 0));
  else
    HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv_slab<float, kSlabIpt>, kNT, 0));
  const std::vector<int> rb = row_blocks(rp, a->nrows, a->nnz, per_cu, kSlabRows);
  const long long nb = (long long)rb.size() - 1;
  const long long K = (a->ncols + (1LL << slab_log2) - 1) >> slab_log2;
  HN_CHECK(nb * K < INT_MAX && nb < INT_MAX, HN_ERR_UNSUPPORTED, "slab layout: too many blocks x slabs");
  HN_TRY(slab_stream(a, rb, slab_log2, K, kSlabColBits, kRowsPerThread, tl.slab_rb, tl.slab_seg, tl.slab_pk,
                     tl.slab_val64, tl.slab_val32, s));
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

namespace {

template <typename V, int IPT, int NT>
void* chunk_kernel() {
  return reinterpret_cast<void*>(&k_spmv_chunk<V, IPT, NT>);
}

template <typename V>
void* chunk_kernel_for(int ipt, int nt) {
  if (ipt == 4) {
    if (nt == 256) return chunk_kernel<V, 4, 256>();
    if (nt == 512) return chunk_kernel<V, 4, 512>();
    return chunk_kernel<V, 4, 1024>();
  }
  if (nt == 256) return chunk_kernel<V, 8, 256>();
  if (nt == 512) return chunk_kernel<V, 8, 512>();
  return chunk_kernel<V, 8, 1024>();
}

}  // namespace

int build_chunks(hn_csr* a, int ipt, int span_log2, int threads, hipStream_t s) {
  Tiling& tl = a->tiling;
  if (ipt == 0) {
    tl.chunk_ipt = 0;
    return HN_OK;
  }
  HN_CHECK(ipt == 4 || ipt == 8, HN_ERR_INVALID, "chunk items_per_thread must be 0, 4 or 8");
  HN_CHECK(span_log2 >= 4 && span_log2 <= kChunkColBits, HN_ERR_INVALID, "chunk span_log2 must be in [4, 19]");
  HN_CHECK(threads == 256 || threads == 512 || threads == 1024, HN_ERR_INVALID,
           "chunk threads must be 256, 512 or 1024");
  if (a->nrows == 0 || a->nnz == 0 || a->ncols == 0) {
    tl.chunk_ipt = 0;
    return HN_OK;
  }
  std::vector<int> rp;
  HN_TRY(download_row_ptr(a, rp, s));
  // Rows per block, summed in LDS: 8 per thread, or more when the blocks
  // would otherwise not all be resident at once (a matrix with many short
  // rows, such as the transpose of a wide one): then the fewest blocks per CU
  // whose LDS still fits, so that every block walks the columns in the same
  // sweep.
  const size_t vb = (a->flags & HN_F64) ? sizeof(double) : sizeof(float);
  const void* kern = (a->flags & HN_F64) ? chunk_kernel_for<double>(ipt, threads) : chunk_kernel_for<float>(ipt, threads);
  const int cus = std::max(1, ctx().info.compute_units);
  auto resident = [&](int p) { return std::max<int64_t>(1, (int64_t)(0.97 * p * cus)); };
  int max_rows = 8 * threads, per_cu = 0;
  HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, vb * max_rows));
  HN_CHECK(per_cu > 0, HN_ERR_UNSUPPORTED, "chunked layout: kernel does not fit a CU");
  if (resident(per_cu) * max_rows < a->nrows) {
    int lds_cap = 0;
    HN_HIP(hipDeviceGetAttribute(&lds_cap, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx().info.device));
    for (int p = per_cu; p >= 1; --p) {
      const int64_t rows = ((a->nrows + resident(p) - 1) / resident(p) + kNT - 1) / kNT * kNT;
      if (rows > kChunkMaxRows || (int64_t)vb * rows > (int64_t)lds_cap) continue;
      int q = 0;
      HN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&q, kern, threads, vb * rows));
      if (q >= p) {
        max_rows = (int)rows;
        per_cu = p;
        break;
      }
    }  // none fits: 8 rows per thread, in more than one round of blocks
  }
  int cb = kChunkColBits;
  while (cb > kChunkMinColBits && (int64_t)max_rows > (int64_t(1) << (32 - cb))) --cb;
  HN_CHECK((int64_t)max_rows <= (int64_t(1) << (32 - cb)), HN_ERR_UNSUPPORTED, "chunked layout: rows exceed the packed word");
  span_log2 = std::min(span_log2, cb);
  const std::vector<int> rb = row_blocks(rp, a->nrows, a->nnz, per_cu, max_rows);
  const long long nb = (long long)rb.size() - 1;
  // slabs only steer the cut: 2^12 columns, or the span if narrower
  const int slab_log2 = std::min(span_log2, 12);
  const long long K = (a->ncols + (1LL << slab_log2) - 1) >> slab_log2;
  const int span_slabs = 1 << (span_log2 - slab_log2);
  const int CH = threads * ipt;
  HN_CHECK(nb * K < INT_MAX, HN_ERR_UNSUPPORTED, "chunked layout: too many blocks x slabs");
  DevBuf drb, seg, spk, sv64, sv32;
  HN_TRY(slab_stream(a, rb, slab_log2, K, cb, max_rows <= kSlabRows ? 8 : (max_rows <= 32 * kNT ? 32 : 80), drb, seg,
                     spk, sv64, sv32, s));
  // Slots per block: each chunk is full (CH), ends a span window (a new
  // window starts at least span_slabs slabs further) or ends the block.
  std::vector<int> cs(nb + 1, 0);
  for (long long b = 0; b < nb; ++b) {
    const long long nz = rp[rb[b + 1]] - rp[rb[b]];
    cs[b + 1] = (int)(cs[b] + nz / CH + (K + span_slabs - 1) / span_slabs + 1);
  }
  const long long slots = cs[nb];
  HN_CHECK(slots < INT_MAX / 2, HN_ERR_UNSUPPORTED, "chunked layout: too many chunks");
  DevBuf dcs, drec, dcnt;
  HN_TRY(dcs.reserve(sizeof(int) * (nb + 1)));
  HN_TRY(drec.reserve(sizeof(int4) * slots));
  HN_TRY(dcnt.reserve(sizeof(int) * nb));
  HN_HIP(hipMemcpyAsync(dcs.p, cs.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_chunk_cut, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, s, seg.as<int>(), (int)nb, (int)K,
                     slab_log2, span_slabs, CH, dcs.as<int>(), drec.as<int4>(), dcnt.as<int>());
  HN_TRY(check_launch("k_chunk_cut"));
  std::vector<int4> rec(slots);
  std::vector<int> cnt(nb);
  HN_HIP(hipMemcpyAsync(rec.data(), drec.p, sizeof(int4) * slots, hipMemcpyDeviceToHost, s));
  HN_HIP(hipMemcpyAsync(cnt.data(), dcnt.p, sizeof(int) * nb, hipMemcpyDeviceToHost, s));
  HN_HIP(hipStreamSynchronize(s));
  // compact the chunk list; every chunk starts at a multiple of ipt
  std::vector<int4> src, out;
  std::vector<int> bc(nb + 1, 0), dst;
  long long padded = 0;
  for (long long b = 0; b < nb; ++b) {
    HN_CHECK(cnt[b] >= 0 && cnt[b] <= cs[b + 1] - cs[b], HN_ERR_UNSUPPORTED, "chunked layout: chunk slots overflow");
    for (int i = 0; i < cnt[b]; ++i) {
      const int4 r = rec[cs[b] + i];
      src.push_back(r);
      dst.push_back((int)padded);
      out.push_back(make_int4((int)padded, r.y, r.z, 0));
      padded += (r.y + ipt - 1) / ipt * ipt;
    }
    bc[b + 1] = (int)src.size();
  }
  HN_CHECK(padded < INT_MAX, HN_ERR_UNSUPPORTED, "chunked layout: too many nonzeros");
  const long long nch = (long long)src.size();
  DevBuf dsrc, ddst;
  HN_TRY(dsrc.reserve(sizeof(int4) * std::max<long long>(nch, 1)));
  HN_TRY(ddst.reserve(sizeof(int) * std::max<long long>(nch, 1)));
  HN_TRY(tl.chunk_rb.reserve(sizeof(int) * (nb + 1)));
  HN_TRY(tl.chunk_bc.reserve(sizeof(int) * (nb + 1)));
  HN_TRY(tl.chunk_rec.reserve(sizeof(int4) * std::max<long long>(nch, 1)));
  HN_TRY(tl.chunk_pk.reserve(sizeof(unsigned) * padded));
  if (a->flags & HN_F64) HN_TRY(tl.chunk_val64.reserve(sizeof(double) * padded));
  if (a->flags & HN_F32) HN_TRY(tl.chunk_val32.reserve(sizeof(float) * padded));
  // padding words read as row 0, column 0 of the chunk (never summed)
  HN_HIP(hipMemsetAsync(tl.chunk_pk.p, 0, sizeof(unsigned) * padded, s));
  if (a->flags & HN_F64) HN_HIP(hipMemsetAsync(tl.chunk_val64.p, 0, sizeof(double) * padded, s));
  if (a->flags & HN_F32) HN_HIP(hipMemsetAsync(tl.chunk_val32.p, 0, sizeof(float) * padded, s));
  HN_HIP(hipMemcpyAsync(tl.chunk_rb.p, rb.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s));
  HN_HIP(hipMemcpyAsync(tl.chunk_bc.p, bc.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, s));
  if (nch) {
    HN_HIP(hipMemcpyAsync(tl.chunk_rec.p, out.data(), sizeof(int4) * nch, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(dsrc.p, src.data(), sizeof(int4) * nch, hipMemcpyHostToDevice, s));
    HN_HIP(hipMemcpyAsync(ddst.p, dst.data(), sizeof(int) * nch, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_chunk_pack, dim3((unsigned)nch), dim3(1024), 0, s, dsrc.as<int4>(), ddst.as<int>(),
                       seg.as<int>(), (int)K, slab_log2, cb, spk.as<unsigned>(),
                       (a->flags & HN_F64) ? sv64.as<double>() : nullptr,
                       (a->flags & HN_F32) ? sv32.as<float>() : nullptr, tl.chunk_pk.as<unsigned>(),
                       (a->flags & HN_F64) ? tl.chunk_val64.as<double>() : nullptr,
                       (a->flags & HN_F32) ? tl.chunk_val32.as<float>() : nullptr);
    HN_TRY(check_launch("k_chunk_pack"));
  }
  HN_HIP(hipStreamSynchronize(s));  // host vectors and the stream buffers end with this frame
  tl.chunk_ipt = ipt;
  tl.chunk_threads = threads;
  tl.chunk_span = span_log2;
  tl.chunk_cb = cb;
  tl.chunk_rows = max_rows;
  tl.chunk_nb = (int)nb;
  tl.chunk_n = (int)nch;
  return HN_OK;
}

template <typename V>
int spmv_chunk_launch(const hn_csr* a, const V* x, V* y, hipStream_t s) {
  const Tiling& tl = a->tiling;
  const V* val = std::is_same<V, double>::value ? (const V*)tl.chunk_val64.p : (const V*)tl.chunk_val32.p;
  const int* rb = tl.chunk_rb.as<int>();
  const int* bc = tl.chunk_bc.as<int>();
  const int4* rec = tl.chunk_rec.as<int4>();
  const unsigned* pk = tl.chunk_pk.as<unsigned>();
  const size_t lds = sizeof(V) * tl.chunk_rows;
  const dim3 g(tl.chunk_nb), t(tl.chunk_threads);
  const int cb = tl.chunk_cb;
#define HN_CHUNK_LAUNCH(IPT, NT) \
  hipLaunchKernelGGL((k_spmv_chunk<V, IPT, NT>), g, t, lds, s, rb, bc, rec, pk, val, x, y, cb)
  if (tl.chunk_ipt == 4) {
    if (tl.chunk_threads == 256)
      HN_CHUNK_LAUNCH(4, 256);
    else if (tl.chunk_threads == 512)
      HN_CHUNK_LAUNCH(4, 512);
    else
      HN_CHUNK_LAUNCH(4, 1024);
  } else {
    if (tl.chunk_threads == 256)
      HN_CHUNK_LAUNCH(8, 256);
    else if (tl.chunk_threads == 512)
      HN_CHUNK_LAUNCH(8, 512);
    else
      HN_CHUNK_LAUNCH(8, 1024);
  }
#undef HN_CHUNK_LAUNCH
  HN_TRY(check_launch("k_spmv_chunk"));
  count_launch();
  return HN_OK;
}

template int spmv_chunk_launch<double>(const hn_csr*, const double*, double*, hipStream_t);
template int spmv_chunk_launch<float>(const hn_csr*, const float*, float*, hipStream_t);

}  // namespace hn
