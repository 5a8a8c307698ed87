// Internal declarations shared by the translation units of libhn_kernels.so.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

#include "hn_kernels.h"

namespace hn {

// ---- error reporting ------------------------------------------------------
// The message of the last failure is kept per host thread; every entry point
// returns a status code and never throws across the C ABI.
void set_error(const std::string& msg);
const std::string& get_error();

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define HN_HIP(call)                                                        \
  do {                                                                      \
    hipError_t hn_e_ = (call);                                              \
    if (hn_e_ != hipSuccess) return ::hn::hip_fail(hn_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define HN_CHECK(cond, code, msg)                                           \
  do {                                                                      \
    if (!(cond)) return ::hn::fail((code), (msg));                          \
  } while (0)

#define HN_TRY(expr)                                                        \
  do {                                                                      \
    int hn_s_ = (expr);                                                     \
    if (hn_s_ != HN_OK) return hn_s_;                                       \
  } while (0)

// ---- library state ----------------------------------------------------------
struct Context {
  bool initialized = false;
  int device = -1;
  hipStream_t stream = nullptr;
  hn_device_info info{};
  std::atomic<uint64_t> launches{0};
};
Context& ctx();

// Makes the library device current on the calling thread; fails if hn_init
// has not succeeded.
int enter();
inline hipStream_t pick_stream(void* s) {
  return s ? static_cast<hipStream_t>(s) : ctx().stream;
}
inline void count_launch(uint64_t n = 1) { ctx().launches.fetch_add(n, std::memory_order_relaxed); }

// Checks the launch that was just enqueued.
int check_launch(const char* what);

// Grow-only device scratch owned by one object (never shrinks until freed).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int reserve(size_t want);
  void release();
  ~DevBuf() { release(); }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Library-wide scratch for the host-buffer entry points (serialised by mtx).
struct HostOpScratch {
  std::mutex mtx;
  DevBuf a, b, c, red;
};
HostOpScratch& host_scratch();
// Frees the verifier's cached device buffers (hn_finalize).
void release_verify_scratch();
// Forgets the batched solver's autotuned table (hn_finalize).
void reset_batch_tuning();

// ---- device CSR -------------------------------------------------------------
constexpr int kSpmvBlock = 256;

struct Tiling {
  int ipt = 0;            // items (row ends + nonzeros) per thread
  int ntiles = 0;
  DevBuf coord;           // int2[ntiles + 1]: (row, nnz) where each tile starts
  DevBuf carry_row;       // int[ntiles]
  DevBuf carry_val64;     // double[ntiles]
  DevBuf carry_val32;     // float[ntiles]
  // Slab layout (used instead of the merge path when slab_log2 > 0): row
  // blocks x column slabs of 2^slab_log2 columns, values copied at build.
  int slab_log2 = 0, slab_nb = 0, slab_K = 0;
  DevBuf slab_rb;         // int[nb + 1]: first row of each block
  DevBuf slab_seg;        // int[nb * K + 1]: start of (block, slab) runs
  DevBuf slab_pk;         // uint32[nnz]: row within the block << 21 | column within the slab
  DevBuf slab_val64;      // double[nnz] (if HN_F64)
  DevBuf slab_val32;      // float[nnz]  (if HN_F32)
  // Chunked stream (used when chunk_ipt > 0, ahead of both layouts above):
  // row blocks of at most chunk_rows rows, each block's nonzeros in
  // column-slab order cut into chunks of at most chunk_threads*chunk_ipt
  // nonzeros spanning at most 2^chunk_span columns, each chunk stored sorted
  // by row.
  int chunk_ipt = 0, chunk_threads = 256, chunk_span = 0, chunk_nb = 0, chunk_n = 0;
  int chunk_cb = 19;      // column bits of the packed word (the row takes the rest)
  int chunk_rows = 0;     // most rows of a block (LDS sums per workgroup)
  DevBuf chunk_rb;        // int[nb + 1]: first row of each block
  DevBuf chunk_bc;        // int[nb + 1]: first chunk of each block
  DevBuf chunk_rec;       // int4[n]: (offset, count, first column, 0); offsets are multiples of chunk_ipt
  DevBuf chunk_pk;        // uint32[padded]: row within the block << chunk_cb | column - first column
  DevBuf chunk_val64;     // double[padded] (if HN_F64)
  DevBuf chunk_val32;     // float[padded]  (if HN_F32)
};

constexpr int kSlabRows = 2048;  // rows per slab-layout block (LDS accumulators)

}  // namespace hn

struct hn_csr {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int flags = 0;
  hn::DevBuf row_ptr;     // int32[nrows + 1]
  hn::DevBuf col;         // int32[nnz]
  hn::DevBuf val64;       // double[nnz]   (if flags & HN_F64)
  hn::DevBuf val32;       // float[nnz]    (if flags & HN_F32)
  hn::Tiling tiling;      // current merge-path tiling
  hn::DevBuf spmm_scratch;
};

namespace hn {

// Builds (or rebuilds) the merge-path tiling for the given items per thread.
int build_tiling(hn_csr* a, int ipt, hipStream_t s);
// Allocates an empty CSR of the given shape (values per flags).
int csr_alloc(hn_csr* a, int64_t nrows, int64_t ncols, int64_t nnz, int flags);

// Builds the slab layout with 2^slab_log2 columns per slab (0: back to the
// merge path).
int build_slab(hn_csr* a, int slab_log2, hipStream_t s);
template <typename V>
int spmv_slab_launch(const hn_csr* a, const V* x, V* y, hipStream_t s);
// Builds the chunked stream for workgroups of `threads` (256, 512, 1024)
// threads, threads*ipt nonzeros per chunk (ipt 4 or 8) and chunks spanning
// at most 2^span_log2 columns (ipt 0: drops it).
int build_chunks(hn_csr* a, int ipt, int span_log2, int threads, hipStream_t s);
template <typename V>
int spmv_chunk_launch(const hn_csr* a, const V* x, V* y, hipStream_t s);

// t = A' (t allocated by csr_alloc with A's shape transposed and flags),
// entries of each row of t in ascending row order of A; no tiling is built.
int transpose_into(const hn_csr* a, hn_csr* t, hipStream_t s);

// y = A x on device buffers (FP64 or FP32) with the handle's current tiling.
int spmv_launch(const hn_csr* a, int prec, const void* x, void* y, hipStream_t s);

// Sorts keys[ptr[s]..ptr[s+1]) ascending for every segment s (duplicates
// allowed); synchronises the stream once.
struct SortScratch {
  DevBuf lists, counters, tmp;
};
int segmented_sort(const int* ptr, int nseg, int* keys, int64_t total, SortScratch& sc, hipStream_t s);
// out[0..n] = exclusive scan of in[0..n-1], out[n] = total (int range).
int exclusive_scan(const int* in, int64_t n, int* out, DevBuf& tmp, hipStream_t s);

// C (+)= A B for a CSR given by device arrays (B row-major ncols x k, C
// row-major nrows x k, FP64 or FP32): a wave per row, or a workgroup per row
// when the mean row holds 32 or more nonzeros; sums in a fixed order.
template <typename V>
int spmm_rows_launch(const int* rowptr, const int* col, const V* val, int64_t nrows, int64_t nnz, const V* B, V* C,
                     int k, bool accumulate, hipStream_t s);
// `batch` products over the same pattern in one launch: product y reads
// val, B and writes C at y * istride bytes past the given pointers.
template <typename V>
int spmm_rows_launch_batch(const int* rowptr, const int* col, const V* val, int64_t nrows, int64_t nnz, const V* B,
                           V* C, int k, bool accumulate, int batch, long long istride, hipStream_t s);

// Two-pass deterministic FP64 reductions into a device scalar.
int dot_launch(const double* a, const double* b, int64_t n, double* out, hipStream_t s);
int norm_inf_launch(const double* v, int64_t n, double* out, hipStream_t s);
int axpy_launch(double alpha, const double* x, double* y, int64_t n, hipStream_t s);
int fill_launch(double* p, double v, int64_t n, hipStream_t s);
// Takes the device address of the allocation-free dot product's scratch
// array on `device` (current at the call); once per device, from hn_init.
int vec_noalloc_setup(int device);

// Learned warm start on the device (hn_gnn.hip).  A job reads one scaled
// standard form (values in the graph's CSR order, b, c) and writes its start
// point (x0, s0 of length n, y0 of length m), all device pointers.
struct GnnJob {
  const double *val, *b, *c;
  double *x0, *y0, *s0;
};
// True when (m, n, rp, ci) is exactly the graph's pattern.
bool gnn_pattern_matches(const hn_gnn_graph* g, int64_t m, int64_t n, const std::vector<int64_t>& rp,
                         const std::vector<int64_t>& ci);
// Runs the network on every job, in chunks of instances that share one
// launch per kernel (gridDim.y = instances), on stream s: work queued on s
// afterwards sees every start point.
int gnn_warm_start_many(hn_gnn* net, hn_gnn_graph* g, const std::vector<GnnJob>& jobs, hipStream_t s);

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace hn
