/*
 * hn_kernels.h - C ABI of libhn_kernels.so, the MI355X (gfx950) kernel library
 * behind the HyperNova compute-backend plugin.
 *
 * The reference solver reaches its GPU through two layers:
 *   - the C++ plugin interface IComputeBackend
 *     (core/execution/gpu_backend.hpp:18-51), and
 *   - a dynamically loaded C-ABI driver table (core/execution/cuda_driver.hpp:41-93,
 *     bound by dlopen/dlsym in core/execution/cuda_driver.cpp:24-48,112-145).
 * This header is the replacement for the second layer.  Every entry point is
 * extern "C", takes plain pointers and sizes, and returns an int status
 * (HN_OK == 0).  On failure hn_last_error() returns a message, mirroring the
 * reference's "last CUresult" slot (cuda_driver.cpp:65-73,356-358).
 *
 * Pointer conventions
 *   - Functions whose names end in _dev take DEVICE pointers and an optional
 *     stream (hipStream_t passed as void*, NULL = the library stream).  They
 *     enqueue work and return without synchronising.
 *   - All other functions take HOST pointers, are synchronous, and behave like
 *     the corresponding IComputeBackend method.
 *
 * Threading: every call re-establishes the device chosen by hn_init(), so any
 * host thread may call.  Calls that use the same matrix handle or the library
 * stream must be serialised by the caller (the reference serialises all calls
 * with one mutex, gpu_backend.cpp:160,205).
 */
#ifndef HN_KERNELS_H
#define HN_KERNELS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define HN_OK 0
#define HN_ERR_INVALID 1      /* bad argument (null pointer, size mismatch)   */
#define HN_ERR_HIP 2          /* a HIP runtime call failed                    */
#define HN_ERR_NOMEM 3        /* device allocation failed                     */
#define HN_ERR_UNSUPPORTED 4  /* shape outside what the kernels support       */
#define HN_ERR_NOT_INIT 5     /* hn_init() has not succeeded                  */
#define HN_ERR_NO_DEVICE 6    /* no usable gfx950 device                      */

/* ---- precision selectors for _dev entry points -------------------------- */
#define HN_F64 1
#define HN_F32 2

/* ---- lifecycle ---------------------------------------------------------- */
typedef struct hn_device_info {
  char name[256];          /* marketing name of the device                    */
  char arch[64];           /* gcnArchName, e.g. "gfx950:sramecc+:xnack-"       */
  uint64_t total_memory;   /* bytes of device memory                          */
  int compute_units;
  int device;              /* ordinal used                                    */
} hn_device_info;

/* Select the device (-1 = the one with most memory, as the reference does in
 * CudaDriverBackend::initialize, gpu_backend.cpp:159-192), create the library
 * stream.  Replaces IComputeBackend::initialize / DriverApi::load+init_context
 * (gpu_backend.hpp:25, cuda_driver.cpp:112-145,204-236).  Idempotent. */
int hn_init(int device, hn_device_info* info);
/* Releases the stream, workspaces and every live matrix handle.
 * Replaces IComputeBackend::finalize (gpu_backend.hpp:26). */
int hn_finalize(void);
/* Copies the message of the last failure on this thread into buf.
 * Replaces DriverApi::last_result/result_string (cuda_driver.hpp:41-93). */
int hn_last_error(char* buf, size_t len);
/* Replaces IComputeBackend::device_memory/free_memory (gpu_backend.hpp:45-46). */
int hn_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Replaces was_gpu_kernel_executed/gpu_kernel_executions/reset_execution_stats
 * (gpu_backend.hpp:48-50): number of hot-path kernel launches so far. */
uint64_t hn_kernel_executions(void);
void hn_reset_execution_stats(void);
/* No counterpart in the reference, whose tests reach their CPU fallbacks by
 * disabling the GPU: for callers that must exercise a per-call fallback
 * (AutoBackend::spmv retries a failed device call on the CPU,
 * gpu_backend.cpp:688-694) with a device present.  The next `count` entry
 * points that use the device fail with `code` (an HN_ERR_* value) before
 * touching it, and hn_last_error() reports an injected failure; an entry
 * point that calls another one inside the library consumes one for each.
 * count 0 clears the injection. */
int hn_inject_failures(uint32_t count, int code);

/* ---- device-resident CSR matrices --------------------------------------- */
typedef struct hn_csr hn_csr;

/* Upload a CSR matrix (64-bit indices as in the reference's std::size_t
 * vectors, gpu_backend.hpp:28-32) once; the handle keeps it resident.
 * Indices are narrowed to int32 on the device (nnz, nrows, ncols < 2^31,
 * otherwise HN_ERR_UNSUPPORTED).  flags: HN_F64 and/or HN_F32 copies of the
 * values to keep.  Replaces the pointer-keyed residency cache
 * CudaDriverBackend::resident_matrix/upload_matrix (gpu_backend.cpp:287-306,
 * 360-393) with an explicit handle. */
int hn_csr_upload(const double* values, const uint64_t* col_indices,
                  const uint64_t* row_ptr, uint64_t nrows, uint64_t ncols,
                  uint64_t nnz, int flags, hn_csr** out);
/* Build the transpose on the device by a stable counting sort: entries of
 * each output row appear in increasing source-row order, exactly the layout
 * of the counting scatter in build_normal_equations
 * (core/lp/interior_point.cpp:153-173).  Duplicates and explicit zeros are
 * kept (the scatter keeps them; SparseMatrix::transpose would drop zero sums,
 * core/numerical/sparse_matrix.cpp:84-96,250-254). */
int hn_csr_transpose(const hn_csr* a, hn_csr** out);
int hn_csr_free(hn_csr* a);
int hn_csr_dims(const hn_csr* a, uint64_t* nrows, uint64_t* ncols, uint64_t* nnz);
/* The sparsity pattern of the normal-equation matrix N = A D A' + r I of an
 * IPM step whose scaling D is positive everywhere, built on the device as
 * the solver builds it once per LP: the union of the column cliques of A
 * and the diagonal, each row's columns ascending.  Replaces the symbolic
 * half of build_normal_equations (core/lp/interior_point.cpp:144-200, whose
 * from_triplets, core/numerical/sparse_matrix.cpp:8-53, merges equal
 * entries): bit-exact with that pattern when no entry sums to exactly zero.
 * *out is a new nrows x nrows handle whose values are zero. */
int hn_csr_adat_pattern(const hn_csr* a, hn_csr** out);

/* ---- sparse LDL' on the exact minimum-degree ordering ------------------- */
typedef struct hn_sldl hn_sldl;
/* SparseCholesky::analyze (core/numerical/factorization.cpp:1240-1245):
 * exact minimum-degree ordering with the lowest index breaking ties, the
 * row patterns of L, the elimination tree and the postorder
 * (compute_md_order_and_pattern :103-191, build_elimination_graph :71-97,
 * compute_elimination_tree / compute_postorder :1327-1362) - the same
 * permutation and patterns as the reference, computed on the host (no device
 * call; usable before hn_init).  The pattern is the n x n CSR (row_ptr,
 * col); values are not needed.  work_limit > 0 stops the ordering once it
 * has merged that many adjacency entries (the fill of dense-filling graphs
 * grows as n^3): then *out is still created and hn_sldl_sizes reports
 * aborted = 1. */
int hn_sldl_analyze(uint64_t n, const uint64_t* row_ptr, const uint64_t* col, double work_limit,
                    hn_sldl** out);
/* n, entries of L (diagonal included), number of dependency levels (the
 * device factor runs one level at a time), aborted. */
int hn_sldl_sizes(const hn_sldl* h, uint64_t* n, uint64_t* l_nnz, uint64_t* levels, int* aborted);
/* SparseCholesky::perm / L_row_ptr / L_col_indices (factorization.hpp:
 * 190-193) and the symbolic parent / postorder; any pointer may be NULL. */
int hn_sldl_symbolic(const hn_sldl* h, uint64_t* perm, uint64_t* l_row_ptr, uint64_t* l_col,
                     uint64_t* parent, uint64_t* postorder);
/* SparseCholesky::factorize / refactorize (factorization.cpp:1230-1238,
 * 1248-1325) on the device: LDL' of P A P' for the device matrix a (FP64
 * values, the analysed pattern, ascending columns per row), bit-identical
 * to the reference's L and D; pivots <= singular_tol are replaced by
 * singular_tol (the reference's 1e-10) and counted in *clamped (may be
 * NULL). */
int hn_sldl_factor(hn_sldl* h, const hn_csr* a, double singular_tol, int* clamped);
/* SparseCholesky::L_values / D_values (factorization.hpp:189-190) copied to
 * the host (l_nnz and n entries; either may be NULL). */
int hn_sldl_factors(const hn_sldl* h, double* l_values, double* d_values);
/* SparseCholesky::solve (factorization.cpp:1364-1396): x = A^{-1} b on
 * device buffers (b, x of n entries, not aliased), bit-identical to the
 * reference's; enqueued on the stream (NULL: the library stream). */
int hn_sldl_solve_dev(const hn_sldl* h, const double* b, double* x, void* stream);
int hn_sldl_free(hn_sldl* h);
/* Copy the device CSR back (row_ptr: nrows+1, col/val: nnz entries). */
int hn_csr_download(const hn_csr* a, double* values, uint64_t* col_indices,
                    uint64_t* row_ptr);
/* Merge-path tiling of SpMV: items (row ends + nonzeros) per workgroup is
 * 256*items_per_thread.  Valid items_per_thread: 4, 8, 16. */
int hn_csr_set_tiling(hn_csr* a, int items_per_thread);
int hn_csr_get_tiling(const hn_csr* a, int* items_per_thread);
/* Slab layout of SpMV, for matrices whose x is larger than L2 (no reference
 * counterpart: a storage choice of this library behind hn_spmv).  Rows are
 * grouped into blocks of at most 2048 rows with balanced nonzeros; inside a
 * block the nonzeros are stored by column slab (2^slab_log2 columns), then
 * row, then column.  One workgroup per block walks the slabs in order with
 * its rows' sums in LDS, so the workgroups of an XCD gather from one
 * L2-resident slab of x at a time.  slab_log2 in [4, 21]; 0 returns to the
 * merge path (as does hn_csr_set_tiling).  The layout holds its own copy of
 * the values. */
int hn_csr_set_slab(hn_csr* a, int slab_log2);
int hn_csr_get_slab(const hn_csr* a, int* slab_log2);
/* Chunked stream of SpMV (no reference counterpart: a storage choice of this
 * library behind hn_spmv; it takes precedence over the two layouts above
 * until hn_csr_set_tiling, hn_csr_set_slab or items_per_thread 0 drops it).
 * Rows are grouped into blocks of at most 8*threads rows with balanced
 * nonzeros, sized so that every block is resident at once; each block's
 * nonzeros, in column order (by slabs of 2^12 columns), are cut into chunks
 * of at most threads*items_per_thread nonzeros spanning at most
 * 2^span_log2 columns, and each chunk is stored sorted by row.  One
 * workgroup of `threads` threads per block takes its chunks in order with
 * the rows' sums in LDS: each thread gathers items_per_thread nonzeros at
 * once while the next chunk streams in, and a chunk needs one barrier.
 * items_per_thread: 0, 4 or 8; span_log2 in [4, 19]; threads 256, 512 or
 * 1024.  Sums are in a fixed order (deterministic). */
int hn_csr_set_chunks(hn_csr* a, int items_per_thread, int span_log2, int threads);
/* items_per_thread, span_log2 and threads of the chunked stream (all 0 when
 * it is not in use). */
int hn_csr_get_chunks(const hn_csr* a, int* items_per_thread, int* span_log2, int* threads);
/* Time every valid layout on this matrix (merge path with each valid
 * items_per_thread; slab layout with 2^14..2^19 columns per slab and chunked
 * stream with 256, 512 or 1024 threads, 4 or 8 items per thread and spans
 * 2^17..2^19, when x spans more than one slab; iters launches each) and keep
 * the fastest.
 * best_us (may be NULL) receives the best time per product in
 * microseconds. */
int hn_csr_autotune(hn_csr* a, int iters, double* best_us);

/* ---- drop-in host-buffer operations (IComputeBackend semantics) -------- */
/* y = A*x with host buffers; x has ncols entries, y receives nrows entries.
 * Replaces IComputeBackend::spmv (gpu_backend.hpp:28-32,
 * CudaDriverBackend::spmv gpu_backend.cpp:200-215). */
int hn_spmv(const hn_csr* a, const double* x, double* y);
/* C = A*B, B row-major ncols x k, C row-major nrows x k.
 * Replaces IComputeBackend::spmm (gpu_backend.hpp:34-39). */
int hn_spmm(const hn_csr* a, const double* B, double* C, uint64_t k);
/* Replace IComputeBackend::dot/axpy/norm_inf (gpu_backend.hpp:41-43). */
int hn_dot(const double* a, const double* b, uint64_t n, double* out);
int hn_axpy(double alpha, const double* x, double* y, uint64_t n);
int hn_norm_inf(const double* v, uint64_t n, double* out);

/* ---- device-resident operations ----------------------------------------- */
/* y = A*x on device buffers in the given precision (HN_F64: double x/y,
 * HN_F32: float x/y with float accumulation). */
int hn_spmv_dev(const hn_csr* a, int prec, const void* x, void* y, void* stream);
/* C = A*B, device buffers, row-major, FP64. */
int hn_spmm_dev(const hn_csr* a, const double* B, double* C, uint64_t k, void* stream);
/* out[0] = a.b, out[0] = max|v|, y += alpha*x on device buffers (FP64);
 * out is a device pointer. */
int hn_dot_dev(const double* a, const double* b, uint64_t n, double* out, void* stream);
int hn_norm_inf_dev(const double* v, uint64_t n, double* out, void* stream);
/* hn_dot_dev that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory - code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * partial sums live in a fixed 8 KiB device array (1,024 doubles) that each
 * GPU holds until the process ends; hn_init takes its address.  Same kernels
 * and grid as hn_dot_dev, so the result is bit-identical.  n up to 2^62
 * (HN_ERR_INVALID above).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams, or different host
 * threads) race and must be serialised by the caller. */
int hn_dot_dev_noalloc(const double* a, const double* b, uint64_t n, double* out, void* stream);
int hn_axpy_dev(double alpha, const double* x, double* y, uint64_t n, void* stream);

/* The Mehrotra ratio test of both step lengths, with the solver's own
 * kernels: out[0] = min(1, 0.99 min_{dx_i < 0} -x_i/dx_i) and out[1] the same
 * for (z, dz) - 0.99 even when no component blocks (max_step,
 * core/lp/interior_point.cpp:134-142, used at :629-630,687-688).  Minima are
 * order-free, so both values are bit-identical to the reference's loop.
 * out is a device pointer to 2 doubles; n < 2^31. */
int hn_ratio_test_dev(const double* x, const double* dx, const double* z, const double* dz, uint64_t n,
                      double* out, void* stream);

/* ---- the interior-point solver ------------------------------------------ */
/* Constraint senses in the order of ConstraintSense (core/model/problem.hpp:24-28). */
#define HN_LE 0
#define HN_GE 1
#define HN_EQ 2

/* Result status in the order of ProblemStatus (core/model/problem.hpp:35-45). */
#define HN_STATUS_UNKNOWN 0
#define HN_STATUS_OPTIMAL 1
#define HN_STATUS_INFEASIBLE 2
#define HN_STATUS_UNBOUNDED 3
#define HN_STATUS_SUBOPTIMAL 4
#define HN_STATUS_TIME_LIMIT 5
#define HN_STATUS_ITER_LIMIT 6
#define HN_STATUS_NUMERICAL_ERROR 7
#define HN_STATUS_INTERRUPTED 8

/* An LP in the shape of the reference's model::Problem
 * (core/model/problem.hpp:79-170): CSR constraint rows, one sense and rhs per
 * row, bounds and objective per variable.  Rows are normalised the way
 * ProblemBuilder::build does (columns sorted, duplicates summed, exact zeros
 * dropped, problem.hpp:265-281). */
typedef struct hn_lp {
  int64_t m, n;
  const int64_t* row_ptr;  /* m+1 */
  const int64_t* col;      /* nnz */
  const double* val;       /* nnz */
  const int8_t* sense;     /* m: HN_LE / HN_GE / HN_EQ */
  const double* rhs;       /* m */
  const double* lb;        /* n (-INFINITY = none) */
  const double* ub;        /* n (+INFINITY = none) */
  const double* obj;       /* n */
  int maximize;
  double obj_offset;
  /* Optional warm start of the scaled standard form (x0, s0: sf_n entries,
   * y0: sf_m; all three or none), used by hn_ipm_solve_batch.  It replaces
   * the standard start point only if it passes the centrality check
   * described at hn_ipm_solve_batch. */
  const double* x0;
  const double* y0;
  const double* s0;
} hn_lp;

/* Normal-equation solver of each Newton step. */
#define HN_NE_AUTO 0   /* dense up to dense_max_m rows, else sparse or PCG   */
#define HN_NE_DENSE 1  /* dense inverse of M D M' + FP64 iterative refinement */
#define HN_NE_PCG 2    /* Jacobi-preconditioned CG on M D M' (matrix-free)   */
#define HN_NE_SPARSE 3 /* the reference's factorisation: sparse LDL' of      */
                       /* M D M' + r I on the exact minimum-degree ordering  */
                       /* (hn_sldl_*), FP64 iterative refinement; FP64 and   */
                       /* HN_PIVOT_CLAMP only.  HN_NE_AUTO (FP64, clamp)     */
                       /* takes it above dense_max_m when N has at most 2e7  */
                       /* candidate entries and the ordering stays within a  */
                       /* work budget (3e8 merged entries, 1e8 entries of    */
                       /* L), else PCG                                       */
/* Arithmetic of the normal-equation solve. */
#define HN_PREC_F64 1   /* plain FP64 everywhere                              */
#define HN_PREC_MIXED 2 /* FP32 bulk, FP64 residuals/refinement, FP64 below mu_switch */
#define HN_PREC_AUTO 3  /* hn_ipm_solve_batch only: per LP, the precision      */
                        /* hn_batch_autotune measured faster for its m bucket  */
                        /* (FP64 where nothing was measured)                   */

/* Pivot rule of the normal-equation factorisation.
 * HN_PIVOT_CLAMP is the reference: N gets r = max(regularization,
 * 1e-12 * max diag N) on its diagonal (interior_point.cpp:188-197) and LDL'
 * pivots <= 1e-10 are replaced by 1e-10 (factorization.cpp:1309-1313).
 * HN_PIVOT_SKIP is not in the reference: r = regularization only, and a
 * pivot below 1e-13 (FP32 factors: 1e-6) times its row's diagonal of N is
 * skipped - its component of the solution is set to zero, the rule of
 * S. Wright, "Modified Cholesky factorizations in interior-point algorithms
 * for linear programming", SIAM J. Optim. 9 (1999).  On degenerate LPs
 * (transportation problems, many of the random batch instances) the
 * reference rule stalls with ||rp|| far above 1e-8 until the iteration limit;
 * the skip rule reaches the reference's 1e-8 stopping test in fewer
 * iterations. */
#define HN_PIVOT_CLAMP 0
#define HN_PIVOT_SKIP 1

typedef int (*hn_interrupt_fn)(void* user);

/* InteriorPointOptions (core/lp/interior_point.hpp:16-27) plus the GPU
 * choices.  Crossover (the CPU simplex polish, :23-24) is not part of this
 * library: the caller runs it on the returned point if it wants one. */
typedef struct hn_ipm_opts {
  int max_iterations;          /* 100 */
  double time_limit_seconds;   /* 0 = none */
  double convergence_tol;      /* 1e-8 */
  double complementarity_tol;  /* 1e-8 */
  double regularization;       /* 1e-8 */
  hn_interrupt_fn interrupt_callback; /* may be NULL */
  void* interrupt_user;
  int solver;                  /* HN_NE_* */
  int precision;               /* HN_PREC_* */
  double mu_switch;            /* mixed: FP64 normal equations once mu < mu_switch (1e-4) */
  int64_t dense_max_m;         /* HN_NE_AUTO threshold (40000) */
  int pcg_max_iter;            /* per solve (2000) */
  const double* x0;            /* optional warm start of the scaled standard form: */
  const double* y0;            /* x0, s0 of length sf_n, y0 of length sf_m (all    */
  const double* s0;            /* three or none; else the LP's own x0/y0/s0); used */
                               /* only if it passes the test at hn_ipm_solve_batch */
  int verbose;                 /* nonzero: per-iteration trace on stderr; batch:  */
                               /* phase times of the call, and with 2 or more the */
                               /* kernel's per-phase workgroup time as well       */
  int pivot_rule;              /* HN_PIVOT_CLAMP (reference, default) or HN_PIVOT_SKIP */
  double mu_target;            /* > 0: stop at the first iterate with mu <= mu_target and
                                  return it (status HN_STATUS_ITER_LIMIT unless the 1e-8
                                  test was met first); 0 = off */
  double gap_tol;              /* > 0: the stopping test (interior_point.cpp:525-530) also
                                  requires the relative gap of the caller's LP (rel_gap of
                                  hn_ipm_result) to be <= gap_tol, so a solve meets a
                                  relative-gap bar and not only the absolute mu bar; 0 = off
                                  (the reference's test).  hn_ipm_solve computes rel_gap
                                  exactly on the host at each iterate that passes the
                                  absolute test; hn_ipm_solve_batch tests
                                  |c_s'x_s - b_s'lam| / (1 + |c'x|) in the kernel, which is
                                  that gap when no column has a finite nonzero bound
                                  shift (otherwise it differs by sum_j shift_j rd_j) */
  double endgame_mu;           /* > 0 (not in the reference; hn_ipm_solve_batch only): at
                                  the first iterate with mu <= endgame_mu and ||rp||, ||rd||
                                  < 1e-4, and again each time mu has halved or 4 more
                                  iterations have passed, project
                                  onto the estimated optimal face B = {j : x_j >= s_j}:
                                  x' = x_B + D_B M'w with N_B w = b - M x_B (x'_N = 0),
                                  y' = y + d with N_B d = M D_B (c - M'y) (least squares
                                  for s'_B = 0), s' = c - M'y', N_B = M D_B M' + r I,
                                  D_B = x_B, factored under the reference's pivot rule.  The projection ends the solve (OPTIMAL,
                                  result.projected = 1) when x' >= -tol and the point
                                  with x', s' clamped at 0 meets the stop test (rp,
                                  rd = max(-s', 0), mu) and, with gap_tol, the gap;
                                  otherwise the iterates continue unchanged.  It finishes
                                  degenerate LPs (transportation problems) whose dual
                                  steps stall near the optimum; 0 = off */
  int loop;                    /* hn_ipm_solve, dense normal equations in FP64 (not in the
                                  reference, which runs on the host): HN_LOOP_GRAPH (default)
                                  keeps each iteration on the device between two reads of
                                  its scalars and replays it as a captured HIP graph;
                                  HN_LOOP_DEVICE the same without graph capture (for tools
                                  that cannot follow graph launches); HN_LOOP_HOST decides
                                  sigma and the refinement sweeps on the host, with three
                                  reads per iteration.  The iterates are the same in all
                                  three. */
} hn_ipm_opts;

#define HN_LOOP_GRAPH 0
#define HN_LOOP_DEVICE 1
#define HN_LOOP_HOST 2

typedef struct hn_ipm_result {
  int status;                  /* HN_STATUS_* */
  int64_t iterations;
  double objective;            /* original space, obj_offset included */
  double solve_ms;             /* wall time of hn_ipm_solve */
  /* caller buffers (each may be NULL): primal has n entries; x, s hold the
   * scaled standard form (capacity cap_n), y capacity cap_m */
  double* primal;
  double* x;
  double* y;
  double* s;
  int64_t cap_m, cap_n;
  int64_t sf_m, sf_n;          /* out: standard-form dimensions */
  /* the reference's stopping quantities on the scaled form (interior_point.cpp:525-530) */
  double rp_norm, rd_norm, mu;
  /* relative KKT measures on the caller's own LP (all its rows and columns,
   * minimisation sense).  The point is mapped back from the scaled standard
   * form: x by to_original_primal (interior_point.cpp:375-395), row duals
   * y_i = +-gamma_i lam_r (0 for a dropped duplicate row), one net bound
   * multiplier z_j = lower - upper per column from the slots' s.  In the
   * standard-form case (EQ rows, x >= 0) they are exactly
   * ||Ax - b||_inf / (1 + ||b||_inf), ||A'y + z - c||_inf / (1 + ||c||_inf)
   * and |c'x - b'y| / (1 + |c'x|); in general the primal one takes row
   * violations by sense and bound violations, the dual one adds the LE/GE
   * rows' sign conditions and the gap adds the bound terms of the dual
   * objective. */
  double rel_primal;
  double rel_dual;
  double rel_gap;
  /* the same three on the scaled standard form, with its rp, rd, c and rhs */
  double scaled_rel_primal;
  double scaled_rel_dual;
  double scaled_rel_gap;
  int64_t ne_solves;           /* normal-equation solves */
  int64_t refine_sweeps;       /* FP64 refinement sweeps */
  int64_t cg_iterations;       /* PCG iterations */
  int64_t fp64_from_iter;      /* first iteration solved in FP64 (-1: never) */
  int warm_started;            /* 1 when x0/y0/s0 were used; 2 when they were used
                                  and then abandoned: mu fell below its tolerance
                                  while a residual stayed above its own for three
                                  iterations, and the solve restarted from the
                                  standard point */
  int projected;               /* 1 when the endgame_mu projection ended the solve */
  int64_t sparse_l_nnz;        /* entries of the sparse LDL' factor (0: not used) */
  int64_t device_steps;        /* iterations run on the device between two reads
                                  (hn_ipm_opts.loop), graph replays included */
  int64_t host_redos;          /* of those, iterations whose solve needed more
                                  refinement sweeps than the device step holds: the step
                                  was undone and run again on the host */
} hn_ipm_result;

int hn_ipm_default_opts(hn_ipm_opts* o);
/* The scaled standard form the IPM iterates on (build_standard_form,
 * interior_point.cpp:202-373, after remove_duplicate_rows :82-132): M as CSR
 * (sf_m x sf_n), rhs, c.  Call once with null buffers to get the sizes, then
 * with buffers of those sizes (row_ptr: sf_m + 1, col/val: nnz). */
int hn_lp_standard_form(const hn_lp* lp, int64_t* sf_m, int64_t* sf_n, int64_t* nnz, int64_t* row_ptr,
                        int64_t* col, double* val, double* rhs, double* c);
/* InteriorPointSolver::solve (core/lp/interior_point.cpp:402-886) with the
 * Newton steps computed on the GPU.  Same standard form, start point, step
 * rule, sigma, termination and status mapping as the reference; the normal
 * equations are solved by the method in opts->solver. */
int hn_ipm_solve(const hn_lp* lp, const hn_ipm_opts* opts, hn_ipm_result* res);
/* Solves `count` independent LPs in one persistent kernel, one workgroup per
 * LP (largest first from a device work queue), each with the same algorithm
 * as hn_ipm_solve and a dense LDL' of its normal equations (standard-form
 * rows <= 20000 per LP).  opts->precision selects FP64 or FP32-factor +
 * FP64-refinement, or per LP by the autotuned table (HN_PREC_AUTO,
 * hn_batch_autotune), opts->pivot_rule the factorisation's pivot rule and
 * opts->mu_target an early stop; solver, time limit, interrupt and the
 * opts->x0/y0/s0 warm start are not used.  An LP's own lps[k].x0/y0/s0 warm
 * start is used when it is finite, x0 > 0, s0 > 0, its mu0 = x0's0/n is below
 * the mu of the standard start, min_j x0_j s0_j >= 1e-3 mu0 (centrality) and
 * ||M x0 - b||_inf is no larger than at the standard start; otherwise the
 * standard start point is used (results[k].warm_started tells which).  results[k] receives LP k (solve_ms is the wall time of the call).
 * No reference counterpart: the reference solves one LP per call
 * (InteriorPointSolver::solve, interior_point.cpp:402). */
int hn_ipm_solve_batch(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, hn_ipm_result* results);

/* The batched solver's choices per size of LP, measured on the device. */
#define HN_TUNE_BUCKETS 6
typedef struct {
  int m_hi[HN_TUNE_BUCKETS];       /* bucket b holds standard-form m <= m_hi[b] (and > m_hi[b-1]) */
  int lps[HN_TUNE_BUCKETS];        /* sample LPs timed in the bucket (0: none, FP64 kept)         */
  int precision[HN_TUNE_BUCKETS];  /* HN_PREC_F64 or HN_PREC_MIXED: the faster one               */
  double ms_f64[HN_TUNE_BUCKETS];  /* batched wall time of the bucket's sample, FP64              */
  double ms_mixed[HN_TUNE_BUCKETS];/* ... with the FP32 factor and FP64 refinement                */
} hn_batch_tuning;
/* Autotuner of the batched solver.  Splits the sample `lps` into m buckets
 * (standard-form rows <= 128, 256, 512, 1024, 2048, larger), solves each
 * bucket's LPs (at most per_bucket of them) as one batch in each precision
 * under *opts (pivot rule, gap and end-game settings as the caller will use
 * them; precision ignored), after one untimed warm-up batch of the whole
 * sample, and keeps the faster precision per bucket.  The table is stored
 * for the device (hn_ipm_solve_batch with HN_PREC_AUTO uses it until the
 * next call or hn_finalize) and copied to *out when out is not null.
 * No reference counterpart: the reference has no batched solver. */
int hn_batch_autotune(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, uint32_t per_bucket,
                      hn_batch_tuning* out);

/* ---- learned warm start ------------------------------------------------- */
/* A bipartite constraint-variable graph network (4 message-passing layers,
 * hidden width 64, BF16 MFMA dense layers with FP32 accumulation) that maps
 * an LP's scaled standard form (M, rhs, c) to a predicted interior point
 * (log x, log s, y), trained on points the IPM itself produced.  No
 * reference counterpart: the reference has no learned component; its
 * warm-start hook is the (x0, y0, s0) start point these predictions fill. */
typedef struct hn_gnn hn_gnn;             /* parameters + Adam state (FP32)  */
typedef struct hn_gnn_graph hn_gnn_graph; /* a sparsity pattern + workspace  */
#define HN_GNN_HIDDEN 64
#define HN_GNN_LAYERS 4
/* Glorot-uniform weights from a splitmix64 stream seeded by `seed`. */
int hn_gnn_create(uint64_t seed, hn_gnn** out);
int hn_gnn_free(hn_gnn* g);
int hn_gnn_num_params(const hn_gnn* g, uint64_t* n);
/* Flat FP32 parameter vector (host); set_params also clears the Adam state. */
int hn_gnn_get_params(const hn_gnn* g, float* params);
int hn_gnn_set_params(hn_gnn* g, const float* params);
/* The pattern of M (CSR, host, 64-bit indices); instances of the same
 * pattern reuse one graph. */
int hn_gnn_graph_create(int64_t m, int64_t n, const int64_t* row_ptr, const int64_t* col, hn_gnn_graph** out);
int hn_gnn_graph_free(hn_gnn_graph* gr);
/* Forward on device inputs (val: nnz FP64 values of M, rhs: m, c: n) into
 * device outputs log_x, log_s (n) and y (m) (each may be NULL). */
int hn_gnn_forward_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs, const double* c,
                       float* log_x, float* log_s, float* y, void* stream);
/* The warm start (x0, y0, s0) of the scaled standard form from one forward
 * pass: x0 = exp(log x), y0 = y, s0 = max(c - M'y, exp(log s)) - dual
 * feasible wherever the predicted reduced cost covers the predicted s.
 * Device FP64 outputs (n, m, n), ready for hn_lp.x0/y0/s0. */
int hn_gnn_warm_start_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs, const double* c,
                          double* x0, double* y0, double* s0, void* stream);
/* Loss and its gradient with respect to every parameter for device targets
 * tx = log x*, ts = log s* (n) and ty = y* (m); grad_host receives
 * hn_gnn_num_params floats, loss_host one float (synchronous). */
int hn_gnn_gradients_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs, const double* c,
                         const float* tx, const float* ts, const float* ty, float* grad_host, float* loss_host);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on one instance; the loss
 * before the step goes to *loss_dev (device, may be NULL). */
int hn_gnn_train_step_dev(hn_gnn* g, hn_gnn_graph* gr, const double* val, const double* rhs, const double* c,
                          const float* tx, const float* ts, const float* ty, float lr, float* loss_dev,
                          void* stream);
/* hn_ipm_solve_batch with learned starts (no reference counterpart: a
 * caller that solves many LPs of one family, e.g. the transportation family,
 * gets its warm starts without a host round trip).  Every LP without its own
 * x0/y0/s0 whose scaled standard form has exactly the graph's pattern is
 * started from the network's prediction, computed on the device from the
 * standard form the solver has packed (several LPs at once on separate
 * streams); the solver tests the prediction as it tests a caller's warm
 * start and falls back to the standard start point when it fails
 * (result.warm_started = 0).  Other LPs start as in hn_ipm_solve_batch.
 * net and graph must not be used by another call at the same time. */
int hn_ipm_solve_batch_gnn(const hn_lp* lps, uint32_t count, const hn_ipm_opts* opts, hn_gnn* net,
                           hn_gnn_graph* graph, hn_ipm_result* results);

/* ---- post-solve solution verification ------------------------------------
 * Replaces SolutionVerifier::verify_detailed
 * (core/validation/solution_verifier.cpp:24-88, declared in
 * solution_verifier.hpp:23-41), which api/cpp/api.cpp:510,523 runs after
 * every solve.  The checks, their tolerances and their order are the
 * reference's; the row activities, reduced costs and objective are summed
 * in the reference's order without fused multiply-adds, so every value,
 * flag and index comes out bit-identical.  The reference's dual check scans
 * all nonzeros once per column (O(n nnz)); here it reads the column of a
 * transpose built on the device. */
typedef struct hn_tolerances {   /* numerical::ToleranceConfig fields read by the verifier */
  double feasibility_tol;        /* 1e-9 (industrial_defaults, tolerance.hpp:60-75) */
  double optimality_tol;         /* 1e-9 */
  double complementarity_tol;    /* 1e-6 */
  double integrality_tol;        /* 1e-6 */
} hn_tolerances;
int hn_default_tolerances(hn_tolerances* t);

/* model::Solution (core/model/problem.hpp) as the verifier reads it.  A
 * dual of another length than lp->m, or reduced costs of another length than
 * lp->n, are treated as absent, as in the reference. */
typedef struct hn_solution {
  const double* primal;          /* n_primal entries */
  uint64_t n_primal;
  const double* dual;            /* n_dual entries (may be NULL with n_dual 0) */
  uint64_t n_dual;
  const double* reduced_costs;   /* n_reduced_costs entries (may be NULL) */
  uint64_t n_reduced_costs;
  double objective_value;
} hn_solution;

/* var_type codes: model::VarType order (problem.hpp:16-22). */
#define HN_VAR_CONTINUOUS 0
#define HN_VAR_INTEGER 1
#define HN_VAR_BINARY 2

typedef struct hn_verify_result {  /* validation::VerificationResult */
  int feasible;
  int optimal;
  double primal_infeasibility;
  double dual_infeasibility;
  double complementarity;
  double integrality_violation;
  double objective_discrepancy;
  /* violated_constraints: column j for a bound violation (listed twice if
   * both bounds are violated), n + i for row i, in the reference's order.
   * n_violated is the full count; the first min(n_violated, capacity)
   * entries go to `violated` when it is not NULL. */
  uint64_t n_violated;
  uint64_t* violated;
  uint64_t violated_capacity;
  char message[160];               /* the reference's message text */
  double device_ms;                /* kernel time of the checks (HIP events) */
} hn_verify_result;

/* var_type: n entries of HN_VAR_* (NULL: all continuous).  lp->x0/y0/s0 are
 * ignored.  With a primal of the wrong length the result is "Primal
 * infeasible: Solution vector size mismatch" with optimal = 0 and no other
 * check (the reference goes on to read past the vector). */
int hn_verify(const hn_lp* lp, const int8_t* var_type, const hn_solution* sol, const hn_tolerances* tol,
              hn_verify_result* out);

/* ---- harness helpers (device buffers, copies, events) ------------------- */
int hn_dev_alloc(size_t bytes, void** ptr);
int hn_dev_free(void* ptr);
int hn_memcpy_h2d(void* dst, const void* src, size_t bytes);
int hn_memcpy_d2h(void* dst, const void* src, size_t bytes);
int hn_stream_sync(void* stream);
/* Plain device-to-device copy kernel (16-byte vector loads/stores); bytes
 * must be a multiple of 16.  Used as the known-size reference when scaling
 * hardware byte counters. */
int hn_copy_dev(void* dst, const void* src, size_t bytes, void* stream);
/* Elapsed milliseconds between two events recorded on a stream. */
typedef struct hn_event hn_event;
int hn_event_create(hn_event** ev);
int hn_event_destroy(hn_event* ev);
int hn_event_record(hn_event* ev, void* stream);
int hn_event_elapsed_ms(const hn_event* start, const hn_event* stop, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* HN_KERNELS_H */
