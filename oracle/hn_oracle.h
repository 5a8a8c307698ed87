/*
 * hn_oracle.h - CPU restatement of the reference's Mehrotra IPM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the timed CPU baseline; the product path never calls it.
 *
 * Every function names the reference file:line it restates.  All arithmetic
 * is plain C, compiled without FMA contraction and without -ffast-math, so
 * elementwise expressions round exactly as the reference's C++ does.
 */
#ifndef HN_ORACLE_H
#define HN_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- sparse kernels ------------------------------------------------------ */
/* y = A x, one sequential sum per row (CPUBackend::spmv,
 * core/execution/gpu_backend.cpp:89-103; matvec, core/lp/interior_point.cpp:52-59). */
void or_spmv(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v,
             const double* x, double* y);
/* out = A' x by row-wise scatter (matvec_t, core/lp/interior_point.cpp:61-69). */
void or_spmv_t(int64_t nrows, int64_t ncols, const int64_t* rp, const int64_t* ci,
               const double* v, const double* x, double* out);
/* C = A B, B row-major ncols x k (CPUBackend::spmm, gpu_backend.cpp:105-122). */
void or_spmm(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v,
             const double* B, double* Cm, int64_t k);
/* Counting-scatter transpose (build_normal_equations,
 * core/lp/interior_point.cpp:153-173): out_rp has ncols+1 entries. */
void or_csr_transpose(int64_t nrows, int64_t ncols, const int64_t* rp, const int64_t* ci,
                      const double* v, int64_t* out_rp, int64_t* out_ci, double* out_v);

/* dot / axpy run to n, norm_inf ignores NaN the way std::max(best,|v|) does
 * (CPUBackend, gpu_backend.cpp:124-136; max_abs, interior_point.cpp:71-75). */
double or_dot(int64_t n, const double* a, const double* b);
void or_axpy(int64_t n, double alpha, const double* x, double* y);
double or_norm_inf(int64_t n, const double* v);
/* Ratio test (max_step, core/lp/interior_point.cpp:134-142). */
double or_max_step(int64_t n, const double* v, const double* d);

/* Pattern of N = M D M' over columns with D > 0, plus the diagonal
 * (build_normal_equations, interior_point.cpp:175-199, before the numeric
 * sums): returns nnz; fills out_rp (m+1) and, when out_ci is not NULL, the
 * sorted column indices.  D == NULL means every column is active. */
int64_t or_adat_pattern(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci,
                        const double* D, int64_t* out_rp, int64_t* out_ci);
/* Dense N = M diag(D) M' + dyn_reg I with the reference's rules: columns with
 * D <= 0 skipped, dyn_reg = max(reg, 1e-12 * max diag) (interior_point.cpp:144-200).
 * N is m x m row-major; returns dyn_reg. */
double or_adat_dense(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci,
                     const double* v, const double* D, double reg, double* N);

/* ---- the IPM ---------------------------------------------------------------- */
/* ProblemStatus (core/model/problem.hpp:35-45). */
enum { OR_UNKNOWN = 0, OR_OPTIMAL, OR_INFEASIBLE, OR_UNBOUNDED, OR_SUBOPTIMAL,
       OR_TIME_LIMIT, OR_ITER_LIMIT, OR_NUMERICAL_ERROR, OR_INTERRUPTED };
/* ConstraintSense (core/model/problem.hpp:24-28). */
enum { OR_LE = 0, OR_GE = 1, OR_EQ = 2 };

typedef struct or_problem {
  int64_t m, n;               /* constraints, variables */
  const int64_t* rp;          /* m+1 */
  const int64_t* ci;          /* nnz, sorted within rows, no duplicates */
  const double* v;
  const int8_t* sense;        /* m */
  const double* rhs;          /* m */
  const double* lb;           /* n */
  const double* ub;           /* n */
  const double* obj;          /* n */
  int maximize;
  double obj_offset;
} or_problem;

typedef struct or_options {   /* InteriorPointOptions, interior_point.hpp:16-27 */
  int max_iterations;         /* 100 */
  double time_limit_seconds;  /* 0 = none */
  double convergence_tol;     /* 1e-8 */
  double complementarity_tol; /* 1e-8 */
  double regularization;      /* 1e-8 */
  int interrupt;              /* nonzero: the interrupt callback returns true */
  int pivot_rule;             /* 0: the reference (r = max(reg, 1e-12 max diag),
                                 pivots <= 1e-10 clamped); 1: not in the reference -
                                 r = reg, pivots <= 1e-13 N(i,i) skipped (Wright) */
  const double* x0;           /* optional start point of the scaled standard form */
  const double* y0;           /* (x0, s0: sf_n entries, y0: sf_m; all or none)   */
  const double* s0;
} or_options;

typedef struct or_result {
  int status;
  int64_t iterations;
  double objective;
  double* primal;             /* caller-allocated, n entries (may be NULL) */
  /* scaled standard form, for KKT checks (sizes reported in sf_m / sf_n) */
  int64_t sf_m, sf_n;
  double* x;                  /* caller-allocated, capacity cap_n (may be NULL) */
  double* lam;                /* capacity cap_m */
  double* z;                  /* capacity cap_n */
  int64_t cap_m, cap_n;
  double rp_norm, rd_norm, mu;
  /* per-iteration trace (HYPERNOVA_IPM_DBG fields, interior_point.cpp:710-714) */
  double* trace;              /* 5 doubles per iteration: mu, rp, rd, ap, ad */
  int64_t trace_cap;
} or_result;

void or_default_options(or_options* o);
/* InteriorPointSolver::solve with crossover disabled (interior_point.cpp:402-886).
 * The normal equations are solved with a dense LDL' in natural order (the
 * reference uses a minimum-degree ordering; both are exact factorisations of
 * the same matrix, so only rounding differs), with the reference's pivot clamp
 * and iterative refinement. Returns 0, or -1 on a bad argument. */
int or_ipm_solve(const or_problem* p, const or_options* o, or_result* r);

/* Standard form only (build_standard_form, interior_point.cpp:202-373, after
 * remove_duplicate_rows :82-132): writes the scaled M (CSR, caller buffers
 * of the capacities returned by a first call with NULL buffers), rhs, c,
 * col_scale and offset.  Returns 0 on success. */
int or_standard_form(const or_problem* p, int64_t* sm, int64_t* sn, int64_t* snnz,
                     int64_t* rp, int64_t* ci, double* v, double* rhs, double* c,
                     double* col_scale, double* offset);

/* ---- post-solve verification ------------------------------------------- */
typedef struct or_verify_out {  /* validation::VerificationResult (solution_verifier.hpp:12-21) */
  int feasible, optimal;
  double primal_infeasibility, dual_infeasibility, complementarity, integrality_violation,
      objective_discrepancy;
  int64_t n_violated;           /* length of violated_constraints */
  int64_t* violated;            /* first min(n_violated, violated_cap) entries (may be NULL) */
  int64_t violated_cap;
  char message[160];
} or_verify_out;

/* SolutionVerifier::verify_detailed (core/validation/solution_verifier.cpp:24-88)
 * with its checks (:92-371) for a problem without quadratic terms.
 * var_type: n entries of model::VarType (0 continuous, 1 integer, 2 binary;
 * NULL = all continuous).  n_dual / n_rc are the lengths of the solution's
 * dual and reduced_costs vectors.  tol = (feasibility, optimality,
 * complementarity, integrality).  With n_primal != n only the size-mismatch
 * result is produced (the reference reads past the vector after it). */
void or_verify(const or_problem* p, const int8_t* var_type, const double* primal, int64_t n_primal,
               const double* dual, int64_t n_dual, const double* rc, int64_t n_rc,
               double objective_value, const double tol[4], or_verify_out* out);

/* Wall seconds of `repeats` products y = A x on `threads` host threads
 * (rows split by nonzero count); a baseline of this repository, not the
 * reference's CPU kernel. */
double or_spmv_mt_seconds(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v, const double* x,
                          double* y, int repeats, int threads);

#ifdef __cplusplus
}
#endif

#endif
