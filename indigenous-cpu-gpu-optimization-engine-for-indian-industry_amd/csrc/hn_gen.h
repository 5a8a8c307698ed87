/*
 * hn_gen.h - seeded synthetic inputs for the LP hot path (host C, libhn_gen.so).
 *
 * Every generator uses one self-contained stream: splitmix64 expands the seed
 * into a xoshiro256** state; uniforms are (next >> 11) * 2^-53; normals come
 * from Box-Muller on two uniforms (u1 mapped to (0,1]), both outputs used in
 * order.  Draw order is documented per generator, so the arrays are identical
 * on every host with the same libm.
 */
#ifndef HN_GEN_H
#define HN_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Constraint senses, in the order of the reference's ConstraintSense
 * (core/model/problem.hpp:24-28). */
#define HN_LE 0
#define HN_GE 1
#define HN_EQ 2

typedef struct hn_gen_lp {
  int64_t m, n, nnz;
  int64_t* row_ptr;   /* m+1, CSR of the constraint matrix, columns sorted */
  int64_t* col;       /* nnz */
  double* val;        /* nnz */
  double* rhs;        /* m */
  int8_t* sense;      /* m, HN_LE/HN_GE/HN_EQ */
  double* lb;         /* n */
  double* ub;         /* n (INFINITY = none) */
  double* obj;        /* n */
  int maximize;       /* 0 = minimise */
  double* x0;         /* n or NULL: generator's interior primal point */
  double* y0;         /* m or NULL: generator's dual point */
  double* s0;         /* n or NULL: generator's dual slacks */
} hn_gen_lp;

/* Random standard-form LP  min c'x, Ax = b, x >= 0  with k distinct uniform
 * rows per column and N(0,1) values (the whole matrix is redrawn while a row
 * is empty).  Draw order: per column j: k row indices (rejecting repeats),
 * then k values; then x0 ~ U(0.5,2) [n], y0 ~ N(0,1) [m], s0 ~ U(0.5,2) [n].
 * b = A x0 and c = A'y0 + s0 make the LP feasible and bounded. */
hn_gen_lp* hn_gen_random_lp(uint64_t seed, int64_t m, int64_t n, int k);

/* Large LP: one row multiplier 10^U(-2,2) per row drawn first [m]; then per
 * column j a count from the discrete power law P(c) ~ c^-alpha on
 * [cmin, cmax], that many distinct uniform rows, N(0,1) values times the row
 * multiplier; then x0, y0, s0 as in hn_gen_random_lp. */
hn_gen_lp* hn_gen_powerlaw_lp(uint64_t seed, int64_t m, int64_t n, double alpha, int cmin,
                              int cmax);

/* Batch instance i of the batched-LP workload: seed 1000+i, m = round(exp(U(ln mlo, ln mhi))),
 * n = round(2.5 m), then hn_gen_random_lp(seed, m, n, k) continuing the same stream. */
hn_gen_lp* hn_gen_batch_lp(uint64_t seed, int64_t mlo, int64_t mhi, int k);

/* Transportation LP: S sources x T sinks.  Draw order: source points [S x 2],
 * sink points [T x 2], supplies U(50,150) [S], raw demands U(50,150) [T]
 * (rescaled to total supply), cost noise N(0,1) [S*T, row-major].
 * Rows: S supply rows sum_j x_ij <= s_i (LE), then T demand rows
 * sum_i x_ij = d_j (EQ); variable x_ij is column i*T + j. */
hn_gen_lp* hn_gen_transport_lp(uint64_t seed, int S, int T);

void hn_gen_lp_free(hn_gen_lp* lp);

/* Power-law SpMV matrix: nrows x ncols, row lengths L_i = clamp(round(lambda Z_i),1,cap)
 * with Z_i ~ Zipf(s, cap), lambda bisected so sum L_i comes closest to nnz_target from
 * below, the remainder added one by one to the longest rows (ties: lower index).
 * Draw order: Z [nrows]; per row, distinct uniform columns (sorted); values N(0,1)
 * [nnz, CSR order]; x ~ U(-1,1) [ncols]; w ~ U(-1,1) [nrows]. */
typedef struct hn_gen_spmv {
  int64_t nrows, ncols, nnz;
  int64_t* row_ptr;
  int32_t* col;
  double* val;
  double* x;
  double* w;
} hn_gen_spmv;

hn_gen_spmv* hn_gen_powerlaw_spmv(uint64_t seed, int64_t nrows, int64_t ncols, int64_t nnz_target,
                                  double s, int64_t cap);
void hn_gen_spmv_free(hn_gen_spmv* g);

/* n doubles from the stream of `seed`: U(lo,hi). */
void hn_gen_uniform(uint64_t seed, double lo, double hi, int64_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif /* HN_GEN_H */
