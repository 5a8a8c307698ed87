/*
 * hn_oracle.c - CPU restatement of the reference's Mehrotra IPM hot path.
 *
 * TEST INFRASTRUCTURE: the checker for the HIP path and the CPU baseline of
 * bench.py.  Never linked into, or called by, the product libraries.
 *
 * Follows core/lp/interior_point.cpp:30-886, core/numerical/refinement.cpp:7-36,
 * core/numerical/factorization.cpp:1248-1396 (numeric LDL' and solve),
 * core/numerical/sparse_matrix.cpp:8-53,230-262 (from_triplets, sum_duplicates)
 * and core/execution/gpu_backend.cpp:80-139 (CPU backend).  Each function
 * cites the lines it restates.
 */
#define _POSIX_C_SOURCE 199309L
#include "hn_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <pthread.h>

/* ---- sparse kernels ------------------------------------------------------ */
void or_spmv(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v,
             const double* x, double* y) {
  /* gpu_backend.cpp:89-103 */
  for (int64_t i = 0; i < nrows; ++i) {
    double sum = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) sum += v[k] * x[ci[k]];
    y[i] = sum;
  }
}

void or_spmv_t(int64_t nrows, int64_t ncols, const int64_t* rp, const int64_t* ci,
               const double* v, const double* x, double* out) {
  /* interior_point.cpp:61-69: out starts at zero, scatter row by row */
  for (int64_t j = 0; j < ncols; ++j) out[j] = 0.0;
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) out[ci[k]] += v[k] * x[i];
}

void or_spmm(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v,
             const double* B, double* Cm, int64_t k) {
  /* gpu_backend.cpp:105-122 */
  for (int64_t i = 0; i < nrows * k; ++i) Cm[i] = 0.0;
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
      const int64_t j = ci[e];
      const double a = v[e];
      for (int64_t c = 0; c < k; ++c) Cm[i * k + c] += a * B[j * k + c];
    }
}

void or_csr_transpose(int64_t nrows, int64_t ncols, const int64_t* rp, const int64_t* ci,
                      const double* v, int64_t* out_rp, int64_t* out_ci, double* out_v) {
  /* interior_point.cpp:153-173: count, prefix, scatter in row order */
  for (int64_t j = 0; j <= ncols; ++j) out_rp[j] = 0;
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) out_rp[ci[k] + 1]++;
  for (int64_t j = 0; j < ncols; ++j) out_rp[j + 1] += out_rp[j];
  int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (ncols > 0 ? ncols : 1));
  for (int64_t j = 0; j < ncols; ++j) cur[j] = out_rp[j];
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t j = ci[k];
      out_ci[cur[j]] = i;
      out_v[cur[j]] = v[k];
      ++cur[j];
    }
  free(cur);
}

double or_dot(int64_t n, const double* a, const double* b) {
  double s = 0.0; /* gpu_backend.cpp:124-128 */
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

void or_axpy(int64_t n, double alpha, const double* x, double* y) {
  for (int64_t i = 0; i < n; ++i) y[i] += alpha * x[i]; /* gpu_backend.cpp:129-131 */
}

double or_norm_inf(int64_t n, const double* v) {
  double best = 0.0; /* interior_point.cpp:71-75: std::max(best, |v|) keeps best on NaN */
  for (int64_t i = 0; i < n; ++i) {
    const double a = fabs(v[i]);
    best = (best < a) ? a : best;
  }
  return best;
}

double or_max_step(int64_t n, const double* v, const double* d) {
  double alpha = 1.0; /* interior_point.cpp:134-142 */
  for (int64_t i = 0; i < n; ++i)
    if (d[i] < 0.0) {
      const double q = -v[i] / d[i];
      alpha = (q < alpha) ? q : alpha;
    }
  const double s = 0.99 * alpha;
  return (s < 1.0) ? s : 1.0;
}

/* ---- triplets -> CSR (from_triplets + sum_duplicates) --------------------- */
typedef struct {
  int64_t row, col;
  double val;
  int64_t seq;
} trip_t;

static int trip_cmp(const void* pa, const void* pb) {
  const trip_t* a = (const trip_t*)pa;
  const trip_t* b = (const trip_t*)pb;
  if (a->row != b->row) return a->row < b->row ? -1 : 1;
  if (a->col != b->col) return a->col < b->col ? -1 : 1;
  return (a->seq < b->seq) ? -1 : (a->seq > b->seq);
}

typedef struct {
  int64_t rows, cols, nnz;
  int64_t* rp;
  int64_t* ci;
  double* v;
} csr_t;

static void csr_free(csr_t* a) {
  free(a->rp);
  free(a->ci);
  free(a->v);
  memset(a, 0, sizeof(*a));
}

/* sparse_matrix.cpp:8-53 and :230-262.  Duplicates are summed in insertion
 * order (the reference's unstable std::sort leaves their order unspecified);
 * exact-zero sums are dropped as the reference does (:250). */
static void csr_from_triplets(int64_t rows, int64_t cols, trip_t* t, int64_t nt, csr_t* out) {
  for (int64_t k = 0; k < nt; ++k) t[k].seq = k;
  qsort(t, (size_t)nt, sizeof(trip_t), trip_cmp);
  out->rows = rows;
  out->cols = cols;
  out->rp = (int64_t*)calloc((size_t)rows + 1, sizeof(int64_t));
  out->ci = (int64_t*)malloc(sizeof(int64_t) * (nt > 0 ? nt : 1));
  out->v = (double*)malloc(sizeof(double) * (nt > 0 ? nt : 1));
  int64_t w = 0, k = 0;
  for (int64_t i = 0; i < rows; ++i) {
    out->rp[i] = w;
    while (k < nt && t[k].row == i) {
      const int64_t c = t[k].col;
      double sum = t[k].val;
      ++k;
      while (k < nt && t[k].row == i && t[k].col == c) {
        sum += t[k].val;
        ++k;
      }
      if (sum != 0.0) {
        out->ci[w] = c;
        out->v[w] = sum;
        ++w;
      }
    }
  }
  out->rp[rows] = w;
  out->nnz = w;
}

/* ---- normal equations ------------------------------------------------------ */
static int cmp_i64(const void* a, const void* b) {
  const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

int64_t or_adat_pattern(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci,
                        const double* D, int64_t* out_rp, int64_t* out_ci) {
  /* interior_point.cpp:175-199: triplets (crows[p], crows[q]) for every pair of
   * rows of each column with D > 0, plus (i, i) for every row; from_triplets
   * merges them, so the pattern is the union of column cliques and the diagonal. */
  int64_t* crp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  int64_t* cri = (int64_t*)malloc(sizeof(int64_t) * (rp[m] > 0 ? rp[m] : 1));
  double* dummy = (double*)malloc(sizeof(double) * (rp[m] > 0 ? rp[m] : 1));
  double* ones = (double*)malloc(sizeof(double) * (rp[m] > 0 ? rp[m] : 1));
  for (int64_t k = 0; k < rp[m]; ++k) ones[k] = 1.0;
  or_csr_transpose(m, n, rp, ci, ones, crp, cri, dummy);
  int64_t* mark = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
  int64_t* row = (int64_t*)malloc(sizeof(int64_t) * (m > 0 ? m : 1));
  for (int64_t i = 0; i < m; ++i) mark[i] = -1;
  int64_t total = 0;
  for (int64_t i = 0; i < m; ++i) {
    int64_t len = 0;
    mark[i] = i;
    row[len++] = i;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t j = ci[k];
      if (D && !(D[j] > 0.0)) continue;
      for (int64_t q = crp[j]; q < crp[j + 1]; ++q) {
        const int64_t r = cri[q];
        if (mark[r] != i) {
          mark[r] = i;
          row[len++] = r;
        }
      }
    }
    qsort(row, (size_t)len, sizeof(int64_t), cmp_i64);
    out_rp[i] = total;
    if (out_ci) memcpy(out_ci + total, row, sizeof(int64_t) * len);
    total += len;
  }
  out_rp[m] = total;
  free(crp);
  free(cri);
  free(dummy);
  free(ones);
  free(mark);
  free(row);
  return total;
}

static double adat_dense_rel(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                             const double* D, double reg, double rel_reg, double* N);

double or_adat_dense(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci,
                     const double* v, const double* D, double reg, double* N) {
  return adat_dense_rel(m, n, rp, ci, v, D, reg, 1e-12, N);
}

/* rel_reg = 1e-12 is the reference's rule; 0 leaves r = reg (the skip rule). */
static double adat_dense_rel(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                             const double* D, double reg, double rel_reg, double* N) {
  /* interior_point.cpp:144-200 */
  int64_t* crp = (int64_t*)malloc(sizeof(int64_t) * (n + 1));
  int64_t* cri = (int64_t*)malloc(sizeof(int64_t) * (rp[m] > 0 ? rp[m] : 1));
  double* cv = (double*)malloc(sizeof(double) * (rp[m] > 0 ? rp[m] : 1));
  or_csr_transpose(m, n, rp, ci, v, crp, cri, cv);
  memset(N, 0, sizeof(double) * (size_t)m * (size_t)m);
  for (int64_t j = 0; j < n; ++j) {
    if (D[j] <= 0.0) continue;
    const double d = D[j];
    for (int64_t p = crp[j]; p < crp[j + 1]; ++p)
      for (int64_t q = crp[j]; q < crp[j + 1]; ++q) N[cri[p] * m + cri[q]] += cv[p] * cv[q] * d;
  }
  double max_diag = 0.0;
  for (int64_t i = 0; i < m; ++i) max_diag = (max_diag < N[i * m + i]) ? N[i * m + i] : max_diag;
  double dyn = rel_reg * max_diag;
  if (dyn < reg) dyn = reg;
  for (int64_t i = 0; i < m; ++i) N[i * m + i] += dyn;
  free(crp);
  free(cri);
  free(cv);
  return dyn;
}

/* ---- dense LDL' with the reference's pivot clamp -------------------------- */
/* factorization.cpp:1273-1318 restated for a dense matrix in natural order:
 * L unit lower, D diagonal, pivots <= singular_tol (1e-10, tolerance.hpp:70)
 * replaced by singular_tol.  L is stored row-major in the lower triangle of F
 * with D on the diagonal. */
/* skip_rel > 0 replaces the clamp by the skip rule of S. Wright (SIAM J.
 * Optim. 9, 1999): a pivot <= skip_rel * N(i,i) becomes 1e128, so its column
 * of L and its component of every solve vanish.  This rule is not in the
 * reference; it backs the HN_PIVOT_SKIP option of the GPU solvers. */
static void ldlt_factor(int64_t m, const double* N, double* F, double singular_tol, double skip_rel) {
  double* w = (double*)malloc(sizeof(double) * (m > 0 ? m : 1));
  for (int64_t i = 0; i < m; ++i) {
    double* Li = F + i * m;
    for (int64_t j = 0; j < i; ++j) {
      const double* Lj = F + j * m;
      double sum = N[i * m + j];
      for (int64_t k = 0; k < j; ++k) sum -= Lj[k] * w[k];
      w[j] = sum;                 /* = L(i,j) * D(j) */
      Li[j] = sum / F[j * m + j];
    }
    double sum = N[i * m + i];
    for (int64_t k = 0; k < i; ++k) sum -= Li[k] * w[k];
    if (skip_rel > 0.0) {
      if (!(sum > skip_rel * N[i * m + i])) sum = 1e128;
    } else if (sum <= singular_tol) {
      sum = singular_tol;
    }
    Li[i] = sum;
  }
  free(w);
}

static void ldlt_solve(int64_t m, const double* F, double* x) {
  /* factorization.cpp:1364-1396 */
  for (int64_t i = 0; i < m; ++i) {
    double sum = x[i];
    const double* Li = F + i * m;
    for (int64_t j = 0; j < i; ++j) sum -= Li[j] * x[j];
    x[i] = sum;
  }
  for (int64_t i = 0; i < m; ++i)
    if (F[i * m + i] != 0.0) x[i] /= F[i * m + i];
  for (int64_t i = m - 1; i >= 0; --i) {
    const double* Li = F + i * m;
    const double xi = x[i];
    for (int64_t j = 0; j < i; ++j) x[j] -= Li[j] * xi;
  }
}

/* refinement.cpp:7-36 */
static void iterative_refinement(int64_t m, const double* F, const double* N, double* x,
                                 const double* rhs, double relative_tol, int max_iter) {
  if (m == 0) return;
  double norm_rhs = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    const double a = fabs(rhs[i]);
    norm_rhs = (norm_rhs < a) ? a : norm_rhs;
  }
  if (norm_rhs == 0.0 || norm_rhs != norm_rhs) return;
  double* ax = (double*)malloc(sizeof(double) * m);
  double* r = (double*)malloc(sizeof(double) * m);
  for (int it = 0; it < max_iter; ++it) {
    for (int64_t i = 0; i < m; ++i) {
      double s = 0.0;
      const double* Ni = N + i * m;
      for (int64_t j = 0; j < m; ++j)
        if (Ni[j] != 0.0) s += Ni[j] * x[j];
      ax[i] = s;
    }
    double rel = 0.0;
    for (int64_t i = 0; i < m; ++i) {
      const double a = fabs(rhs[i] - ax[i]);
      rel = (rel < a) ? a : rel;
    }
    rel /= norm_rhs;
    if (rel <= relative_tol) break;
    for (int64_t i = 0; i < m; ++i) r[i] = rhs[i] - ax[i];
    ldlt_solve(m, F, r);
    for (int64_t i = 0; i < m; ++i) x[i] += r[i];
  }
  free(ax);
  free(r);
}

/* ---- problem normalisation and standard form ---------------------------- */
typedef struct {
  int is_fixed;
  double fixed;
  int ns;
  int64_t slot[2];
  double sign[2];
  double shift[2];
} varinfo_t;

typedef struct {
  csr_t M;
  double* rhs;
  double* c;
  double* col_scale;
  double offset;
  int64_t m, n;
  varinfo_t* vi;
  int64_t nvar;
} stdform_t;

static void sf_free(stdform_t* f) {
  csr_free(&f->M);
  free(f->rhs);
  free(f->c);
  free(f->col_scale);
  free(f->vi);
  memset(f, 0, sizeof(*f));
}

static uint64_t dbits(double d) {
  uint64_t u;
  memcpy(&u, &d, sizeof(u));
  return u;
}

/* Row signature hash for remove_duplicate_rows (exact comparison follows). */
static uint64_t row_hash(const or_problem* p, int64_t i) {
  uint64_t h = 1469598103934665603ULL;
#define MIX(x)                         \
  do {                                 \
    h ^= (uint64_t)(x);                \
    h *= 1099511628211ULL;             \
    h ^= h >> 29;                      \
  } while (0)
  for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
    MIX(p->ci[k]);
    MIX(dbits(p->v[k]));
  }
  MIX(p->sense[i] + 7);
  MIX(dbits(p->rhs[i]));
#undef MIX
  return h;
}

static int rows_equal(const or_problem* p, int64_t a, int64_t b) {
  if (p->sense[a] != p->sense[b]) return 0;
  if (dbits(p->rhs[a]) != dbits(p->rhs[b])) return 0;
  const int64_t la = p->rp[a + 1] - p->rp[a], lb = p->rp[b + 1] - p->rp[b];
  if (la != lb) return 0;
  for (int64_t u = 0; u < la; ++u) {
    if (p->ci[p->rp[a] + u] != p->ci[p->rp[b] + u]) return 0;
    if (dbits(p->v[p->rp[a] + u]) != dbits(p->v[p->rp[b] + u])) return 0;
  }
  return 1;
}

typedef struct {
  uint64_t h;
  int64_t row;
} hrow_t;

static int hrow_cmp(const void* pa, const void* pb) {
  const hrow_t* a = (const hrow_t*)pa;
  const hrow_t* b = (const hrow_t*)pb;
  if (a->h != b->h) return a->h < b->h ? -1 : 1;
  return (a->row > b->row) - (a->row < b->row);
}

/* interior_point.cpp:82-132: a row whose (sorted entries, sense, rhs bits)
 * equal those of an earlier row is dropped.  keep[i] = 1 for kept rows. */
static int64_t remove_duplicate_rows(const or_problem* p, char* keep) {
  const int64_t m = p->m;
  hrow_t* hr = (hrow_t*)malloc(sizeof(hrow_t) * (m > 0 ? m : 1));
  for (int64_t i = 0; i < m; ++i) {
    hr[i].h = row_hash(p, i);
    hr[i].row = i;
    keep[i] = 1;
  }
  qsort(hr, (size_t)m, sizeof(hrow_t), hrow_cmp);
  int64_t kept = m;
  for (int64_t g = 0; g < m;) {
    int64_t e = g;
    while (e < m && hr[e].h == hr[g].h) ++e;
    for (int64_t a = g + 1; a < e; ++a)
      for (int64_t b = g; b < a; ++b)
        if (keep[hr[b].row] && rows_equal(p, hr[a].row, hr[b].row)) {
          keep[hr[a].row] = 0;
          --kept;
          break;
        }
    g = e;
  }
  free(hr);
  return kept;
}

#define FEAS_TOL 1e-9 /* ToleranceConfig::industrial_defaults, tolerance.hpp:60-75 */
#define OPT_TOL 1e-9
#define SINGULAR_TOL 1e-10

/* interior_point.cpp:202-373 on the rows with keep[i] != 0 */
static void build_standard_form(const or_problem* p, const char* keep, stdform_t* f) {
  const int64_t n = p->n;
  memset(f, 0, sizeof(*f));
  double* cbase = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int64_t j = 0; j < n; ++j) cbase[j] = p->maximize ? -p->obj[j] : p->obj[j];
  f->vi = (varinfo_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(varinfo_t));
  f->nvar = n;
  int64_t nslots = 0, ntle = 0, m_ub = 0;
  for (int64_t j = 0; j < n; ++j) {
    const double lb = p->lb[j], ub = p->ub[j];
    varinfo_t* v = &f->vi[j];
    if (isfinite(lb) && isfinite(ub) && fabs(ub - lb) <= FEAS_TOL) {
      v->is_fixed = 1;
      v->fixed = lb;
      f->offset += cbase[j] * lb;
    } else if (isfinite(lb)) {
      v->ns = 1;
      v->slot[0] = nslots++;
      v->sign[0] = 1.0;
      v->shift[0] = lb;
      f->offset += cbase[j] * lb;
      if (isfinite(ub)) ++m_ub;
    } else if (isfinite(ub)) {
      v->ns = 1;
      v->slot[0] = nslots++;
      v->sign[0] = -1.0;
      v->shift[0] = ub;
      f->offset += cbase[j] * ub;
    } else {
      v->ns = 2;
      v->slot[0] = nslots++;
      v->sign[0] = 1.0;
      v->shift[0] = 0.0;
      v->slot[1] = nslots++;
      v->sign[1] = -1.0;
      v->shift[1] = 0.0;
    }
  }
  /* surviving rows, renumbered in order */
  int64_t norig = 0;
  int64_t* orig = (int64_t*)malloc(sizeof(int64_t) * (p->m > 0 ? p->m : 1));
  for (int64_t i = 0; i < p->m; ++i)
    if (keep[i]) orig[norig++] = i;
  for (int64_t r = 0; r < norig; ++r) {
    const int s = p->sense[orig[r]];
    if (s == OR_LE || s == OR_GE) ++ntle;
  }
  const int64_t m_rows = norig + m_ub;
  const int64_t ncols = nslots + ntle + m_ub;
  f->m = m_rows;
  f->n = ncols;
  f->rhs = (double*)calloc((size_t)(m_rows > 0 ? m_rows : 1), sizeof(double));
  f->c = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
  for (int64_t j = 0; j < n; ++j)
    for (int s = 0; s < f->vi[j].ns; ++s)
      f->c[f->vi[j].slot[s]] = f->vi[j].sign[s] > 0 ? cbase[j] : -cbase[j];

  int64_t cap = p->rp[p->m] * 2 + 2 * m_rows + 2 * n + 4;
  trip_t* t = (trip_t*)malloc(sizeof(trip_t) * cap);
  int64_t nt = 0;
  int64_t next_slack = nslots;
  for (int64_t r = 0; r < norig; ++r) {
    const int64_t i = orig[r];
    double rowmax = 0.0;
    for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
      const double a = fabs(p->v[k]);
      rowmax = (rowmax < a) ? a : rowmax;
    }
    const double gamma = (rowmax > 0.0) ? 1.0 / rowmax : 1.0;
    double const_i = 0.0;
    const int flip = (p->sense[i] == OR_GE);
    for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
      const int64_t j = p->ci[k];
      const double a = gamma * p->v[k];
      const varinfo_t* v = &f->vi[j];
      if (v->is_fixed) {
        const_i += a * v->fixed;
      } else {
        for (int s = 0; s < v->ns; ++s) {
          t[nt].row = r;
          t[nt].col = v->slot[s];
          t[nt].val = flip ? -a * v->sign[s] : a * v->sign[s];
          ++nt;
          const_i += a * v->shift[s];
        }
      }
    }
    if (p->sense[i] == OR_LE || p->sense[i] == OR_GE) {
      t[nt].row = r;
      t[nt].col = next_slack++;
      t[nt].val = gamma;
      ++nt;
    }
    f->rhs[r] = flip ? (-gamma * p->rhs[i] + const_i) : (gamma * p->rhs[i] - const_i);
  }
  int64_t next_ub_row = norig;
  for (int64_t j = 0; j < n; ++j) {
    const double lb = p->lb[j], ub = p->ub[j];
    if (isfinite(lb) && isfinite(ub) && fabs(ub - lb) > FEAS_TOL) {
      t[nt].row = next_ub_row;
      t[nt].col = f->vi[j].slot[0];
      t[nt].val = 1.0;
      ++nt;
      t[nt].row = next_ub_row;
      t[nt].col = next_slack++;
      t[nt].val = 1.0;
      ++nt;
      f->rhs[next_ub_row] = ub - lb;
      ++next_ub_row;
    }
  }
  csr_t M0;
  csr_from_triplets(m_rows, ncols, t, nt, &M0);
  /* column scaling, interior_point.cpp:345-370 */
  f->col_scale = (double*)malloc(sizeof(double) * (ncols > 0 ? ncols : 1));
  double* colmax = (double*)calloc((size_t)(ncols > 0 ? ncols : 1), sizeof(double));
  for (int64_t j = 0; j < ncols; ++j) f->col_scale[j] = 1.0;
  for (int64_t k = 0; k < M0.nnz; ++k) {
    const double a = fabs(M0.v[k]);
    colmax[M0.ci[k]] = (colmax[M0.ci[k]] < a) ? a : colmax[M0.ci[k]];
  }
  for (int64_t k = 0; k < M0.nnz; ++k)
    if (colmax[M0.ci[k]] > 0.0) f->col_scale[M0.ci[k]] = 1.0 / colmax[M0.ci[k]];
  nt = 0;
  for (int64_t i = 0; i < m_rows; ++i)
    for (int64_t k = M0.rp[i]; k < M0.rp[i + 1]; ++k) {
      t[nt].row = i;
      t[nt].col = M0.ci[k];
      t[nt].val = M0.v[k] * f->col_scale[M0.ci[k]];
      ++nt;
    }
  csr_from_triplets(m_rows, ncols, t, nt, &f->M);
  for (int64_t j = 0; j < ncols; ++j) f->c[j] *= f->col_scale[j];
  csr_free(&M0);
  free(colmax);
  free(t);
  free(orig);
  free(cbase);
}

/* interior_point.cpp:375-395 */
static void to_original_primal(const stdform_t* f, const double* y, double* primal) {
  for (int64_t j = 0; j < f->nvar; ++j) {
    const varinfo_t* v = &f->vi[j];
    if (v->is_fixed) {
      primal[j] = v->fixed;
    } else {
      double val = 0.0;
      for (int s = 0; s < v->ns; ++s) {
        const int64_t gi = v->slot[s];
        if (gi < 0 || gi >= f->n) continue;
        val += v->sign[s] * f->col_scale[gi] * y[gi];
        if (s == 0) val += v->shift[s];
      }
      primal[j] = val;
    }
  }
}

/* ProblemBuilder::build (problem.hpp:265-281): columns sorted within rows and
 * duplicates summed, exact-zero sums dropped. */
static void normalize(const or_problem* in, or_problem* out, int64_t** rp, int64_t** ci, double** v) {
  const int64_t m = in->m;
  const int64_t nz = in->rp[m];
  trip_t* t = (trip_t*)malloc(sizeof(trip_t) * (nz > 0 ? nz : 1));
  for (int64_t i = 0; i < m; ++i)
    for (int64_t k = in->rp[i]; k < in->rp[i + 1]; ++k) {
      t[k].row = i;
      t[k].col = in->ci[k];
      t[k].val = in->v[k];
    }
  csr_t A;
  csr_from_triplets(m, in->n, t, nz, &A);
  free(t);
  *out = *in;
  out->rp = A.rp;
  out->ci = A.ci;
  out->v = A.v;
  *rp = A.rp;
  *ci = A.ci;
  *v = A.v;
}

void or_default_options(or_options* o) {
  o->max_iterations = 100;
  o->time_limit_seconds = 0.0;
  o->convergence_tol = 1e-8;
  o->complementarity_tol = 1e-8;
  o->regularization = 1e-8;
  o->interrupt = 0;
  o->pivot_rule = 0;
  o->x0 = o->y0 = o->s0 = NULL;
}

int or_standard_form(const or_problem* p, int64_t* sm, int64_t* sn, int64_t* snnz, int64_t* rp,
                     int64_t* ci, double* v, double* rhs, double* c, double* col_scale,
                     double* offset) {
  if (!p || p->m < 0 || p->n < 0) return -1;
  int64_t *nrp, *nci;
  double* nv;
  or_problem q;
  normalize(p, &q, &nrp, &nci, &nv);
  char* keep = (char*)malloc(q.m > 0 ? q.m : 1);
  remove_duplicate_rows(&q, keep);
  stdform_t f;
  build_standard_form(&q, keep, &f);
  *sm = f.m;
  *sn = f.n;
  *snnz = f.M.nnz;
  if (rp) memcpy(rp, f.M.rp, sizeof(int64_t) * (f.m + 1));
  if (ci) memcpy(ci, f.M.ci, sizeof(int64_t) * f.M.nnz);
  if (v) memcpy(v, f.M.v, sizeof(double) * f.M.nnz);
  if (rhs) memcpy(rhs, f.rhs, sizeof(double) * f.m);
  if (c) memcpy(c, f.c, sizeof(double) * f.n);
  if (col_scale) memcpy(col_scale, f.col_scale, sizeof(double) * f.n);
  if (offset) *offset = f.offset;
  sf_free(&f);
  free(keep);
  free(nrp);
  free(nci);
  free(nv);
  return 0;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int all_finite(int64_t n, const double* a) {
  for (int64_t i = 0; i < n; ++i)
    if (!isfinite(a[i])) return 0;
  return 1;
}

int or_ipm_solve(const or_problem* p_in, const or_options* o, or_result* r) {
  if (!p_in || !o || !r || p_in->m < 0 || p_in->n < 0) return -1;
  const double t0 = now_s();
  r->status = OR_UNKNOWN;
  r->iterations = 0;
  r->objective = 0.0;
  r->sf_m = r->sf_n = 0;
  r->rp_norm = r->rd_norm = r->mu = 0.0;
  /* interior_point.cpp:406-410 */
  if (p_in->n == 0) {
    r->status = OR_OPTIMAL;
    r->objective = p_in->obj_offset;
    return 0;
  }
  int64_t *nrp, *nci;
  double* nv;
  or_problem pp;
  normalize(p_in, &pp, &nrp, &nci, &nv);
  const or_problem* p = &pp;
  char* keep = (char*)malloc(p->m > 0 ? p->m : 1);
  remove_duplicate_rows(p, keep);
  stdform_t f;
  build_standard_form(p, keep, &f);
  const int64_t m = f.m, n = f.n;
  r->sf_m = m;
  r->sf_n = n;
  double* primal = (double*)calloc((size_t)(p->n > 0 ? p->n : 1), sizeof(double));

  if (n == 0) { /* :416-421 */
    r->status = OR_OPTIMAL;
    r->objective = f.offset;
    to_original_primal(&f, NULL, primal);
    goto finish_primal;
  }
  if (m == 0) { /* :423-441 */
    int unbounded = 0;
    for (int64_t j = 0; j < n; ++j)
      if (f.c[j] < -OPT_TOL) {
        unbounded = 1;
        break;
      }
    r->status = unbounded ? OR_UNBOUNDED : OR_OPTIMAL;
    double* y0 = (double*)calloc((size_t)n, sizeof(double));
    to_original_primal(&f, y0, primal);
    free(y0);
    double obj = p->obj_offset;
    for (int64_t j = 0; j < p->n; ++j) obj += p->obj[j] * primal[j];
    r->objective = obj;
    goto finish_primal;
  }
  {
    /* start point :443-453 */
    double b_norm = 1.0, c_norm = 1.0;
    for (int64_t i = 0; i < m; ++i) b_norm = (b_norm < fabs(f.rhs[i])) ? fabs(f.rhs[i]) : b_norm;
    for (int64_t j = 0; j < n; ++j) c_norm = (c_norm < fabs(f.c[j])) ? fabs(f.c[j]) : c_norm;
    double* lam = (double*)calloc((size_t)m, sizeof(double));
    double* x = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    for (int64_t j = 0; j < n; ++j) {
      x[j] = b_norm + 1.0;
      const double a = c_norm + 1.0, b = fabs(f.c[j]);
      z[j] = (a < b) ? b : a;
    }
    if (o->x0 && o->y0 && o->s0) { /* caller's warm start of the scaled standard form */
      memcpy(x, o->x0, sizeof(double) * n);
      memcpy(z, o->s0, sizeof(double) * n);
      memcpy(lam, o->y0, sizeof(double) * m);
    }
    double* rp = (double*)calloc((size_t)m, sizeof(double));
    double* rd = (double*)calloc((size_t)n, sizeof(double));
    double* D = (double*)malloc(sizeof(double) * n);
    double* tmp_m = (double*)malloc(sizeof(double) * m);
    double* tmp_n = (double*)malloc(sizeof(double) * n);
    double* dx = (double*)malloc(sizeof(double) * n);
    double* dz = (double*)malloc(sizeof(double) * n);
    double* dlam = (double*)malloc(sizeof(double) * m);
    double* rhs_norm = (double*)malloc(sizeof(double) * m);
    double* mdrd = (double*)malloc(sizeof(double) * m);
    double* mx = (double*)malloc(sizeof(double) * m);
    double* mz = (double*)malloc(sizeof(double) * m);
    double* zs = (double*)malloc(sizeof(double) * n);
    double* N = (double*)malloc(sizeof(double) * (size_t)m * (size_t)m);
    double* F = (double*)malloc(sizeof(double) * (size_t)m * (size_t)m);
    int done = 0, time_up = 0, interrupted = 0, numerical_error = 0;
    const int64_t max_it = o->max_iterations;
    int64_t iter;
#define DEADLINE() (o->time_limit_seconds > 0.0 && now_s() - t0 >= o->time_limit_seconds)
    for (iter = 0; iter < max_it; ++iter) {
      if (o->interrupt) { interrupted = 1; break; }                 /* :490-493 */
      if (DEADLINE()) { time_up = 1; break; }                       /* :494-501 */
      or_spmv(m, f.M.rp, f.M.ci, f.M.v, x, tmp_m);                  /* :507-511 */
      for (int64_t i = 0; i < m; ++i) rp[i] = tmp_m[i] - f.rhs[i];
      or_spmv_t(m, n, f.M.rp, f.M.ci, f.M.v, lam, tmp_n);           /* :514-515 */
      for (int64_t j = 0; j < n; ++j) rd[j] = tmp_n[j] + z[j] - f.c[j];
      double mu = 0.0;                                              /* :517-523 */
      for (int64_t j = 0; j < n; ++j) mu += x[j] * z[j];
      mu /= (double)n;
      if (or_norm_inf(m, rp) < o->convergence_tol && or_norm_inf(n, rd) < o->convergence_tol &&
          mu < o->complementarity_tol) {                            /* :525-530 */
        done = 1;
        break;
      }
      for (int64_t j = 0; j < n; ++j) D[j] = x[j] / z[j];           /* :532-539 */
      for (int64_t j = 0; j < n; ++j)
        if (!(D[j] >= 0.0) || !isfinite(D[j]) || D[j] > 1e10) D[j] = 1e10;
      adat_dense_rel(m, n, f.M.rp, f.M.ci, f.M.v, D, o->regularization,       /* :541 */
                     o->pivot_rule ? 0.0 : 1e-12, N);
      if (DEADLINE()) { time_up = 1; break; }
      ldlt_factor(m, N, F, SINGULAR_TOL, o->pivot_rule ? 1e-13 : 0.0); /* :558-576 */
      if (DEADLINE()) { time_up = 1; break; }
      for (int64_t j = 0; j < n; ++j) tmp_n[j] = D[j] * rd[j];      /* :584 */
      or_spmv(m, f.M.rp, f.M.ci, f.M.v, tmp_n, mdrd);               /* :585-590 */
      or_spmv(m, f.M.rp, f.M.ci, f.M.v, x, mx);                     /* :591-596 */
      for (int64_t i = 0; i < m; ++i) rhs_norm[i] = -rp[i] - mdrd[i] + mx[i];
      memcpy(dlam, rhs_norm, sizeof(double) * m);                   /* :602-604 */
      ldlt_solve(m, F, dlam);
      iterative_refinement(m, F, N, dlam, rhs_norm, 1e-10, 5);
      or_spmv_t(m, n, f.M.rp, f.M.ci, f.M.v, dlam, tmp_n);          /* :607-615 */
      for (int64_t j = 0; j < n; ++j) {
        dx[j] = D[j] * (tmp_n[j] + rd[j]) - x[j];
        dz[j] = -(tmp_n[j] + rd[j]);
      }
      if (!all_finite(n, dx) || !all_finite(n, dz) || !all_finite(m, dlam)) {
        numerical_error = 1;                                        /* :617-627 */
        break;
      }
      const double ap_a = or_max_step(n, x, dx);                    /* :629-630 */
      const double ad_a = or_max_step(n, z, dz);
      double mu_aff = 0.0;                                          /* :632-636 */
      for (int64_t j = 0; j < n; ++j) mu_aff += (x[j] + ap_a * dx[j]) * (z[j] + ad_a * dz[j]);
      mu_aff /= (double)n;
      double sigma;                                                 /* :638-643 */
      if (!isfinite(mu_aff) || mu_aff < 0.0 || !(mu > 0.0)) {
        sigma = 0.5;
      } else {
        const double den = (mu < 1e-300) ? 1e-300 : mu;
        sigma = pow(mu_aff / den, 3.0);
        sigma = sigma < 0.1 ? 0.1 : (0.9 < sigma ? 0.9 : sigma);
      }
      for (int64_t j = 0; j < n; ++j) {                             /* :645-650 */
        const double zinv = 1.0 / z[j];
        zs[j] = (sigma * mu - dx[j] * dz[j]) * zinv;
      }
      or_spmv(m, f.M.rp, f.M.ci, f.M.v, zs, mz);                    /* :651-656 */
      for (int64_t i = 0; i < m; ++i) rhs_norm[i] = -rp[i] - mdrd[i] + mx[i] - mz[i];
      memcpy(dlam, rhs_norm, sizeof(double) * m);                   /* :661-663 */
      ldlt_solve(m, F, dlam);
      iterative_refinement(m, F, N, dlam, rhs_norm, 1e-10, 5);
      or_spmv_t(m, n, f.M.rp, f.M.ci, f.M.v, dlam, tmp_n);          /* :666-673 */
      for (int64_t j = 0; j < n; ++j) {
        dx[j] = D[j] * (tmp_n[j] + rd[j]) - x[j] + zs[j];
        dz[j] = -(tmp_n[j] + rd[j]);
      }
      if (!all_finite(n, dx) || !all_finite(n, dz) || !all_finite(m, dlam)) {
        numerical_error = 1;                                        /* :675-685 */
        break;
      }
      const double ap = or_max_step(n, x, dx);                      /* :687-688 */
      const double ad = or_max_step(n, z, dz);
      for (int64_t j = 0; j < n; ++j) {                             /* :690-696 */
        x[j] += ap * dx[j];
        z[j] += ad * dz[j];
      }
      for (int64_t i = 0; i < m; ++i) lam[i] += ad * dlam[i];
      int bad = 0;                                                  /* :698-708 */
      for (int64_t j = 0; j < n && !bad; ++j)
        if (x[j] < 0.0 || z[j] < 0.0 || !isfinite(x[j]) || !isfinite(z[j])) bad = 1;
      if (bad) {
        numerical_error = 1;
        break;
      }
      if (r->trace && iter < r->trace_cap) {                        /* :710-714 */
        r->trace[5 * iter + 0] = mu;
        r->trace[5 * iter + 1] = or_norm_inf(m, rp);
        r->trace[5 * iter + 2] = or_norm_inf(n, rd);
        r->trace[5 * iter + 3] = ap;
        r->trace[5 * iter + 4] = ad;
      }
      r->iterations = iter + 1;                                     /* :725 */
    }
#undef DEADLINE
    double mu_last = 0.0;
    if (!done && !interrupted) {                                    /* :728-753 */
      or_spmv(m, f.M.rp, f.M.ci, f.M.v, x, tmp_m);
      for (int64_t i = 0; i < m; ++i) rp[i] = tmp_m[i] - f.rhs[i];
      or_spmv_t(m, n, f.M.rp, f.M.ci, f.M.v, lam, tmp_n);
      for (int64_t j = 0; j < n; ++j) rd[j] = tmp_n[j] + z[j] - f.c[j];
      double mu = 0.0;
      for (int64_t j = 0; j < n; ++j) mu += x[j] * z[j];
      mu /= (double)n;
      mu_last = mu;
      if (or_norm_inf(m, rp) < 1e-4 && or_norm_inf(n, rd) < 1e-4 && mu < 1e-4) done = 1;
    } else {
      for (int64_t j = 0; j < n; ++j) mu_last += x[j] * z[j];
      mu_last /= (double)n;
    }
    to_original_primal(&f, x, primal);                              /* :757 */
    double obj = 0.0;                                               /* :759-769 */
    int obj_ok = isfinite(p->obj_offset);
    for (int64_t j = 0; j < p->n && obj_ok; ++j) {
      const double term = p->obj[j] * primal[j];
      if (!isfinite(term)) obj_ok = 0;
      obj += term;
    }
    if (obj_ok) obj += p->obj_offset;
    if (!obj_ok) numerical_error = 1;
    r->objective = obj_ok ? obj : 0.0;
    const double rpn = or_norm_inf(m, rp);
    const int feasible_incumbent = rpn <= FEAS_TOL;                 /* :773-780 */
    r->status = numerical_error ? OR_NUMERICAL_ERROR
                : interrupted   ? OR_INTERRUPTED
                : done          ? OR_OPTIMAL
                : feasible_incumbent ? OR_SUBOPTIMAL
                : time_up       ? OR_TIME_LIMIT
                                : OR_ITER_LIMIT;
    r->rp_norm = rpn;
    r->rd_norm = or_norm_inf(n, rd);
    r->mu = mu_last;
    if (r->x && r->cap_n >= n) memcpy(r->x, x, sizeof(double) * n);
    if (r->z && r->cap_n >= n) memcpy(r->z, z, sizeof(double) * n);
    if (r->lam && r->cap_m >= m) memcpy(r->lam, lam, sizeof(double) * m);
    free(lam); free(x); free(z); free(rp); free(rd); free(D); free(tmp_m); free(tmp_n);
    free(dx); free(dz); free(dlam); free(rhs_norm); free(mdrd); free(mx); free(mz); free(zs);
    free(N); free(F);
  }
finish_primal:
  if (r->primal) memcpy(r->primal, primal, sizeof(double) * p->n);
  free(primal);
  sf_free(&f);
  free(keep);
  free(nrp);
  free(nci);
  free(nv);
  return 0;
}

/* ---- post-solve verification ---------------------------------------------
 * SolutionVerifier (core/validation/solution_verifier.cpp).  std::max(a, b)
 * is (a < b) ? b : a and std::min(a, b) is (b < a) ? b : a, which is how a
 * NaN stays out of every maximum below. */
static double vmax(double a, double b) { return (a < b) ? b : a; }

static void v_push(or_verify_out* out, int64_t idx) {
  if (out->violated && out->n_violated < out->violated_cap) out->violated[out->n_violated] = idx;
  ++out->n_violated;
}

void or_verify(const or_problem* p, const int8_t* var_type, const double* primal, int64_t n_primal,
               const double* dual, int64_t n_dual, const double* rc, int64_t n_rc,
               double objective_value, const double tol[4], or_verify_out* out) {
  const int64_t m = p->m, n = p->n;
  const double feas = tol[0], otol = tol[1], ctol = tol[2], itol = tol[3];
  char pmsg[128] = "";
  int primal_ok = 1;
  double max_violation = 0.0;
  int64_t n_int = 0;
  out->feasible = 1;
  out->optimal = 1;
  out->primal_infeasibility = out->dual_infeasibility = out->complementarity = 0.0;
  out->integrality_violation = out->objective_discrepancy = 0.0;
  out->n_violated = 0;
  out->message[0] = 0;
  if (var_type)
    for (int64_t j = 0; j < n; ++j) n_int += (var_type[j] == 1 || var_type[j] == 2);

  /* check_primal_feasibility (:92-191) */
  if (n_primal != n) {
    out->feasible = 0;
    out->optimal = 0;
    snprintf(out->message, sizeof(out->message), "Primal infeasible: Solution vector size mismatch");
    return;
  }
  {
    double* col_scale = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    const double ftol = feas * 10.0;
    int64_t nonfinite = 0;
    for (int64_t j = 0; j < n; ++j) col_scale[j] = 1.0;
    for (int64_t k = 0; k < p->rp[m]; ++k) {  /* :108-114 */
      const int64_t j = p->ci[k];
      if (j < n) col_scale[j] = vmax(col_scale[j], fabs(p->v[k]));
    }
    for (int64_t j = 0; j < n; ++j) {  /* :115-121 */
      col_scale[j] = vmax(col_scale[j], fabs(p->obj[j]));
      if (isfinite(p->lb[j])) col_scale[j] = vmax(col_scale[j], fabs(p->lb[j]));
      if (isfinite(p->ub[j])) col_scale[j] = vmax(col_scale[j], fabs(p->ub[j]));
    }
    for (int64_t j = 0; j < n; ++j) {  /* :125-146 */
      const double val = primal[j];
      if (!isfinite(val)) {  /* early return: no maximum recorded (:119-123) */
        nonfinite = 1;
        snprintf(pmsg, sizeof(pmsg), "Solution contains NaN or Inf values");
        break;
      }
      if (val < p->lb[j] - ftol * col_scale[j]) {
        max_violation = vmax(max_violation, p->lb[j] - val);
        v_push(out, j);
      }
      if (val > p->ub[j] + ftol * col_scale[j]) {
        max_violation = vmax(max_violation, val - p->ub[j]);
        v_push(out, j);
      }
    }
    free(col_scale);
    if (nonfinite) {
      primal_ok = 0;
    } else {
      for (int64_t i = 0; i < m; ++i) {  /* :148-184 */
        double activity = 0.0, activity_abs = 0.0;
        for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
          const double term = p->v[k] * primal[p->ci[k]];
          activity += term;
          activity_abs += fabs(term);
        }
        const double scale = 1.0 + fabs(p->rhs[i]) + activity_abs;
        double violation = 0.0, excess;
        if (p->sense[i] == OR_LE) {
          excess = activity - p->rhs[i] - ftol * scale;
          if (excess > 0.0) violation = activity - p->rhs[i];
        } else if (p->sense[i] == OR_GE) {
          excess = p->rhs[i] - activity - ftol * scale;
          if (excess > 0.0) violation = p->rhs[i] - activity;
        } else {
          excess = fabs(activity - p->rhs[i]) - ftol * scale;
          if (excess > 0.0) violation = fabs(activity - p->rhs[i]);
        }
        if (excess > 0.0) {
          max_violation = vmax(max_violation, violation);
          v_push(out, n + i);
        }
      }
      out->primal_infeasibility = max_violation;
      if (out->n_violated > 0) {
        primal_ok = 0;
        snprintf(pmsg, sizeof(pmsg), "Max primal violation: %f", max_violation);
      }
    }
  }
  if (!primal_ok) {  /* verify_detailed (:32-37) */
    out->feasible = 0;
    snprintf(out->message, sizeof(out->message), "Primal infeasible: %s", pmsg);
  }

  /* check_dual_feasibility (:193-296): reduced cost c_j - sum_k a_k y_row(k)
   * over the nonzeros of column j in storage order, accumulated here for all
   * columns in one pass over the rows (each column's sum sees the same terms
   * in the same order as the reference's per-column scan). */
  double dual_inf = 0.0;
  if (!(n_dual != m || (n_int > 0 && n_rc != n))) {
    const int use_rc = n_rc == n;
    double* acc = NULL;
    if (!use_rc) {
      acc = (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
      for (int64_t i = 0; i < m; ++i)
        for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) acc[p->ci[k]] += p->v[k] * dual[i];
    }
    for (int64_t j = 0; j < n; ++j) {
      double r = use_rc ? rc[j] : p->obj[j] - acc[j];
      if (p->maximize) r = -r;
      const double val = primal[j];
      double violation;
      if (val <= p->lb[j] + feas)
        violation = (r < 0.0) ? r : 0.0;
      else if (val >= p->ub[j] - feas)
        violation = (0.0 < r) ? r : 0.0;
      else
        violation = fabs(r);
      violation = fabs(violation) - otol;
      if (violation > 0.0) dual_inf = vmax(dual_inf, violation);
    }
    free(acc);
  }
  out->dual_infeasibility = dual_inf;

  /* check_complementarity (:298-347) */
  double comp = 0.0;
  if (n_dual == m) {
    for (int64_t i = 0; i < m; ++i) {
      double activity = 0.0, activity_abs = 0.0;
      for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) {
        const double term = p->v[k] * primal[p->ci[k]];
        activity += term;
        activity_abs += fabs(term);
      }
      double slack = 0.0;
      if (p->sense[i] == OR_LE)
        slack = p->rhs[i] - activity;
      else if (p->sense[i] == OR_GE)
        slack = activity - p->rhs[i];
      double scale = 1.0;  /* std::max({1.0, |rhs|, activity_abs}) */
      if (scale < fabs(p->rhs[i])) scale = fabs(p->rhs[i]);
      if (scale < activity_abs) scale = activity_abs;
      const double dual_scale = vmax(1.0, fabs(dual[i]));
      comp = vmax(comp, fabs(slack) / scale * (fabs(dual[i]) / dual_scale));
    }
  }
  out->complementarity = comp;
  if (n_int == 0) {  /* :48-60 */
    if (dual_inf > otol) {
      out->optimal = 0;
      if (out->feasible)
        snprintf(out->message, sizeof(out->message), "Dual infeasible: Max dual violation: %f", dual_inf);
    }
    if (comp > ctol) {
      out->optimal = 0;
      if (out->feasible)
        snprintf(out->message, sizeof(out->message),
                 "Complementarity violation: Max complementarity violation: %f", comp);
    }
  }

  /* check_integrality (:349-371) */
  double integ = 0.0;
  if (var_type)
    for (int64_t j = 0; j < n; ++j)
      if (var_type[j] == 1 || var_type[j] == 2) integ = vmax(integ, fabs(primal[j] - round(primal[j])));
  out->integrality_violation = integ;
  if (integ > itol) {
    out->feasible = 0;
    snprintf(out->message, sizeof(out->message), "Integrality violation: Max integrality violation: %f", integ);
  }

  /* objective (:70-87) */
  double obj = p->obj_offset;
  for (int64_t j = 0; j < n; ++j) obj += p->obj[j] * primal[j];
  out->objective_discrepancy = fabs(obj - objective_value);
  if (out->objective_discrepancy > 1e-4 * (1.0 + fabs(obj))) {
    out->optimal = 0;
    if (out->feasible)
      snprintf(out->message, sizeof(out->message), "Objective discrepancy: %f", out->objective_discrepancy);
  }
}

/* ---- all-core SpMV timing (a baseline of this repository, not the
 * reference's kernel): rows cut into `threads` ranges of equal nonzero
 * count, each summed left to right like or_spmv. ------------------------- */
typedef struct {
  const int64_t *rp, *ci;
  const double *v, *x;
  double* y;
  int64_t r0, r1;
  int repeats;
} or_spmv_part;

static void* or_spmv_worker(void* arg) {
  const or_spmv_part* p = (const or_spmv_part*)arg;
  for (int r = 0; r < p->repeats; ++r)
    for (int64_t i = p->r0; i < p->r1; ++i) {
      double s = 0.0;
      for (int64_t k = p->rp[i]; k < p->rp[i + 1]; ++k) s += p->v[k] * p->x[p->ci[k]];
      p->y[i] = s;
    }
  return NULL;
}

double or_spmv_mt_seconds(int64_t nrows, const int64_t* rp, const int64_t* ci, const double* v, const double* x,
                          double* y, int repeats, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  or_spmv_part parts[256];
  pthread_t tid[256];
  const int64_t nnz = rp[nrows];
  int64_t row = 0;
  for (int t = 0; t < threads; ++t) {
    const int64_t target = (t + 1 == threads) ? nnz : nnz / threads * (t + 1);
    int64_t e = row;
    while (e < nrows && rp[e] < target) ++e;
    if (t + 1 == threads) e = nrows;
    parts[t] = (or_spmv_part){rp, ci, v, x, y, row, e, repeats};
    row = e;
  }
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, or_spmv_worker, &parts[t]);
  for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}
