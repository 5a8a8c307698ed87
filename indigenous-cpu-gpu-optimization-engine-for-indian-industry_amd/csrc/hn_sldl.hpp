// Sparse LDL' of the normal equations: host symbolic analysis (exact minimum
// degree) and the device factor and solve built on it.
#pragma once

#include <cstdint>
#include <vector>

#include "hn_internal.hpp"

namespace hn {

// Result of the symbolic analysis of a symmetric pattern (n x n).  All
// positions are in the permuted system B = P A P' (row r of B is original
// row perm[r]).
struct SldlSymbolic {
  int64_t n = 0;
  std::vector<int64_t> perm, inv_perm;  // perm[r]: original vertex eliminated at step r
  std::vector<int64_t> Lrp;             // n + 1: row pointer of L (diagonal last in each row)
  std::vector<int32_t> Lci;             // columns of L, ascending within a row
  std::vector<int64_t> parent;          // elimination tree (n: a root)
  std::vector<int64_t> postorder;
  // rows grouped by level: a row's level is one more than the highest level
  // of the rows its off-diagonal entries name (0 when it has none), so the
  // rows of one level depend only on rows of lower levels
  std::vector<int32_t> level_rows;
  std::vector<int64_t> level_ptr;
  // strict lower part of L by column: for column j, the rows i > j with an
  // entry (i, j) in descending order, and that entry's position in L
  std::vector<int64_t> Ucp;
  std::vector<int32_t> Uri;
  std::vector<int64_t> Upos;
  double work = 0.0;     // adjacency elements merged during the ordering
  // sequential steps on the factor's critical path: per level, the most merge
  // steps one row takes (sum over its entries j of the lengths of row j and
  // of its own prefix), summed over the levels
  double crit_steps = 0.0;
  bool aborted = false;  // the work or size limit was reached; nothing else is valid
};

// Exact minimum-degree ordering of the pattern (rp, ci) and the row patterns
// of its LDL' factor (compute_md_order_and_pattern and
// build_elimination_graph, core/numerical/factorization.cpp:71-191), then
// the elimination tree, the postorder (:1327-1362), the levels and the
// column lists.  Stops with S.aborted once the merge work exceeds work_limit
// or the factor exceeds max_lnnz entries (<= 0: no limit).
int sldl_symbolic(int64_t n, const int64_t* rp, const int64_t* ci, double work_limit, double max_lnnz,
                  SldlSymbolic& S);

}  // namespace hn

struct hn_sldl {
  hn::SldlSymbolic S;
  bool on_device = false;
  hn::DevBuf Lrp, Lci, Ucp, Uri, Upos, rows, perm;  // symbolic arrays
  hn::DevBuf bidx;                                   // int64 per entry of L: its value's index in A (-1: zero)
  hn::DevBuf L, D;                                   // the factor
  hn::DevBuf xin, xf, xd;                            // solve workspace
  hn::DevBuf sing;                                   // int: pivots clamped by the last factorisation
  const hn_csr* bound = nullptr;                     // matrix whose pattern bidx was built for
  int64_t bound_nnz = -1;
};

namespace hn {

// Handle = symbolic analysis + device factor (hn_sparse_ldl.hip).  create
// runs the analysis on the host (no device call); check h->S.aborted.
int sldl_create(int64_t n, const int64_t* rp, const int64_t* ci, double work_limit, double max_lnnz, hn_sldl** out);
// Numeric factorisation of a (device CSR, FP64 values, the analysed pattern
// with ascending columns per row), enqueued on s.  The first call with a
// matrix uploads the analysis and maps its entries (synchronises s).
int sldl_factor(hn_sldl* h, const hn_csr* a, double singular_tol, hipStream_t s);
// x = A^{-1} b (accumulate: x += A^{-1} b) by the factor, enqueued on s;
// b and x must not alias.
int sldl_solve(const hn_sldl* h, const double* b, double* x, bool accumulate, hipStream_t s);
void sldl_destroy(hn_sldl* h);

}  // namespace hn
