// ref_driver.cpp - C entry points around the reference's own IPM, CPU backend
// and sparse containers, compiled together with the reference sources into
// oracle/_ref/libhnref.so (see oracle/Makefile).  TEST INFRASTRUCTURE: used
// only to pin the restated oracle and as the reference CPU baseline.
//
// Only the reference's public API is used: ProblemBuilder
// (core/model/problem.hpp:199-292), InteriorPointSolver
// (core/lp/interior_point.hpp:16-54), ComputeBackendFactory
// (core/execution/gpu_backend.hpp:53-58) and SparseMatrix
// (core/numerical/sparse_matrix.hpp:30-89), SolutionVerifier
// (core/validation/solution_verifier.hpp:23-41) and SparseCholesky
// (core/numerical/factorization.hpp:162-205).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <iostream>
#include <map>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "execution/gpu_backend.hpp"
#include "lp/interior_point.hpp"
#include "lp/simplex.hpp"
#include "model/lp_parser.hpp"
#include "model/problem.hpp"
#include "numerical/factorization.hpp"
#include "numerical/sparse_matrix.hpp"
#include "numerical/tolerance.hpp"
#include "presolve/presolve.hpp"
#include "validation/solution_verifier.hpp"

using namespace hypernova;

extern "C" {

// Solves with InteriorPointSolver (crossover off).  primal (n) may be NULL.
// trace receives up to trace_cap rows of (mu, rp, rd, ap, ad) parsed from the
// solver's own HYPERNOVA_IPM_DBG lines; *trace_len is the number of rows.
int ref_ipm_solve(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                  const int8_t* sense, const double* rhs, const double* lb, const double* ub,
                  const double* obj, int maximize, double obj_offset, int max_iterations,
                  double time_limit, int interrupt, int* status, int64_t* iterations,
                  double* objective, double* primal, double* trace, int64_t trace_cap,
                  int64_t* trace_len, double* solve_ms) {
  model::ProblemBuilder b("ref");
  for (int64_t j = 0; j < n; ++j) b.add_variable(lb[j], ub[j], model::VarType::CONTINUOUS);
  for (int64_t i = 0; i < m; ++i) {
    std::vector<std::pair<std::size_t, double>> coeffs;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) coeffs.push_back({(std::size_t)ci[k], v[k]});
    b.add_constraint(coeffs, static_cast<model::ConstraintSense>(sense[i]), rhs[i]);
  }
  std::vector<std::pair<std::size_t, double>> o;
  for (int64_t j = 0; j < n; ++j) o.push_back({(std::size_t)j, obj[j]});
  b.set_objective(o, maximize ? model::ObjectiveSense::MAXIMIZE : model::ObjectiveSense::MINIMIZE);
  model::Problem prob = b.build();
  prob.obj_offset = obj_offset;

  lp::InteriorPointOptions opts;
  opts.max_iterations = max_iterations;
  opts.time_limit_seconds = time_limit;
  opts.crossover = false;
  if (interrupt) opts.interrupt_callback = []() { return true; };
  lp::InteriorPointSolver solver(numerical::ToleranceConfig::industrial_defaults(), opts);

  std::ostringstream captured;
  std::streambuf* old = nullptr;
  if (trace) {
    setenv("HYPERNOVA_IPM_DBG", "1", 1);
    old = std::cerr.rdbuf(captured.rdbuf());
  }
  lp::InteriorPointResult res = solver.solve(prob);
  if (trace) {
    std::cerr.rdbuf(old);
    unsetenv("HYPERNOVA_IPM_DBG");
    std::istringstream in(captured.str());
    std::string line;
    int64_t rows = 0;
    while (std::getline(in, line) && rows < trace_cap) {
      long it;
      double mu, r1, r2, ap, ad;
      if (std::sscanf(line.c_str(), "IPM it=%ld mu=%lf rp=%lf rd=%lf ap=%lf ad=%lf", &it, &mu, &r1, &r2,
                      &ap, &ad) == 6) {
        double* t = trace + 5 * rows;
        t[0] = mu;
        t[1] = r1;
        t[2] = r2;
        t[3] = ap;
        t[4] = ad;
        ++rows;
      }
    }
    if (trace_len) *trace_len = rows;
  }
  *status = static_cast<int>(res.status);
  *iterations = (int64_t)res.iterations;
  *objective = res.objective_value;
  if (solve_ms) *solve_ms = res.solve_time_ms;
  if (primal)
    for (int64_t j = 0; j < n && j < (int64_t)res.primal.size(); ++j) primal[j] = res.primal[j];
  return 0;
}

// y = A x through ComputeBackendFactory::create_backend(CPU)->spmv.
int ref_cpu_spmv(int64_t nrows, int64_t nnz, const uint64_t* rp, const uint64_t* ci, const double* v,
                 int64_t ncols, const double* x, double* y, int repeats) {
  auto be = execution::ComputeBackendFactory::create_backend(execution::ComputeBackendType::CPU);
  std::vector<double> vals(v, v + nnz);
  std::vector<std::size_t> cols(ci, ci + nnz), rows(rp, rp + nrows + 1);
  std::vector<double> xs(x, x + ncols), ys;
  for (int r = 0; r < (repeats > 0 ? repeats : 1); ++r) be->spmv(vals, cols, rows, xs, ys);
  std::memcpy(y, ys.data(), sizeof(double) * ys.size());
  return 0;
}

// Same, timing only the spmv calls (the std::vector copies are made once).
double ref_cpu_spmv_time(int64_t nrows, int64_t nnz, const uint64_t* rp, const uint64_t* ci,
                         const double* v, int64_t ncols, const double* x, int repeats) {
  auto be = execution::ComputeBackendFactory::create_backend(execution::ComputeBackendType::CPU);
  std::vector<double> vals(v, v + nnz);
  std::vector<std::size_t> cols(ci, ci + nnz), rows(rp, rp + nrows + 1);
  std::vector<double> xs(x, x + ncols), ys;
  be->spmv(vals, cols, rows, xs, ys);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < repeats; ++r) be->spmv(vals, cols, rows, xs, ys);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

// Wall seconds of `repeats` SparseMatrix::transpose_multiply calls
// (sparse_matrix.cpp:161-184, the row-wise scatter the IPM's matvec_t also
// uses), on a SparseMatrix filled directly from the CSR arrays.
double ref_cpu_spmv_t_time(int64_t nrows, int64_t ncols, int64_t nnz, const uint64_t* rp, const uint64_t* ci,
                           const double* v, const double* w, int repeats) {
  numerical::SparseMatrix A(nrows, ncols);
  A.mutable_values().assign(v, v + nnz);
  A.mutable_col_indices().assign(ci, ci + nnz);
  A.mutable_row_ptr().assign(rp, rp + nrows + 1);
  std::vector<double> ws(w, w + nrows);
  std::vector<double> out = A.transpose_multiply(ws);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int r = 0; r < repeats; ++r) out = A.transpose_multiply(ws);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

// SparseMatrix::from_triplets(rows, cols, triplets).transposed(); returns nnz,
// and writes the CSR of the transpose when the output pointers are not NULL.
int64_t ref_transpose(int64_t nrows, int64_t ncols, const int64_t* rp, const int64_t* ci,
                      const double* v, int64_t* out_rp, int64_t* out_ci, double* out_v) {
  std::vector<numerical::Triplet> t;
  for (int64_t i = 0; i < nrows; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) t.emplace_back((std::size_t)i, (std::size_t)ci[k], v[k]);
  numerical::SparseMatrix A = numerical::SparseMatrix::from_triplets(nrows, ncols, t);
  numerical::SparseMatrix At = A.transposed();
  if (out_rp)
    for (int64_t i = 0; i <= ncols; ++i) out_rp[i] = (int64_t)At.row_ptr()[i];
  if (out_ci)
    for (std::size_t k = 0; k < At.nnz(); ++k) out_ci[k] = (int64_t)At.col_indices()[k];
  if (out_v)
    for (std::size_t k = 0; k < At.nnz(); ++k) out_v[k] = At.values()[k];
  return (int64_t)At.nnz();
}

namespace {

model::Problem build_problem(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                             const int8_t* sense, const double* rhs, const double* lb, const double* ub,
                             const double* obj, int maximize, double obj_offset) {
  model::ProblemBuilder b("ref");
  for (int64_t j = 0; j < n; ++j) b.add_variable(lb[j], ub[j], model::VarType::CONTINUOUS);
  for (int64_t i = 0; i < m; ++i) {
    std::vector<std::pair<std::size_t, double>> coeffs;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) coeffs.push_back({(std::size_t)ci[k], v[k]});
    b.add_constraint(coeffs, static_cast<model::ConstraintSense>(sense[i]), rhs[i]);
  }
  std::vector<std::pair<std::size_t, double>> o;
  for (int64_t j = 0; j < n; ++j) o.push_back({(std::size_t)j, obj[j]});
  b.set_objective(o, maximize ? model::ObjectiveSense::MAXIMIZE : model::ObjectiveSense::MINIMIZE);
  model::Problem prob = b.build();
  prob.obj_offset = obj_offset;
  return prob;
}

// True when two rows of the built problem carry the same entries, sense and
// rhs (the rows remove_duplicate_rows, interior_point.cpp:82-132, would drop).
bool has_duplicate_rows(const model::Problem& p) {
  std::map<std::string, int> seen;
  const auto& A = p.constraint_matrix;
  for (std::size_t i = 0; i < p.constraints.size(); ++i) {
    std::ostringstream key;
    key.precision(17);
    key << static_cast<int>(p.constraints[i].sense) << ' ' << p.constraints[i].rhs;
    for (std::size_t k = A.row_ptr()[i]; k < A.row_ptr()[i + 1]; ++k)
      key << ' ' << A.col_indices()[k] << ':' << A.values()[k];
    if (!seen.emplace(key.str(), 1).second) return true;
  }
  return false;
}

// Silences the solvers' progress prints on std::cerr for one scope.
struct QuietCerr {
  std::ostringstream sink;
  std::streambuf* old;
  QuietCerr() : old(std::cerr.rdbuf(sink.rdbuf())) {}
  ~QuietCerr() { std::cerr.rdbuf(old); }
};

void copy_out(const std::vector<double>& src, double* dst, int64_t cap) {
  if (!dst) return;
  for (int64_t q = 0; q < cap; ++q) dst[q] = q < (int64_t)src.size() ? src[q] : 0.0;
}

}  // namespace

// The reference's whole solve with crossover on (InteriorPointSolver::solve,
// interior_point.cpp:402-886): status, objective, iterations, primal (n) and
// dual (m).
int ref_ipm_solve_crossover(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                            const int8_t* sense, const double* rhs, const double* lb, const double* ub,
                            const double* obj, int maximize, double obj_offset, int max_iterations,
                            int crossover_max_iter, int* status, int64_t* iterations, double* objective,
                            double* primal, double* dual) {
  model::Problem prob = build_problem(m, n, rp, ci, v, sense, rhs, lb, ub, obj, maximize, obj_offset);
  lp::InteriorPointOptions opts;
  opts.max_iterations = max_iterations;
  opts.crossover = true;
  opts.crossover_max_iter = crossover_max_iter;
  lp::InteriorPointSolver solver(numerical::ToleranceConfig::industrial_defaults(), opts);
  QuietCerr quiet;
  lp::InteriorPointResult res = solver.solve(prob);
  *status = static_cast<int>(res.status);
  *iterations = (int64_t)res.iterations;
  *objective = res.objective_value;
  copy_out(res.primal, primal, n);
  copy_out(res.dual, dual, m);
  return 0;
}

// The crossover block of InteriorPointSolver::solve (interior_point.cpp:786-878)
// run on the reference's SimplexSolver from a barrier endpoint given by the
// caller (for example the GPU IPM's): barrier status / objective / iterations /
// primal in, final ones plus the dual out.  No time limit and no interrupt,
// as in a default solve.  The reference polishes the problem after
// remove_duplicate_rows; a problem with duplicate rows returns -2 here.
int ref_crossover(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
                  const int8_t* sense, const double* rhs, const double* lb, const double* ub, const double* obj,
                  int maximize, double obj_offset, int barrier_status, double barrier_objective,
                  int64_t barrier_iterations, const double* barrier_primal, int crossover_max_iter, int* status,
                  int64_t* iterations, double* objective, double* primal, double* dual) {
  model::Problem cleaned = build_problem(m, n, rp, ci, v, sense, rhs, lb, ub, obj, maximize, obj_offset);
  if (has_duplicate_rows(cleaned)) return -2;
  model::ProblemStatus st = static_cast<model::ProblemStatus>(barrier_status);
  const bool done = st == model::ProblemStatus::OPTIMAL;
  double objective_value = barrier_objective;
  int64_t iters = barrier_iterations;
  std::vector<double> x(barrier_primal, barrier_primal + n), y((std::size_t)m, 0.0);
  auto adopt = [&](const lp::SimplexResult& sim) {
    if (sim.status == model::ProblemStatus::OPTIMAL) {
      st = model::ProblemStatus::OPTIMAL;
      objective_value = sim.objective_value;
      x = sim.primal;
      y = sim.dual;
      iters = (int64_t)sim.iterations;
    } else if (!done && sim.status == model::ProblemStatus::INFEASIBLE) {
      st = model::ProblemStatus::INFEASIBLE;
    } else if (sim.status == model::ProblemStatus::UNBOUNDED) {
      st = model::ProblemStatus::UNBOUNDED;
    }
  };
  const numerical::ToleranceConfig tol = numerical::ToleranceConfig::industrial_defaults();
  lp::SimplexOptions sim_opts;
  sim_opts.time_limit_seconds = 0.0;
  std::vector<int> var_status(cleaned.variables.size(), 0);
  for (std::size_t j = 0; j < cleaned.variables.size(); ++j) {
    const double l = cleaned.variables[j].lower_bound, u = cleaned.variables[j].upper_bound;
    const double at_tol = 1e-6 * (1.0 + std::max(std::abs(l), std::abs(u)));
    if (barrier_primal[j] <= l + at_tol)
      var_status[j] = -1;
    else if (barrier_primal[j] >= u - at_tol)
      var_status[j] = 1;
  }
  bool resolved = false;
  QuietCerr quiet;
  try {
    lp::SimplexOptions warm_opts = sim_opts;
    warm_opts.max_iterations = crossover_max_iter;
    lp::SimplexSolver polisher(tol, warm_opts);
    lp::SimplexResult sim = polisher.solve_with_basis(cleaned, var_status);
    adopt(sim);
    resolved = sim.status == model::ProblemStatus::OPTIMAL || sim.status == model::ProblemStatus::UNBOUNDED ||
               (sim.status == model::ProblemStatus::INFEASIBLE && !done);
  } catch (...) {
  }
  if (!resolved) {
    try {
      lp::SimplexSolver polisher(tol, sim_opts);
      adopt(polisher.solve(cleaned));
    } catch (...) {
    }
  }
  *status = static_cast<int>(st);
  *iterations = iters;
  *objective = objective_value;
  copy_out(x, primal, n);
  copy_out(y, dual, m);
  return 0;
}

// ---- presolve / postsolve (presolve::Presolver, core/presolve/presolve.hpp:49-85)
struct RefPresolve {
  model::Problem original;
  presolve::PresolveResult res;
  presolve::Presolver presolver;
  RefPresolve(model::Problem p, const presolve::PresolveOptions& o) : original(std::move(p)), presolver(o) {}
};

// Presolves as Solver::solve does (api/cpp/api.cpp:384-393; level 2 =
// AGGRESSIVE, the default, adds singleton removal, coefficient strengthening
// and dual reductions; level 1 = CONSERVATIVE).  dims = (m, n, nnz) of the
// reduced problem; returns -3 if its matrix is not stored by rows.
int ref_presolve(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v, const int8_t* sense,
                 const double* rhs, const double* lb, const double* ub, const double* obj, int maximize,
                 double obj_offset, int level, void** handle, int64_t* dims, double* ms) {
  presolve::PresolveOptions popts;
  popts.remove_singletons = level == 2;
  popts.coefficient_strengthening = level == 2;
  popts.dual_reductions = level == 2;
  auto* h = new RefPresolve(build_problem(m, n, rp, ci, v, sense, rhs, lb, ub, obj, maximize, obj_offset), popts);
  const auto t0 = std::chrono::steady_clock::now();
  h->res = h->presolver.presolve(h->original);
  const auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  const model::Problem& r = h->res.reduced_problem;
  if (r.constraint_matrix.order() != model::StorageOrder::CSR) {
    delete h;
    return -3;
  }
  dims[0] = (int64_t)r.constraints.size();
  dims[1] = (int64_t)r.variables.size();
  dims[2] = (int64_t)r.constraint_matrix.nnz();
  *handle = h;
  return 0;
}

// The reduced problem in the hn_lp layout (buffers sized by ref_presolve's dims).
int ref_presolve_reduced(void* handle, int64_t* rp, int64_t* ci, double* v, int8_t* sense, double* rhs, double* lb,
                         double* ub, double* obj, int* maximize, double* obj_offset) {
  const model::Problem& r = static_cast<RefPresolve*>(handle)->res.reduced_problem;
  const auto& A = r.constraint_matrix;
  for (std::size_t i = 0; i <= r.constraints.size(); ++i) rp[i] = i < A.row_ptr().size() ? (int64_t)A.row_ptr()[i] : 0;
  for (std::size_t k = 0; k < A.nnz(); ++k) {
    ci[k] = (int64_t)A.col_indices()[k];
    v[k] = A.values()[k];
  }
  for (std::size_t i = 0; i < r.constraints.size(); ++i) {
    sense[i] = static_cast<int8_t>(r.constraints[i].sense);
    rhs[i] = r.constraints[i].rhs;
  }
  for (std::size_t j = 0; j < r.variables.size(); ++j) {
    lb[j] = r.variables[j].lower_bound;
    ub[j] = r.variables[j].upper_bound;
    obj[j] = r.variables[j].objective_coeff;
  }
  *maximize = r.obj_sense == model::ObjectiveSense::MAXIMIZE;
  *obj_offset = r.obj_offset;
  return 0;
}

// Presolver::postsolve of a reduced solution (primal n2, dual m2 entries),
// then the objective as Solver::solve recomputes it (api.cpp:92-106).
int ref_postsolve(void* handle, const double* primal, int64_t n2, const double* dual, int64_t m2,
                  double* primal_out, double* dual_out, double* objective, double* ms) {
  auto* h = static_cast<RefPresolve*>(handle);
  model::Solution red;
  red.primal.assign(primal, primal + n2);
  red.dual.assign(dual, dual + m2);
  const auto t0 = std::chrono::steady_clock::now();
  model::Solution full = h->presolver.postsolve(h->original, red, h->res);
  const auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  double o = h->original.obj_offset;
  for (std::size_t j = 0; j < h->original.variables.size() && j < full.primal.size(); ++j)
    o += h->original.variables[j].objective_coeff * full.primal[j];
  *objective = o;
  copy_out(full.primal, primal_out, (int64_t)h->original.variables.size());
  copy_out(full.dual, dual_out, (int64_t)h->original.constraints.size());
  return 0;
}

void ref_presolve_free(void* handle) { delete static_cast<RefPresolve*>(handle); }

// SolutionVerifier::verify_detailed on the problem ProblemBuilder builds
// from the arguments.  var_type: n model::VarType codes (NULL: continuous);
// n_dual / n_rc: the lengths given to the solution's dual and reduced_costs
// (entries are read only when a vector is not NULL).  flags = (feasible,
// optimal); vals = (primal, dual, complementarity, integrality, objective
// discrepancy); *seconds = wall time of verify_detailed.
int ref_verify(int64_t m, int64_t n, const int64_t* rp, const int64_t* ci, const double* v,
               const int8_t* sense, const double* rhs, const double* lb, const double* ub,
               const double* obj, int maximize, double obj_offset, const int8_t* var_type,
               const double* primal, int64_t n_primal, const double* dual, int64_t n_dual,
               const double* rc, int64_t n_rc, double objective_value, const double* tol, int* flags,
               double* vals, int64_t* violated, int64_t cap, int64_t* n_violated, char* msg,
               int64_t msg_cap, double* seconds) {
  model::ProblemBuilder b("ref");
  for (int64_t j = 0; j < n; ++j)
    b.add_variable(lb[j], ub[j], var_type ? static_cast<model::VarType>(var_type[j]) : model::VarType::CONTINUOUS);
  for (int64_t i = 0; i < m; ++i) {
    std::vector<std::pair<std::size_t, double>> coeffs;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) coeffs.push_back({(std::size_t)ci[k], v[k]});
    b.add_constraint(coeffs, static_cast<model::ConstraintSense>(sense[i]), rhs[i]);
  }
  std::vector<std::pair<std::size_t, double>> o;
  for (int64_t j = 0; j < n; ++j) o.push_back({(std::size_t)j, obj[j]});
  b.set_objective(o, maximize ? model::ObjectiveSense::MAXIMIZE : model::ObjectiveSense::MINIMIZE);
  model::Problem prob = b.build();
  prob.obj_offset = obj_offset;

  model::Solution sol;
  sol.objective_value = objective_value;
  sol.primal.assign(primal, primal + n_primal);
  if (dual)
    sol.dual.assign(dual, dual + n_dual);
  else
    sol.dual.assign((std::size_t)n_dual, 0.0);
  if (rc)
    sol.reduced_costs.assign(rc, rc + n_rc);
  else
    sol.reduced_costs.assign((std::size_t)n_rc, 0.0);
  numerical::ToleranceConfig t = numerical::ToleranceConfig::industrial_defaults();
  if (tol) t.feasibility_tol(tol[0]).optimality_tol(tol[1]).complementarity_tol(tol[2]).integrality_tol(tol[3]);
  validation::SolutionVerifier verifier(t);
  const auto w0 = std::chrono::steady_clock::now();
  validation::VerificationResult r = verifier.verify_detailed(prob, sol);
  const auto w1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(w1 - w0).count();
  flags[0] = r.feasible;
  flags[1] = r.optimal;
  vals[0] = r.primal_infeasibility;
  vals[1] = r.dual_infeasibility;
  vals[2] = r.complementarity;
  vals[3] = r.integrality_violation;
  vals[4] = r.objective_discrepancy;
  *n_violated = (int64_t)r.violated_constraints.size();
  for (int64_t q = 0; q < cap && q < *n_violated; ++q) violated[q] = (int64_t)r.violated_constraints[q];
  if (msg && msg_cap > 0) std::snprintf(msg, (size_t)msg_cap, "%s", r.message.c_str());
  return 0;
}


// ---- LP-format files (model::LPParser and model::write_lp, core/model/lp_parser.cpp)
struct RefLpFile {
  model::LPParseResult res;
};

// LPParser::parse_string (lp_parser.cpp:25-111).  dims = (m, n, nnz,
// number of errors, success); returns -3 if the matrix is not stored by rows.
int ref_lp_parse(const char* text, int keep_names, void** handle, int64_t* dims) {
  model::LPParser parser;
  parser.set_keep_names(keep_names != 0);
  auto* h = new RefLpFile{parser.parse_string(text)};
  const model::Problem& p = h->res.problem;
  if (h->res.success && p.constraint_matrix.order() != model::StorageOrder::CSR) {
    delete h;
    return -3;
  }
  dims[0] = (int64_t)p.constraints.size();
  dims[1] = (int64_t)p.variables.size();
  dims[2] = (int64_t)p.constraint_matrix.nnz();
  dims[3] = (int64_t)h->res.errors.size();
  dims[4] = h->res.success ? 1 : 0;
  *handle = h;
  return 0;
}

// The parsed problem in the hn_lp layout (buffers sized by ref_lp_parse's
// dims), the column types (VarType) and the names, each ended by '\n'.
int ref_lp_problem(void* handle, int64_t* rp, int64_t* ci, double* v, int8_t* sense, double* rhs, double* lb,
                   double* ub, double* obj, int8_t* vtype, int* maximize, char* names, int64_t names_cap) {
  const model::Problem& r = static_cast<RefLpFile*>(handle)->res.problem;
  const auto& A = r.constraint_matrix;
  for (std::size_t i = 0; i <= r.constraints.size(); ++i) rp[i] = i < A.row_ptr().size() ? (int64_t)A.row_ptr()[i] : 0;
  for (std::size_t k = 0; k < A.nnz(); ++k) {
    ci[k] = (int64_t)A.col_indices()[k];
    v[k] = A.values()[k];
  }
  std::string nm;
  for (std::size_t i = 0; i < r.constraints.size(); ++i) {
    sense[i] = static_cast<int8_t>(r.constraints[i].sense);
    rhs[i] = r.constraints[i].rhs;
  }
  for (std::size_t j = 0; j < r.variables.size(); ++j) {
    lb[j] = r.variables[j].lower_bound;
    ub[j] = r.variables[j].upper_bound;
    obj[j] = r.variables[j].objective_coeff;
    vtype[j] = static_cast<int8_t>(r.variables[j].type);
    nm += r.variables[j].name + "\n";
  }
  for (const auto& c : r.constraints) nm += c.name + "\n";
  *maximize = r.obj_sense == model::ObjectiveSense::MAXIMIZE;
  if ((int64_t)nm.size() + 1 > names_cap) return -4;
  std::memcpy(names, nm.c_str(), nm.size() + 1);
  return 0;
}

// The parser's errors, each ended by '\n'.
int ref_lp_errors(void* handle, char* buf, int64_t cap) {
  std::string all;
  for (const auto& e : static_cast<RefLpFile*>(handle)->res.errors) all += e + "\n";
  if ((int64_t)all.size() + 1 > cap) return -4;
  std::memcpy(buf, all.c_str(), all.size() + 1);
  return 0;
}

// write_lp (lp_parser.cpp:603-714) of the parsed problem.
int ref_lp_write(void* handle, const char* path) {
  try {
    model::write_lp(static_cast<RefLpFile*>(handle)->res.problem, path);
  } catch (...) {
    return -1;
  }
  return 0;
}

void ref_lp_free(void* handle) { delete static_cast<RefLpFile*>(handle); }

// SparseCholesky (analyze + factorize with the industrial tolerances) of the
// n x n CSR given as is (rows sorted, no duplicates expected).  *l_nnz and
// *singular receive the factor size and the singular flag.
void* ref_chol_create(int64_t n, const uint64_t* rp, const uint64_t* ci, const double* v, int64_t* l_nnz,
                      int* singular) {
  numerical::SparseMatrix A((std::size_t)n, (std::size_t)n);
  A.mutable_row_ptr().assign(rp, rp + n + 1);
  A.mutable_col_indices().assign(ci, ci + rp[n]);
  A.mutable_values().assign(v, v + rp[n]);
  auto* c = new numerical::SparseCholesky(A, numerical::ToleranceConfig::industrial_defaults());
  *l_nnz = (int64_t)c->L_row_ptr()[n];
  *singular = c->stats().singular ? 1 : 0;
  return c;
}

// perm (n), L_row_ptr (n+1), L_col_indices and L_values (l_nnz), D (n);
// any pointer may be NULL.
int ref_chol_get(void* h, uint64_t* perm, uint64_t* lrp, uint64_t* lci, double* lv, double* d) {
  auto* c = static_cast<numerical::SparseCholesky*>(h);
  const std::size_t n = c->perm().size();
  if (perm) for (std::size_t i = 0; i < n; ++i) perm[i] = c->perm()[i];
  if (lrp) for (std::size_t i = 0; i <= n; ++i) lrp[i] = c->L_row_ptr()[i];
  const std::size_t nnz = c->L_row_ptr()[n];
  if (lci) for (std::size_t k = 0; k < nnz; ++k) lci[k] = c->L_col_indices()[k];
  if (lv) for (std::size_t k = 0; k < nnz; ++k) lv[k] = c->L_values()[k];
  if (d) for (std::size_t i = 0; i < n; ++i) d[i] = c->D_values()[i];
  return 0;
}

// x = A^{-1} b by SparseCholesky::solve.
int ref_chol_solve(void* h, const double* b, double* x) {
  auto* c = static_cast<numerical::SparseCholesky*>(h);
  std::vector<double> t(b, b + c->perm().size());
  c->solve(t);
  std::memcpy(x, t.data(), sizeof(double) * t.size());
  return 0;
}

void ref_chol_free(void* h) { delete static_cast<numerical::SparseCholesky*>(h); }

}  // extern "C"
