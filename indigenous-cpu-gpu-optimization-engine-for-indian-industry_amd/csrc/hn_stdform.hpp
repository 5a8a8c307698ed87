// Host-side preparation of an LP for the device IPM: the reference's problem
// normalisation, duplicate-row removal and scaled standard form, and the map
// back to the original variables.
#pragma once

#include <cstdint>
#include <vector>

#include "hn_kernels.h"

namespace hn {

struct VarMap {
  bool fixed = false;
  double fixed_value = 0.0;
  int nslots = 0;
  int64_t slot[2] = {0, 0};
  double sign[2] = {0, 0};
  double shift[2] = {0, 0};
};

// Scaled standard form  min c'x, M x = rhs, x >= 0  (interior_point.cpp:30-50).
struct StandardForm {
  int64_t m = 0, n = 0;
  std::vector<int64_t> rp, ci;  // CSR of M, columns sorted, no explicit zeros
  std::vector<double> v;
  std::vector<double> rhs, c, col_scale;
  double offset = 0.0;
  std::vector<VarMap> vars;     // one per original variable
  std::vector<int64_t> kept_rows;  // original row of each of the first rows
};

// Builds the standard form of `lp` (normalisation as ProblemBuilder::build,
// problem.hpp:265-281; remove_duplicate_rows, interior_point.cpp:82-132;
// build_standard_form, interior_point.cpp:202-373).  Returns HN_OK or an
// HN_ERR_* code with the error message set.
int build_standard_form(const hn_lp& lp, StandardForm& f);

// interior_point.cpp:375-395
void to_original_primal(const StandardForm& f, const double* x, double* primal);

// True when (x0, y0, s0) may replace the standard start point: finite,
// x0 > 0, s0 > 0, mu0 below the standard start's mu, min_j x0_j s0_j >=
// 1e-3 mu0, and ||M x0 - b||_inf no larger than at the standard start.
bool warm_start_acceptable(const StandardForm& f, const double* x0, const double* y0, const double* s0);

}  // namespace hn
