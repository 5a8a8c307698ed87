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
  std::vector<double> row_gamma;   // row scale of each kept row
  std::vector<int64_t> row_slack;  // standard-form column of each kept row's slack (-1: EQ row)
  std::vector<int64_t> ub_row;     // per original variable: its upper-bound row (-1: none)
  std::vector<int64_t> ub_slack;   // per original variable: that row's slack column (-1: none)
};

// KKT measures of an iterate on the caller's own problem (minimisation
// sense, all original rows and columns).  The scaled standard-form point
// (x, lam, s) is mapped back: primal by to_original_primal, row duals
// y_i = +-gamma_i lam_r (0 for a dropped duplicate row), and one net bound
// multiplier per column z_j = (lower multiplier) - (upper multiplier), with
// lower = s_k / k_k and upper = s_w / k_w (k: column scale) for the column's
// slot k and its bound row's slack w.
//   primal_res = max(row violations by sense, bound violations)
//   dual_res   = max(|(A'y + z - c)_j| over non-fixed columns,
//                    |(+-y_i) + v_i| over LE/GE rows, v_i the row slack's multiplier)
//   primal_obj = c'x, dual_obj = b'y + sum_j (lb_j lower_j - ub_j upper_j) + fixed terms
// and the relative forms divide by 1 + ||b||_inf, 1 + ||c||_inf and
// 1 + |c'x|.  In the standard-form case (EQ rows, x >= 0) this is
// ||Ax - b||, ||A'y + s - c|| and |c'x - b'y| with x = K x_s, y = G lam,
// s = K^-1 s_s.
struct OriginalKkt {
  double primal_res = 0.0, dual_res = 0.0, primal_obj = 0.0, dual_obj = 0.0;
  double rel_primal = 0.0, rel_dual = 0.0, rel_gap = 0.0;
};
void original_kkt(const hn_lp& lp, const StandardForm& f, const double* x, const double* lam, const double* s,
                  OriginalKkt& out);
// The same three relative measures on the scaled standard form (the space
// the reference's stopping test works in): ||rp||_inf / (1 + ||rhs||_inf),
// ||rd||_inf / (1 + ||c_s||_inf), |c_s'x - rhs'lam| / (1 + |c_s'x|), into
// res.scaled_rel_*.
void scaled_kkt(const StandardForm& f, const double* x, const double* lam, double rp_norm, double rd_norm,
                hn_ipm_result& res);

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
