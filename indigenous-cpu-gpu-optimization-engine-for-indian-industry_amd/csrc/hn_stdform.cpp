// Host-side LP preparation for the device IPM.  See hn_stdform.hpp.
//
// The three steps follow the reference exactly because the device iterates
// on their output and must reach the same point:
//   - rows normalised like ProblemBuilder::build (problem.hpp:265-281);
//   - a row whose sorted entries, sense and rhs are bit-identical to an
//     earlier row is dropped (remove_duplicate_rows, interior_point.cpp:82-132);
//   - bound shifts, free splits, LE/GE slacks, upper-bound rows, row scale
//     1/max|a_ij| and column scale 1/max|m_ij| (interior_point.cpp:202-373).
#include "hn_stdform.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "hn_internal.hpp"

namespace hn {

namespace {

constexpr double kFeasTol = 1e-9;  // ToleranceConfig::industrial_defaults (tolerance.hpp:60-75)

struct Trip {
  int64_t r, c;
  double v;
};

// CSR from triplets: sort by (row, col, insertion), sum duplicates, drop
// exact-zero sums (SparseMatrix::from_triplets/sum_duplicates,
// sparse_matrix.cpp:8-53,230-262).
void csr_from_triplets(int64_t rows, std::vector<Trip>& t, std::vector<int64_t>& rp,
                       std::vector<int64_t>& ci, std::vector<double>& v) {
  const auto less = [](const Trip& a, const Trip& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; };
  // triplets usually arrive in order (rows in turn, columns increasing);
  // the stable sort of an ordered sequence is the sequence itself
  if (!std::is_sorted(t.begin(), t.end(), less)) std::stable_sort(t.begin(), t.end(), less);
  rp.assign(rows + 1, 0);
  ci.clear();
  v.clear();
  ci.reserve(t.size());
  v.reserve(t.size());
  size_t k = 0;
  for (int64_t i = 0; i < rows; ++i) {
    rp[i] = (int64_t)ci.size();
    while (k < t.size() && t[k].r == i) {
      const int64_t c = t[k].c;
      double s = t[k].v;
      ++k;
      while (k < t.size() && t[k].r == i && t[k].c == c) s += t[k++].v;
      if (s != 0.0) {
        ci.push_back(c);
        v.push_back(s);
      }
    }
  }
  rp[rows] = (int64_t)ci.size();
}

uint64_t bits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, sizeof u);
  return u;
}

}  // namespace

int build_standard_form(const hn_lp& lp, StandardForm& f) {
  const int64_t m0 = lp.m, n = lp.n;
  if (m0 < 0 || n < 0) return fail(HN_ERR_INVALID, "hn_ipm_solve: negative dimensions");
  if ((m0 > 0 && (!lp.row_ptr || !lp.sense || !lp.rhs)) || (n > 0 && (!lp.lb || !lp.ub || !lp.obj)))
    return fail(HN_ERR_INVALID, "hn_ipm_solve: null problem array");
  const int64_t nz0 = m0 > 0 ? lp.row_ptr[m0] : 0;
  if (m0 > 0 && lp.row_ptr[0] != 0) return fail(HN_ERR_INVALID, "hn_ipm_solve: row_ptr[0] must be 0");
  if (nz0 > 0 && (!lp.col || !lp.val)) return fail(HN_ERR_INVALID, "hn_ipm_solve: null col/val");
  for (int64_t i = 0; i < m0; ++i) {
    if (lp.row_ptr[i + 1] < lp.row_ptr[i]) return fail(HN_ERR_INVALID, "hn_ipm_solve: row_ptr decreasing");
    if (lp.sense[i] < HN_LE || lp.sense[i] > HN_EQ) return fail(HN_ERR_INVALID, "hn_ipm_solve: bad sense");
  }
  for (int64_t k = 0; k < nz0; ++k)
    if (lp.col[k] < 0 || lp.col[k] >= n) return fail(HN_ERR_INVALID, "hn_ipm_solve: column out of range");

  // 1. normalise rows
  std::vector<Trip> t;
  t.reserve(nz0);
  for (int64_t i = 0; i < m0; ++i)
    for (int64_t k = lp.row_ptr[i]; k < lp.row_ptr[i + 1]; ++k) t.push_back({i, lp.col[k], lp.val[k]});
  std::vector<int64_t> arp, aci;
  std::vector<double> av;
  csr_from_triplets(m0, t, arp, aci, av);

  // 2. duplicate rows (first occurrence kept)
  std::vector<char> keep(m0, 1);
  {
    std::vector<std::pair<uint64_t, int64_t>> h(m0);
    for (int64_t i = 0; i < m0; ++i) {
      uint64_t x = 1469598103934665603ULL;
      auto mix = [&](uint64_t y) {
        x ^= y;
        x *= 1099511628211ULL;
        x ^= x >> 29;
      };
      for (int64_t k = arp[i]; k < arp[i + 1]; ++k) {
        mix((uint64_t)aci[k]);
        mix(bits(av[k]));
      }
      mix((uint64_t)lp.sense[i] + 7);
      mix(bits(lp.rhs[i]));
      h[i] = {x, i};
    }
    std::sort(h.begin(), h.end());
    auto same = [&](int64_t a, int64_t b) {
      if (lp.sense[a] != lp.sense[b] || bits(lp.rhs[a]) != bits(lp.rhs[b])) return false;
      if (arp[a + 1] - arp[a] != arp[b + 1] - arp[b]) return false;
      for (int64_t u = 0; u < arp[a + 1] - arp[a]; ++u)
        if (aci[arp[a] + u] != aci[arp[b] + u] || bits(av[arp[a] + u]) != bits(av[arp[b] + u])) return false;
      return true;
    };
    for (int64_t g = 0; g < m0;) {
      int64_t e = g;
      while (e < m0 && h[e].first == h[g].first) ++e;
      for (int64_t a = g + 1; a < e; ++a)
        for (int64_t b = g; b < a; ++b)
          if (keep[h[b].second] && same(h[a].second, h[b].second)) {
            keep[h[a].second] = 0;
            break;
          }
      g = e;
    }
  }

  // 3. standard form
  std::vector<double> cbase(n);
  for (int64_t j = 0; j < n; ++j) cbase[j] = lp.maximize ? -lp.obj[j] : lp.obj[j];
  f = StandardForm();
  f.vars.assign(n, VarMap());
  int64_t nslots = 0, ntle = 0, m_ub = 0;
  for (int64_t j = 0; j < n; ++j) {
    const double lb = lp.lb[j], ub = lp.ub[j];
    VarMap& vm = f.vars[j];
    if (std::isfinite(lb) && std::isfinite(ub) && std::abs(ub - lb) <= kFeasTol) {
      vm.fixed = true;
      vm.fixed_value = lb;
      f.offset += cbase[j] * lb;
    } else if (std::isfinite(lb)) {
      vm.nslots = 1;
      vm.slot[0] = nslots++;
      vm.sign[0] = 1.0;
      vm.shift[0] = lb;
      f.offset += cbase[j] * lb;
      if (std::isfinite(ub)) ++m_ub;
    } else if (std::isfinite(ub)) {
      vm.nslots = 1;
      vm.slot[0] = nslots++;
      vm.sign[0] = -1.0;
      vm.shift[0] = ub;
      f.offset += cbase[j] * ub;
    } else {
      vm.nslots = 2;
      vm.slot[0] = nslots++;
      vm.sign[0] = 1.0;
      vm.shift[0] = 0.0;
      vm.slot[1] = nslots++;
      vm.sign[1] = -1.0;
      vm.shift[1] = 0.0;
    }
  }
  for (int64_t i = 0; i < m0; ++i)
    if (keep[i]) {
      f.kept_rows.push_back(i);
      if (lp.sense[i] == HN_LE || lp.sense[i] == HN_GE) ++ntle;
    }
  const int64_t norig = (int64_t)f.kept_rows.size();
  const int64_t m = norig + m_ub, nc = nslots + ntle + m_ub;
  f.m = m;
  f.n = nc;
  f.rhs.assign(m, 0.0);
  f.c.assign(nc, 0.0);
  for (int64_t j = 0; j < n; ++j)
    for (int s = 0; s < f.vars[j].nslots; ++s) f.c[f.vars[j].slot[s]] = f.vars[j].sign[s] > 0 ? cbase[j] : -cbase[j];
  t.clear();
  f.row_gamma.assign(norig, 1.0);
  f.row_slack.assign(norig, -1);
  f.ub_row.assign(n, -1);
  f.ub_slack.assign(n, -1);
  int64_t next_slack = nslots;
  for (int64_t r = 0; r < norig; ++r) {
    const int64_t i = f.kept_rows[r];
    double rowmax = 0.0;
    for (int64_t k = arp[i]; k < arp[i + 1]; ++k) rowmax = std::max(rowmax, std::abs(av[k]));
    const double gamma = rowmax > 0.0 ? 1.0 / rowmax : 1.0;
    const bool flip = lp.sense[i] == HN_GE;
    double const_i = 0.0;
    for (int64_t k = arp[i]; k < arp[i + 1]; ++k) {
      const VarMap& vm = f.vars[aci[k]];
      const double a = gamma * av[k];
      if (vm.fixed) {
        const_i += a * vm.fixed_value;
      } else {
        for (int s = 0; s < vm.nslots; ++s) {
          t.push_back({r, vm.slot[s], flip ? -a * vm.sign[s] : a * vm.sign[s]});
          const_i += a * vm.shift[s];
        }
      }
    }
    f.row_gamma[r] = gamma;
    if (lp.sense[i] == HN_LE || lp.sense[i] == HN_GE) {
      f.row_slack[r] = next_slack;
      t.push_back({r, next_slack++, gamma});
    }
    f.rhs[r] = flip ? (-gamma * lp.rhs[i] + const_i) : (gamma * lp.rhs[i] - const_i);
  }
  int64_t next_ub = norig;
  for (int64_t j = 0; j < n; ++j) {
    const double lb = lp.lb[j], ub = lp.ub[j];
    if (std::isfinite(lb) && std::isfinite(ub) && std::abs(ub - lb) > kFeasTol) {
      f.ub_row[j] = next_ub;
      f.ub_slack[j] = next_slack;
      t.push_back({next_ub, f.vars[j].slot[0], 1.0});
      t.push_back({next_ub, next_slack++, 1.0});
      f.rhs[next_ub] = ub - lb;
      ++next_ub;
    }
  }
  std::vector<int64_t> mrp, mci;
  std::vector<double> mv;
  csr_from_triplets(m, t, mrp, mci, mv);
  f.col_scale.assign(nc, 1.0);
  std::vector<double> colmax(nc, 0.0);
  for (size_t k = 0; k < mci.size(); ++k) colmax[mci[k]] = std::max(colmax[mci[k]], std::abs(mv[k]));
  for (size_t k = 0; k < mci.size(); ++k)
    if (colmax[mci[k]] > 0.0) f.col_scale[mci[k]] = 1.0 / colmax[mci[k]];
  t.clear();
  for (int64_t i = 0; i < m; ++i)
    for (int64_t k = mrp[i]; k < mrp[i + 1]; ++k) t.push_back({i, mci[k], mv[k] * f.col_scale[mci[k]]});
  csr_from_triplets(m, t, f.rp, f.ci, f.v);
  for (int64_t j = 0; j < nc; ++j) f.c[j] *= f.col_scale[j];
  return HN_OK;
}

void to_original_primal(const StandardForm& f, const double* x, double* primal) {
  for (size_t j = 0; j < f.vars.size(); ++j) {
    const VarMap& vm = f.vars[j];
    if (vm.fixed) {
      primal[j] = vm.fixed_value;
      continue;
    }
    double val = 0.0;
    for (int s = 0; s < vm.nslots; ++s) {
      const int64_t gi = vm.slot[s];
      if (gi < 0 || gi >= f.n) continue;
      val += vm.sign[s] * f.col_scale[gi] * x[gi];
      if (s == 0) val += vm.shift[s];
    }
    primal[j] = val;
  }
}

void original_kkt(const hn_lp& lp, const StandardForm& f, const double* x, const double* lam, const double* s,
                  OriginalKkt& out) {
  out = OriginalKkt();
  const int64_t m0 = lp.m, n = lp.n;
  std::vector<double> primal(n), y(m0, 0.0), aty(n, 0.0);
  to_original_primal(f, x, primal.data());
  const int64_t norig = (int64_t)f.kept_rows.size();
  for (int64_t r = 0; r < norig; ++r) {
    const int64_t i = f.kept_rows[r];
    const double yi = f.row_gamma[r] * lam[r];
    y[i] = lp.sense[i] == HN_GE ? -yi : yi;
  }
  double bn = 0.0, cn = 0.0, pres = 0.0, dres = 0.0, pobj = 0.0, dobj = 0.0;
  for (int64_t i = 0; i < m0; ++i) {
    double ax = 0.0;
    for (int64_t k = lp.row_ptr[i]; k < lp.row_ptr[i + 1]; ++k) {
      ax += lp.val[k] * primal[lp.col[k]];
      aty[lp.col[k]] += lp.val[k] * y[i];
    }
    const double d = ax - lp.rhs[i];
    const double v = lp.sense[i] == HN_EQ ? std::abs(d) : lp.sense[i] == HN_LE ? std::max(0.0, d) : std::max(0.0, -d);
    pres = std::max(pres, v);
    bn = std::max(bn, std::abs(lp.rhs[i]));
    dobj += lp.rhs[i] * y[i];
  }
  for (int64_t r = 0; r < norig; ++r) {  // row slacks: (+-y_i) + v_i = 0
    const int64_t w = f.row_slack[r];
    if (w < 0) continue;
    const int64_t i = f.kept_rows[r];
    const double v = s[w] / f.col_scale[w];
    dres = std::max(dres, std::abs((lp.sense[i] == HN_GE ? -y[i] : y[i]) + v));
  }
  for (int64_t j = 0; j < n; ++j) {
    const double cj = lp.maximize ? -lp.obj[j] : lp.obj[j];
    const double lb = lp.lb[j], ub = lp.ub[j], xj = primal[j];
    cn = std::max(cn, std::abs(lp.obj[j]));
    pobj += cj * xj;
    pres = std::max(pres, std::max(std::max(0.0, lb - xj), std::max(0.0, xj - ub)));
    const VarMap& vm = f.vars[j];
    if (vm.fixed) {  // its bound multiplier is free: c_j - (A'y)_j, no residual
      dobj += vm.fixed_value * (cj - aty[j]);
      continue;
    }
    double lower = 0.0, upper = 0.0;
    if (vm.nslots == 1) {
      const int64_t k = vm.slot[0];
      const double mult = s[k] / f.col_scale[k];
      if (vm.sign[0] > 0) lower = mult; else upper = mult;
      if (f.ub_slack[j] >= 0) upper = s[f.ub_slack[j]] / f.col_scale[f.ub_slack[j]];
    }
    if (lower != 0.0) dobj += lb * lower;
    if (upper != 0.0) dobj -= ub * upper;
    dres = std::max(dres, std::abs(aty[j] + (lower - upper) - cj));
  }
  out.primal_res = pres;
  out.dual_res = dres;
  out.primal_obj = pobj;
  out.dual_obj = dobj;
  out.rel_primal = pres / (1.0 + bn);
  out.rel_dual = dres / (1.0 + cn);
  out.rel_gap = std::abs(pobj - dobj) / (1.0 + std::abs(pobj));
}

void scaled_kkt(const StandardForm& f, const double* x, const double* lam, double rp_norm, double rd_norm,
                hn_ipm_result& res) {
  double bn = 0.0, cn = 0.0, cx = 0.0, by = 0.0;
  for (int64_t i = 0; i < f.m; ++i) {
    bn = std::max(bn, std::abs(f.rhs[i]));
    by += f.rhs[i] * lam[i];
  }
  for (int64_t j = 0; j < f.n; ++j) {
    cn = std::max(cn, std::abs(f.c[j]));
    cx += f.c[j] * x[j];
  }
  res.scaled_rel_primal = rp_norm / (1.0 + bn);
  res.scaled_rel_dual = rd_norm / (1.0 + cn);
  res.scaled_rel_gap = std::abs(cx - by) / (1.0 + std::abs(cx));
}

// The warm-start acceptance test (no counterpart in the reference, whose
// solver always starts at interior_point.cpp:443-453): finite, strictly
// positive x0 and s0; mu0 below the standard start's mu; every product
// x0_j s0_j at least 1e-3 mu0; ||M x0 - b||_inf no larger than at the
// standard start.
bool warm_start_acceptable(const StandardForm& f, const double* x0, const double* y0, const double* s0) {
  if (!x0 || !y0 || !s0) return false;
  const int64_t m = f.m, n = f.n;
  if (n == 0) return false;
  double bn = 1.0, cn = 1.0;
  for (double v : f.rhs) bn = std::max(bn, std::abs(v));
  for (double v : f.c) cn = std::max(cn, std::abs(v));
  double mu_cold = 0.0, mu0 = 0.0, pmin = HUGE_VAL;
  for (int64_t j = 0; j < n; ++j) {
    if (!(x0[j] > 0.0) || !(s0[j] > 0.0) || !std::isfinite(x0[j]) || !std::isfinite(s0[j])) return false;
    mu_cold += (bn + 1.0) * std::max(cn + 1.0, std::abs(f.c[j]));
    const double p = x0[j] * s0[j];
    mu0 += p;
    pmin = std::min(pmin, p);
  }
  for (int64_t i = 0; i < m; ++i)
    if (!std::isfinite(y0[i])) return false;
  mu_cold /= (double)n;
  mu0 /= (double)n;
  if (!(mu0 < mu_cold) || !(pmin >= 1e-3 * mu0)) return false;
  double rw = 0.0, rc = 0.0;
  for (int64_t i = 0; i < m; ++i) {
    double sw = 0.0, sc = 0.0;
    for (int64_t k = f.rp[i]; k < f.rp[i + 1]; ++k) {
      sw += f.v[k] * x0[f.ci[k]];
      sc += f.v[k] * (bn + 1.0);
    }
    rw = std::max(rw, std::abs(sw - f.rhs[i]));
    rc = std::max(rc, std::abs(sc - f.rhs[i]));
  }
  return rw <= rc;
}

}  // namespace hn

extern "C" int hn_lp_standard_form(const hn_lp* lp, int64_t* sf_m, int64_t* sf_n, int64_t* nnz, int64_t* row_ptr,
                                   int64_t* col, double* val, double* rhs, double* c) {
  HN_CHECK(lp && sf_m && sf_n && nnz, HN_ERR_INVALID, "hn_lp_standard_form: null argument");
  hn::StandardForm f;
  HN_TRY(hn::build_standard_form(*lp, f));
  *sf_m = f.m;
  *sf_n = f.n;
  *nnz = (int64_t)f.ci.size();
  if (row_ptr) std::memcpy(row_ptr, f.rp.data(), sizeof(int64_t) * f.rp.size());
  if (col) std::memcpy(col, f.ci.data(), sizeof(int64_t) * f.ci.size());
  if (val) std::memcpy(val, f.v.data(), sizeof(double) * f.v.size());
  if (rhs) std::memcpy(rhs, f.rhs.data(), sizeof(double) * f.rhs.size());
  if (c) std::memcpy(c, f.c.data(), sizeof(double) * f.c.size());
  return HN_OK;
}
