// Seeded synthetic inputs for the LP hot path.  See hn_gen.h for the exact
// distributions and draw order.
#include "hn_gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  bool have_spare = false;
  double spare = 0.0;

  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  int64_t below(int64_t n) { return (int64_t)(uniform() * (double)n); }
  double normal() {
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 2.0 * M_PI * u2;
    spare = r * std::sin(th);
    have_spare = true;
    return r * std::cos(th);
  }
};

template <typename T>
T* dup(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(sizeof(T) * std::max<size_t>(v.size(), 1)));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

// Column-wise triplets -> CSR with columns sorted within each row (stable).
void columns_to_csr(int64_t m, int64_t n, const std::vector<int64_t>& colptr,
                    const std::vector<int64_t>& rows, const std::vector<double>& vals,
                    std::vector<int64_t>& rp, std::vector<int64_t>& ci, std::vector<double>& cv) {
  const int64_t nnz = (int64_t)rows.size();
  rp.assign(m + 1, 0);
  for (int64_t k = 0; k < nnz; ++k) rp[rows[k] + 1]++;
  for (int64_t i = 0; i < m; ++i) rp[i + 1] += rp[i];
  std::vector<int64_t> cur(rp.begin(), rp.end() - 1);
  ci.assign(nnz, 0);
  cv.assign(nnz, 0.0);
  for (int64_t j = 0; j < n; ++j)
    for (int64_t k = colptr[j]; k < colptr[j + 1]; ++k) {
      const int64_t p = cur[rows[k]]++;
      ci[p] = j;
      cv[p] = vals[k];
    }
}

hn_gen_lp* finish_std_lp(Rng& rng, int64_t m, int64_t n, std::vector<int64_t>& rp,
                         std::vector<int64_t>& ci, std::vector<double>& cv) {
  std::vector<double> x0(n), y0(m), s0(n);
  for (auto& v : x0) v = rng.uniform(0.5, 2.0);
  for (auto& v : y0) v = rng.normal();
  for (auto& v : s0) v = rng.uniform(0.5, 2.0);
  std::vector<double> b(m, 0.0), c(s0);
  for (int64_t i = 0; i < m; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += cv[k] * x0[ci[k]];
    b[i] = s;
  }
  std::vector<double> aty(n, 0.0);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) aty[ci[k]] += cv[k] * y0[i];
  for (int64_t j = 0; j < n; ++j) c[j] = aty[j] + s0[j];
  hn_gen_lp* lp = static_cast<hn_gen_lp*>(std::calloc(1, sizeof(hn_gen_lp)));
  lp->m = m;
  lp->n = n;
  lp->nnz = (int64_t)ci.size();
  lp->row_ptr = dup(rp);
  lp->col = dup(ci);
  lp->val = dup(cv);
  lp->rhs = dup(b);
  lp->sense = dup(std::vector<int8_t>(m, HN_EQ));
  lp->lb = dup(std::vector<double>(n, 0.0));
  lp->ub = dup(std::vector<double>(n, std::numeric_limits<double>::infinity()));
  lp->obj = dup(c);
  lp->maximize = 0;
  lp->x0 = dup(x0);
  lp->y0 = dup(y0);
  lp->s0 = dup(s0);
  return lp;
}

// k distinct rows in [0,m) by rejection of repeats, in draw order.
void distinct_rows(Rng& rng, int64_t m, int k, std::vector<int64_t>& out) {
  out.clear();
  while ((int)out.size() < k) {
    const int64_t r = rng.below(m);
    if (std::find(out.begin(), out.end(), r) == out.end()) out.push_back(r);
  }
}

hn_gen_lp* random_lp(Rng& rng, int64_t m, int64_t n, int k) {
  if (m <= 0 || n <= 0 || k <= 0 || k > m) return nullptr;
  std::vector<int64_t> colptr(n + 1), rows, r;
  std::vector<double> vals;
  for (;;) {
    rows.clear();
    vals.clear();
    for (int64_t j = 0; j < n; ++j) {
      colptr[j] = (int64_t)rows.size();
      distinct_rows(rng, m, k, r);
      for (int u = 0; u < k; ++u) rows.push_back(r[u]);
      for (int u = 0; u < k; ++u) vals.push_back(rng.normal());
    }
    colptr[n] = (int64_t)rows.size();
    std::vector<char> seen(m, 0);
    for (int64_t v : rows) seen[v] = 1;
    if (std::all_of(seen.begin(), seen.end(), [](char c) { return c != 0; })) break;
  }
  std::vector<int64_t> rp, ci;
  std::vector<double> cv;
  columns_to_csr(m, n, colptr, rows, vals, rp, ci, cv);
  return finish_std_lp(rng, m, n, rp, ci, cv);
}

}  // namespace

extern "C" hn_gen_lp* hn_gen_random_lp(uint64_t seed, int64_t m, int64_t n, int k) {
  Rng rng(seed);
  return random_lp(rng, m, n, k);
}

extern "C" hn_gen_lp* hn_gen_batch_lp(uint64_t seed, int64_t mlo, int64_t mhi, int k) {
  Rng rng(seed);
  const double lm = rng.uniform(std::log((double)mlo), std::log((double)mhi));
  const int64_t m = (int64_t)std::llround(std::exp(lm));
  const int64_t n = (int64_t)std::llround(2.5 * (double)m);
  return random_lp(rng, m, n, k);
}

extern "C" hn_gen_lp* hn_gen_powerlaw_lp(uint64_t seed, int64_t m, int64_t n, double alpha, int cmin,
                                         int cmax) {
  if (m <= 0 || n <= 0 || cmin < 1 || cmax < cmin || cmax > m) return nullptr;
  Rng rng(seed);
  std::vector<double> rowmul(m);
  for (auto& v : rowmul) v = std::pow(10.0, rng.uniform(-2.0, 2.0));
  std::vector<double> cdf(cmax - cmin + 1);
  double acc = 0.0;
  for (int c = cmin; c <= cmax; ++c) {
    acc += std::pow((double)c, -alpha);
    cdf[c - cmin] = acc;
  }
  for (auto& v : cdf) v /= acc;
  std::vector<int64_t> colptr(n + 1), rows, r;
  std::vector<double> vals;
  for (int64_t j = 0; j < n; ++j) {
    colptr[j] = (int64_t)rows.size();
    const double u = rng.uniform();
    const int cnt = cmin + (int)(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    distinct_rows(rng, m, std::min(cnt, cmax), r);
    for (int64_t v : r) rows.push_back(v);
    for (size_t u2 = 0; u2 < r.size(); ++u2) vals.push_back(rng.normal() * rowmul[r[u2]]);
  }
  colptr[n] = (int64_t)rows.size();
  std::vector<int64_t> rp, ci;
  std::vector<double> cv;
  columns_to_csr(m, n, colptr, rows, vals, rp, ci, cv);
  return finish_std_lp(rng, m, n, rp, ci, cv);
}

extern "C" hn_gen_lp* hn_gen_transport_lp(uint64_t seed, int S, int T) {
  if (S <= 0 || T <= 0) return nullptr;
  Rng rng(seed);
  std::vector<double> p(2 * S), q(2 * T), sup(S), dem(T), cost((size_t)S * T);
  for (auto& v : p) v = rng.uniform();
  for (auto& v : q) v = rng.uniform();
  for (auto& v : sup) v = rng.uniform(50.0, 150.0);
  for (auto& v : dem) v = rng.uniform(50.0, 150.0);
  const double ssum = std::accumulate(sup.begin(), sup.end(), 0.0);
  const double dsum = std::accumulate(dem.begin(), dem.end(), 0.0);
  for (auto& v : dem) v *= ssum / dsum;
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < T; ++j) {
      const double dx = p[2 * i] - q[2 * j], dy = p[2 * i + 1] - q[2 * j + 1];
      cost[(size_t)i * T + j] = std::sqrt(dx * dx + dy * dy) * (1.0 + 0.2 * rng.normal());
    }
  const int64_t m = S + T, n = (int64_t)S * T;
  std::vector<int64_t> rp(m + 1), ci;
  std::vector<double> cv;
  ci.reserve(2 * n);
  cv.reserve(2 * n);
  for (int i = 0; i < S; ++i) {
    rp[i] = (int64_t)ci.size();
    for (int j = 0; j < T; ++j) {
      ci.push_back((int64_t)i * T + j);
      cv.push_back(1.0);
    }
  }
  for (int j = 0; j < T; ++j) {
    rp[S + j] = (int64_t)ci.size();
    for (int i = 0; i < S; ++i) {
      ci.push_back((int64_t)i * T + j);
      cv.push_back(1.0);
    }
  }
  rp[m] = (int64_t)ci.size();
  std::vector<double> rhs(m);
  std::vector<int8_t> sense(m);
  for (int i = 0; i < S; ++i) {
    rhs[i] = sup[i];
    sense[i] = HN_LE;
  }
  for (int j = 0; j < T; ++j) {
    rhs[S + j] = dem[j];
    sense[S + j] = HN_EQ;
  }
  hn_gen_lp* lp = static_cast<hn_gen_lp*>(std::calloc(1, sizeof(hn_gen_lp)));
  lp->m = m;
  lp->n = n;
  lp->nnz = (int64_t)ci.size();
  lp->row_ptr = dup(rp);
  lp->col = dup(ci);
  lp->val = dup(cv);
  lp->rhs = dup(rhs);
  lp->sense = dup(sense);
  lp->lb = dup(std::vector<double>(n, 0.0));
  lp->ub = dup(std::vector<double>(n, std::numeric_limits<double>::infinity()));
  lp->obj = dup(cost);
  lp->maximize = 0;
  return lp;
}

extern "C" void hn_gen_lp_free(hn_gen_lp* lp) {
  if (!lp) return;
  std::free(lp->row_ptr);
  std::free(lp->col);
  std::free(lp->val);
  std::free(lp->rhs);
  std::free(lp->sense);
  std::free(lp->lb);
  std::free(lp->ub);
  std::free(lp->obj);
  std::free(lp->x0);
  std::free(lp->y0);
  std::free(lp->s0);
  std::free(lp);
}

extern "C" hn_gen_spmv* hn_gen_powerlaw_spmv(uint64_t seed, int64_t nrows, int64_t ncols, int64_t nnz_target,
                                             double s, int64_t cap) {
  if (nrows <= 0 || ncols <= 0 || cap < 1 || cap > ncols || nnz_target < nrows || nnz_target > nrows * cap)
    return nullptr;
  Rng rng(seed);
  std::vector<double> cdf(cap);
  double acc = 0.0;
  for (int64_t k = 1; k <= cap; ++k) {
    acc += std::pow((double)k, -s);
    cdf[k - 1] = acc;
  }
  for (auto& v : cdf) v /= acc;
  std::vector<int32_t> Z(nrows);
  for (auto& z : Z) z = 1 + (int32_t)(std::lower_bound(cdf.begin(), cdf.end(), rng.uniform()) - cdf.begin());
  auto total = [&](double lam) {
    int64_t t = 0;
    for (int32_t z : Z) t += std::min<int64_t>(cap, std::max<int64_t>(1, std::llround(lam * z)));
    return t;
  };
  double lo = 0.0, hi = (double)cap;
  for (int it = 0; it < 100; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (total(mid) <= nnz_target) lo = mid; else hi = mid;
  }
  std::vector<int64_t> L(nrows);
  int64_t sum = 0;
  for (int64_t i = 0; i < nrows; ++i) {
    L[i] = std::min<int64_t>(cap, std::max<int64_t>(1, std::llround(lo * Z[i])));
    sum += L[i];
  }
  if (sum < nnz_target) {
    std::vector<int64_t> order(nrows);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return L[a] > L[b]; });
    int64_t rem = nnz_target - sum;
    while (rem > 0) {
      for (int64_t idx = 0; idx < nrows && rem > 0; ++idx) {
        if (L[order[idx]] < cap) {
          ++L[order[idx]];
          --rem;
        }
      }
    }
  }
  hn_gen_spmv* g = static_cast<hn_gen_spmv*>(std::calloc(1, sizeof(hn_gen_spmv)));
  g->nrows = nrows;
  g->ncols = ncols;
  g->nnz = nnz_target;
  g->row_ptr = static_cast<int64_t*>(std::malloc(sizeof(int64_t) * (nrows + 1)));
  g->col = static_cast<int32_t*>(std::malloc(sizeof(int32_t) * nnz_target));
  g->val = static_cast<double*>(std::malloc(sizeof(double) * nnz_target));
  g->x = static_cast<double*>(std::malloc(sizeof(double) * ncols));
  g->w = static_cast<double*>(std::malloc(sizeof(double) * nrows));
  g->row_ptr[0] = 0;
  for (int64_t i = 0; i < nrows; ++i) g->row_ptr[i + 1] = g->row_ptr[i] + L[i];
  std::vector<int32_t> buf;
  for (int64_t i = 0; i < nrows; ++i) {
    const int64_t len = L[i];
    buf.clear();
    while ((int64_t)buf.size() < len) {
      const int64_t need = len - (int64_t)buf.size();
      for (int64_t u = 0; u < need; ++u) buf.push_back((int32_t)rng.below(ncols));
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    }
    std::memcpy(g->col + g->row_ptr[i], buf.data(), sizeof(int32_t) * len);
  }
  for (int64_t k = 0; k < nnz_target; ++k) g->val[k] = rng.normal();
  for (int64_t j = 0; j < ncols; ++j) g->x[j] = rng.uniform(-1.0, 1.0);
  for (int64_t i = 0; i < nrows; ++i) g->w[i] = rng.uniform(-1.0, 1.0);
  return g;
}

extern "C" void hn_gen_spmv_free(hn_gen_spmv* g) {
  if (!g) return;
  std::free(g->row_ptr);
  std::free(g->col);
  std::free(g->val);
  std::free(g->x);
  std::free(g->w);
  std::free(g);
}

extern "C" void hn_gen_uniform(uint64_t seed, double lo, double hi, int64_t n, double* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = rng.uniform(lo, hi);
}
