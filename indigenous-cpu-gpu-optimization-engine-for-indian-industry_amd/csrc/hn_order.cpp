// Exact minimum-degree ordering and the symbolic LDL' factor (host).
//
// The reference orders the normal equations by exact minimum degree, the
// lowest vertex index breaking ties (compute_md_order_and_pattern,
// core/numerical/factorization.cpp:103-191): at each step the live vertex of
// least degree in the elimination graph is eliminated and its live
// neighbours become a clique.  The rule here is the same, so the permutation
// and the factor pattern are the same; the data structures are chosen so
// that it scales:
//   * a vertex keeps its live neighbours only, as a sorted list.  Eliminating
//     v replaces each live neighbour a's list by the sorted union of its own
//     list (v removed) and v's list (a removed), so a live vertex's degree is
//     its list's length.  The reference keeps dead vertices in its lists and
//     counts the same degree with a running counter.
//   * the pattern of row r of L is the set of steps that eliminated a
//     neighbour of the vertex eliminated at step r; each elimination appends
//     its step to its live neighbours' sets, so the sets come out sorted.
//     The reference reads the same set back from its lists.
//   * the next vertex, least (degree, index), comes from a segment tree over
//     the vertices instead of a scan of all of them.
#include "hn_sldl.hpp"

#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

namespace hn {

int sldl_symbolic(int64_t n, const int64_t* rp, const int64_t* ci, double work_limit, double max_lnnz,
                  SldlSymbolic& S) {
  S = SldlSymbolic();
  S.n = n;
  HN_CHECK(n >= 0 && n < INT32_MAX, HN_ERR_UNSUPPORTED, "sparse LDL': dimension must be below 2^31");
  HN_CHECK(n == 0 || (rp && (rp[n] == 0 || ci)), HN_ERR_INVALID, "sparse LDL': null pattern");
  S.perm.assign(n, n);
  S.inv_perm.assign(n, n);
  S.Lrp.assign(n + 1, 0);
  S.parent.assign(n, n);
  S.postorder.assign(n, 0);
  S.level_ptr.assign(1, 0);
  S.Ucp.assign(n + 1, 0);
  if (n == 0) return HN_OK;

  // elimination graph: the symmetric closure of the stored pattern without
  // self loops (build_elimination_graph, :71-97); columns outside [0, n)
  // are ignored as there
  std::vector<std::vector<int32_t>> adj(n);
  for (int64_t i = 0; i < n; ++i) {
    HN_CHECK(rp[i + 1] >= rp[i], HN_ERR_INVALID, "sparse LDL': row pointer decreases");
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int64_t j = ci[k];
      if (j < 0 || j >= n || j == i) continue;
      adj[i].push_back((int32_t)j);
      adj[j].push_back((int32_t)i);
    }
  }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }

  int64_t P = 1;
  while (P < n) P <<= 1;
  std::vector<uint64_t> tree(2 * P, UINT64_MAX);
  auto key = [&](int64_t v) { return ((uint64_t)adj[v].size() << 32) | (uint64_t)v; };
  auto update = [&](int64_t v, uint64_t k) {
    int64_t t = P + v;
    tree[t] = k;
    for (t >>= 1; t > 0; t >>= 1) tree[t] = std::min(tree[2 * t], tree[2 * t + 1]);
  };
  for (int64_t v = 0; v < n; ++v) tree[P + v] = key(v);
  for (int64_t t = P - 1; t > 0; --t) tree[t] = std::min(tree[2 * t], tree[2 * t + 1]);

  std::vector<std::vector<int32_t>> dead(n);  // steps that eliminated a neighbour
  std::vector<int32_t> merged;
  S.Lci.reserve((size_t)std::min<int64_t>(4 * n, INT32_MAX));
  double work = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t best = (int64_t)(tree[1] & 0xffffffffull);
    S.perm[r] = best;
    S.inv_perm[best] = r;
    for (const int32_t st : dead[best]) S.Lci.push_back(st);
    S.Lci.push_back((int32_t)r);
    S.Lrp[r + 1] = (int64_t)S.Lci.size();
    std::vector<int32_t>().swap(dead[best]);
    if (max_lnnz > 0.0 && (double)S.Lci.size() > max_lnnz) {
      S.aborted = true;
      break;
    }
    std::vector<int32_t> nb;
    nb.swap(adj[best]);
    update(best, UINT64_MAX);
    for (const int32_t a : nb) {
      const std::vector<int32_t>& la = adj[a];
      merged.clear();
      merged.reserve(la.size() + nb.size());
      size_t p = 0, q = 0;
      while (p < la.size() || q < nb.size()) {
        int32_t v;
        if (q == nb.size() || (p < la.size() && la[p] < nb[q])) {
          v = la[p++];
        } else if (p == la.size() || nb[q] < la[p]) {
          v = nb[q++];
        } else {
          v = la[p++];
          ++q;
        }
        if (v != best && v != a) merged.push_back(v);
      }
      work += (double)(la.size() + nb.size());
      adj[a].swap(merged);
      dead[a].push_back((int32_t)r);
      update(a, key(a));
    }
    if (work_limit > 0.0 && work > work_limit) {
      S.aborted = true;
      break;
    }
  }
  S.work = work;
  if (S.aborted) return HN_OK;

  // elimination tree and postorder (compute_elimination_tree and
  // compute_postorder, :1327-1362): parent[j] is the first row below j with
  // an entry in column j; the postorder is the reference's depth-first
  // search over the row patterns, without recursion
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = S.Lrp[i]; k < S.Lrp[i + 1]; ++k) {
      const int64_t j = S.Lci[k];
      if (j < i && S.parent[j] == n) S.parent[j] = i;
    }
  {
    std::vector<char> visited(n, 0);
    std::vector<std::pair<int64_t, int64_t>> stack;  // (vertex, next entry of its row)
    int64_t idx = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (visited[i]) continue;
      visited[i] = 1;
      stack.emplace_back(i, S.Lrp[i]);
      while (!stack.empty()) {
        const int64_t u = stack.back().first;
        int64_t& k = stack.back().second;
        int64_t next = -1;
        for (; k < S.Lrp[u + 1]; ++k) {
          const int64_t v = S.Lci[k];
          if (v < u && !visited[v]) {
            next = v;
            ++k;
            break;
          }
        }
        if (next >= 0) {
          visited[next] = 1;
          stack.emplace_back(next, S.Lrp[next]);
        } else {
          S.postorder[idx++] = u;
          stack.pop_back();
        }
      }
    }
  }

  // levels
  std::vector<int32_t> level(n, 0);
  int32_t top = 0;
  for (int64_t i = 0; i < n; ++i) {
    int32_t lv = 0;
    for (int64_t k = S.Lrp[i]; k < S.Lrp[i + 1]; ++k) {
      const int64_t j = S.Lci[k];
      if (j < i) lv = std::max(lv, level[j] + 1);
    }
    level[i] = lv;
    top = std::max(top, lv);
  }
  {
    std::vector<double> lmax((size_t)top + 1, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      double st = 0.0;
      for (int64_t k = S.Lrp[i]; k < S.Lrp[i + 1]; ++k) {
        const int64_t j = S.Lci[k];
        st += (double)(S.Lrp[j + 1] - S.Lrp[j]) + (double)(k - S.Lrp[i]);
      }
      lmax[level[i]] = std::max(lmax[level[i]], st);
    }
    for (const double v : lmax) S.crit_steps += v;
  }
  S.level_ptr.assign((size_t)top + 2, 0);
  for (int64_t i = 0; i < n; ++i) ++S.level_ptr[level[i] + 1];
  for (int32_t l = 0; l <= top; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
  S.level_rows.assign(n, 0);
  {
    std::vector<int64_t> cur(S.level_ptr.begin(), S.level_ptr.end() - 1);
    for (int64_t i = 0; i < n; ++i) S.level_rows[cur[level[i]]++] = (int32_t)i;
  }

  // strict lower part by column, rows descending
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = S.Lrp[i]; k < S.Lrp[i + 1]; ++k)
      if (S.Lci[k] < i) ++S.Ucp[S.Lci[k] + 1];
  for (int64_t j = 0; j < n; ++j) S.Ucp[j + 1] += S.Ucp[j];
  S.Uri.assign(S.Ucp[n], 0);
  S.Upos.assign(S.Ucp[n], 0);
  {
    std::vector<int64_t> cur(S.Ucp.begin(), S.Ucp.end() - 1);
    for (int64_t i = n - 1; i >= 0; --i)
      for (int64_t k = S.Lrp[i]; k < S.Lrp[i + 1]; ++k) {
        const int64_t j = S.Lci[k];
        if (j < i) {
          S.Uri[cur[j]] = (int32_t)i;
          S.Upos[cur[j]] = k;
          ++cur[j];
        }
      }
  }
  return HN_OK;
}

}  // namespace hn
