// oracle_core.cpp -- CPU restatement of the GiSBDW colex codec,
// EliminationTree, DynamicCPU, DynamicCPUImprov and the table DP.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and bench.py's
// cpu_baseline leg as the checker.  The product library never links it.
//
// Citations are to the reference tree (Treedepth/...).
#include <omp.h>

#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <unordered_map>
#include <utility>
#include <vector>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace oracle {

// ---------------------------------------------------------------------------
// colex codec
// ---------------------------------------------------------------------------
// NChooseK: multiplicative loop with a division per step, exactly as
// set_encoder.cu:6-17 (k mirrored to n-k when k > n/2).
uint64_t NChooseK(uint64_t n, uint64_t k) {
  if (k > n) return 0;
  if (k > n / 2) k = n - k;
  uint64_t ret = 1;
  for (uint64_t i = 1; i <= k; ++i, --n) {
    ret *= n;
    ret /= i;
  }
  return ret;
}

// Encode(sorted_set, size): sum_i C(s_i, i+1)   (set_encoder.hpp:40-46)
uint64_t EncodeSorted(const int32_t* s, int32_t k) {
  uint64_t r = 0;
  for (int32_t i = 0; i < k; ++i) r += NChooseK(s[i], i + 1);
  return r;
}

// Encode(sorted_set, size, exclude): rank of S minus its exclude-th element
// (set_encoder.hpp:48-59).
uint64_t EncodeExclude(const int32_t* s, int32_t k, int32_t ex) {
  uint64_t r = 0;
  int32_t i = -1;
  while (++i < ex) r += NChooseK(s[i], i + 1);
  while (++i < k) r += NChooseK(s[i], i);
  return r;
}

// Decode: greedy from the top (set_encoder.hpp:61-73).
void Decode(int32_t nverts, int32_t k, uint64_t code, int32_t* out) {
  while (k > 0) {
    uint64_t nk = NChooseK(--nverts, k);
    if (code >= nk) {
      out[--k] = nverts;
      code -= nk;
    }
  }
}

uint64_t EncodeMask(uint64_t mask) {
  uint64_t r = 0;
  int i = 0;
  while (mask) {
    int v = __builtin_ctzll(mask);
    mask &= mask - 1;
    r += NChooseK(v, ++i);
  }
  return r;
}

// ---------------------------------------------------------------------------
// graph helpers
// ---------------------------------------------------------------------------
uint64_t ComponentOf(const uint64_t* adj, uint64_t set, int v) {
  uint64_t comp = 1ull << v, frontier = comp;
  while (frontier) {
    int u = __builtin_ctzll(frontier);
    frontier &= frontier - 1;
    uint64_t nb = adj[u] & set & ~comp;
    comp |= nb;
    frontier |= nb;
  }
  return comp;
}

std::vector<uint64_t> Components(const uint64_t* adj, uint64_t set) {
  std::vector<uint64_t> out;
  while (set) {
    uint64_t c = ComponentOf(adj, set, __builtin_ctzll(set));
    out.push_back(c);
    set &= ~c;
  }
  return out;
}

int EdgesIn(const uint64_t* adj, uint64_t set) {
  int e2 = 0;
  for (uint64_t s = set; s; s &= s - 1)
    e2 += __builtin_popcountll(adj[__builtin_ctzll(s)] & set);
  return e2 / 2;
}

uint64_t FullMask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }

// ---------------------------------------------------------------------------
// EliminationTree restatement (elimination_tree.hpp:20-215,
// elimination_tree.cpp:24-108).  A component is its vertex set (the induced
// adjacency follows from the graph), its depth and its creator (the vertex
// whose elimination produced it, -1 for the initial component).  Components
// are ordered by their lowest vertex (ComponentCmp, elimination_tree.hpp:67).
// ---------------------------------------------------------------------------
ElimTree::ElimTree(int n, const uint64_t* adj) : n_(n), adj_(adj, adj + n) {
  node_depth_.assign(n, -1);
  node_parent_.assign(n, -1);
  uint64_t all = FullMask(n);
  // The reference requires a connected input; the restatement also accepts a
  // disconnected one (one initial component each) for the heuristics' use.
  for (uint64_t c : Components(adj, all)) comps_[Low(c)] = Comp{c, 0, -1};
}

std::vector<uint64_t> ElimTree::Eliminate(int v) {
  // find the component holding v
  auto it = comps_.begin();
  for (; it != comps_.end(); ++it)
    if (it->second.mask >> v & 1) break;
  Comp c = it->second;
  comps_.erase(it);
  node_depth_[v] = c.depth;
  node_parent_[v] = c.creator;
  order_.push_back(v);
  stack_.push_back({v, c});
  // split the rest into components, lowest vertex first (elimination_tree.cpp
  // :38-67 extracts begin() and grows it by BFS)
  std::vector<uint64_t> made;
  uint64_t rest = c.mask & ~(1ull << v);
  while (rest) {
    uint64_t nc = ComponentOf(adj_.data(), rest, __builtin_ctzll(rest));
    rest &= ~nc;
    comps_[Low(nc)] = Comp{nc, c.depth + 1, v};
    made.push_back(nc);
  }
  return made;
}

int ElimTree::Merge() {
  auto [v, c] = stack_.back();
  stack_.pop_back();
  order_.pop_back();
  // remove the components v's elimination produced (all still uneliminated
  // under LIFO use, elimination_tree.cpp:80-92) and restore c
  uint64_t rest = c.mask & ~(1ull << v);
  for (auto it = comps_.begin(); it != comps_.end();) {
    if (it->second.mask & rest)
      it = comps_.erase(it);
    else
      ++it;
  }
  comps_[Low(c.mask)] = c;
  node_depth_[v] = -1;
  node_parent_[v] = -1;
  return v;
}

// Decompose (elimination_tree.hpp:193-215): tree edges {v, parent}, the
// treedepth as the number of BFS levels, the root = first eliminated vertex.
ElimTree::Result ElimTree::Decompose() const {
  Result r;
  r.parent = node_parent_;
  r.treedepth = 0;
  for (int d : node_depth_) r.treedepth = std::max(r.treedepth, d + 1);
  r.root = order_.empty() ? -1 : order_.front();
  return r;
}

// ---------------------------------------------------------------------------
// DynamicCPU restatement (dynamic_cpu.cpp:14-107, dynamic_cpu.hpp:46-81)
// ---------------------------------------------------------------------------
struct DynCPU {
  int n;
  const uint64_t* adj;
  // history_[i] : code -> (treedepth, root), connected i-subsets only
  std::vector<std::map<uint64_t, std::pair<uint64_t, int>>> history;

  // SearchGraphOnVertices: DFS restricted to `verts`, ascending neighbours
  // (mapS) -- dynamic_cpu.cpp:14-24.  Returns the visited set.
  void Search(const std::set<int>& verts, std::set<int>* visited, int cur,
              std::vector<bool>* total) {
    if (!verts.count(cur) || visited->count(cur)) return;
    visited->insert(cur);
    if (total) (*total)[cur] = true;
    for (uint64_t m = adj[cur]; m; m &= m - 1)
      Search(verts, visited, __builtin_ctzll(m), total);
  }

  static uint64_t EncodeSet(const std::set<int>& s) {
    uint64_t r = 0;
    uint64_t i = 0;
    for (int v : s) r += NChooseK(v, ++i);
    return r;
  }

  void Run() {
    history.clear();
    history.resize(n + 1);
    history[0][0] = {0, -1};
    for (int i = 1; i <= n; ++i) {
      uint64_t nk = NChooseK(n, i);
      std::vector<int32_t> dec(i);
      for (uint64_t code = 0; code < nk; ++code) {
        Decode(n, i, code, dec.data());
        std::set<int> set(dec.begin(), dec.end());
        std::set<int> visited;
        Search(set, &visited, *set.begin(), nullptr);
        if (visited.size() != set.size()) continue;  // :53-64
        uint64_t treedepth = UINT64_MAX;
        int root = -1;
        for (auto it = set.begin(); it != set.end(); ++it) {
          int v = *it;
          it = set.erase(it);
          uint64_t comp_td = 0;
          std::vector<bool> total(n, false);
          total[v] = true;
          for (uint64_t m = adj[v]; m; m &= m - 1) {
            int nb = __builtin_ctzll(m);
            if (total[nb]) continue;
            visited.clear();
            Search(set, &visited, nb, &total);
            comp_td = std::max(
                comp_td, history[visited.size()][EncodeSet(visited)].first);
          }
          if (comp_td + 1 < treedepth) {  // strict <, ascending v (:96)
            treedepth = comp_td + 1;
            root = v;
          }
          it = set.insert(it, v);
        }
        history[i][code] = {treedepth, root};
      }
    }
  }

  // GetTDDecomp(code = 0, g) on the whole graph (dynamic_cpu.hpp:60-81)
  uint32_t Decomp(int32_t* parent, int32_t* root) {
    auto& top = history[n];
    if (top.find(0) == top.end()) return UINT32_MAX;
    ElimTree et(n, adj);
    while (et.HasComponents()) {
      uint64_t m = et.FirstComponent().mask;
      auto& lvl = history[__builtin_popcountll(m)];
      et.Eliminate(lvl.find(EncodeMask(m))->second.second);
    }
    auto r = et.Decompose();
    for (int v = 0; v < n; ++v) parent[v] = r.parent[v];
    *root = r.root;
    return r.treedepth;
  }
};

// ---------------------------------------------------------------------------
// DynamicCPUImprov restatement (dynamic_cpu_improv.cpp:13-65,
// dynamic_cpu_improv.hpp:53-85); bitmask keys (n <= 64 instead of 128).
// ---------------------------------------------------------------------------
struct DynImprov {
  int n;
  const uint64_t* adj;
  ElimTree* et;
  std::vector<std::unordered_map<uint64_t, std::pair<uint64_t, int>>> history;

  uint64_t Run(uint64_t comp_mask) {
    int sz = __builtin_popcountll(comp_mask);
    auto found = history[sz].find(comp_mask);
    if (found != history[sz].end()) return found->second.first;
    history[sz][comp_mask] = {UINT64_MAX, -1};
    std::pair<uint64_t, int> p = {UINT64_MAX, -1};
    for (uint64_t m = comp_mask; m; m &= m - 1) {  // ascending map order
      int v = __builtin_ctzll(m);
      auto comps = et->Eliminate(v);
      uint64_t tdcomp = 0;
      for (uint64_t c : comps) tdcomp = std::max(tdcomp, Run(c));
      ++tdcomp;
      if (p.first > tdcomp) p = {tdcomp, v};  // strict >, :38
      et->Merge();
    }
    history[sz][comp_mask] = p;
    return p.first;
  }
};

// ---------------------------------------------------------------------------
// fast flat-table DP: level tables indexed by colex rank, uint8 values.
// Recurrence of dynamic_gpu.cu:40-113: connected S -> 1 + min_v td(S\v);
// disconnected S -> max over components, each looked up at its own level.
// ---------------------------------------------------------------------------
struct Pascal {
  uint64_t c[65][65];
  Pascal() {
    for (int i = 0; i <= 64; ++i)
      for (int j = 0; j <= 64; ++j) c[i][j] = 0;
    for (int i = 0; i <= 64; ++i) {
      c[i][0] = 1;
      for (int j = 1; j <= i; ++j) c[i][j] = c[i - 1][j - 1] + c[i - 1][j];
    }
  }
};
static const Pascal kPascal;

static inline uint64_t RankMask(uint64_t m) {
  uint64_t r = 0;
  int i = 0;
  while (m) {
    int v = __builtin_ctzll(m);
    m &= m - 1;
    r += kPascal.c[v][++i];
  }
  return r;
}

static inline uint64_t UnrankMask(int n, int k, uint64_t code) {
  uint64_t m = 0;
  int v = n;
  while (k > 0) {
    uint64_t nk = kPascal.c[--v][k];
    if (code >= nk) {
      m |= 1ull << v;
      code -= nk;
      --k;
    }
  }
  return m;
}

static inline uint64_t Gosper(uint64_t x) {
  uint64_t c = x & (0 - x);
  uint64_t r = x + c;
  return (((r ^ x) >> 2) / c) | r;
}

struct FastDP {
  int n;
  const uint64_t* adj;
  std::vector<std::vector<uint8_t>> lv;

  uint8_t Td(uint64_t m) const {
    return lv[__builtin_popcountll(m)][RankMask(m)];
  }

  void Run(int threads) {
    lv.assign(n + 1, {});
    for (int k = 0; k <= n; ++k) lv[k].assign(kPascal.c[n][k], 0);
    lv[0][0] = 0;
    for (int k = 1; k <= n; ++k) {
      int64_t total = (int64_t)kPascal.c[n][k];
      const int64_t chunk = 4096;
      int64_t nchunks = (total + chunk - 1) / chunk;
      uint8_t* out = lv[k].data();
      const uint8_t* prev = lv[k - 1].data();
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
      for (int64_t ci = 0; ci < nchunks; ++ci) {
        int64_t r0 = ci * chunk, r1 = std::min(total, r0 + chunk);
        uint64_t s = UnrankMask(n, k, r0);
        for (int64_t r = r0; r < r1; ++r) {
          uint64_t first = ComponentOf(adj, s, __builtin_ctzll(s));
          uint8_t td;
          if (first == s) {
            unsigned best = 255;
            // rank(S\v) for every v, by prefix/suffix sums
            uint64_t rs = 0;
            int idx = 0;
            uint64_t pre[64], suf_shift[65];
            int el[64];
            for (uint64_t m = s; m; m &= m - 1) el[idx++] = __builtin_ctzll(m);
            suf_shift[k] = 0;
            for (int i = k - 1; i >= 0; --i)
              suf_shift[i] = suf_shift[i + 1] + kPascal.c[el[i]][i];
            for (int i = 0; i < k; ++i) {
              pre[i] = rs;
              rs += kPascal.c[el[i]][i + 1];
            }
            for (int i = 0; i < k; ++i) {
              unsigned t = prev[pre[i] + suf_shift[i + 1]];
              if (t < best) best = t;
            }
            td = (uint8_t)(best + 1);
          } else {
            unsigned mx = 0;
            uint64_t rest = s;
            while (rest) {
              uint64_t c = ComponentOf(adj, rest, __builtin_ctzll(rest));
              rest &= ~c;
              unsigned t = lv[__builtin_popcountll(c)][RankMask(c)];
              if (t > mx) mx = t;
            }
            td = (uint8_t)mx;
          }
          out[r] = td;
          if (r + 1 < r1) s = Gosper(s);
        }
      }
    }
  }

  // canonical tree: root(C) = min v with td(C\v) = td(C)-1; children are
  // the components of C\root.  Roots of the input's components get -1.
  void Tree(int8_t* parent) const {
    std::vector<std::pair<uint64_t, int>> stack;
    for (uint64_t c : Components(adj, FullMask(n))) stack.push_back({c, -1});
    while (!stack.empty()) {
      auto [c, par] = stack.back();
      stack.pop_back();
      unsigned tdc = Td(c);
      int root = -1;
      for (uint64_t m = c; m; m &= m - 1) {
        int v = __builtin_ctzll(m);
        if (Td(c & ~(1ull << v)) + 1u == tdc) {
          root = v;
          break;
        }
      }
      parent[root] = (int8_t)par;
      for (uint64_t cc : Components(adj, c & ~(1ull << root)))
        stack.push_back({cc, root});
    }
  }
};

}  // namespace oracle

using namespace oracle;  // NOLINT

extern "C" {

uint64_t or_nchoosek(uint64_t n, uint64_t k) { return NChooseK(n, k); }
uint64_t or_encode(const int32_t* s, int32_t k) { return EncodeSorted(s, k); }
uint64_t or_encode_exclude(const int32_t* s, int32_t k, int32_t e) {
  return EncodeExclude(s, k, e);
}
void or_decode(int32_t n, int32_t k, uint64_t code, int32_t* out) {
  Decode(n, k, code, out);
}

int or_elimtree_replay(int32_t n, const uint64_t* adj, const int32_t* order,
                       uint64_t* comp_masks, int32_t* comp_depths,
                       int32_t* comp_counts, int32_t* parent,
                       int32_t* treedepth, int32_t* root) {
  ElimTree et(n, adj);
  auto dump = [&](int step) {
    int i = 0;
    for (auto& c : et.AllComponents()) {
      comp_masks[step * n + i] = c.mask;
      comp_depths[step * n + i] = c.depth;
      ++i;
    }
    comp_counts[step] = i;
  };
  dump(0);
  for (int i = 0; i < n; ++i) {
    et.Eliminate(order[i]);
    dump(i + 1);
  }
  auto r = et.Decompose();
  for (int v = 0; v < n; ++v) parent[v] = r.parent[v];
  *treedepth = r.treedepth;
  *root = r.root;
  return 0;
}

int or_elimtree_merge_check(int32_t n, const uint64_t* adj,
                            const int32_t* order) {
  ElimTree et(n, adj);
  std::vector<std::vector<std::pair<uint64_t, int>>> down;
  auto snap = [&]() {
    std::vector<std::pair<uint64_t, int>> s;
    for (auto& c : et.AllComponents()) s.push_back({c.mask, c.depth});
    return s;
  };
  down.push_back(snap());
  for (int i = 0; i < n; ++i) {
    et.Eliminate(order[i]);
    down.push_back(snap());
  }
  for (int i = n; i > 0; --i) {
    if (snap() != down[i]) return 1;
    if (et.Merge() != order[i - 1]) return 2;
  }
  return snap() == down[0] ? 0 : 3;
}

uint32_t or_dyncpu(int32_t n, const uint64_t* adj, int32_t* parent,
                   int32_t* root) {
  DynCPU d{n, adj, {}};
  d.Run();
  return d.Decomp(parent, root);
}

uint64_t or_dyncpu_subset_td(int32_t n, const uint64_t* adj, int32_t k,
                             uint64_t code) {
  DynCPU d{n, adj, {}};
  d.Run();
  auto it = d.history[k].find(code);
  return it == d.history[k].end() ? UINT64_MAX : it->second.first;
}

uint32_t or_dyncpu_improv(int32_t n, const uint64_t* adj, int32_t* parent,
                          int32_t* root) {
  if (Components(adj, FullMask(n)).size() != 1) return UINT32_MAX;
  ElimTree et(n, adj);
  DynImprov d{n, adj, &et, {}};
  d.history.resize(n + 1);
  d.history[0][0] = {0, -1};
  for (int i = 0; i < n; ++i) d.history[1][1ull << i] = {1, i};
  d.Run(FullMask(n));
  // GetTDDecompImpl (dynamic_cpu_improv.cpp:48-65)
  ElimTree out(n, adj);
  while (out.HasComponents()) {
    uint64_t m = out.FirstComponent().mask;
    out.Eliminate(d.history[__builtin_popcountll(m)][m].second);
  }
  auto r = out.Decompose();
  for (int v = 0; v < n; ++v) parent[v] = r.parent[v];
  *root = r.root;
  return r.treedepth;
}

int32_t or_dp_fast(int32_t n, const uint64_t* adj, uint8_t* levels_out,
                   int8_t* parent, int32_t threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  FastDP dp{n, adj, {}};
  dp.Run(threads);
  if (levels_out) {
    size_t off = 0;
    for (int k = 0; k <= n; ++k) {
      std::memcpy(levels_out + off, dp.lv[k].data(), dp.lv[k].size());
      off += dp.lv[k].size();
    }
  }
  if (parent) dp.Tree(parent);
  return dp.lv[n][0];
}

int or_certify(int32_t n, const uint64_t* adj, const int8_t* parent,
               int32_t td) {
  std::vector<int> depth(n, 0);
  for (int v = 0; v < n; ++v) {
    if (parent[v] < -1 || parent[v] >= n || parent[v] == v) return 0;
    int d = 1, u = v;
    while (parent[u] != -1) {
      u = parent[u];
      if (++d > n) return 0;  // cycle
    }
    depth[v] = d;
  }
  auto is_anc = [&](int a, int b) {  // a ancestor-or-self of b
    while (b != -1) {
      if (b == a) return true;
      b = parent[b];
    }
    return false;
  };
  for (int v = 0; v < n; ++v)
    for (uint64_t m = adj[v]; m; m &= m - 1) {
      int u = __builtin_ctzll(m);
      if (!is_anc(u, v) && !is_anc(v, u)) return 0;
    }
  int mx = 0;
  for (int d : depth) mx = std::max(mx, d);
  return mx == td ? 1 : 0;
}

// Brute force by the definition: minimum height over every parent array
// (rooted forest) whose ancestor/descendant closure contains every edge.
int32_t or_brute_td(int32_t n, const uint64_t* adj) {
  if (n > 8) return -1;
  std::vector<int> par(n, -1);
  int best = n + 1;
  // enumerate parent choice p[v] in {-1, 0..n-1}\{v}
  std::vector<int> choice(n, 0);
  uint64_t total = 1;
  for (int i = 0; i < n; ++i) total *= (uint64_t)n;  // n options each
  for (uint64_t code = 0; code < total; ++code) {
    uint64_t c = code;
    bool ok = true;
    for (int v = 0; v < n; ++v) {
      int o = (int)(c % n);
      c /= n;
      // option o: o == v means root (-1), otherwise parent o
      par[v] = (o == v) ? -1 : o;
    }
    // acyclic + heights
    int h = 0;
    for (int v = 0; v < n && ok; ++v) {
      int d = 1, u = v;
      while (par[u] != -1) {
        u = par[u];
        if (++d > n) {
          ok = false;
          break;
        }
      }
      h = std::max(h, d);
    }
    if (!ok || h >= best) continue;
    for (int v = 0; v < n && ok; ++v)
      for (uint64_t m = adj[v]; m && ok; m &= m - 1) {
        int u = __builtin_ctzll(m);
        bool anc = false;
        for (int b = v; b != -1; b = par[b])
          if (b == u) anc = true;
        for (int b = u; b != -1 && !anc; b = par[b])
          if (b == v) anc = true;
        if (!anc) ok = false;
      }
    if (ok) best = h;
  }
  return best;
}

}  // extern "C"
