// oracle_bu.cpp -- CPU restatement of BottomUpHeuristicGPU
// (Treedepth/src/branch_and_bound/heuristics/bottom_up_heuristic.hpp,
// bottom_up_heuristic_gpu_algorithm.{hpp,cu}) and of the union-find it uses
// (Treedepth/src/union_find/ext_array_union_find.hpp).
//
// TEST INFRASTRUCTURE ONLY: used by tests/ as the checker and by bench.py's
// CPU legs.  The product library never links it.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"
#include "oracle_internal.hpp"

namespace oracle {
namespace {

// ext_array_union_find::Find (ext_array_union_find.hpp:10-21): follow the
// parent links to the root (uf[x] < 0), then point the path at the root.
int8_t UfFind(int8_t* uf, int8_t elem) {
  int8_t root = elem;
  while (uf[root] >= 0) root = uf[root];
  while (elem != root) {
    const int8_t prev = uf[elem];
    uf[elem] = root;
    elem = prev;
  }
  return root;
}

// ext_array_union_find::Union (ext_array_union_find.hpp:23-34): s1 becomes
// the parent of s2; a root stores -height; uf[val_ix] keeps the largest
// height seen.
void UfUnion(int8_t* uf, int8_t s1, int8_t s2, int val_ix) {
  if (s1 == s2) return;
  int new_val = -uf[s1];
  if (new_val <= -uf[s2]) new_val = -uf[s2] + 1;
  if (uf[val_ix] < new_val) uf[val_ix] = static_cast<int8_t>(new_val);
  uf[s1] = static_cast<int8_t>(-new_val);
  uf[s2] = s1;
}

uint64_t LevelOffset(int n, int k) {
  uint64_t off = 0;
  for (int j = 0; j < k; ++j) off += NChooseK(n, j);
  return off;
}

}  // namespace

// Run (bottom_up_heuristic_gpu_algorithm.cu:203-259) with the per-subset
// step of DynamicStepKernel (:57-95): the empty set holds uf = -1 everywhere
// and height 1; a k-subset S tries its vertices in ascending order as the
// last one added: it takes the state of S \ v, records v, joins v with the
// roots of its neighbours inside S (ascending), and keeps the first state of
// strictly smallest height (starting from k + 1).  history_[0] is (-1, 0)
// (:232-234); history_[k] holds (last vertex, height) per subset.
int32_t BottomUp(int n, const uint64_t* adj, int iterations, int8_t* last_out,
                 uint8_t* height_out) {
  const int W = n + 2;  // SetPlaceholderSize (:189-192)
  std::vector<int8_t> prev(W, -1), next;
  prev[n + 1] = 1;
  if (last_out) last_out[0] = -1;
  if (height_out) height_out[0] = 0;
  int32_t full = -1;
  std::vector<int32_t> set(n + 1);
  std::vector<int8_t> uf(W);
  for (int k = 1; k < iterations; ++k) {
    const uint64_t cnt = NChooseK(n, k);
    next.assign(cnt * W, 0);
    const uint64_t off = LevelOffset(n, k);
    for (uint64_t code = 0; code < cnt; ++code) {
      Decode(n, k, code, set.data());
      std::sort(set.begin(), set.begin() + k);
      uint64_t smask = 0;
      for (int i = 0; i < k; ++i) smask |= 1ull << set[i];
      int best = k + 1;
      for (int i = 0; i < k; ++i) {
        const uint64_t c2 = EncodeExclude(set.data(), k, i);
        std::memcpy(uf.data(), prev.data() + c2 * W, W);
        const int v = set[i];
        uf[n] = static_cast<int8_t>(v);
        for (uint64_t m = adj[v] & smask; m; m &= m - 1) {
          const int w = __builtin_ctzll(m);
          UfUnion(uf.data(), static_cast<int8_t>(v),
                  UfFind(uf.data(), static_cast<int8_t>(w)), n + 1);
        }
        if (uf[n + 1] < best) {
          best = uf[n + 1];
          std::memcpy(next.data() + code * W, uf.data(), W);
        }
      }
      const int8_t* rec = next.data() + code * W;
      if (last_out) last_out[off + code] = rec[n];
      if (height_out) height_out[off + code] = static_cast<uint8_t>(rec[n + 1]);
      if (k == n) full = rec[n + 1];
    }
    prev.swap(next);
  }
  return full;
}

// The same heuristic stated on components instead of union-find arrays:
// the state of S is a height for each connected component of G[S]; adding v
// to S \ v merges the components v touches into one of height 1 + their
// largest height (1 if none), and the set's height is the largest component
// height (1 for the empty set).  Independent code for cross-checking BottomUp.
int32_t BottomUpComponents(int n, const uint64_t* adj, int iterations,
                           int8_t* last_out, uint8_t* height_out) {
  // per subset: height of each vertex's component (0 outside S)
  std::vector<uint8_t> prev(32, 0), next;
  std::vector<uint8_t> prev_h(1, 1), next_h;
  if (last_out) last_out[0] = -1;
  if (height_out) height_out[0] = 0;
  int32_t full = -1;
  for (int k = 1; k < iterations; ++k) {
    const uint64_t cnt = NChooseK(n, k);
    next.assign(cnt * 32, 0);
    next_h.assign(cnt, 0);
    const uint64_t off = LevelOffset(n, k);
    // walk the k-subsets as increasing bitmasks (colex order)
    uint64_t S = (1ull << k) - 1;
    for (uint64_t code = 0; code < cnt; ++code) {
      int best = k + 1, bv = -1, bnew = 0;
      uint64_t bprev = 0;
      for (uint64_t m = S; m; m &= m - 1) {
        const int v = __builtin_ctzll(m);
        const uint64_t p = EncodeMask(S & ~(1ull << v));
        const uint8_t* hv = prev.data() + p * 32;
        int top = 0;
        for (uint64_t a = adj[v] & S; a; a &= a - 1) top = std::max<int>(top, hv[__builtin_ctzll(a)]);
        const int newh = top + 1;
        const int h = std::max<int>(prev_h[p], newh);
        if (h < best) {
          best = h;
          bv = v;
          bnew = newh;
          bprev = p;
        }
      }
      const uint64_t comp = ComponentOf(adj, S, bv);
      uint8_t* out = next.data() + code * 32;
      for (int w = 0; w < n; ++w)
        if (S >> w & 1) out[w] = (comp >> w & 1) ? static_cast<uint8_t>(bnew)
                                                  : prev[bprev * 32 + w];
      next_h[code] = static_cast<uint8_t>(best);
      if (last_out) last_out[off + code] = static_cast<int8_t>(bv);
      if (height_out) height_out[off + code] = static_cast<uint8_t>(best);
      if (k == n) full = best;
      // next k-subset in colex order (Gosper)
      const uint64_t c = S & (0 - S), r = S + c;
      S = r | (((S ^ r) >> 2) / c);
    }
    prev.swap(next);
    prev_h.swap(next_h);
  }
  return full;
}

}  // namespace oracle

extern "C" {

int32_t or_bottom_up(int32_t n, const uint64_t* adj, int32_t iterations,
                     int8_t* last_out, uint8_t* height_out, int32_t* order,
                     int32_t* parent, int32_t* root) {
  if (n < 1 || n > 32 || iterations < 1 || iterations > n + 1) return -2;
  const bool full_run = iterations == n + 1;
  std::vector<int8_t> last(last_out || order ? (1ull << n) : 0);
  int8_t* lp = last_out ? last_out : (last.empty() ? nullptr : last.data());
  const int32_t td = oracle::BottomUp(n, adj, iterations, lp, height_out);
  if (!full_run) return -1;  // UINT_MAX in BottomUpHeuristicGPU::Get (:24-31)
  if (order) {
    // GetElimination(n, n, 0) (bottom_up_heuristic_gpu_algorithm.hpp:92-108):
    // the stored last vertex of the current set, removed, n times
    uint64_t S = oracle::FullMask(n);
    for (int i = 0; i < n; ++i) {
      const int k = __builtin_popcountll(S);
      uint64_t off = 0;
      for (int j = 0; j < k; ++j) off += oracle::NChooseK(n, j);
      const int v = lp[off + oracle::EncodeMask(S)];
      order[i] = v;
      S &= ~(1ull << v);
    }
    if (parent) {
      // BottomUpHeuristicGPU::Get: EliminationTree::Eliminate in that order,
      // then Decompose (bottom_up_heuristic.hpp:17-31)
      oracle::ElimTree t(n, adj);
      for (int i = 0; i < n; ++i) t.Eliminate(order[i]);
      const auto res = t.Decompose();
      for (int v = 0; v < n; ++v) parent[v] = res.parent[v];
      if (root) *root = res.root;
    }
  }
  return td;
}

int32_t or_bottom_up_components(int32_t n, const uint64_t* adj, int32_t iterations,
                                int8_t* last_out, uint8_t* height_out) {
  if (n < 1 || n > 32 || iterations < 1 || iterations > n + 1) return -2;
  return oracle::BottomUpComponents(n, adj, iterations, last_out, height_out);
}

}  // extern "C"
