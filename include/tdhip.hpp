// tdhip.hpp -- C++ host adapters over the C ABI of libtdhip.so, with the
// method names of the reference's classes so callers of
//   td::DynamicGPU  (Treedepth/src/dynamic_gpu/dynamic_gpu.hpp:15-64)
//   td::BnBGPU      (Treedepth/src/bnb_gpu/bnb_gpu.hpp:8-31)
//   td::EliminationTree::Result (Treedepth/src/elimination_tree/
//                                elimination_tree.hpp:26-30)
// switch by changing the graph type they pass.  Boost is not a dependency:
// a graph is td::Graph (n + bitmask rows), with AddEdge like boost::add_edge.
// Errors become td::Error exceptions on this side of the boundary only.
#ifndef TDHIP_HPP_
#define TDHIP_HPP_

#include <cstdint>
#include <limits>
#include <list>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "tdhip.h"

namespace td {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Undirected simple graph on vertices 0..n-1 (n <= 64).
struct Graph {
  int n = 0;
  std::vector<uint64_t> adj;
  explicit Graph(int nverts = 0) : n(nverts), adj(nverts, 0) {}
  void AddEdge(int a, int b) {
    if (a == b) throw Error(TD_E_INVALID, "Self loops are not allowed");
    adj.at(a) |= 1ull << b;
    adj.at(b) |= 1ull << a;
  }
  std::size_t NumEdges() const {
    std::size_t e = 0;
    for (uint64_t r : adj) e += static_cast<std::size_t>(__builtin_popcountll(r));
    return e / 2;
  }
};

// A decomposition as the reference returns it: tree edges, treedepth, root.
// parent[v] = -1 for a root (several roots for a disconnected input).
struct Result {
  std::vector<std::pair<int, int>> td_decomp;
  unsigned treedepth = 0;
  std::size_t root = 0;
  std::vector<int> parent;
  static Result FromParent(const std::vector<int>& parent, unsigned treedepth) {
    Result r;
    r.parent = parent;
    r.treedepth = treedepth;
    for (std::size_t v = 0; v < parent.size(); ++v) {
      if (parent[v] >= 0) r.td_decomp.emplace_back(static_cast<int>(v), parent[v]);
    }
    std::size_t u = 0;
    while (!parent.empty() && parent[u] >= 0) u = static_cast<std::size_t>(parent[u]);
    r.root = u;
    return r;
  }
};

// Owns a td_ctx (one per host thread).
class Context {
 public:
  explicit Context(int device = 0, uint64_t mem_budget_bytes = 0) {
    td_ctx* c = nullptr;
    const int rc = td_ctx_create(&c, device, mem_budget_bytes);
    if (rc != TD_OK) throw Error(rc, "td_ctx_create failed");
    ctx_.reset(c);
  }
  td_ctx* get() const { return ctx_.get(); }
  void Check(int rc) const {
    if (rc != TD_OK) throw Error(rc, td_last_error(ctx_.get()));
  }
  void SetParam(const char* key, int64_t v) { Check(td_ctx_set_param(get(), key, v)); }

 private:
  struct Del {
    void operator()(td_ctx* c) const { td_ctx_destroy(c); }
  };
  std::unique_ptr<td_ctx, Del> ctx_;
};

// td::DynamicGPU: operator()(g) runs the subset DP; the level tables stay on
// the device for GetElimination / GetTreedepth.
class DynamicGPU {
 public:
  explicit DynamicGPU(std::shared_ptr<Context> ctx = std::make_shared<Context>())
      : ctx_(std::move(ctx)) {}

  void operator()(const Graph& g) {
    uint8_t td = 0;
    std::vector<int8_t> par(static_cast<std::size_t>(g.n));
    ctx_->Check(td_dp_solve(ctx_->get(), g.n, g.adj.data(), &td, par.data(), &stats_));
    nverts_ = g.n;
    treedepth_ = td;
    parent_.assign(par.begin(), par.end());
  }

  template <typename VertexType = std::size_t>
  std::list<VertexType> GetElimination(std::size_t nverts, std::size_t subset_size,
                                       std::size_t subset_code) {
    std::vector<int32_t> order(subset_size ? subset_size : 1);
    ctx_->Check(td_dp_get_elimination(ctx_->get(), static_cast<int32_t>(nverts),
                                      static_cast<int32_t>(subset_size), subset_code,
                                      order.data()));
    std::list<VertexType> out;
    for (std::size_t i = 0; i < subset_size; ++i)
      out.push_back(static_cast<VertexType>(order[i]));
    return out;
  }

  // DynamicCPU::GetTreedepth counterpart (any subset, colex code).
  std::size_t GetTreedepth(std::size_t /*nverts*/, std::size_t subset_size,
                           std::size_t subset_code) const {
    uint32_t v = 0;
    ctx_->Check(td_dp_get_treedepth(ctx_->get(), static_cast<int32_t>(subset_size),
                                    subset_code, &v));
    return v;
  }

  // The decomposition the reference builds from GetElimination(n, n, 0)
  // through EliminationTree::Decompose (Treedepth/main.cpp:187-195).
  Result GetTDDecomp() const { return Result::FromParent(parent_, treedepth_); }
  const td_dp_stats& Stats() const { return stats_; }

 private:
  std::shared_ptr<Context> ctx_;
  int nverts_ = 0;
  unsigned treedepth_ = 0;
  std::vector<int> parent_;
  td_dp_stats stats_{};
};

// td::BnBGPU: int operator()(g, heur_td) like the reference; Solve(g) also
// returns the decomposition and whether the search completed.
class BnBGPU {
 public:
  explicit BnBGPU(std::shared_ptr<Context> ctx = std::make_shared<Context>())
      : ctx_(std::move(ctx)) {}

  int operator()(const Graph& g, std::size_t heur_td) {
    int32_t td = 0;
    uint8_t cert = 0;
    std::vector<int8_t> par(static_cast<std::size_t>(g.n));
    ctx_->Check(td_bnb_solve(ctx_->get(), g.n, g.adj.data(), static_cast<int32_t>(heur_td),
                             &td, par.data(), &cert));
    parent_.assign(par.begin(), par.end());
    certified_ = cert != 0;
    return td;
  }

  // The search from an optional upper-bound seed (heur_td, 0 = none): the
  // decomposition and whether the search completed.
  std::pair<Result, bool> Solve(const Graph& g, std::size_t heur_td = 0) {
    const int td = (*this)(g, heur_td);
    return {Result::FromParent(parent_, static_cast<unsigned>(td)), certified_};
  }

 private:
  std::shared_ptr<Context> ctx_;
  std::vector<int> parent_;
  bool certified_ = false;
};

// The reference's Heuristic plugins (Treedepth/src/branch_and_bound/
// heuristics/) on the GPU: Get(g) like BranchAndBound::Heuristic::Get
// (branch_and_bound.hpp:38-42).
class HeuristicGPU {
 public:
  enum Kind {
    kChain = 1,
    kHighestDegree = 2,
    kSpanningTree = 3,
    kVariance = 4,
    kChainBottomUp = 5,  // HighestDegree(SpanningTree(Variance(BottomUp)))
    kChainBnBGPU = 6     // HighestDegree(Variance(BottomUp))
  };
  explicit HeuristicGPU(Kind kind = kChain,
                        std::shared_ptr<Context> ctx = std::make_shared<Context>())
      : kind_(kind), ctx_(std::move(ctx)) {}

  Result Get(const Graph& g) {
    int32_t n = g.n, td = 0;
    std::vector<int8_t> par(static_cast<std::size_t>(g.n));
    ctx_->Check(td_heuristic_batch(ctx_->get(), 1, &n, g.adj.data(), kind_, &td, par.data()));
    return Result::FromParent(std::vector<int>(par.begin(), par.end()),
                              static_cast<unsigned>(td));
  }

 private:
  Kind kind_;
  std::shared_ptr<Context> ctx_;
};

// td::BottomUpHeuristicGPUAlgorithm (Treedepth/src/branch_and_bound/
// heuristics/bottom_up_heuristic_gpu_algorithm.hpp:15-65): operator()(g, k)
// runs k levels (default: as many as memory holds), then the queries.
class BottomUpHeuristicGPUAlgorithm {
 public:
  explicit BottomUpHeuristicGPUAlgorithm(std::shared_ptr<Context> ctx = std::make_shared<Context>())
      : ctx_(std::move(ctx)) {}

  void operator()(const Graph& g, int iterations = -1) {
    parent_.assign(static_cast<std::size_t>(g.n), 0);
    int32_t it = 0;
    ctx_->Check(td_bottom_up(ctx_->get(), g.n, g.adj.data(), iterations, &treedepth_,
                             parent_.data(), &it));
    iterations_ = static_cast<std::size_t>(it);
  }
  std::size_t GetIterationsPerformed() const { return iterations_; }
  unsigned GetTreedepth(std::size_t /*nverts*/, std::size_t subset_size,
                        std::size_t subset_code) const {
    uint32_t v = 0;
    ctx_->Check(td_bottom_up_get_treedepth(ctx_->get(), static_cast<int32_t>(subset_size),
                                           subset_code, &v));
    return v;
  }
  template <typename VertexType = std::size_t>
  std::list<VertexType> GetElimination(std::size_t /*nverts*/, std::size_t subset_size,
                                       std::size_t subset_code) const {
    std::vector<int32_t> order(subset_size ? subset_size : 1);
    ctx_->Check(td_bottom_up_get_elimination(ctx_->get(), static_cast<int32_t>(subset_size),
                                             subset_code, order.data()));
    std::list<VertexType> out;
    for (std::size_t i = 0; i < subset_size; ++i) out.push_back(static_cast<VertexType>(order[i]));
    return out;
  }
  // the decomposition td_bottom_up built (valid after a full run)
  int32_t Treedepth() const { return treedepth_; }
  const std::vector<int8_t>& Parent() const { return parent_; }

 private:
  std::shared_ptr<Context> ctx_;
  int32_t treedepth_ = -1;
  std::vector<int8_t> parent_;
  std::size_t iterations_ = 0;
};

// td::BottomUpHeuristicGPU::Get (bottom_up_heuristic.hpp:17-31): the
// decomposition, or treedepth UINT_MAX when not every level fitted.
class BottomUpHeuristicGPU {
 public:
  explicit BottomUpHeuristicGPU(std::shared_ptr<Context> ctx = std::make_shared<Context>())
      : algo_(std::move(ctx)) {}
  Result Get(const Graph& g) {
    if (g.n > 32) {  // its levels never fit: the reference returns UINT_MAX
      Result r;
      r.treedepth = std::numeric_limits<unsigned>::max();
      return r;
    }
    algo_(g);
    if (algo_.Treedepth() < 0) {
      Result r;
      r.treedepth = std::numeric_limits<unsigned>::max();
      return r;
    }
    const auto& p = algo_.Parent();
    return Result::FromParent(std::vector<int>(p.begin(), p.end()),
                              static_cast<unsigned>(algo_.Treedepth()));
  }

 private:
  BottomUpHeuristicGPUAlgorithm algo_;
};

}  // namespace td

#endif  // TDHIP_HPP_
