// heur.hip -- the reference's upper-bound heuristics on the GPU, one wave per
// graph, lanes = vertices (n <= 64):
//   HighestDegreeHeuristic (Treedepth/src/branch_and_bound/heuristics/
//     highest_degree_heuristic.hpp:19-45),
//   VarianceHeuristic (variance_heuristic.hpp:24-78),
//   SpanningTreeHeuristic (spanning_tree_heuristic.hpp:20-83) over the
//     union-find spanning forest of common/utils/graph_gen.hpp:47-62,
//   and the chains HighestDegree(SpanningTree(Variance(1.0, 0.2, 0.8))),
//   HighestDegree(SpanningTree(Variance(BottomUp))) and
//   HighestDegree(Variance(BottomUp)) (Treedepth/main.cpp:172-186), whose
//   decorators keep the inner result only when it is strictly shallower; the
//   bottom-up link comes from td_bottom_up (bottomup.hip).
// Each heuristic eliminates one vertex at a time from the first component
// (lowest vertex first, the EliminationTree's component order) and the
// result is the EliminationTree decomposition: the parent of v is the vertex
// whose elimination created v's component.  The Variance score is computed
// in the reference's arithmetic: double sums and square root, float E times
// the component count, no fused multiply-add.
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace tdhip {
namespace {

constexpr int kMaxHeurN = 64;

__device__ __forceinline__ uint64_t HBallot(bool p) {
  return static_cast<uint64_t>(__ballot(p));
}

// component of `from` inside X by ballot BFS (lane u holds row u)
__device__ __forceinline__ uint64_t HComponent(uint64_t row, uint64_t X, uint64_t from) {
  uint64_t comp = from, front = from;
  for (;;) {
    const uint64_t nb = HBallot((row & front) != 0ull) & X & ~comp;
    if (!nb) break;
    comp |= nb;
    front = nb;
  }
  return comp;
}

// Elimination-tree bookkeeping of one wave: lane u keeps the creator of its
// current component and, once eliminated, its depth.
struct Elim {
  int creator = -1;
  int depth = 0;  // of this lane's vertex once eliminated
  int8_t parent = -1;
  uint64_t R = 0;  // vertices not eliminated yet
};

// eliminate v (uniform) from the component C of R that holds it
__device__ __forceinline__ void Eliminate(Elim& E, uint64_t C, int v) {
  const int lane = threadIdx.x & 63;
  const int cv = __builtin_amdgcn_readlane(E.creator, v);
  const int dcv = cv >= 0 ? __builtin_amdgcn_readlane(E.depth, cv) : 0;
  if (lane == v) {
    E.parent = static_cast<int8_t>(cv);
    E.depth = dcv + 1;
  }
  if (((C >> lane) & 1ull) && lane != v) E.creator = v;
  E.R &= ~(1ull << v);
}

__device__ __forceinline__ int WaveMaxI(int x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x = max(x, __shfl_xor(x, off));
  return x;
}

// HighestDegreeHeuristic: the vertex of largest degree inside the first
// component, the lowest one on ties ("first strictly greater wins").
__device__ int HighestDegree(uint64_t row, int n, int8_t* parent) {
  const int lane = threadIdx.x & 63;
  Elim E;
  E.R = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
  while (E.R) {
    const uint64_t C = HComponent(row, E.R, E.R & (0ull - E.R));
    const bool in = (C >> lane) & 1ull;
    const int key = in ? (__builtin_popcountll(row & C) << 6) | (63 - lane) : -1;
    const int best = WaveMaxI(key);
    Eliminate(E, C, 63 - (best & 63));
  }
  if (lane < n) parent[lane] = E.parent;
  return WaveMaxI(lane < n ? E.depth : 0);
}

// VarianceHeuristic's score of eliminating p from C: the components of
// C \ p, in lowest-vertex order, of sizes s_i (m of them):
//   avg = sum s_i / m, gamma = sum (s_i - avg)^2,
//   coef = D / sqrt(gamma + A) + E * m   (E * m in float).
#pragma clang fp contract(off)
__device__ double VarianceCoef(uint64_t C, int p, const uint64_t* s_adj, uint8_t* sizes,
                               float A, float D, float E) {
  uint64_t rest = C & ~(1ull << p);
  int m = 0;
  double sum = 0.0;
  while (rest) {
    uint64_t comp = rest & (0ull - rest), front = comp;
    while (front) {
      uint64_t nb = 0;
      for (uint64_t f = front; f; f &= f - 1) nb |= s_adj[__builtin_ctzll(f)];
      nb &= rest & ~comp;
      comp |= nb;
      front = nb;
    }
    const int s = __builtin_popcountll(comp);
    sizes[m++] = static_cast<uint8_t>(s);
    sum += static_cast<double>(s);
    rest &= ~comp;
  }
  const double avg = sum / static_cast<double>(m);
  double gamma = 0.0;
  for (int i = 0; i < m; ++i) {
    const double d = static_cast<double>(sizes[i]) - avg;
    gamma += d * d;
  }
  const float em = E * static_cast<float>(m);
  return static_cast<double>(D) / __builtin_sqrt(gamma + static_cast<double>(A)) +
         static_cast<double>(em);
}
#pragma clang fp contract(on)

// VarianceHeuristic: every vertex p of the first component is scored; the
// largest score wins, then the larger degree, then the lower vertex.
__device__ int Variance(uint64_t row, int n, const uint64_t* s_adj, uint8_t* s_sizes,
                        float A, float D, float E, int8_t* parent) {
  const int lane = threadIdx.x & 63;
  Elim E_;
  E_.R = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
  while (E_.R) {
    const uint64_t C = HComponent(row, E_.R, E_.R & (0ull - E_.R));
    const bool in = (C >> lane) & 1ull;
    double coef = -2.0;
    int deg = -1;
    if (in) {
      coef = VarianceCoef(C, lane, s_adj, s_sizes + 64 * lane, A, D, E);
      deg = __builtin_popcountll(row & C);
    }
    // lexicographic max of (coef, deg, -lane) over the wave
    int best = lane;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double oc = __shfl_xor(coef, off);
      const int od = __shfl_xor(deg, off);
      const int ob = __shfl_xor(best, off);
      const bool take = oc > coef || (oc == coef && (od > deg || (od == deg && ob < best)));
      if (take) {
        coef = oc;
        deg = od;
        best = ob;
      }
    }
    Eliminate(E_, C, __builtin_amdgcn_readfirstlane(best));
  }
  if (lane < n) parent[lane] = E_.parent;
  return WaveMaxI(lane < n ? E_.depth : 0);
}

// SpanningTreeHeuristic: take the first component of the spanning forest
// (of the vertices left); below 3 vertices eliminate its lowest vertex,
// otherwise peel leaves until at most 2 vertices are left and eliminate the
// lowest of them; the same vertex is eliminated in the graph's own tree.
// s_st must hold the spanning forest's rows.
__device__ int SpanningTree(uint64_t row, int n, const uint64_t* s_st, int8_t* parent) {
  const int lane = threadIdx.x & 63;
  const uint64_t strow = lane < n ? s_st[lane] : 0ull;
  Elim E;
  E.R = n >= 64 ? ~0ull : ((1ull << n) - 1ull);
  while (E.R) {
    const uint64_t v0 = E.R & (0ull - E.R);
    const uint64_t c = HComponent(strow, E.R, v0);
    const int sz = __builtin_popcountll(c);
    int v = __builtin_ctzll(v0);
    if (sz >= 3) {
      const bool in = (c >> lane) & 1ull;
      uint64_t taken = HBallot(in && __builtin_popcountll(strow & c) == 1);
      // a tree of >= 3 vertices always has a leaf left to peel; the cap
      // only bounds the loop
      for (int it = 0; it < 64 && __builtin_popcountll(taken) + 2 < sz; ++it) {
        const bool free_v = in && !((taken >> lane) & 1ull);
        taken |= HBallot(free_v && __builtin_popcountll(strow & c & ~taken) == 1);
      }
      const uint64_t left = c & ~taken;
      v = __builtin_ctzll(left ? left : c);
    }
    Eliminate(E, HComponent(row, E.R, 1ull << v), v);
  }
  if (lane < n) parent[lane] = E.parent;
  return WaveMaxI(lane < n ? E.depth : 0);
}

// Spanning forest of common/utils/graph_gen.hpp:47-62: union-find over the
// edges in (lower endpoint, higher endpoint) order, keeping each edge that
// joins two sets.  One lane, n <= 64.
__device__ void SpanningForest(const uint64_t* s_adj, int n, uint64_t* s_st, int* s_uf) {
  const int lane = threadIdx.x & 63;
  s_st[lane] = 0ull;
  s_uf[lane] = lane;
  __syncthreads();
  if (lane == 0) {
    for (int i = 0; i < n; ++i) {
      for (uint64_t m = s_adj[i] & ~((2ull << i) - 1ull); m; m &= m - 1) {
        const int j = __builtin_ctzll(m);
        int a = i, b = j;
        while (s_uf[a] != a) a = s_uf[a];
        while (s_uf[b] != b) b = s_uf[b];
        if (a != b) {
          s_uf[b] = a;
          s_st[i] |= 1ull << j;
          s_st[j] |= 1ull << i;
        }
      }
    }
  }
  __syncthreads();
}

struct HeurTree {
  int td;
  int8_t parent[64];
};

// kind: 1 = HighestDegree(SpanningTree(Variance(1.0, 0.2, 0.8))), 2 =
// HighestDegree, 3 = SpanningTree, 4 = Variance(1.0, 0.2, 0.8).
__global__ __launch_bounds__(64) void HeuristicKernel(const int32_t* __restrict__ ns,
                                                      const uint64_t* __restrict__ adj64,
                                                      int kind,
                                                      const int32_t* __restrict__ inner_td,
                                                      const int8_t* __restrict__ inner_par,
                                                      int32_t* __restrict__ td_out,
                                                      int8_t* __restrict__ parent_out) {
  __shared__ uint64_t s_adj[64], s_st[64];
  __shared__ int s_uf[64];
  __shared__ uint8_t s_sizes[64 * 64];
  __shared__ int8_t s_par[4][64];
  const int gi = blockIdx.x, lane = threadIdx.x;
  const int n = ns[gi];
  const uint64_t row = lane < n ? adj64[static_cast<size_t>(gi) * 64 + lane] : 0ull;
  s_adj[lane] = row;
  __syncthreads();
  const float A = 1.0f, D = 0.2f, E = 0.8f;
  int td = 0;
  int8_t* out = s_par[0];
  if (kind == 2) {
    td = HighestDegree(row, n, s_par[0]);
  } else if (kind == 3) {
    SpanningForest(s_adj, n, s_st, s_uf);
    td = SpanningTree(row, n, s_st, s_par[0]);
  } else if (kind == 4) {
    td = Variance(row, n, s_adj, s_sizes, A, D, E, s_par[0]);
  } else if (kind == 6) {
    // HighestDegree(Variance(inner)): decorators keep the inner result only
    // when strictly shallower; a missing inner result (-1) never is
    int tv = Variance(row, n, s_adj, s_sizes, A, D, E, s_par[0]);
    int8_t* vt = s_par[0];
    const int ti = inner_td[gi];
    if (ti >= 0 && ti < tv) {
      tv = ti;
      s_par[3][lane] = lane < n ? inner_par[static_cast<size_t>(gi) * 64 + lane] : -1;
      vt = s_par[3];
    }
    __syncthreads();
    const int th = HighestDegree(row, n, s_par[2]);
    td = th;
    out = s_par[2];
    if (tv < th) {
      td = tv;
      out = vt;
    }
  } else {
    // kind 1: HighestDegree(SpanningTree(Variance)); kind 5: the same over an
    // inner result.  Decorators keep the inner result only when strictly
    // shallower.
    int tv = Variance(row, n, s_adj, s_sizes, A, D, E, s_par[0]);
    if (kind == 5) {
      const int ti = inner_td[gi];
      if (ti >= 0 && ti < tv) {
        tv = ti;
        s_par[0][lane] = lane < n ? inner_par[static_cast<size_t>(gi) * 64 + lane] : -1;
      }
      __syncthreads();
    }
    SpanningForest(s_adj, n, s_st, s_uf);
    int ts = SpanningTree(row, n, s_st, s_par[1]);
    int8_t* st = s_par[1];
    if (tv < ts) {
      ts = tv;
      st = s_par[0];
    }
    const int th = HighestDegree(row, n, s_par[2]);
    td = th;
    out = s_par[2];
    if (ts < th) {
      td = ts;
      out = st;
    }
  }
  __syncthreads();
  if (lane < n) parent_out[static_cast<size_t>(gi) * 64 + lane] = out[lane];
  if (lane == 0) td_out[gi] = td;
}

template <typename F>
int HGuard(td_ctx* ctx, F&& f) {
  if (!ctx) return TD_E_INVALID;
  try {
    TD_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return TD_OK;
  } catch (const Error& e) {
    ctx->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return TD_E_INVALID;
  }
}

}  // namespace
}  // namespace tdhip

using namespace tdhip;  // NOLINT

extern "C" {

int td_heuristic_batch(td_ctx* ctx, int32_t ngraphs, const int32_t* n,
                       const uint64_t* adj_concat, int32_t kind, int32_t* td_out,
                       int8_t* parent_concat) {
  return HGuard(ctx, [&] {
    TD_REQUIRE(ngraphs >= 1 && n && adj_concat && td_out, TD_E_INVALID,
               "empty batch or null pointer");
    TD_REQUIRE(kind >= 1 && kind <= 6, TD_E_INVALID, "kind must be in [1, 6]");
    std::vector<size_t> off(static_cast<size_t>(ngraphs) + 1, 0);
    for (int i = 0; i < ngraphs; ++i) {
      TD_REQUIRE(n[i] >= 1 && n[i] <= kMaxHeurN, TD_E_INVALID, "graphs need n in [1, 64]");
      off[i + 1] = off[i] + static_cast<size_t>(n[i]);
    }
    std::vector<uint64_t> rows(static_cast<size_t>(ngraphs) * 64, 0ull);
    for (int i = 0; i < ngraphs; ++i) {
      ValidateGraph(n[i], adj_concat + off[i], kMaxHeurN);
      std::memcpy(rows.data() + static_cast<size_t>(i) * 64, adj_concat + off[i],
                  sizeof(uint64_t) * n[i]);
    }
    auto* dn = static_cast<int32_t*>(ctx->Buffer("heur.n", 4ull * ngraphs));
    auto* dadj = static_cast<uint64_t*>(ctx->Buffer("heur.adj", rows.size() * 8));
    auto* dtd = static_cast<int32_t*>(ctx->Buffer("heur.td", 4ull * ngraphs));
    auto* dpar = static_cast<int8_t*>(ctx->Buffer("heur.parent", 64ull * ngraphs));
    TD_HIP(hipMemcpyAsync(dn, n, 4ull * ngraphs, hipMemcpyHostToDevice, ctx->stream));
    TD_HIP(hipMemcpyAsync(dadj, rows.data(), rows.size() * 8, hipMemcpyHostToDevice,
                          ctx->stream));
    // kinds 5 and 6: the bottom-up heuristic is the innermost link; it runs
    // for graphs of at most 32 vertices (a larger graph's levels never fit,
    // so the reference's link returns UINT_MAX there)
    std::vector<int32_t> itd(static_cast<size_t>(ngraphs), -1);
    std::vector<int8_t> ipar(kind >= 5 ? 64ull * ngraphs : 1, -1);
    if (kind >= 5) {
      for (int i = 0; i < ngraphs; ++i) {
        if (n[i] > 32) continue;
        int32_t t = -1;
        const int rc = td_bottom_up(ctx, n[i], adj_concat + off[i], -1, &t,
                                    ipar.data() + 64ull * i, nullptr);
        if (rc != TD_OK) throw Error{rc, ctx->err};
        itd[i] = t;
      }
    }
    auto* ditd = static_cast<int32_t*>(ctx->Buffer("heur.inner_td", 4ull * ngraphs));
    auto* dipar = static_cast<int8_t*>(ctx->Buffer("heur.inner_parent", ipar.size()));
    TD_HIP(hipMemcpyAsync(ditd, itd.data(), 4ull * ngraphs, hipMemcpyHostToDevice, ctx->stream));
    TD_HIP(hipMemcpyAsync(dipar, ipar.data(), ipar.size(), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(HeuristicKernel, dim3(ngraphs), dim3(64), 0, ctx->stream, dn, dadj,
                       kind, ditd, dipar, dtd, dpar);
    TD_HIP(hipGetLastError());
    std::vector<int8_t> par(64ull * ngraphs);
    TD_HIP(hipMemcpyAsync(td_out, dtd, 4ull * ngraphs, hipMemcpyDeviceToHost, ctx->stream));
    TD_HIP(hipMemcpyAsync(par.data(), dpar, par.size(), hipMemcpyDeviceToHost, ctx->stream));
    TD_HIP(hipStreamSynchronize(ctx->stream));
    if (parent_concat)
      for (int i = 0; i < ngraphs; ++i)
        std::memcpy(parent_concat + off[i], par.data() + 64ull * i, n[i]);
  });
}

}  // extern "C"
