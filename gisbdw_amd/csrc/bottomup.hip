// bottomup.hip -- the bottom-up union-find heuristic on gfx950.
//
// Reference: BottomUpHeuristicGPU (Treedepth/src/branch_and_bound/heuristics/
// bottom_up_heuristic.hpp:11-35) over BottomUpHeuristicGPUAlgorithm
// (bottom_up_heuristic_gpu_algorithm.{hpp,cu}).  Level k holds, for every
// k-subset S of the vertices, the elimination forest of G[S] that the
// heuristic keeps: the empty set has height 1; S tries its vertices v in
// ascending order as the one added last (v becomes the parent of the roots of
// the components of its neighbours in S, DynamicStepKernel :57-95 with the
// union-find of Treedepth/src/union_find/ext_array_union_find.hpp) and keeps
// the first forest of strictly smallest height, starting from |S| + 1.
//
// Only the component heights matter for later levels (Union reads the heights
// of the two roots; Find only the root), and the components of the kept forest
// are those of G[S].  So a subset's record here holds 32 slots, one per
// vertex: slot w is the height of w's component for w in S, and the height of
// the whole forest for w outside S.  Adding v to S \ v then needs one record:
// slot v gives the height of S \ v, v's neighbours' slots the heights that v
// joins (the new component's height is 1 + their largest, 1 if none), and the
// forest height is the larger of the two.  Records are bit-sliced: plane j is
// a 32-bit mask holding bit j of every slot, so the largest slot over v's
// neighbours is a descent over the planes with no per-neighbour loop.  A
// level's heights are at most k and at most 1 + the largest height of the
// level below, so a level whose heights must stay below 16 stores 4 planes
// (16-byte records), any other level 8 (32 bytes), instead of the reference's
// n + 2 union-find bytes; each launch finds the width from the largest heights
// the launches before it recorded.  The history keeps, per subset, the vertex
// added last and the height (history_ in the reference), 2 bytes per subset
// over all levels; the records live for two levels only.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace tdhip {
namespace {

constexpr int kBuStride = 33;  // s_cb[x * 33 + j] = C(x, j)
constexpr int kRecBytes = 32;  // largest per-subset record (byte slots)
constexpr int kFourPlaneLevels = 15; // levels k <= 15 always fit four planes
constexpr int kBuBlock = 256;
// The binomial buffer continues with O_k = sum_{j<k} C(16, j) (k = 0..16)
// and, from word kBuTailWord, the colex k-subsets of {0..15} as 16-bit masks
// (entry O_k + r is the subset of rank r).
constexpr int kBuTailBits = 16;
constexpr int kBuOffs = 32 * kBuStride;
constexpr int kBuCbWords = kBuOffs + kBuTailBits + 1;
constexpr int kBuTailWord = (kBuCbWords + 31) / 32 * 32;
constexpr int kBuPascalWords = kBuTailWord + (1 << kBuTailBits) / 2;

// Colex unrank of r among the k-subsets of {0..n-1} (set_encoder.hpp:61-73):
// greedy from the top down to vertex 16, then one table read for the rest
// (the remainder is then below C(16, k)); needs r < C(n, k).
__device__ __forceinline__ uint32_t BuUnrank(uint32_t r, int k, int n, const uint32_t* s_cb,
                                             const uint16_t* __restrict__ tail) {
  uint32_t S = 0;
  for (int x = n - 1; x >= kBuTailBits; --x) {
    const uint32_t c = s_cb[x * kBuStride + k];
    const uint32_t take = r >= c ? 1u : 0u;
    r = take ? r - c : r;
    S = 2u * S + take;
    k -= static_cast<int>(take);
  }
  return (S << kBuTailBits) | tail[s_cb[kBuOffs + k] + r];
}

// Connected component of v in G[S] (breadth-first on bitmask rows).
__device__ __forceinline__ uint32_t BuComponent(const uint32_t* s_adj, uint32_t S, int v) {
  uint32_t comp = 1u << v, front = comp;
  while (front) {
    uint32_t nb = 0;
    for (uint32_t m = front; m; m &= m - 1) nb |= s_adj[__builtin_ctz(m)];
    front = nb & S & ~comp;
    comp |= front;
  }
  return comp;
}

// A record in registers, bit-sliced: plane j (p0..p7) holds bit j of every
// slot, slot w at bit w of each plane.  Levels whose heights stay below 16
// use planes 0-3 (16 bytes), the others planes 0-7 (32 bytes).  Named fields,
// not an array, so that plane selects stay in registers.
struct BuRec {
  uint32_t p0, p1, p2, p3, p4, p5, p6, p7;
};

template <int P>
__device__ __forceinline__ BuRec BuLoad(const uint8_t* ptr) {
  BuRec q;
  const uint4 x = reinterpret_cast<const uint4*>(ptr)[0];
  q.p0 = x.x;
  q.p1 = x.y;
  q.p2 = x.z;
  q.p3 = x.w;
  q.p4 = q.p5 = q.p6 = q.p7 = 0;
  if (P == 8) {
    const uint4 y = reinterpret_cast<const uint4*>(ptr)[1];
    q.p4 = y.x;
    q.p5 = y.y;
    q.p6 = y.z;
    q.p7 = y.w;
  }
  return q;
}

__device__ __forceinline__ uint32_t BuPlane(const BuRec& q, int j) {  // j compile-time
  return j == 0 ? q.p0 : j == 1 ? q.p1 : j == 2 ? q.p2 : j == 3 ? q.p3
       : j == 4 ? q.p4 : j == 5 ? q.p5 : j == 6 ? q.p6 : q.p7;
}

// Slot w.
template <int P>
__device__ __forceinline__ int BuSlot(const BuRec& q, int w) {
  int x = 0;
#pragma unroll
  for (int j = 0; j < P; ++j) x |= static_cast<int>((BuPlane(q, j) >> w) & 1u) << j;
  return x;
}

// Largest slot among the vertices of M (0 for an empty M), from the top
// plane down: keep the candidates that have the bit whenever any does.
template <int P>
__device__ __forceinline__ int BuMaxSlot(const BuRec& q, uint32_t M) {
  int top = 0;
#pragma unroll
  for (int j = P - 1; j >= 0; --j) {
    const uint32_t t = M & BuPlane(q, j);
    top |= (t != 0u) << j;
    M = t ? t : M;
  }
  return top;
}

// Levels whose heights all stay below 16 keep four planes: level k's
// heights are at most 1 + the largest height of level k - 1 (and at most k).
__device__ __forceinline__ bool BuFourPlanes(int k, const uint32_t* maxh) {
  return k <= kFourPlaneLevels || maxh[k - 1] <= 14u;
}

// Every k-subset of rank r < count at one level; returns the largest height
// this lane wrote.
template <int PIn, int POut>
__device__ __forceinline__ uint32_t BuLevel(const uint32_t* s_cb, const uint32_t* s_adj,
                                            const uint16_t* __restrict__ tail, int n, int k,
                                            uint32_t count, const uint8_t* __restrict__ prev_rec,
                                            uint8_t* __restrict__ cur_rec,
                                            int8_t* __restrict__ cur_last,
                                            uint8_t* __restrict__ cur_height) {
  constexpr int kIn = 4 * PIn, kOut = 4 * POut;
  uint32_t lane_max = 0;
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count;
       r += gridDim.x * blockDim.x) {
    const uint32_t S = BuUnrank(r, k, n, s_cb, tail);
    // Elements y_0 < ... < y_{k-1} of S from the top: rank(S \ y_i) =
    // sum_{j<i} C(y_j, j+1) + sum_{j>i} C(y_j, j) (set_encoder.hpp:48-59),
    // where the first sum is r - sum_{j>=i} C(y_j, j+1).  Walking down with
    // "<=" keeps the smallest vertex of least height, as the ascending
    // strict "<" of the reference does.  The records of kGroup removals are
    // loaded together; every lane has k elements, so the groups (and the
    // `ok` flags) are the same across the wave.
    constexpr int kGroup = PIn == 4 ? 4 : 2;
    uint32_t sufa = 0, sufb = 0, brank = 0;
    int best = k + 1, bv = 0, bnew = 0;
    int i = k - 1;
    for (uint32_t m = S; m;) {
      int vs[kGroup];
      uint32_t r2s[kGroup];
      bool ok[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        ok[u] = m != 0;
        const int v = ok[u] ? 31 - __builtin_clz(m) : 0;
        vs[u] = v;
        r2s[u] = 0;  // record 0 of the previous level: a safe address
        if (ok[u]) {
          m ^= 1u << v;
          sufb += s_cb[v * kBuStride + i + 1];
          r2s[u] = r - sufb + sufa;
          sufa += s_cb[v * kBuStride + i];
          --i;
        }
      }
      BuRec qs[kGroup];
#pragma unroll
      for (int u = 0; u < kGroup; ++u)
        qs[u] = BuLoad<PIn>(prev_rec + static_cast<size_t>(r2s[u]) * kIn);
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        if (!ok[u]) break;
        const int v = vs[u];
        // the joined component's height: 1 + the largest neighbour slot
        const int newh = BuMaxSlot<PIn>(qs[u], s_adj[v] & S) + 1;
        // slot v of S \ v holds its forest height (0 for the empty set,
        // where the reference's union-find state holds 1; newh >= 1 makes
        // the two agree here)
        const int h = max(BuSlot<PIn>(qs[u], v), newh);
        if (h <= best) {
          best = h;
          bv = v;
          bnew = newh;
          brank = r2s[u];
        }
      }
    }
    const uint32_t comp = BuComponent(s_adj, S, bv);
    // a level with four planes holds no height above 15, so the planes a
    // narrower record lacks are zero and those a wider one drops are too
    const BuRec q = BuLoad<PIn>(prev_rec + static_cast<size_t>(brank) * kIn);
    // kept slots: S outside v's component; v's component: bnew; outside S:
    // the forest height
    const uint32_t keep = S & ~comp, out_of_s = ~S;
    uint32_t w[8];
#pragma unroll
    for (int j = 0; j < POut; ++j)
      w[j] = (BuPlane(q, j) & keep) | ((bnew >> j) & 1 ? comp : 0u) |
             ((best >> j) & 1 ? out_of_s : 0u);
    uint4* dst = reinterpret_cast<uint4*>(cur_rec + static_cast<size_t>(r) * kOut);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    if (POut == 8) dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    cur_last[r] = static_cast<int8_t>(bv);
    cur_height[r] = static_cast<uint8_t>(best);
    lane_max = max(lane_max, static_cast<uint32_t>(best));
  }
  return lane_max;
}

// One level k (1 <= k <= n).  maxh[j] is the largest height of level j
// (maxh[0] = 0, maxh[k] cleared before the launch); it picks the width of
// level k - 1's records and of this level's, and gets level k's maximum.
__global__ __launch_bounds__(kBuBlock) void BottomUpLevelKernel(
    const uint32_t* __restrict__ g_adj, const uint32_t* __restrict__ g_cb, int n, int k,
    uint32_t count, const uint8_t* __restrict__ prev_rec, uint8_t* __restrict__ cur_rec,
    int8_t* __restrict__ cur_last, uint8_t* __restrict__ cur_height,
    uint32_t* __restrict__ maxh) {
  __shared__ uint32_t s_cb[kBuCbWords];
  __shared__ uint32_t s_adj[32];
  __shared__ uint32_t s_max;
  for (int i = threadIdx.x; i < kBuCbWords; i += blockDim.x) s_cb[i] = g_cb[i];
  if (threadIdx.x < 32) s_adj[threadIdx.x] = threadIdx.x < static_cast<unsigned>(n) ? g_adj[threadIdx.x] : 0u;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  const uint16_t* tail = reinterpret_cast<const uint16_t*>(g_cb + kBuTailWord);
  const bool four_in = BuFourPlanes(k - 1, maxh), four_out = BuFourPlanes(k, maxh);
  uint32_t lane_max;
  if (four_in && four_out)
    lane_max = BuLevel<4, 4>(s_cb, s_adj, tail, n, k, count, prev_rec, cur_rec, cur_last,
                             cur_height);
  else if (four_in)
    lane_max = BuLevel<4, 8>(s_cb, s_adj, tail, n, k, count, prev_rec, cur_rec, cur_last,
                             cur_height);
  else if (four_out)
    lane_max = BuLevel<8, 4>(s_cb, s_adj, tail, n, k, count, prev_rec, cur_rec, cur_last,
                             cur_height);
  else
    lane_max = BuLevel<8, 8>(s_cb, s_adj, tail, n, k, count, prev_rec, cur_rec, cur_last,
                             cur_height);
  atomicMax(&s_max, lane_max);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(&maxh[k], s_max);
}

// GetElimination (bottom_up_heuristic_gpu_algorithm.hpp:92-108): from set S,
// the vertex stored as added last, removed, `steps` times.  One lane; every
// step is one table read.
__global__ __launch_bounds__(64) void BottomUpWalkKernel(
    const uint32_t* __restrict__ g_cb, const int8_t* __restrict__ last,
    const uint64_t* __restrict__ level_off, uint32_t S, int steps,
    int32_t* __restrict__ out) {
  __shared__ uint32_t s_cb[32 * kBuStride];
  for (int i = threadIdx.x; i < 32 * kBuStride; i += blockDim.x) s_cb[i] = g_cb[i];
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < steps; ++i) {
    uint32_t r = 0;
    int j = 0;
    for (uint32_t m = S; m; m &= m - 1) r += s_cb[__builtin_ctz(m) * kBuStride + (++j)];
    const int v = last[level_off[__builtin_popcount(S)] + r];
    out[i] = v;
    S &= ~(1u << (v & 31));
  }
}

template <typename F>
int BGuard(td_ctx* ctx, F&& f) {
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

const uint32_t* BuPascal(td_ctx* ctx) {
  auto it = ctx->bufs.find("bu.pascal");
  if (it != ctx->bufs.end() && it->second.p) return static_cast<const uint32_t*>(it->second.p);
  std::vector<uint32_t> cb(kBuPascalWords, 0);
  for (int x = 0; x < 32; ++x)
    for (int j = 0; j < kBuStride; ++j) cb[x * kBuStride + j] = static_cast<uint32_t>(Binom(x, j));
  uint16_t* tail = reinterpret_cast<uint16_t*>(cb.data() + kBuTailWord);
  uint32_t off = 0;
  for (int k = 0; k <= kBuTailBits; ++k) {
    cb[kBuOffs + k] = off;
    // the k-subsets of {0..15} in colex order = increasing integers
    for (uint32_t mask = 0; mask < (1u << kBuTailBits); ++mask)
      if (__builtin_popcount(mask) == k) tail[off++] = static_cast<uint16_t>(mask);
  }
  void* d = ctx->Buffer("bu.pascal", cb.size() * 4);
  TD_HIP(hipMemcpy(d, cb.data(), cb.size() * 4, hipMemcpyHostToDevice));
  return static_cast<const uint32_t*>(d);
}

// Device bytes for levels 0..K-1: two record buffers sized for the largest
// level below K, the history (2 bytes per subset), offsets, adjacency, the
// binomials and the walk's output.
uint64_t BuBytes(int n, int K) {
  uint64_t maxc = 1, hist = 0;
  for (int k = 0; k < K; ++k) {
    maxc = std::max<uint64_t>(maxc, Binom(n, k));
    hist += Binom(n, k);
  }
  return 2 * maxc * kRecBytes + 2 * ((hist + 255) / 256 * 256) + 8 * 40 + 4 * 32 +
         4 * kBuPascalWords + 4 * 32 + 4 * 40;
}

uint64_t BuHeld(td_ctx* ctx) {
  uint64_t b = 0;
  for (const char* k : {"bu.rec0", "bu.rec1", "bu.last", "bu.height", "bu.off", "bu.adj",
                        "bu.pascal", "bu.walk", "bu.maxh"}) {
    auto it = ctx->bufs.find(k);
    if (it != ctx->bufs.end()) b += it->second.bytes;
  }
  return b;
}

// Walks `steps` removals from S on the device; returns the vertices.
std::vector<int32_t> BuWalk(td_ctx* ctx, uint32_t S, int steps) {
  int32_t* dout = static_cast<int32_t*>(ctx->Buffer("bu.walk", 4 * 32));
  hipLaunchKernelGGL(BottomUpWalkKernel, dim3(1), dim3(64), 0, ctx->stream, BuPascal(ctx),
                     static_cast<const int8_t*>(ctx->bufs["bu.last"].p),
                     static_cast<const uint64_t*>(ctx->bufs["bu.off"].p), S, steps, dout);
  TD_HIP(hipGetLastError());
  std::vector<int32_t> order(std::max(steps, 1));
  TD_HIP(hipMemcpyAsync(order.data(), dout, 4 * steps, hipMemcpyDeviceToHost, ctx->stream));
  TD_HIP(hipStreamSynchronize(ctx->stream));
  order.resize(steps);
  return order;
}

}  // namespace
}  // namespace tdhip

using namespace tdhip;  // NOLINT

extern "C" {

int td_bottom_up(td_ctx* ctx, int32_t n, const uint64_t* adj, int32_t iterations,
                 int32_t* treedepth_out, int8_t* parent_out, int32_t* iterations_out) {
  return BGuard(ctx, [&] {
    ValidateGraph(n, adj, 32);
    TD_REQUIRE(treedepth_out != nullptr, TD_E_INVALID, "treedepth_out is null");
    TD_REQUIRE(iterations == -1 || (iterations >= 1 && iterations <= n + 1), TD_E_INVALID,
               "iterations must be -1 or in [1, n + 1]");
    ctx->have_bu = false;
    // GetMaxIterations (bottom_up_heuristic_gpu_algorithm.cu:152-175): as
    // many levels as the memory holds, here the context budget or the
    // device's free memory
    const uint64_t held = BuHeld(ctx);
    uint64_t avail;
    if (ctx->budget) {
      avail = ctx->budget - std::min(ctx->budget, ctx->allocated - held);
    } else {
      size_t free_b = 0, total_b = 0;
      TD_HIP(hipMemGetInfo(&free_b, &total_b));
      avail = static_cast<uint64_t>((free_b + held) * 0.9);
    }
    int K = iterations;
    if (K < 0) {
      K = 1;
      while (K < n + 1 && BuBytes(n, K + 1) <= avail) ++K;
    }
    TD_REQUIRE(BuBytes(n, K) <= avail, TD_E_NOMEM,
               "bottom-up tables for " + std::to_string(K) + " levels exceed the memory");
    uint64_t maxc = 1, hist = 0;
    std::vector<uint64_t> off(n + 2, 0);
    for (int k = 0; k <= n; ++k) {
      off[k] = hist;
      if (k < K) {
        maxc = std::max<uint64_t>(maxc, Binom(n, k));
        hist += Binom(n, k);
      }
    }
    off[n + 1] = hist;
    uint8_t* rec[2] = {static_cast<uint8_t*>(ctx->Buffer("bu.rec0", maxc * kRecBytes)),
                       static_cast<uint8_t*>(ctx->Buffer("bu.rec1", maxc * kRecBytes))};
    int8_t* last = static_cast<int8_t*>(ctx->Buffer("bu.last", (hist + 255) / 256 * 256));
    uint8_t* height = static_cast<uint8_t*>(ctx->Buffer("bu.height", (hist + 255) / 256 * 256));
    uint64_t* doff = static_cast<uint64_t*>(ctx->Buffer("bu.off", 8 * 40));
    uint32_t* dadj = static_cast<uint32_t*>(ctx->Buffer("bu.adj", 4 * 32));
    uint32_t* dmaxh = static_cast<uint32_t*>(ctx->Buffer("bu.maxh", 4 * 40));
    const uint32_t* cb = BuPascal(ctx);
    uint32_t hadj[32] = {};
    for (int v = 0; v < n; ++v) hadj[v] = static_cast<uint32_t>(adj[v]);
    TD_HIP(hipMemcpyAsync(dadj, hadj, sizeof(hadj), hipMemcpyHostToDevice, ctx->stream));
    TD_HIP(hipMemcpyAsync(doff, off.data(), 8 * off.size(), hipMemcpyHostToDevice, ctx->stream));
    // level 0: the empty set, history (-1, 0) (:232-234), record all zero
    // (every slot is outside the set and holds its height, 0)
    TD_HIP(hipMemsetAsync(rec[0], 0, kRecBytes, ctx->stream));
    TD_HIP(hipMemsetAsync(last, 0xFF, 1, ctx->stream));
    TD_HIP(hipMemsetAsync(height, 0, 1, ctx->stream));
    TD_HIP(hipMemsetAsync(dmaxh, 0, 4 * 40, ctx->stream));  // largest height per level
    for (int k = 1; k < K; ++k) {
      const uint32_t count = static_cast<uint32_t>(Binom(n, k));
      const int grid = static_cast<int>(std::min<uint64_t>(
          (count + kBuBlock - 1) / kBuBlock,
          static_cast<uint64_t>(ctx->prop.multiProcessorCount) * 32));
      hipLaunchKernelGGL(BottomUpLevelKernel, dim3(grid), dim3(kBuBlock), 0, ctx->stream, dadj,
                         cb, n, k, count, rec[(k - 1) & 1], rec[k & 1], last + off[k],
                         height + off[k], dmaxh);
      TD_HIP(hipGetLastError());
    }
    ctx->bu_n = n;
    ctx->bu_iters = K;
    ctx->bu_off = off;
    if (iterations_out) *iterations_out = K;
    if (K < n + 1) {  // BottomUpHeuristicGPU::Get returns UINT_MAX treedepth
      TD_HIP(hipStreamSynchronize(ctx->stream));
      ctx->have_bu = true;
      *treedepth_out = -1;
      return;
    }
    // GetElimination(n, n, 0), then the elimination forest of that order
    // (EliminationTree::Eliminate ... Decompose, bottom_up_heuristic.hpp:23-28);
    // the walk follows the levels on the same stream and synchronizes
    const uint32_t full = n == 32 ? 0xFFFFFFFFu : ((1u << n) - 1u);
    const std::vector<int32_t> order = BuWalk(ctx, full, n);
    ctx->have_bu = true;
    std::vector<int> parent(n, -1);
    std::vector<char> added(n, 0);
    for (int i = n - 1; i >= 0; --i) {
      const int v = order[i];
      TD_REQUIRE(v >= 0 && v < n && !added[v], TD_E_HIP, "bottom-up history is inconsistent");
      for (uint64_t m = adj[v]; m; m &= m - 1) {
        const int w = __builtin_ctzll(m);
        if (!added[w]) continue;
        int r = w;
        while (parent[r] >= 0) r = parent[r];
        if (r != v) parent[r] = v;
      }
      added[v] = 1;
    }
    int td = 0;
    for (int v = 0; v < n; ++v) {
      int d = 1;
      for (int u = v; parent[u] >= 0; u = parent[u]) ++d;
      td = std::max(td, d);
      if (parent_out) parent_out[v] = static_cast<int8_t>(parent[v]);
    }
    *treedepth_out = td;
  });
}

int td_bottom_up_get_treedepth(td_ctx* ctx, int32_t k, uint64_t code, uint32_t* td_out) {
  return BGuard(ctx, [&] {
    TD_REQUIRE(ctx->have_bu, TD_E_STATE, "no td_bottom_up result to query");
    TD_REQUIRE(td_out != nullptr, TD_E_INVALID, "td_out is null");
    TD_REQUIRE(k >= 0 && k <= ctx->bu_n && code < Binom(ctx->bu_n, k), TD_E_INVALID,
               "subset code out of range");
    if (k >= ctx->bu_iters) {  // GetTreedepth (:181-188): 0 beyond the levels run
      *td_out = 0;
      return;
    }
    uint8_t h = 0;
    TD_HIP(hipMemcpy(&h,
                     static_cast<const uint8_t*>(ctx->bufs["bu.height"].p) + ctx->bu_off[k] + code,
                     1, hipMemcpyDeviceToHost));
    *td_out = h;
  });
}

int td_bottom_up_get_level(td_ctx* ctx, int32_t k, int8_t* last_out, uint8_t* height_out,
                           uint64_t out_len) {
  return BGuard(ctx, [&] {
    TD_REQUIRE(ctx->have_bu, TD_E_STATE, "no td_bottom_up result to query");
    TD_REQUIRE(k >= 0 && k < ctx->bu_iters, TD_E_STATE, "level not run");
    const uint64_t cnt = Binom(ctx->bu_n, k);
    TD_REQUIRE(out_len >= cnt, TD_E_INVALID, "output shorter than the level");
    if (last_out)
      TD_HIP(hipMemcpy(last_out, static_cast<const int8_t*>(ctx->bufs["bu.last"].p) +
                                     ctx->bu_off[k], cnt, hipMemcpyDeviceToHost));
    if (height_out)
      TD_HIP(hipMemcpy(height_out, static_cast<const uint8_t*>(ctx->bufs["bu.height"].p) +
                                       ctx->bu_off[k], cnt, hipMemcpyDeviceToHost));
  });
}

int td_bottom_up_get_elimination(td_ctx* ctx, int32_t k, uint64_t code, int32_t* order_out) {
  return BGuard(ctx, [&] {
    TD_REQUIRE(ctx->have_bu, TD_E_STATE, "no td_bottom_up result to query");
    TD_REQUIRE(order_out != nullptr || k == 0, TD_E_INVALID, "order_out is null");
    TD_REQUIRE(k >= 0 && k <= ctx->bu_n && code < Binom(ctx->bu_n, k), TD_E_INVALID,
               "subset code out of range");
    TD_REQUIRE(k < ctx->bu_iters, TD_E_STATE, "subset size beyond the levels run");
    if (k == 0) return;
    const std::vector<int32_t> order =
        BuWalk(ctx, static_cast<uint32_t>(UnrankColex(ctx->bu_n, k, code)), k);
    std::memcpy(order_out, order.data(), 4 * k);
  });
}

}  // extern "C"
