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
// forest height is the larger of the two.  A level's heights are at most k and
// at most 1 + the largest height of the level below, so a level whose heights
// must stay below 16 stores 4-bit slots (16-byte records), any other level
// bytes (32-byte records), instead of the reference's n + 2 union-find bytes;
// each launch finds the width from the largest heights the launches before it
// recorded.  The history keeps, per subset, the vertex added last and the
// height (history_ in the reference), 2 bytes per subset over all levels; the
// records live for two levels only.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ctx.hpp"

namespace tdhip {
namespace {

constexpr int kBuStride = 33;  // s_cb[x * 33 + j] = C(x, j)
constexpr int kRecBytes = 32;  // largest per-subset record (byte slots)
constexpr int kNibLevels = 15; // levels k <= 15 always fit 4-bit slots
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

// A record in registers: a, b (4-bit slots, slot w at bits 4w of the
// 128-bit value b:a) or a..d (byte slots, slot w at bits 8w of d:c:b:a).
// Named fields, not an array, so that slot selects stay register selects.
struct BuRec {
  uint64_t a, b, c, d;
};

template <bool kNib>
__device__ __forceinline__ BuRec BuLoad(const uint8_t* p) {
  BuRec q;
  const uint4 x = reinterpret_cast<const uint4*>(p)[0];
  q.a = x.x | static_cast<uint64_t>(x.y) << 32;
  q.b = x.z | static_cast<uint64_t>(x.w) << 32;
  q.c = q.d = 0;
  if (!kNib) {
    const uint4 y = reinterpret_cast<const uint4*>(p)[1];
    q.c = y.x | static_cast<uint64_t>(y.y) << 32;
    q.d = y.z | static_cast<uint64_t>(y.w) << 32;
  }
  return q;
}

// Slot w; the words come as values (a reference would let the compiler turn
// the selects into a dynamically indexed load from private memory).
template <bool kNib>
__device__ __forceinline__ int BuSlot(uint64_t a, uint64_t b, uint64_t c, uint64_t d, int w) {
  if (kNib) {
    const uint64_t x = (w & 16) ? b : a;
    return static_cast<int>(x >> (4 * (w & 15))) & 15;
  }
  const uint64_t lo = (w & 8) ? b : a;
  const uint64_t hi = (w & 8) ? d : c;
  return static_cast<int>(((w & 16) ? hi : lo) >> (8 * (w & 7))) & 255;
}

// 32-bit word d (compile-time after unrolling) of a record
__device__ __forceinline__ uint32_t BuWord(uint64_t a, uint64_t b, uint64_t c, uint64_t d,
                                           int i) {
  const uint64_t x = i < 2 ? a : i < 4 ? b : i < 6 ? c : d;
  return static_cast<uint32_t>(x >> (32 * (i & 1)));
}

// bit i of the low 8 bits -> 0xF in nibble i (fills the LDS table)
__device__ __forceinline__ uint32_t BuSpreadNib(uint32_t x) {
  x &= 0xFFu;
  x = (x | (x << 12)) & 0x000F000Fu;
  x = (x | (x << 6)) & 0x03030303u;
  x = (x | (x << 3)) & 0x11111111u;
  return x * 15u;
}

// bit i of the low 4 bits -> 0xFF in byte i
__device__ __forceinline__ uint32_t BuSpreadByte(uint32_t x) {
  x &= 0xFu;
  x = (x | (x << 14)) & 0x00030003u;
  x = (x | (x << 7)) & 0x01010101u;
  return x * 255u;
}

// 8 nibbles -> 8 bytes (nibble i to byte i)
__device__ __forceinline__ uint64_t BuNibToBytes(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  return x;
}

// 8 bytes, each below 16 -> 8 nibbles (byte i to nibble i)
__device__ __forceinline__ uint32_t BuBytesToNib(uint64_t x) {
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return static_cast<uint32_t>(x);
}

// Levels whose heights all stay below 16 keep 4-bit slots: level k's
// heights are at most 1 + the largest height of level k - 1 (and at most k).
__device__ __forceinline__ bool BuNibLevel(int k, const uint32_t* maxh) {
  return k <= kNibLevels || maxh[k - 1] <= 14u;
}

// Every k-subset of rank r < count at one level; returns the largest height
// this lane wrote.
template <bool kNibIn, bool kNibOut>
__device__ __forceinline__ uint32_t BuLevel(const uint32_t* s_cb, const uint32_t* s_adj,
                                            const uint32_t* s_spread,
                                            const uint16_t* __restrict__ tail, int n, int k,
                                            uint32_t count, const uint8_t* __restrict__ prev_rec,
                                            uint8_t* __restrict__ cur_rec,
                                            int8_t* __restrict__ cur_last,
                                            uint8_t* __restrict__ cur_height) {
  constexpr int kIn = kNibIn ? 16 : 32, kOut = kNibOut ? 16 : 32;
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
    constexpr int kGroup = kNibIn ? 4 : 2;
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
        qs[u] = BuLoad<kNibIn>(prev_rec + static_cast<size_t>(r2s[u]) * kIn);
#pragma unroll
      for (int u = 0; u < kGroup; ++u) {
        if (!ok[u]) break;
        const int v = vs[u];
        const BuRec& q = qs[u];
        int top = 0;
        for (uint32_t a = s_adj[v] & S; a; a &= a - 1)
          top = max(top, BuSlot<kNibIn>(q.a, q.b, q.c, q.d, __builtin_ctz(a)));
        const int newh = top + 1;
        // slot v of S \ v holds its forest height (0 for the empty set,
        // where the reference's union-find state holds 1; newh >= 1 makes
        // the two agree here)
        const int h = max(BuSlot<kNibIn>(q.a, q.b, q.c, q.d, v), newh);
        if (h <= best) {
          best = h;
          bv = v;
          bnew = newh;
          brank = r2s[u];
        }
      }
    }
    const uint32_t comp = BuComponent(s_adj, S, bv);
    BuRec q = BuLoad<kNibIn>(prev_rec + static_cast<size_t>(brank) * kIn);
    if (kNibIn && !kNibOut) {  // widen 4-bit slots to bytes
      const uint64_t lo = q.a, hi = q.b;
      q.a = BuNibToBytes(static_cast<uint32_t>(lo));
      q.b = BuNibToBytes(static_cast<uint32_t>(lo >> 32));
      q.c = BuNibToBytes(static_cast<uint32_t>(hi));
      q.d = BuNibToBytes(static_cast<uint32_t>(hi >> 32));
    } else if (!kNibIn && kNibOut) {  // narrow bytes (all below 15 here)
      q.a = BuBytesToNib(q.a) | static_cast<uint64_t>(BuBytesToNib(q.b)) << 32;
      q.b = BuBytesToNib(q.c) | static_cast<uint64_t>(BuBytesToNib(q.d)) << 32;
    }
    // kept slots: S outside v's component; v's component: bnew; outside S:
    // the forest height
    const uint32_t out_of_s = ~S;
    uint32_t w[8];
    if (kNibOut) {
      const uint32_t rep_new = static_cast<uint32_t>(bnew) * 0x11111111u;
      const uint32_t rep_best = static_cast<uint32_t>(best) * 0x11111111u;
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        const uint32_t in = BuWord(q.a, q.b, q.c, q.d, d);
        const uint32_t cm = s_spread[(comp >> (8 * d)) & 0xFFu];
        const uint32_t nm = s_spread[(out_of_s >> (8 * d)) & 0xFFu];
        w[d] = (in & ~(cm | nm)) | (rep_new & cm) | (rep_best & nm);
      }
    } else {
      const uint32_t rep_new = static_cast<uint32_t>(bnew) * 0x01010101u;
      const uint32_t rep_best = static_cast<uint32_t>(best) * 0x01010101u;
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const uint32_t in = BuWord(q.a, q.b, q.c, q.d, d);
        const uint32_t cm = BuSpreadByte(comp >> (4 * d)), nm = BuSpreadByte(out_of_s >> (4 * d));
        w[d] = (in & ~(cm | nm)) | (rep_new & cm) | (rep_best & nm);
      }
    }
    uint4* dst = reinterpret_cast<uint4*>(cur_rec + static_cast<size_t>(r) * kOut);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    if (!kNibOut) dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
    cur_last[r] = static_cast<int8_t>(bv);
    cur_height[r] = static_cast<uint8_t>(best);
    lane_max = max(lane_max, static_cast<uint32_t>(best));
  }
  return lane_max;
}

// One level k (1 <= k <= n).  maxh[j] is the largest height of level j
// (maxh[0] = 0, maxh[k] cleared before the launch); it picks the slot width
// of level k - 1's records and of this level's, and gets level k's maximum.
__global__ __launch_bounds__(kBuBlock) void BottomUpLevelKernel(
    const uint32_t* __restrict__ g_adj, const uint32_t* __restrict__ g_cb, int n, int k,
    uint32_t count, const uint8_t* __restrict__ prev_rec, uint8_t* __restrict__ cur_rec,
    int8_t* __restrict__ cur_last, uint8_t* __restrict__ cur_height,
    uint32_t* __restrict__ maxh) {
  __shared__ uint32_t s_cb[kBuCbWords];
  __shared__ uint32_t s_adj[32];
  __shared__ uint32_t s_spread[256];
  __shared__ uint32_t s_max;
  for (int i = threadIdx.x; i < kBuCbWords; i += blockDim.x) s_cb[i] = g_cb[i];
  for (int i = threadIdx.x; i < 256; i += blockDim.x) s_spread[i] = BuSpreadNib(i);
  if (threadIdx.x < 32) s_adj[threadIdx.x] = threadIdx.x < static_cast<unsigned>(n) ? g_adj[threadIdx.x] : 0u;
  if (threadIdx.x == 0) s_max = 0;
  __syncthreads();
  const uint16_t* tail = reinterpret_cast<const uint16_t*>(g_cb + kBuTailWord);
  const bool nib_in = BuNibLevel(k - 1, maxh), nib_out = BuNibLevel(k, maxh);
  uint32_t lane_max;
  if (nib_in && nib_out)
    lane_max = BuLevel<true, true>(s_cb, s_adj, s_spread, tail, n, k, count, prev_rec, cur_rec,
                                   cur_last, cur_height);
  else if (nib_in)
    lane_max = BuLevel<true, false>(s_cb, s_adj, s_spread, tail, n, k, count, prev_rec,
                                    cur_rec, cur_last, cur_height);
  else if (nib_out)
    lane_max = BuLevel<false, true>(s_cb, s_adj, s_spread, tail, n, k, count, prev_rec,
                                    cur_rec, cur_last, cur_height);
  else
    lane_max = BuLevel<false, false>(s_cb, s_adj, s_spread, tail, n, k, count, prev_rec,
                                     cur_rec, cur_last, cur_height);
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
