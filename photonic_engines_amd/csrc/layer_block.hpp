// Block-structured pull kernel for one SLOS layer (layer K-1 -> K), the hot loop of the forward
// pass (replaces the per-slab gather / scatter_add_ of photonic_engines/slos/kernel.py:216-227).
// The index arithmetic (blocks, aligned prefix streams, suffix-tuple table) is in block_index.hpp.
//
// A wave owns a contiguous range of children and walks the blocks it meets in order. Block data
// (start, bases of the prefix streams, their U coefficients) is wave-uniform and lives in scalar
// registers; each step covers 64 consecutive children of one block, so the child writes and every
// prefix stream are fully coalesced. Nothing is written twice and nothing is atomic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "block_index.hpp"

namespace slos {

constexpr int kBlockThreads = 256;

struct BlockLayer {
  int K, m, col;
  const float2* U;       // [B][m][m]
  const float2* par;     // [B][S_par]
  uint64_t S_par;
  void* out;             // [B][S]
  uint64_t S;
  float scale;
  unsigned flags;
  uint64_t per_wave;     // children per wave
};

// wave-uniform copy of a value the compiler cannot prove uniform (keeps it in scalar registers)
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ float2 uniform_f2(float2 v) {
  return make_float2(__builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v.x))),
                     __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v.y))));
}

struct BlockAcc {
  float re, im, cre, cim;
};

__device__ __forceinline__ void bk_two_sum(float& s, float& c, float x) {
  const float t = s + x;
  const float bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

// acc += u * v; COMP: products split exactly by fma, sums by two-sum, errors kept in cre / cim
template <bool COMP>
__device__ __forceinline__ void bk_cmac(BlockAcc& a, float2 u, float2 v) {
  if constexpr (!COMP) {
    a.re = fmaf(u.x, v.x, a.re);
    a.re = fmaf(-u.y, v.y, a.re);
    a.im = fmaf(u.x, v.y, a.im);
    a.im = fmaf(u.y, v.x, a.im);
  } else {
    float p, e;
    p = u.x * v.x; e = fmaf(u.x, v.x, -p); bk_two_sum(a.re, a.cre, p); a.cre += e;
    p = -u.y * v.y; e = fmaf(-u.y, v.y, -p); bk_two_sum(a.re, a.cre, p); a.cre += e;
    p = u.x * v.y; e = fmaf(u.x, v.y, -p); bk_two_sum(a.im, a.cim, p); a.cim += e;
    p = u.y * v.x; e = fmaf(u.y, v.x, -p); bk_two_sum(a.im, a.cim, p); a.cim += e;
  }
}

// KIND 0: intermediate layer (complex64), 1: probabilities (float32), 2: scaled amplitudes
// The counting tables are read with wave-uniform indices from read-only (__restrict__) kernel
// arguments, so the per-block arithmetic runs on the scalar unit (s_load / s_add) and the stream
// bases sit in scalar registers; only the U column goes to LDS (per-lane indexed by suffix modes).
template <int KIND, bool COMP, int MAXD>
__global__ __launch_bounds__(kBlockThreads) void k_layer_block(const uint64_t* __restrict__ cg_t,
                                                               const int* __restrict__ lo_t,
                                                               const TupleEntry* __restrict__ tuples,
                                                               int me, int stride, int multiset,
                                                               const float2* __restrict__ par_all,
                                                               void* __restrict__ out_all, BlockLayer a) {
  __shared__ float2 ucol[256];
  const int K = a.K;
  const int64_t b = blockIdx.y;
  for (int j = threadIdx.x; j < a.m; j += blockDim.x)
    ucol[j] = a.U[(size_t)b * a.m * a.m + (size_t)j * a.m + a.col];
  __syncthreads();
  const BlockTables t{cg_t, lo_t, tuples, me, stride, multiset};

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t wg_l = xcd_remap(blockIdx.x, gridDim.x);
  uint64_t g = uniform_u64(((uint64_t)wg_l * (kBlockThreads / 64) + wv) * a.per_wave);
  if (g >= a.S) return;
  const uint64_t g_end = (g + a.per_wave < a.S) ? g + a.per_wave : a.S;
  const int d = K - kSuf;
  const float2* par = par_all + (size_t)b * a.S_par;

  int q[MAXD];
  uint64_t u0 = unrank_prefix<MAXD>(t, K, d, g, q);
  BlockInfo bi;
  Streams<MAXD> st;
  while (true) {
    block_info<MAXD>(t, K, d, q, bi, st);
    // stream slots: inactive ones (repeated modes, positions >= d) re-read an active stream's
    // address with a zero coefficient, so the chunk body below has no branches and all of its
    // loads are in flight together
    float2 coef[MAXD];
    const float2* sp[MAXD];
    const float2* sp_any = par;
#pragma unroll
    for (int s = 0; s < MAXD; ++s)
      if (st.active[s]) sp_any = par + st.base[s];
#pragma unroll
    for (int s = 0; s < MAXD; ++s) {
      coef[s] = st.active[s] ? uniform_f2(ucol[st.mode[s]]) : make_float2(0.f, 0.f);
      sp[s] = st.active[s] ? par + st.base[s] : sp_any;
    }
    const float2* lp = par + bi.LBp;
    const TupleEntry* tb = t.tuples + bi.toff;
    const uint64_t avail = g_end - g;
    const uint32_t u_stop = (uint32_t)(((uint64_t)bi.L - u0 < avail) ? (uint64_t)bi.L : u0 + avail);
    for (uint32_t uc = (uint32_t)u0; uc < u_stop; uc += 64) {
      const uint32_t u_raw = uc + lane;
      const bool in = u_raw < u_stop;
      const uint32_t u = in ? u_raw : u_stop - 1;  // tail lanes re-read a valid child, store nothing
      float2 sv[MAXD];
#pragma unroll
      for (int s = 0; s < MAXD; ++s) sv[s] = sp[s][u];
      const TupleEntry e = tb[u];
      float2 lv[kSuf];
#pragma unroll
      for (int j = 0; j < kSuf; ++j) lv[j] = lp[e.lr[j]];  // every lr_j indexes a real parent
      BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < MAXD; ++s) bk_cmac<COMP>(acc, coef[s], sv[s]);
#pragma unroll
      for (int j = 0; j < kSuf; ++j) {
        const int xj = (int)((e.pos >> (8 * j)) & 0xffu);
        bool on = (e.flags >> j) & 1u;
        if (j == 0) on = on && !(bi.skip0 && xj == bi.lo);
        const float2 cu = ucol[xj];
        bk_cmac<COMP>(acc, on ? cu : make_float2(0.f, 0.f), lv[j]);
      }
      float re = acc.re + acc.cre, im = acc.im + acc.cim;
      const uint64_t c = bi.CB + u;
      if constexpr (KIND == 0) {
        if (in) reinterpret_cast<float2*>(out_all)[(size_t)b * a.S + c] = make_float2(re, im);
      } else {
        const float f = child_norm(bi, e, t.multiset);
        re *= a.scale;
        im *= a.scale;
        if constexpr (KIND == 1) {
          if (in) reinterpret_cast<float*>(out_all)[(size_t)b * a.S + c] = (re * re + im * im) * f;
        } else {
          const float s = (a.flags & 2u) ? sqrtf(f) : 1.f;
          if (in) reinterpret_cast<float2*>(out_all)[(size_t)b * a.S + c] = make_float2(re * s, im * s);
        }
      }
    }
    g += u_stop - u0;
    if (g >= g_end) break;
    u0 = 0;
    if (!next_prefix<MAXD>(t, K, d, q)) break;
  }
}

}  // namespace slos
