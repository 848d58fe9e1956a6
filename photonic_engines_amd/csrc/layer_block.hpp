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
template <int KIND, bool COMP, int MAXD>
__global__ __launch_bounds__(kBlockThreads) void k_layer_block(BlockTables t, BlockLayer a) {
  __shared__ float2 ucol[256];
  const int64_t b = blockIdx.y;
  for (int j = threadIdx.x; j < a.m; j += blockDim.x)
    ucol[j] = a.U[(size_t)b * a.m * a.m + (size_t)j * a.m + a.col];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t wg_l = xcd_remap(blockIdx.x, gridDim.x);
  uint64_t g = ((uint64_t)wg_l * (kBlockThreads / 64) + wv) * a.per_wave;
  if (g >= a.S) return;
  const uint64_t g_end = (g + a.per_wave < a.S) ? g + a.per_wave : a.S;

  const int K = a.K;
  const int d = K - kSuf;
  const float2* par = a.par + (size_t)b * a.S_par;

  int q[MAXD];
  uint64_t u0 = unrank_prefix<MAXD>(t, K, d, g, q);
  BlockInfo bi;
  Streams<MAXD> st;
  while (true) {
    block_info<MAXD>(t, K, d, q, bi, st);
    float2 coef[MAXD];
    const float2* sp[MAXD];
#pragma unroll
    for (int s = 0; s < MAXD; ++s)
      if (s < st.count) {
        coef[s] = uniform_f2(ucol[st.mode[s]]);
        sp[s] = par + st.base[s];
      }
    const float2* lp = par + bi.LBp;
    const uint64_t avail = g_end - g;
    const uint32_t u_stop = (uint32_t)(((uint64_t)bi.L - u0 < avail) ? (uint64_t)bi.L : u0 + avail);
    for (uint32_t uc = (uint32_t)u0; uc < u_stop; uc += 64) {
      const uint32_t u = uc + lane;
      if (u < u_stop) {
        BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < MAXD; ++s)
          if (s < st.count) bk_cmac<COMP>(acc, coef[s], sp[s][u]);
        const TupleEntry e = t.tuples[bi.toff + u];
        const int x0 = e.w0 & 0xff, x1 = (e.w0 >> 8) & 0xff, x2 = (e.w0 >> 16) & 0xff;
        const uint32_t valid = (e.w0 >> 24) & 7u;
        if ((valid & 1u) && !(bi.skip0 && x0 == bi.lo)) bk_cmac<COMP>(acc, ucol[x0], lp[e.lr0]);
        if (valid & 2u) bk_cmac<COMP>(acc, ucol[x1], lp[e.lr1]);
        if (valid & 4u) bk_cmac<COMP>(acc, ucol[x2], lp[e.lr2]);
        float re = acc.re + acc.cre, im = acc.im + acc.cim;
        const uint64_t c = bi.CB + u;
        if constexpr (KIND == 0) {
          reinterpret_cast<float2*>(a.out)[(size_t)b * a.S + c] = make_float2(re, im);
        } else {
          const float f = child_norm(bi, e.w0, t.multiset);
          re *= a.scale;
          im *= a.scale;
          if constexpr (KIND == 1) {
            reinterpret_cast<float*>(a.out)[(size_t)b * a.S + c] = (re * re + im * im) * f;
          } else {
            const float s = (a.flags & 2u) ? sqrtf(f) : 1.f;
            reinterpret_cast<float2*>(a.out)[(size_t)b * a.S + c] = make_float2(re * s, im * s);
          }
        }
      }
    }
    g += u_stop - u0;
    if (g >= g_end) break;
    u0 = 0;
    if (!next_prefix(t, K, d, q)) break;
  }
}

}  // namespace slos
