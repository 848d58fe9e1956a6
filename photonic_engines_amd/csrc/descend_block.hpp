// Block-structured pull over children for one step of the uniform-loss recurrence
// (photonic_engines/imperfections.py:254-276): q_{K-1}(t) = sum_a q_K(t + e_a) (t_a + 1) / K, for
// multiset spaces (lo[x] = x). The index arithmetic (segment bases, tail offsets) is in
// block_index.hpp.
//
// Work split, as in the forward block kernel: a wave owns a contiguous range of parents of layer
// K-1. It computes the data of 64 consecutive parent blocks at once (lane i: block r + i, with the
// counting tables in LDS) into its block table in LDS, then takes the blocks one at a time with
// wave-uniform LDS reads and sweeps their parents 64 at a time, one per lane. For each mode a in
// ascending order a lane adds q_K(child) * copies: modes below the block's last prefix position
// reach a child at a block-uniform base plus the lane's offset (a scalar base, coalesced loads);
// the others reach it at a per-lane offset from the first child of block (P, last(P)). The sums
// run in the same order, with the same fused multiply-adds, as the thread-per-state kernel, so the
// two give bit-identical results. No atomics; each parent is written by one lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layer_block.hpp"

namespace slos {

// per-wave descent block table: CBp (2), CB0 (2), L, toff, lastP | lastrun << 8, the prefix
// positions (4 per word), the runs (4 per word), the segment bases (2 each)
__host__ __device__ constexpr int desc_words(int maxd) { return 7 + 2 * ((maxd + 3) / 4) + 2 * maxd; }
__host__ __device__ inline size_t desc_table_offset(int K, int me) { return (block_lds_bytes(K, me) + 15) & ~(size_t)15; }
__host__ __device__ inline size_t desc_dyn_lds(int K, int me, int maxd) {
  return desc_table_offset(K, me) + (size_t)kWaves * kBlockBatch * desc_words(maxd) * 4;
}

// per-block state of the parent sweep (wave-uniform except the lane-held tables)
struct DescCtx {
  const float* qk;        // child layer
  float* qout;            // parent layer
  const TupleWord* tb;    // tuple words of the block's parents
  const uint64_t* cgl;    // counting table (LDS)
  uint64_t CBp, CB0;
  uint64_t sb;            // lane a: base of mode a's child stream
  float scnt;             // lane a < last(P): copies of a in the children
  uint32_t tA0, tA1, tA2, tA3;  // lane a: cg[4][a+1], cg[3][a+1], cg[2][a+1], cg[1][a]
  uint32_t u_stop;
  int W, me, lastP, lastrun, lane;
  float invk, scale;
};

// Modes are taken in chunks of kDescChunk: a chunk's child loads are all issued before its sums
// (ascending modes throughout, so the sums and their order are those of the thread kernel). A chunk
// keeps only its own per-mode wave-uniform values (stream bases, rank terms, copies) in scalar
// registers; loading every mode's before the first sum needs ~7 scalars a mode and spills them.
constexpr int kDescChunk = 8;

// q_K of one child (mode a) of each parent held by a lane, one parent per lane
template <int ME>
__device__ __forceinline__ void descend_step1(const DescCtx& c, uint32_t uc) {
  const uint32_t u = uc + (uint32_t)c.lane;
  const bool in = u < c.u_stop;
  const uint32_t uu = in ? u : c.u_stop - 1u;
  const TupleWord w = c.tb[uu];
  const int sx[kSuf] = {(int)(w & 63u), (int)((w >> 8) & 63u), (int)((w >> 16) & 63u), (int)((w >> 24) & 63u)};
  DescLane dl;
  desc_lane_init(c.cgl, c.W, c.lastP, sx, dl);
  float acc = 0.f;
#pragma unroll
  for (int a0 = 0; a0 < ME; a0 += kDescChunk) {
    if (a0 >= c.me) break;
    float qv[kDescChunk];
#pragma unroll
    for (int k = 0; k < kDescChunk; ++k) {
      const int a = a0 + k;
      qv[k] = 0.f;
      if (a < c.me) {
        const uint64_t sba = rl64(c.sb, a);
        if (a < c.lastP) {
          qv[k] = at_bytes(c.qk + sba, uu * 4u);
        } else {
          const uint32_t loc = desc_local_offset(dl, sx, a, rl(c.tA0, a), rl(c.tA1, a), rl(c.tA2, a), rl(c.tA3, a));
          const uint64_t idx = a <= sx[0] ? sba + uu : c.CB0 + loc;
          qv[k] = c.qk[idx];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kDescChunk; ++k) {
      const int a = a0 + k;
      if (a < c.me) {
        const float cnt = a < c.lastP ? __builtin_bit_cast(float, rl(__builtin_bit_cast(uint32_t, c.scnt), a))
                                      : (float)desc_tail_copies(sx, a, c.lastP, c.lastrun);
        acc = fmaf(qv[k], cnt, acc);
      }
    }
  }
  if (in) c.qout[c.CBp + u] = (acc * c.invk) * c.scale;
}

// 128 parents from uc (uc + 128 <= u_stop), two adjacent per lane. A mode below last(P), or not
// above the step's smallest s0 (lane 0's first parent: s0 never decreases inside a block), reaches
// both children through one 8-byte load of its stream; the other modes load per parent.
template <int ME>
__device__ __forceinline__ void descend_step2(const DescCtx& c, uint32_t uc) {
  const uint32_t u = uc + 2u * (uint32_t)c.lane;
  uint2 w2;
  __builtin_memcpy(&w2, reinterpret_cast<const char*>(c.tb) + u * 4u, 8);
  const int s0x[kSuf] = {(int)(w2.x & 63u), (int)((w2.x >> 8) & 63u), (int)((w2.x >> 16) & 63u), (int)((w2.x >> 24) & 63u)};
  const int s1x[kSuf] = {(int)(w2.y & 63u), (int)((w2.y >> 8) & 63u), (int)((w2.y >> 16) & 63u), (int)((w2.y >> 24) & 63u)};
  const int s0min = (int)rfl((uint32_t)s0x[0]);
  DescLane d0, d1;
  desc_lane_init(c.cgl, c.W, c.lastP, s0x, d0);
  desc_lane_init(c.cgl, c.W, c.lastP, s1x, d1);
  float acc0 = 0.f, acc1 = 0.f;
#pragma unroll
  for (int a0 = 0; a0 < ME; a0 += kDescChunk) {
    if (a0 >= c.me) break;
    float q0[kDescChunk], q1[kDescChunk];
#pragma unroll
    for (int k = 0; k < kDescChunk; ++k) {
      const int a = a0 + k;
      q0[k] = q1[k] = 0.f;
      if (a < c.me) {
        const uint64_t sba = rl64(c.sb, a);
        if (a < c.lastP || a <= s0min) {
          float2 v;
          __builtin_memcpy(&v, reinterpret_cast<const char*>(c.qk + sba) + u * 4u, 8);
          q0[k] = v.x;
          q1[k] = v.y;
        } else {
          const uint32_t A0 = rl(c.tA0, a), A1 = rl(c.tA1, a), A2 = rl(c.tA2, a), A3 = rl(c.tA3, a);
          const uint64_t i0 = a <= s0x[0] ? sba + u : c.CB0 + desc_local_offset(d0, s0x, a, A0, A1, A2, A3);
          const uint64_t i1 = a <= s1x[0] ? sba + u + 1u : c.CB0 + desc_local_offset(d1, s1x, a, A0, A1, A2, A3);
          q0[k] = c.qk[i0];
          q1[k] = c.qk[i1];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kDescChunk; ++k) {
      const int a = a0 + k;
      if (a < c.me) {
        if (a < c.lastP) {
          const float cnt = __builtin_bit_cast(float, rl(__builtin_bit_cast(uint32_t, c.scnt), a));
          acc0 = fmaf(q0[k], cnt, acc0);
          acc1 = fmaf(q1[k], cnt, acc1);
        } else {
          acc0 = fmaf(q0[k], (float)desc_tail_copies(s0x, a, c.lastP, c.lastrun), acc0);
          acc1 = fmaf(q1[k], (float)desc_tail_copies(s1x, a, c.lastP, c.lastrun), acc1);
        }
      }
    }
  }
  const float2 o = make_float2((acc0 * c.invk) * c.scale, (acc1 * c.invk) * c.scale);
  __builtin_memcpy(c.qout + c.CBp + u, &o, 8);
}

// The block walk of the pulls over children: the wave's parents [g, g + per_wave) of layer K-1,
// block by block (64 blocks' data computed at once into the wave's table `tbl`, then read back one
// block at a time); for each block, sweep(c, u0, u_stop) runs over the block's parents
// [u0, u_stop) with the block's child-index context c.
template <int MAXD, class Sweep>
__device__ __forceinline__ void desc_block_walk(const BlockTables& tg, const BlockTables& tl, int K, const float* qk,
                                                float* qout, uint64_t Sp, uint64_t S_pre, uint64_t per_wave,
                                                uint32_t* tbl, uint64_t g, float scale, Sweep&& sweep) {
  constexpr int kPW = (MAXD + 3) / 4;
  constexpr int kBW = 7 + 2 * kPW;
  const int lane = threadIdx.x & 63;
  const int D = K - 1, dp = D - kSuf;
  const int W = tg.stride, me = tg.me;
  const uint64_t* cgl = tl.cg;
  if (g >= Sp) return;
  const uint64_t g_end = (g + per_wave < Sp) ? g + per_wave : Sp;
  const float invk = 1.f / (float)K;
  // lane a holds the per-mode rank terms of mode a: cg[4][a+1], cg[3][a+1], cg[2][a+1], cg[1][a], cg[5][a+1]
  uint32_t tA0 = 0, tA1 = 0, tA2 = 0, tA3 = 0, tC5 = 0;
  if (lane < me) {
    tA0 = (uint32_t)cgl[4 * W + lane + 1];
    tA1 = (uint32_t)cgl[3 * W + lane + 1];
    tA2 = (uint32_t)cgl[2 * W + lane + 1];
    tA3 = (uint32_t)cgl[W + lane];
    tC5 = (uint32_t)cgl[5 * W + lane + 1];
  }
  uint64_t r, u0;
  {
    int q[MAXD];
    u0 = unrank_prefix<MAXD>(tg, D, dp, g, q);
    r = prefix_rank<MAXD>(tg, dp, q);
  }
  while (g < g_end) {
    if (r >= S_pre) break;  // cannot happen while g < Sp; keeps the loop bounded regardless
    {
      DescBlock<MAXD> db;
      const uint64_t ri = r + (uint64_t)lane;
      if (ri < S_pre) {
        int q[MAXD];
        const SpaceView sv{tl.cg, tl.lo, me, W};
        sv.unrank(ri, dp, q);
        descend_block_info<MAXD>(tl, K, dp, q, db);
      } else {
        db.CBp = db.CB0 = 0;
        db.L = 0;
        db.toff = 0;
        db.lastP = db.lastrun = 0;
#pragma unroll
        for (int i = 0; i < MAXD; ++i) {
          db.p[i] = db.run[i] = 0;
          db.B[i] = 0;
        }
      }
      wave_sync();  // every lane is done reading the previous batch
      uint32_t* col = tbl + lane;
      col[0 * kBlockBatch] = (uint32_t)db.CBp;
      col[1 * kBlockBatch] = (uint32_t)(db.CBp >> 32);
      col[2 * kBlockBatch] = (uint32_t)db.CB0;
      col[3 * kBlockBatch] = (uint32_t)(db.CB0 >> 32);
      col[4 * kBlockBatch] = db.L;
      col[5 * kBlockBatch] = db.toff;
      col[6 * kBlockBatch] = (uint32_t)db.lastP | ((uint32_t)db.lastrun << 8);
#pragma unroll
      for (int w = 0; w < kPW; ++w) {
        uint32_t pw = 0, rw = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int i = 4 * w + k;
          if (i < MAXD) {
            pw |= ((uint32_t)db.p[i < MAXD ? i : 0] & 0xffu) << (8 * k);
            rw |= ((uint32_t)db.run[i < MAXD ? i : 0] & 0xffu) << (8 * k);
          }
        }
        col[(7 + w) * kBlockBatch] = pw;
        col[(7 + kPW + w) * kBlockBatch] = rw;
      }
#pragma unroll
      for (int i = 0; i < MAXD; ++i) {
        col[(kBW + 2 * i) * kBlockBatch] = (uint32_t)db.B[i];
        col[(kBW + 2 * i + 1) * kBlockBatch] = (uint32_t)(db.B[i] >> 32);
      }
      wave_sync();
    }
    const int nb = (int)((S_pre - r < kBlockBatch) ? S_pre - r : kBlockBatch);
    for (int j = 0; j < nb && g < g_end; ++j) {
      const uint32_t* bj = tbl + j;
      auto word = [&](int f) { return rfl(bj[f * kBlockBatch]); };
      const uint32_t L = word(4);
      if (L == 0) continue;
      const uint64_t CBp = (uint64_t)word(0) | ((uint64_t)word(1) << 32);
      const uint64_t CB0 = (uint64_t)word(2) | ((uint64_t)word(3) << 32);
      const uint32_t toff = word(5);
      const uint32_t w6 = word(6);
      const int lastP = (int)(w6 & 0xffu), lastrun = (int)(w6 >> 8);
      const uint32_t c5l = (uint32_t)cgl[5 * W + lastP + 1];
      // lane a: the base of mode a's child stream (child = sb + u while a <= s0) and, below
      // last(P), the copies of a in its children (the same for every parent of the block)
      uint64_t sb = 0;
      float scnt = 1.f;
      {
        const int a = lane;
        if (a < lastP) {
          int i = 0, lo_i = 0, run_i = 0;
          uint64_t Bi = (uint64_t)word(kBW) | ((uint64_t)word(kBW + 1) << 32);
#pragma unroll
          for (int l = 1; l < MAXD; ++l)
            if (l < dp) {
              const int pl = (int)((word(7 + (l - 1) / 4) >> (8 * ((l - 1) % 4))) & 0xffu);
              if (pl <= a) {
                i = l;
                lo_i = pl;
                run_i = (int)((word(7 + kPW + l / 4) >> (8 * (l % 4))) & 0xffu);
                Bi = (uint64_t)word(kBW + 2 * l) | ((uint64_t)word(kBW + 2 * l + 1) << 32);
              }
            }
          sb = Bi - cgl[(K - i) * W + a + 1];
          scnt = (i > 0 && a == lo_i) ? (float)(run_i + 1) : 1.f;
        } else if (a < me) {
          sb = CB0 + c5l - tC5;
        }
      }
      const uint64_t avail = g_end - g;
      const uint32_t u_stop = (uint32_t)(((uint64_t)L - u0 < avail) ? (uint64_t)L : u0 + avail);
      DescCtx c{qk, qout, static_cast<const TupleWord*>(tg.tuples) + toff, cgl, CBp, CB0, sb, scnt, tA0, tA1, tA2, tA3, u_stop, W, me, lastP, lastrun,
                lane, invk, scale};
      sweep(c, (uint32_t)u0, u_stop);
      g += u_stop - u0;
      u0 = 0;
    }
    r += kBlockBatch;
  }
}

// ME: a compile-time bound on the mode count (me <= ME <= 64), so that the mode loops unroll and
// every child load of a step is issued before the first sum.
template <int MAXD, int ME>
__device__ __forceinline__ void descend_wave(const BlockTables& tg, const BlockTables& tl, int K, const float* __restrict__ qk,
                                             float* __restrict__ qout, uint64_t Sp, uint64_t S_pre, uint64_t per_wave,
                                             uint32_t* tbl, uint64_t g, float scale) {
  desc_block_walk<MAXD>(tg, tl, K, qk, qout, Sp, S_pre, per_wave, tbl, g, scale,
                        [](const DescCtx& c, uint32_t uc, uint32_t u_stop) {
                          for (; uc + 128u <= u_stop; uc += 128u) descend_step2<ME>(c, uc);
                          for (; uc < u_stop; uc += 64u) descend_step1<ME>(c, uc);
                        });
}

template <int MAXD, int ME>
__global__ __launch_bounds__(kBlockThreads) void k_descend_block(const uint64_t* __restrict__ cg_t, const int* __restrict__ lo_t,
                                                                 const TupleWord* __restrict__ tuples, int me, int stride,
                                                                 int K, const float* __restrict__ qk_all, uint64_t Sk,
                                                                 float* __restrict__ qout_all, uint64_t Sp, uint64_t S_pre,
                                                                 uint64_t per_wave, float scale, uint32_t xcd_chunk) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int ncg = (K + 1) * stride;
  uint64_t* cg_l = reinterpret_cast<uint64_t*>(smem);
  int* lo_l = reinterpret_cast<int*>(cg_l + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg_l[i] = cg_t[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo_l[i] = lo_t[i];
  __syncthreads();
  const BlockTables tg{cg_t, lo_t, tuples, me, stride, 1};
  const BlockTables tl{cg_l, lo_l, tuples, me, stride, 1};
  const int64_t b = blockIdx.y;
  const int wv = (int)rfl(threadIdx.x >> 6);
  // work groups in XCD chunks, as the forward block kernel
  const uint64_t g = uniform_u64(((uint64_t)xcd_chunk_index(blockIdx.x, xcd_chunk) * kWaves + wv) * per_wave);
  uint32_t* tbl = reinterpret_cast<uint32_t*>(smem + desc_table_offset(K, me)) + (size_t)wv * kBlockBatch * desc_words(MAXD);
  descend_wave<MAXD, ME>(tg, tl, K, qk_all + (size_t)b * Sk, qout_all + (size_t)b * Sp, Sp, S_pre, per_wave, tbl, g, scale);
}


// ------------------------------------------------------------------------------------------
// Adjoint of one forward layer on the same block structure (the backward of the recurrence,
// photonic_engines/slos/kernel.py:63-80, as a pull over children): for every parent t of layer
// K-1 and every mode a, the child t + e_a is an edge with coefficient U[a, col], so
//   grad_prev[t] = sum_a conj(U[a, col]) g_K[t + e_a]          (ascending a)
//   dL/dU[a, col] += sum_t conj(psi_{K-1}[t]) g_K[t + e_a].
// One parent per lane; the children are found exactly as in the loss descent. The U-gradient is
// kept per lane and per mode in registers (mode a is wave-uniform in the unrolled mode loop),
// summed over the wave's lanes in FP64 and over the work group's waves, and written as one FP64
// partial per work group and mode; a second kernel sums the partials in FP64. No atomics.
// ------------------------------------------------------------------------------------------
struct AdjCtx {
  const float2* gk;     // g_K (child layer)
  float2* gout;         // grad_prev (parent layer), or nullptr
  const float2* psi;    // psi_{K-1} (parent layer)
  uint32_t ure, uim;    // lane a: U[a, col]
};

template <int ME>
__device__ __forceinline__ void adjoint_step1(const DescCtx& c, const AdjCtx& x, uint32_t uc, float2 (&gacc)[ME]) {
  const uint32_t u = uc + (uint32_t)c.lane;
  const bool in = u < c.u_stop;
  const uint32_t uu = in ? u : c.u_stop - 1u;
  const TupleWord w = c.tb[uu];
  const int sx[kSuf] = {(int)(w & 63u), (int)((w >> 8) & 63u), (int)((w >> 16) & 63u), (int)((w >> 24) & 63u)};
  DescLane dl;
  desc_lane_init(c.cgl, c.W, c.lastP, sx, dl);
  const float2 ps = in ? x.psi[c.CBp + uu] : make_float2(0.f, 0.f);  // lanes past the block add nothing
  float gr = 0.f, gi = 0.f;
  constexpr int kChunk = 8;
#pragma unroll
  for (int a0 = 0; a0 < ME; a0 += kChunk) {
    float2 gv[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int a = a0 + i;
      gv[i] = make_float2(0.f, 0.f);
      if (a < c.me) {
        const uint64_t sba = rl64(c.sb, a);
        uint64_t idx;
        if (a < c.lastP) {
          idx = sba + uu;
        } else {
          const uint32_t loc = desc_local_offset(dl, sx, a, rl(c.tA0, a), rl(c.tA1, a), rl(c.tA2, a), rl(c.tA3, a));
          idx = a <= sx[0] ? sba + uu : c.CB0 + loc;
        }
        gv[i] = x.gk[idx];
      }
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int a = a0 + i;
      if (a < c.me) {
        const float ur = __builtin_bit_cast(float, rl(x.ure, a)), ui = __builtin_bit_cast(float, rl(x.uim, a));
        gr = fmaf(ur, gv[i].x, gr);   // conj(u) * g
        gr = fmaf(ui, gv[i].y, gr);
        gi = fmaf(ur, gv[i].y, gi);
        gi = fmaf(-ui, gv[i].x, gi);
        gacc[a].x = fmaf(ps.x, gv[i].x, gacc[a].x);   // conj(psi) * g
        gacc[a].x = fmaf(ps.y, gv[i].y, gacc[a].x);
        gacc[a].y = fmaf(ps.x, gv[i].y, gacc[a].y);
        gacc[a].y = fmaf(-ps.y, gv[i].x, gacc[a].y);
      }
    }
  }
  if (x.gout && in) x.gout[c.CBp + u] = make_float2(gr, gi);
}

// partial[b][work group][a] (double2), the U-gradient sums of the work group's parents
template <int MAXD, int ME>
__global__ __launch_bounds__(kBlockThreads) void k_adjoint_block(const uint64_t* __restrict__ cg_t, const int* __restrict__ lo_t,
                                                                 const TupleWord* __restrict__ tuples, int me, int stride,
                                                                 int K, int m, int col, const float2* __restrict__ U,
                                                                 const float2* __restrict__ psi_all,
                                                                 const float2* __restrict__ gk_all, uint64_t Sk,
                                                                 float2* __restrict__ gout_all, uint64_t Sp, uint64_t S_pre,
                                                                 uint64_t per_wave, uint32_t xcd_chunk,
                                                                 double2* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double2 wsum[kWaves][ME];
  const int ncg = (K + 1) * stride;
  uint64_t* cg_l = reinterpret_cast<uint64_t*>(smem);
  int* lo_l = reinterpret_cast<int*>(cg_l + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg_l[i] = cg_t[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo_l[i] = lo_t[i];
  __syncthreads();
  const BlockTables tg{cg_t, lo_t, tuples, me, stride, 1};
  const BlockTables tl{cg_l, lo_l, tuples, me, stride, 1};
  const int64_t b = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int wv = (int)rfl(threadIdx.x >> 6);
  AdjCtx x;
  x.gk = gk_all + (size_t)b * Sk;
  x.gout = gout_all ? gout_all + (size_t)b * Sp : nullptr;
  x.psi = psi_all + (size_t)b * Sp;
  const float2 ua = lane < m ? U[(size_t)b * m * m + (size_t)lane * m + col] : make_float2(0.f, 0.f);
  x.ure = __builtin_bit_cast(uint32_t, ua.x);
  x.uim = __builtin_bit_cast(uint32_t, ua.y);
  float2 gacc[ME];
#pragma unroll
  for (int a = 0; a < ME; ++a) gacc[a] = make_float2(0.f, 0.f);
  const uint64_t g = uniform_u64(((uint64_t)xcd_chunk_index(blockIdx.x, xcd_chunk) * kWaves + wv) * per_wave);
  uint32_t* tbl = reinterpret_cast<uint32_t*>(smem + desc_table_offset(K, me)) + (size_t)wv * kBlockBatch * desc_words(MAXD);
  desc_block_walk<MAXD>(tg, tl, K, nullptr, nullptr, Sp, S_pre, per_wave, tbl, g, 1.f,
                        [&](const DescCtx& c, uint32_t uc, uint32_t u_stop) {
                          for (; uc < u_stop; uc += 64u) adjoint_step1<ME>(c, x, uc, gacc);
                        });
  // wave sums in FP64 (butterfly over the 64 lanes), then the work group's four waves in order
#pragma unroll
  for (int a = 0; a < ME; ++a) {
    double sr = (double)gacc[a].x, si = (double)gacc[a].y;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sr += __shfl_xor(sr, o, 64);
      si += __shfl_xor(si, o, 64);
    }
    if (lane == 0) wsum[wv][a] = make_double2(sr, si);
  }
  __syncthreads();
  for (int a = threadIdx.x; a < me; a += blockDim.x) {
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      sr += wsum[w][a].x;
      si += wsum[w][a].y;
    }
    partial[((size_t)b * gridDim.x + blockIdx.x) * me + a] = make_double2(sr, si);
  }
}

}  // namespace slos
