// Block-structured pull kernel for one SLOS layer (layer K-1 -> K), the hot loop of the forward
// pass (replaces the per-slab gather / scatter_add_ of photonic_engines/slos/kernel.py:216-227).
// The index arithmetic (blocks, aligned prefix streams, suffix-tuple table) is in block_index.hpp.
//
// Work split. A wave owns a contiguous range of children. It walks the blocks it meets in batches
// of 64 consecutive prefixes: lane i unranks prefix r + i and computes that block's data (start,
// size, stream bases, coefficients) with the counting tables in LDS, all lanes in parallel, and
// writes it to the wave's block table in LDS. The wave then takes the blocks one at a time, reading
// the block's data with wave-uniform LDS reads into scalar registers, and sweeps the block's
// children 128 at a time (two adjacent children per lane). In that sweep every address is a scalar
// base plus a 32-bit lane offset, the prefix streams and the child writes are fully coalesced, all
// loads of a step are in flight together (no branches), and nothing is written twice or
// atomically. A child's suffix edges come from its 4-byte tuple word and the suffix table in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "block_index.hpp"

namespace slos {

constexpr int kBlockThreads = 256;
constexpr int kWaves = kBlockThreads / 64;

struct BlockLayer {
  int K, m, col;
  const float2* U;       // [B][m][m]
  uint64_t S_par;        // parent layer size
  uint64_t S;            // child layer size
  uint64_t S_pre;        // number of d-prefixes (layer d of the space)
  float scale;
  unsigned flags;
  uint64_t per_wave;     // children per wave
  uint32_t xcd_chunk;    // work groups per XCD chunk (see k_layer_block); 1: plain launch order
  // compensated sums only: element offsets of the low-order parts of the parent layer and of a
  // complex64 output layer from their high parts (0: none; a hi/lo layer is [2][B][S])
  int64_t par_lo, out_lo;
};

__device__ __forceinline__ uint32_t rfl(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t rl(uint32_t v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ uint64_t rl64(uint64_t v, int lane) {
  return ((uint64_t)rl((uint32_t)(v >> 32), lane) << 32) | rl((uint32_t)v, lane);
}
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  return ((uint64_t)rfl((uint32_t)(v >> 32)) << 32) | rfl((uint32_t)v);
}
__device__ __forceinline__ float2 uniform_f2(float2 v) {
  return make_float2(__builtin_bit_cast(float, rfl(__builtin_bit_cast(uint32_t, v.x))),
                     __builtin_bit_cast(float, rfl(__builtin_bit_cast(uint32_t, v.y))));
}
// element `off` bytes past a scalar base: lets the compiler use base-in-SGPR addressing
template <class T>
__device__ __forceinline__ const T& at_bytes(const T* base, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

// |amp|^2 * f with each product and the sum rounded on their own (no fused multiply-add), so that
// every kernel writes the same float for the same amplitude whatever the compiler fuses elsewhere
__device__ __forceinline__ float prob_readout(float re, float im, float f) {
#pragma clang fp contract(off)
  return (re * re + im * im) * f;
}

struct BlockAcc {
  float re, im, cre, cim;
};

// Compensated arithmetic must not be contracted: a product fused into the next addition would make
// the two-sum's error terms wrong, so these bodies turn contraction off.
__device__ __forceinline__ void bk_two_sum(float& s, float& c, float x) {
#pragma clang fp contract(off)
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
#pragma clang fp contract(off)
    float p, e;
    p = u.x * v.x; e = fmaf(u.x, v.x, -p); bk_two_sum(a.re, a.cre, p); a.cre += e;
    p = -u.y * v.y; e = fmaf(-u.y, v.y, -p); bk_two_sum(a.re, a.cre, p); a.cre += e;
    p = u.x * v.y; e = fmaf(u.x, v.y, -p); bk_two_sum(a.im, a.cim, p); a.cim += e;
    p = u.y * v.x; e = fmaf(u.y, v.x, -p); bk_two_sum(a.im, a.cim, p); a.cim += e;
  }
}

// the low-order part w of a hi/lo parent value: acc.c += u * w (plain products into the correction)
__device__ __forceinline__ void bk_lo_mac(BlockAcc& a, float2 u, float2 w) {
  a.cre = fmaf(u.x, w.x, a.cre);
  a.cre = fmaf(-u.y, w.y, a.cre);
  a.cim = fmaf(u.x, w.y, a.cim);
  a.cim = fmaf(u.y, w.x, a.cim);
}

// One row of the per-launch suffix table (LDS, kSufRows rows, row = tuple-word byte): the U
// coefficient of the position (zero for marked rows, x | 64) and its rank terms, packed:
//   SUF 4: t0 = cg[3][x+1], t1 = cg[2][x+1] | cg[1][x] << 16;
//   SUF 5: t0 = cg[4][x+1] | cg[1][x] << 24, t1 = cg[3][x+1] | cg[2][x+1] << 16
// (the plan checks that every term fits its field).
struct SufRow {
  float re, im;
  uint32_t t0, t1;
};

// rows for U[b, :, col] (Ucol = &U[b][0][col], stride m), positions < me; cg: the global table
template <int SUF>
__device__ __forceinline__ void fill_suffix_rows(SufRow* rows, const float2* Ucol, int m, const uint64_t* cg, int stride,
                                                 int me) {
  for (int i = threadIdx.x; i < kSufRows; i += blockDim.x) {
    const int x = i & (kBlockModes - 1);
    const float2 u = (i < kBlockModes && x < m) ? Ucol[(size_t)x * m] : make_float2(0.f, 0.f);
    SufRow r;
    r.re = u.x;
    r.im = u.y;
    r.t0 = r.t1 = 0u;
    if (x < me) {
      if constexpr (SUF == 4) {
        r.t0 = (uint32_t)cg[3 * stride + x + 1];
        r.t1 = (uint32_t)cg[2 * stride + x + 1] | ((uint32_t)cg[stride + x] << 16);
      } else {
        r.t0 = (uint32_t)cg[4 * stride + x + 1] | ((uint32_t)cg[stride + x] << 24);
        r.t1 = (uint32_t)cg[3 * stride + x + 1] | ((uint32_t)cg[2 * stride + x + 1] << 16);
      }
    }
    rows[i] = r;
  }
}

// Readout norms by (row, code): row 0 is prod of tau's run factorials for the (SUF-1)-bit mark
// code; row 1 + r multiplies it by binom(r + lead, lead) (tau's leading run joins a last prefix
// run of r copies). kNormRows covers r <= 16, the widest prefix of the block kernel.
constexpr int kNormRows = 18;
constexpr int kNormCodes = 1 << (kSufMax - 1);
template <int SUF>
__device__ __forceinline__ void fill_norm_rows(float* rn) {
  constexpr int NC = 1 << (SUF - 1);
  for (int i = threadIdx.x; i < kNormRows * NC; i += blockDim.x) {
    const int row = i / NC;
    const uint32_t c = (uint32_t)(i % NC);
    const float f = (float)code_fct<SUF>(c);
    rn[i] = row == 0 ? f : f * lead_binom(row - 1, code_lead<SUF>(c));
  }
}

// per-block state of the chunk sweep (all wave-uniform except `lane`)
template <int MAXD, int SUF>
struct ChunkCtx {
  const float2* const (&sp)[MAXD];  // prefix-stream bases (parent of child u at sp[s] + u)
  const float2 (&coef)[MAXD];       // their U coefficients (zero for inactive slots)
  const float2* lp;                 // local parents: lp + lr_j
  const TupleWordT<SUF>* tb;        // tuple words of the block's children
  char* ob;                         // output of the block's first child
  const SufRow* rows;               // suffix table (LDS)
  const float* rn;                  // readout norms (LDS)
  uint32_t C0;                      // cg[SUF-1][0]
  uint32_t u_stop;
  int lane;
  float fQ;
  int lastrun, skip0, lo, multiset;
  float scale;
  unsigned flags;
  int kind;  // KIND 3 only: 1 probabilities, 0 raw amplitudes
  int64_t plo = 0, olo = 0;  // compensated sums: low-order parts of the parents / the output (elements)
};

// the suffix edges of one child from its tuple word: coefficients c[j] (edge 0 of a tau that
// continues the prefix's last run is no edge: zero) and local-parent byte offsets off[j]
template <int MAXD, int SUF>
__device__ __forceinline__ void suffix_edges(const ChunkCtx<MAXD, SUF>& cx, TupleWordT<SUF> w, float2 (&c)[SUF],
                                             uint32_t (&off)[SUF]) {
  uint32_t lr[SUF];
  if constexpr (SUF == 4) {
    uint32_t a3[4], b2[4], b1[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const SufRow r = cx.rows[tw_byte(w, j)];
      c[j] = make_float2(r.re, r.im);
      a3[j] = r.t0;
      b2[j] = r.t1 & 0xffffu;
      b1[j] = r.t1 >> 16;
    }
    suffix_parent_ranks(cx.C0, a3, b2, b1, lr);
  } else {
    const uint32_t lo32 = (uint32_t)w, hi32 = (uint32_t)(w >> 32);
    uint32_t t4[5], t3[5], t2[5], t1[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      const SufRow r = cx.rows[j < 4 ? (lo32 >> (8 * j)) & 0xffu : hi32 & 0xffu];
      c[j] = make_float2(r.re, r.im);
      t4[j] = r.t0 & 0xffffffu;
      t1[j] = r.t0 >> 24;
      t3[j] = r.t1 & 0xffffu;
      t2[j] = r.t1 >> 16;
    }
    suffix_parent_ranks5(cx.C0, t4, t3, t2, t1, lr);
  }
  if (cx.skip0 && (int)((uint32_t)w & 0xffu) == cx.lo) c[0] = make_float2(0.f, 0.f);
#pragma unroll
  for (int j = 0; j < SUF; ++j) off[j] = lr[j] * 8u;
}

// prod t_j! of a child from its tuple word and the block's prefix runs
template <int MAXD, int SUF>
__device__ __forceinline__ float child_norm_w(const ChunkCtx<MAXD, SUF>& cx, TupleWordT<SUF> w) {
  if (!cx.multiset) return 1.f;
  const uint32_t row = (cx.skip0 && (int)((uint32_t)w & 0xffu) == cx.lo) ? (uint32_t)cx.lastrun + 1u : 0u;
  return cx.fQ * cx.rn[row * (1u << (SUF - 1)) + tuple_code<SUF>(w)];
}

// NC consecutive 64-child chunks starting at uc: every load of the NC chunks is issued before any
// result is needed. Lanes past u_stop re-read the last valid child and store nothing.
template <int KIND, bool COMP, int MAXD, int NC, int SUF>
__device__ __forceinline__ void sweep(const ChunkCtx<MAXD, SUF>& cx, uint32_t uc) {
  uint32_t u[NC];
  bool in[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const uint32_t ur = uc + 64u * c + (uint32_t)cx.lane;
    in[c] = ur < cx.u_stop;
    u[c] = in[c] ? ur : cx.u_stop - 1u;
  }
  TupleWordT<SUF> tw[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) tw[c] = at_bytes(cx.tb, u[c] * (uint32_t)sizeof(TupleWordT<SUF>));
  float2 sv[NC][MAXD];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int s = 0; s < MAXD; ++s) sv[c][s] = at_bytes(cx.sp[s], u[c] * 8u);
  float2 lc[NC][SUF], lv[NC][SUF];
  uint32_t loff[NC][SUF];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    suffix_edges(cx, tw[c], lc[c], loff[c]);
#pragma unroll
    for (int jj = 0; jj < SUF; ++jj) lv[c][jj] = at_bytes(cx.lp, loff[c][jj]);
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < MAXD; ++s) bk_cmac<COMP>(acc, cx.coef[s], sv[c][s]);
#pragma unroll
    for (int jj = 0; jj < SUF; ++jj) bk_cmac<COMP>(acc, lc[c][jj], lv[c][jj]);
    if constexpr (COMP) {
      if (cx.plo) {  // the parents' low-order parts, into the correction terms (plain products)
#pragma unroll
        for (int s = 0; s < MAXD; ++s) bk_lo_mac(acc, cx.coef[s], at_bytes(cx.sp[s] + cx.plo, u[c] * 8u));
#pragma unroll
        for (int jj = 0; jj < SUF; ++jj) bk_lo_mac(acc, lc[c][jj], at_bytes(cx.lp + cx.plo, loff[c][jj]));
      }
    }
    // (plain sums: the correction terms are zero and are not added)
    float re = COMP ? acc.re + acc.cre : acc.re, im = COMP ? acc.im + acc.cim : acc.im;
    if (KIND == 0 || (KIND == 3 && cx.kind == 0)) {
      if (in[c]) *reinterpret_cast<float2*>(cx.ob + u[c] * 8u) = make_float2(re, im);
      if constexpr (COMP) {
        // the rounding error of the stored value (fast two-sum: |acc.re| >= |acc.cre| as the
        // correction is the accumulated rounding error of the sum)
        if (cx.olo && in[c])
          *reinterpret_cast<float2*>(cx.ob + u[c] * 8u + cx.olo * 8) =
              make_float2(acc.cre - (re - acc.re), acc.cim - (im - acc.im));
      }
    } else {
      const float f = child_norm_w(cx, tw[c]);
      re *= cx.scale;
      im *= cx.scale;
      if constexpr (KIND == 1 || KIND == 3) {
        if (in[c]) *reinterpret_cast<float*>(cx.ob + u[c] * 4u) = prob_readout(re, im, f);
      } else {
        const float s = (cx.flags & 2u) ? sqrtf(f) : 1.f;
        if (in[c]) *reinterpret_cast<float2*>(cx.ob + u[c] * 8u) = make_float2(re * s, im * s);
      }
    }
  }
}

// two adjacent complex64 values from an 8-byte aligned address (one 16-byte load)
__device__ __forceinline__ float4 load_pair(const float2* base, uint32_t off_bytes) {
  float4 v;
  __builtin_memcpy(&v, reinterpret_cast<const char*>(base) + off_bytes, 16);
  return v;
}

// the tuple words of lane i's two adjacent children uc + 2i and uc + 2i + 1 (one 8- or 16-byte load)
template <int SUF>
struct TuplePair {
  TupleWordT<SUF> a, b;
};
template <int MAXD, int SUF>
__device__ __forceinline__ TuplePair<SUF> load_tuple_pair(const ChunkCtx<MAXD, SUF>& cx, uint32_t uc) {
  TuplePair<SUF> w;
  __builtin_memcpy(&w, reinterpret_cast<const char*>(cx.tb) + (uc + 2u * (uint32_t)cx.lane) * sizeof(TupleWordT<SUF>),
                   sizeof(w));
  return w;
}

// 128 consecutive children from uc, all inside the block (uc + 128 <= u_stop): lane i takes the
// adjacent children uc + 2i and uc + 2i + 1, so each prefix stream is one 16-byte load per lane and
// the store is one 8- (probabilities) or 16-byte (amplitudes) store per lane. `tn` holds this
// step's tuple words on entry; if `next`, the next step's (uc + 128, inside the block) are loaded
// into it behind this step's loads, so the suffix-parent loads of a step never wait for a tuple load.
template <int KIND, bool COMP, int MAXD, int SUF>
__device__ __forceinline__ void sweep_pairs(const ChunkCtx<MAXD, SUF>& cx, uint32_t uc, TuplePair<SUF>& tn, bool next) {
  const uint32_t u = uc + 2u * (uint32_t)cx.lane;
  const TupleWordT<SUF> tw[2] = {tn.a, tn.b};
  float4 sv[MAXD];
#pragma unroll
  for (int s = 0; s < MAXD; ++s) sv[s] = load_pair(cx.sp[s], u * 8u);
  float2 lc[2][SUF], lv[2][SUF];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    uint32_t off[SUF];
    suffix_edges(cx, tw[c], lc[c], off);
#pragma unroll
    for (int jj = 0; jj < SUF; ++jj) lv[c][jj] = at_bytes(cx.lp, off[jj]);
  }
  if (next) tn = load_tuple_pair(cx, uc + 128u);
  float res[2][2];
  float f[2] = {1.f, 1.f};
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < MAXD; ++s)
      bk_cmac<COMP>(acc, cx.coef[s], c == 0 ? make_float2(sv[s].x, sv[s].y) : make_float2(sv[s].z, sv[s].w));
#pragma unroll
    for (int jj = 0; jj < SUF; ++jj) bk_cmac<COMP>(acc, lc[c][jj], lv[c][jj]);
    res[c][0] = COMP ? acc.re + acc.cre : acc.re;
    res[c][1] = COMP ? acc.im + acc.cim : acc.im;
    if (!(KIND == 0 || (KIND == 3 && cx.kind == 0))) f[c] = child_norm_w(cx, tw[c]);
  }
  if (KIND == 0 || (KIND == 3 && cx.kind == 0)) {
    const float4 v = make_float4(res[0][0], res[0][1], res[1][0], res[1][1]);
    __builtin_memcpy(reinterpret_cast<float2*>(cx.ob) + u, &v, 16);
  } else {
    float re0 = res[0][0] * cx.scale, im0 = res[0][1] * cx.scale;
    float re1 = res[1][0] * cx.scale, im1 = res[1][1] * cx.scale;
    if constexpr (KIND == 1 || KIND == 3) {
      const float2 v = make_float2(prob_readout(re0, im0, f[0]), prob_readout(re1, im1, f[1]));
      __builtin_memcpy(reinterpret_cast<float*>(cx.ob) + u, &v, 8);
    } else {
      const float s0 = (cx.flags & 2u) ? sqrtf(f[0]) : 1.f, s1 = (cx.flags & 2u) ? sqrtf(f[1]) : 1.f;
      const float4 v = make_float4(re0 * s0, im0 * s0, re1 * s1, im1 * s1);
      __builtin_memcpy(reinterpret_cast<float2*>(cx.ob) + u, &v, 16);
    }
  }
}

// Dynamic LDS of the block kernels: counting-table rows 0..K and lo[] (per-lane block data is
// computed from them); the suffix table and the readout norms are static.
__host__ __device__ inline size_t block_lds_bytes(int K, int me) { return (size_t)(K + 1) * (me + 1) * 8 + (size_t)me * 4; }
constexpr size_t kBlockStaticLds = kSufRows * sizeof(SufRow) + kNormRows * kNormCodes * sizeof(float);  // rows, rn
// Per-wave block table in LDS: the data of kBlockBatch consecutive blocks, one column per block,
// computed by all lanes at once (lane i: block r + i) and then read back one block at a time with
// wave-uniform LDS reads, so none of it occupies vector registers during the child sweeps. Words:
// CB (2), LBp (2), L, toff, lo | skip0 << 8 | lastrun << 16, fQ, the slot modes (4 per word,
// kNoEdgeRow for inactive slots), the slot bases (2 each; inactive slots carry slot 0's base).
constexpr int kBlockBatch = 64;
__host__ __device__ constexpr int block_words(int maxd) { return 8 + (maxd + 3) / 4 + 2 * maxd; }
__host__ __device__ inline size_t block_table_offset(int K, int me) { return (block_lds_bytes(K, me) + 15) & ~(size_t)15; }
__host__ __device__ inline size_t block_dyn_lds(int K, int me, int maxd) {
  return block_table_offset(K, me) + (size_t)kWaves * kBlockBatch * block_words(maxd) * 4;
}

// Work group `bid` of a 1-D grid in XCD chunks of C: the dispatcher deals work groups
// round-robin to the 8 XCDs, and work group i = 8 t + x (XCD x) takes range 8 C (t / C) + C x + t % C,
// so each XCD sweeps runs of C consecutive ranges while the eight XCDs stay in one region. The work
// groups past the last whole multiple of 8 C keep their own index (a bijection of [0, gridDim.x)).
__device__ __forceinline__ uint32_t xcd_chunk_index(uint32_t bid, uint32_t C) {
  if (C <= 1) return bid;
  const uint32_t full = (gridDim.x / (8u * C)) * (8u * C);
  if (bid >= full) return bid;
  const uint32_t x = bid & 7u, t = bid >> 3;
  return (t / C) * 8u * C + x * C + (t % C);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the wave-uniform data of one block's sweep (see block_children)
struct BlockSweep {
  const float2* lp;     // local parents: lp + lr_j
  const void* tb;       // tuple words of the block's children
  char* ob;             // output of the block's first child
  const SufRow* rows;
  const float* rn;
  uint32_t C0, u0, u_stop;
  int lane;
  float fQ;
  int lastrun, skip0, lo, multiset;
  float scale;
  unsigned flags;
  int kind;
  int64_t plo, olo;
};

// Children [u0, u_stop) of one block with its first NS stream slots of the block table column `bj`
// (slots past the block's active count have zero coefficients).
template <int KIND, bool COMP, int MAXD, int NS, int SUF>
__device__ __forceinline__ void block_children(const uint32_t* bj, const float2* par, const BlockSweep& bs) {
  constexpr int kBaseW = 8 + (MAXD + 3) / 4;
  auto word = [&](int f) { return rfl(bj[f * kBlockBatch]); };
  float2 coef[NS];
  const float2* sp[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const uint32_t md = (word(8 + s / 4) >> (8 * (s % 4))) & 0xffu;
    const SufRow rw = bs.rows[md];
    coef[s] = uniform_f2(make_float2(rw.re, rw.im));
    sp[s] = par + ((uint64_t)word(kBaseW + 2 * s) | ((uint64_t)word(kBaseW + 2 * s + 1) << 32));
  }
  ChunkCtx<NS, SUF> cx{sp,       coef,    bs.lp,       static_cast<const TupleWordT<SUF>*>(bs.tb),
                       bs.ob,    bs.rows, bs.rn,       bs.C0,    bs.u_stop, bs.lane, bs.fQ, bs.lastrun,
                       bs.skip0, bs.lo,   bs.multiset, bs.scale, bs.flags,  bs.kind, bs.plo, bs.olo};
  const uint32_t u_stop = bs.u_stop;
  uint32_t uc = bs.u0;
  if constexpr (!COMP) {
    // full groups of 128 children as adjacent pairs per lane, the rest 64 children per step
    if (uc + 128u <= u_stop) {
      TuplePair<SUF> tn = load_tuple_pair(cx, uc);
      for (; uc + 128u <= u_stop; uc += 128u) sweep_pairs<KIND, COMP, NS, SUF>(cx, uc, tn, uc + 256u <= u_stop);
    }
    if (uc + 64u < u_stop) sweep<KIND, COMP, NS, 2, SUF>(cx, uc);
    else if (uc < u_stop) sweep<KIND, COMP, NS, 1, SUF>(cx, uc);
  } else {
    // the compensated sums keep one child per lane (register pressure)
    for (; uc + 64u < u_stop; uc += 128u) sweep<KIND, COMP, NS, 2, SUF>(cx, uc);
    if (uc < u_stop) sweep<KIND, COMP, NS, 1, SUF>(cx, uc);
  }
}

// One wave's share of a layer slice: the children [g, g + a.per_wave) of one instance, whose parent
// layer is `par` and whose output layer starts at `out`. `rows` is the suffix table and `rn` the
// readout norms (LDS), `tl` the counting tables in LDS, `tg` the same tables in global memory (for
// the wave-uniform scalar unrank of g), `tbl` this wave's block table (kBlockBatch *
// block_words(MAXD) words of LDS).
// KIND 3 decides between probabilities (kind 1) and raw amplitudes (kind 0) at run time.
template <int KIND, bool COMP, int MAXD, int SUF>
__device__ __forceinline__ void block_wave(const BlockTables& tg, const BlockTables& tl, const float2* __restrict__ par,
                                           void* __restrict__ out, const SufRow* rows, const float* rn,
                                           uint32_t* tbl, const BlockLayer& a, uint64_t g, int kind = 0) {
  constexpr int kModeW = (MAXD + 3) / 4;
  constexpr int kBaseW = 8 + kModeW;
  const int lane = threadIdx.x & 63;
  const int K = a.K;
  const int me = tg.me, stride = tg.stride, multiset = tg.multiset;
  const uint64_t* cg_l = tl.cg;
  const int* lo_l = tl.lo;
  const TupleWordT<SUF>* tuples = static_cast<const TupleWordT<SUF>*>(tg.tuples);
  const uint32_t C0 = (uint32_t)tg.cg[(SUF - 1) * stride];
  if (g >= a.S) return;
  const uint64_t g_end = (g + a.per_wave < a.S) ? g + a.per_wave : a.S;
  const int d = K - SUF;

  // the block holding g: its prefix rank r and g's offset inside it (scalar path)
  uint64_t r, u0;
  {
    int q[MAXD];
    u0 = unrank_prefix<MAXD>(tg, K, d, g, q);
    r = prefix_rank<MAXD>(tg, d, q);
  }
  while (g < g_end) {
    if (r >= a.S_pre) break;  // cannot happen while g < S; keeps the loop bounded regardless
    // ---- lane i: data of the block of prefix r + i, into column i of the table ----
    {
      BlockInfo bi;
      Streams<MAXD> st;
      int q[MAXD];
      const uint64_t ri = r + (uint64_t)lane;
      const SpaceView sv{cg_l, lo_l, me, stride};
      if (ri < a.S_pre) {
        sv.unrank(ri, d, q);
        block_info<MAXD, SUF>(tl, K, d, q, bi, st);
      } else {
        bi.L = 0;
        bi.CB = bi.LBp = 0;
        bi.toff = 0;
        bi.lo = 0;
        bi.skip0 = 0;
        bi.lastrun = 0;
        bi.fQ = 1.f;
#pragma unroll
        for (int s = 0; s < MAXD; ++s) {
          st.base[s] = 0;
          st.mode[s] = kNoEdgeRow;
          st.active[s] = false;
        }
      }
      wave_sync();  // every lane is done reading the previous batch
      uint32_t* col = tbl + lane;
      col[0 * kBlockBatch] = (uint32_t)bi.CB;
      col[1 * kBlockBatch] = (uint32_t)(bi.CB >> 32);
      col[2 * kBlockBatch] = (uint32_t)bi.LBp;
      col[3 * kBlockBatch] = (uint32_t)(bi.LBp >> 32);
      col[4 * kBlockBatch] = bi.L;
      col[5 * kBlockBatch] = bi.toff;
      // active slots compacted to the front in ascending prefix position (the summation order of
      // the thread-per-state kernel); slots past the active count carry slot 0's base and no edge
#pragma unroll
      for (int w = 0; w < kModeW; ++w) col[(8 + w) * kBlockBatch] = kNoEdgeRow * 0x01010101u;
#pragma unroll
      for (int t = 0; t < MAXD; ++t) {
        col[(kBaseW + 2 * t) * kBlockBatch] = (uint32_t)st.base[0];
        col[(kBaseW + 2 * t + 1) * kBlockBatch] = (uint32_t)(st.base[0] >> 32);
      }
      uint32_t nact = 0;
#pragma unroll
      for (int s = 0; s < MAXD; ++s)
        if (st.active[s]) {
          reinterpret_cast<uint8_t*>(col + (8 + (nact >> 2)) * kBlockBatch)[nact & 3u] = (uint8_t)st.mode[s];
          col[(kBaseW + 2 * nact) * kBlockBatch] = (uint32_t)st.base[s];
          col[(kBaseW + 2 * nact + 1) * kBlockBatch] = (uint32_t)(st.base[s] >> 32);
          ++nact;
        }
      col[6 * kBlockBatch] = (uint32_t)bi.lo | ((uint32_t)bi.skip0 << 8) | ((uint32_t)bi.lastrun << 16) | (nact << 24);
      col[7 * kBlockBatch] = __builtin_bit_cast(uint32_t, bi.fQ);
      wave_sync();
    }
    const int nb = (int)((a.S_pre - r < kBlockBatch) ? a.S_pre - r : kBlockBatch);
    for (int j = 0; j < nb && g < g_end; ++j) {
      const uint32_t* bj = tbl + j;
      auto word = [&](int f) { return rfl(bj[f * kBlockBatch]); };
      const uint32_t L = word(4);
      if (L == 0) continue;                       // empty block (set masks only)
      const uint64_t CB = (uint64_t)word(0) | ((uint64_t)word(1) << 32);
      const uint64_t LBp = (uint64_t)word(2) | ((uint64_t)word(3) << 32);
      const uint32_t toff = word(5);
      const uint32_t w6 = word(6);
      const int blo = (int)(w6 & 0xffu);
      const int skip0 = (int)((w6 >> 8) & 1u);
      const int nact = (int)(w6 >> 24);
      float fQ = 1.f;
      int lastrun = 0;
      if constexpr (KIND != 0) {
        fQ = __builtin_bit_cast(float, word(7));
        lastrun = (int)((w6 >> 16) & 0xffu);
      }
      const bool probs = KIND == 1 || (KIND == 3 && kind == 1);
      const uint64_t avail = g_end - g;
      const uint32_t u_stop = (uint32_t)(((uint64_t)L - u0 < avail) ? (uint64_t)L : u0 + avail);
      const BlockSweep bs{par + LBp, tuples + toff, reinterpret_cast<char*>(out) + CB * (probs ? 4 : 8), rows, rn, C0,
                          (uint32_t)u0, u_stop, lane, fQ, lastrun, skip0, blo, multiset, a.scale, a.flags, kind,
                          COMP ? a.par_lo : 0, COMP ? a.out_lo : 0};
      // one sweep per count of active prefix streams, so inactive slots cost no loads or FMAs
      if constexpr (MAXD > 8) {
        block_children<KIND, COMP, MAXD, MAXD, SUF>(bj, par, bs);
      } else {
        constexpr int N0 = MAXD > 3 ? MAXD - 3 : 1;
        if (nact <= N0) block_children<KIND, COMP, MAXD, N0, SUF>(bj, par, bs);
        else if (nact == N0 + 1) block_children<KIND, COMP, MAXD, (N0 + 1 <= MAXD ? N0 + 1 : MAXD), SUF>(bj, par, bs);
        else if (nact == N0 + 2) block_children<KIND, COMP, MAXD, (N0 + 2 <= MAXD ? N0 + 2 : MAXD), SUF>(bj, par, bs);
        else block_children<KIND, COMP, MAXD, MAXD, SUF>(bj, par, bs);
      }
      g += u_stop - u0;
      u0 = 0;
    }
    r += kBlockBatch;
  }
}


// KIND 0: intermediate layer (complex64), 1: probabilities (float32), 2: scaled amplitudes;
// SUF: suffix length (4, or 5 with the 8-byte tuple words). Occupancy: 5 waves per SIMD for the
// plain 4-suffix sweep; 4 for the 5-suffix sweep (its widest prefixes would spill at 96 registers,
// and 4 measured slightly faster at n = 11, m = 32 as well); 3 for compensated sums and wide prefixes.
template <int KIND, bool COMP, int MAXD, int SUF>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(COMP || MAXD > 8 ? 3 : SUF == 5 ? 4 : 5))) void k_layer_block(const uint64_t* __restrict__ cg_t,
                                                               const int* __restrict__ lo_t,
                                                               const void* __restrict__ tuples,
                                                               int me, int stride, int multiset,
                                                               const float2* __restrict__ par_all,
                                                               void* __restrict__ out_all, BlockLayer a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ SufRow rows[kSufRows];
  __shared__ float rn[kNormRows * kNormCodes];
  const int K = a.K;
  const int ncg = (K + 1) * stride;
  uint64_t* cg_l = reinterpret_cast<uint64_t*>(smem);
  int* lo_l = reinterpret_cast<int*>(cg_l + ncg);
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg_l[i] = cg_t[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo_l[i] = lo_t[i];
  fill_suffix_rows<SUF>(rows, a.U + (size_t)b * a.m * a.m + a.col, a.m, cg_t, stride, me);
  fill_norm_rows<SUF>(rn);
  __syncthreads();
  const BlockTables tg{cg_t, lo_t, tuples, me, stride, multiset};  // global: scalar loads
  const BlockTables tl{cg_l, lo_l, tuples, me, stride, multiset};  // LDS: per-lane lookups

  // Work groups in XCD chunks (xcd_chunk_index) of about 1M children: a block's local parents are
  // fetched into one XCD's L2, not eight, while the eight XCDs stay in one region of the layer, so
  // the far prefix streams one XCD reads are found by the others in the shared last-level cache.
  // At n = 11, m = 32 this measured 28.7 ms for the last layer against 30.3 ms in plain launch
  // order (C = 1) and 30.4 ms with C = 4096; contiguous eighths of the layer were slower still.
  const uint32_t bid = xcd_chunk_index(blockIdx.x, a.xcd_chunk);
  const int wv = (int)rfl(threadIdx.x >> 6);
  const uint64_t g = uniform_u64(((uint64_t)bid * kWaves + wv) * a.per_wave);
  char* out_b = reinterpret_cast<char*>(out_all) + (size_t)b * a.S * (KIND == 1 ? 4 : 8);
  uint32_t* tbl = reinterpret_cast<uint32_t*>(smem + block_table_offset(K, me)) +
                  (size_t)wv * kBlockBatch * block_words(MAXD);
  block_wave<KIND, COMP, MAXD, SUF>(tg, tl, par_all + (size_t)b * a.S_par, out_b, rows, rn, tbl, a, g);
}

// ------------------------------------------------------------------------------------------
// The last layer in group order (5-position suffixes, multiset spaces, prefixes of 2..6 positions;
// items from group_items in block_index.hpp). A wave takes one item: a group R and a range of
// suffix positions u. Lane a computes block (a, R) into the wave's group table (32 words a block,
// read back with wave-uniform 16-byte LDS reads). Each 128-child step then loads the parent block
// R's pair of values (the prefix stream that drops the first photon, whose reuse distance spans
// the layer), the tuple words and the local-parent offsets once, and sweeps the step for every
// block of the group with its own prefix streams, local parents, coefficients and readout. Every
// child sums its parents in the order of the block kernel (ascending photon position), so both
// give bit-identical probabilities.
// ------------------------------------------------------------------------------------------
constexpr int kGroupWords = 32;  // CB(2) LBp(2) w6 fQ base0(2) coef0(2) 5 x {base(2), coef(2)} toff L
__host__ __device__ inline size_t group_table_offset(int K, int me) { return (block_lds_bytes(K, me) + 15) & ~(size_t)15; }
__host__ __device__ inline size_t group_dyn_lds(int K, int me) {
  return group_table_offset(K, me) + (size_t)kWaves * me * kGroupWords * 4;
}

// one block of the group for the pair step: NS stream slots (slot 0 shared), both children;
// KIND 1 writes probabilities (float32), KIND 0 the raw complex64 amplitudes of a middle layer
// COMP: compensated sums as the block kernel's sweep (two-sum accumulation; with plo the parents'
// low-order parts at element offset plo, slot 0's in v0lo; with olo the stored value's rounding
// error at element offset olo of a complex64 output).
template <int NS, int KIND, bool COMP = false>
__device__ __forceinline__ void group_block_pairs(const uint4 (&bw)[8], const float2* par, const float4& v0,
                                                  const float2 (&lc)[2][5], const uint32_t (&off)[2][5], uint32_t u,
                                                  const uint32_t (&code)[2], const bool (&lead)[2], const float* rn,
                                                  float scale, void* out, const float4& v0lo = float4{},
                                                  int64_t plo = 0, int64_t olo = 0) {
  const uint64_t CB = (uint64_t)bw[0].x | ((uint64_t)bw[0].y << 32);
  const uint64_t LBp = (uint64_t)bw[0].z | ((uint64_t)bw[0].w << 32);
  const uint32_t w6 = bw[1].x;
  const float fQ = __builtin_bit_cast(float, bw[1].y);
  const uint32_t* wd = reinterpret_cast<const uint32_t*>(bw);
  float4 sv[NS];
  float2 coef[NS];
  sv[0] = v0;
  coef[0] = make_float2(__builtin_bit_cast(float, wd[8]), __builtin_bit_cast(float, wd[9]));
#pragma unroll
  for (int s = 1; s < NS; ++s) {
    const uint64_t base = (uint64_t)wd[10 + 4 * (s - 1)] | ((uint64_t)wd[11 + 4 * (s - 1)] << 32);
    sv[s] = load_pair(par + base, u * 8u);
    coef[s] = make_float2(__builtin_bit_cast(float, wd[12 + 4 * (s - 1)]), __builtin_bit_cast(float, wd[13 + 4 * (s - 1)]));
  }
  const float2* lp = par + LBp;
  float2 lv[2][5];
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int j = 0; j < 5; ++j) lv[c][j] = at_bytes(lp, off[c][j]);
  const uint32_t lastrow = ((w6 >> 16) & 0xffu) + 1u;
  float res[2][2];
  BlockAcc accs[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < NS; ++s)
      bk_cmac<COMP>(acc, coef[s], c == 0 ? make_float2(sv[s].x, sv[s].y) : make_float2(sv[s].z, sv[s].w));
#pragma unroll
    for (int j = 0; j < 5; ++j) bk_cmac<COMP>(acc, lc[c][j], lv[c][j]);
    accs[c] = acc;
  }
  if constexpr (COMP) {
    if (plo) {  // the parents' low-order parts, into the correction terms (plain products)
      float4 sl[NS];
      sl[0] = v0lo;
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        const uint64_t base = (uint64_t)wd[10 + 4 * (s - 1)] | ((uint64_t)wd[11 + 4 * (s - 1)] << 32);
        sl[s] = load_pair(par + plo + base, u * 8u);
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
          bk_lo_mac(accs[c], coef[s], c == 0 ? make_float2(sl[s].x, sl[s].y) : make_float2(sl[s].z, sl[s].w));
#pragma unroll
        for (int j = 0; j < 5; ++j) bk_lo_mac(accs[c], lc[c][j], at_bytes(lp + plo, off[c][j]));
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    res[c][0] = COMP ? accs[c].re + accs[c].cre : accs[c].re;
    res[c][1] = COMP ? accs[c].im + accs[c].cim : accs[c].im;
  }
  if constexpr (KIND == 0) {
    const float4 v = make_float4(res[0][0], res[0][1], res[1][0], res[1][1]);
    __builtin_memcpy(static_cast<float2*>(out) + CB + u, &v, 16);
    if constexpr (COMP) {
      // the rounding error of the stored value (fast two-sum, as the block kernel)
      if (olo) {
        const float4 e = make_float4(accs[0].cre - (res[0][0] - accs[0].re), accs[0].cim - (res[0][1] - accs[0].im),
                                     accs[1].cre - (res[1][0] - accs[1].re), accs[1].cim - (res[1][1] - accs[1].im));
        __builtin_memcpy(static_cast<float2*>(out) + olo + CB + u, &e, 16);
      }
    }
  } else {
    float pr[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const float f = fQ * rn[(lead[c] ? lastrow : 0u) * 16u + code[c]];
      pr[c] = prob_readout(res[c][0] * scale, res[c][1] * scale, f);
    }
    const float2 v = make_float2(pr[0], pr[1]);
    __builtin_memcpy(static_cast<float*>(out) + CB + u, &v, 8);
  }
}

// the same for one child per lane (the item's last partial step)
template <int NS, int KIND, bool COMP = false>
__device__ __forceinline__ void group_block_one(const uint4 (&bw)[8], const float2* par, const float2& v0,
                                                const float2 (&lc)[5], const uint32_t (&off)[5], uint32_t u, bool in,
                                                uint32_t code, bool lead, const float* rn, float scale, void* out,
                                                const float2& v0lo = float2{}, int64_t plo = 0, int64_t olo = 0) {
  const uint64_t CB = (uint64_t)bw[0].x | ((uint64_t)bw[0].y << 32);
  const uint64_t LBp = (uint64_t)bw[0].z | ((uint64_t)bw[0].w << 32);
  const uint32_t w6 = bw[1].x;
  const float fQ = __builtin_bit_cast(float, bw[1].y);
  const uint32_t* wd = reinterpret_cast<const uint32_t*>(bw);
  float2 sv[NS], coef[NS];
  sv[0] = v0;
  coef[0] = make_float2(__builtin_bit_cast(float, wd[8]), __builtin_bit_cast(float, wd[9]));
#pragma unroll
  for (int s = 1; s < NS; ++s) {
    const uint64_t base = (uint64_t)wd[10 + 4 * (s - 1)] | ((uint64_t)wd[11 + 4 * (s - 1)] << 32);
    sv[s] = at_bytes(par + base, u * 8u);
    coef[s] = make_float2(__builtin_bit_cast(float, wd[12 + 4 * (s - 1)]), __builtin_bit_cast(float, wd[13 + 4 * (s - 1)]));
  }
  const float2* lp = par + LBp;
  float2 lv[5];
#pragma unroll
  for (int j = 0; j < 5; ++j) lv[j] = at_bytes(lp, off[j]);
  BlockAcc acc{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < NS; ++s) bk_cmac<COMP>(acc, coef[s], sv[s]);
#pragma unroll
  for (int j = 0; j < 5; ++j) bk_cmac<COMP>(acc, lc[j], lv[j]);
  if constexpr (COMP) {
    if (plo) {
      bk_lo_mac(acc, coef[0], v0lo);
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        const uint64_t base = (uint64_t)wd[10 + 4 * (s - 1)] | ((uint64_t)wd[11 + 4 * (s - 1)] << 32);
        bk_lo_mac(acc, coef[s], at_bytes(par + plo + base, u * 8u));
      }
#pragma unroll
      for (int j = 0; j < 5; ++j) bk_lo_mac(acc, lc[j], at_bytes(lp + plo, off[j]));
    }
  }
  const float re = COMP ? acc.re + acc.cre : acc.re, im = COMP ? acc.im + acc.cim : acc.im;
  if constexpr (KIND == 0) {
    if (in) static_cast<float2*>(out)[CB + u] = make_float2(re, im);
    if constexpr (COMP) {
      if (olo && in) static_cast<float2*>(out)[olo + CB + u] = make_float2(acc.cre - (re - acc.re), acc.cim - (im - acc.im));
    }
  } else {
    const float f = fQ * rn[(lead ? ((w6 >> 16) & 0xffu) + 1u : 0u) * 16u + code];
    if (in) static_cast<float*>(out)[CB + u] = prob_readout(re * scale, im * scale, f);
  }
}

// the group table entry of block g (wave-uniform 16-byte LDS reads)
__device__ __forceinline__ void group_words(const uint32_t* tbl, int g, uint4 (&bw)[8]) {
  const uint4* src = reinterpret_cast<const uint4*>(tbl + g * kGroupWords);
#pragma unroll
  for (int i = 0; i < 8; ++i) bw[i] = src[i];
}

template <int MAXD, class F>
__device__ __forceinline__ void group_dispatch(uint32_t nact, F&& f) {
  constexpr int N0 = MAXD > 3 ? MAXD - 3 : 1;
  if (nact <= (uint32_t)N0) f(std::integral_constant<int, N0>{});
  else if (nact == (uint32_t)N0 + 1) f(std::integral_constant<int, (N0 + 1 <= MAXD ? N0 + 1 : MAXD)>{});
  else if (nact == (uint32_t)N0 + 2) f(std::integral_constant<int, (N0 + 2 <= MAXD ? N0 + 2 : MAXD)>{});
  else f(std::integral_constant<int, MAXD>{});
}

template <int MAXD, int KIND, bool COMP = false>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_layer_group(
    const uint64_t* __restrict__ cg_t, const int* __restrict__ lo_t, const uint64_t* __restrict__ tuples5, int me,
    int stride, const float2* __restrict__ par_all, void* __restrict__ out_all, BlockLayer a,
    const uint4* __restrict__ items, uint32_t n_items) {
  static_assert(MAXD >= 2 && MAXD <= 6, "group table holds 6 stream slots");
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ SufRow rows[kSufRows];
  __shared__ float rn[kNormRows * kNormCodes];
  const int K = a.K;
  const int d = K - 5;
  const int ncg = (K + 1) * stride;
  uint64_t* cg_l = reinterpret_cast<uint64_t*>(smem);
  int* lo_l = reinterpret_cast<int*>(cg_l + ncg);
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg_l[i] = cg_t[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo_l[i] = lo_t[i];
  fill_suffix_rows<5>(rows, a.U + (size_t)b * a.m * a.m + a.col, a.m, cg_t, stride, me);
  fill_norm_rows<5>(rn);
  __syncthreads();
  const int wv = (int)rfl(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const uint32_t item = xcd_chunk_index(blockIdx.x, a.xcd_chunk) * kWaves + (uint32_t)wv;
  if (item >= n_items) return;
  const uint4 it = items[item];
  const uint32_t r = rfl(it.x), u_b = rfl(it.y), u_e = rfl(it.z);
  const BlockTables tl{cg_l, lo_l, tuples5, me, stride, 1};
  const SpaceView sv{cg_l, lo_l, me, stride};
  int R[MAXD];
  sv.unrank(r, d - 1, R);
  const int G = (int)rfl((uint32_t)R[0]) + 1;  // blocks (a, R), a = 0..R_0 (lo[x] = x)
  uint32_t* tbl = reinterpret_cast<uint32_t*>(smem + group_table_offset(K, me)) + (size_t)wv * me * kGroupWords;
  const float2* par = par_all + (size_t)b * a.S_par;
  void* out = static_cast<char*>(out_all) + (size_t)b * a.S * (KIND == 0 ? 8 : 4);
  if (lane < G) {
    int q[MAXD];
    q[0] = lane;
#pragma unroll
    for (int i = 1; i < MAXD; ++i) q[i] = i < d ? R[i - 1] : 0;
    BlockInfo bi;
    Streams<MAXD> st;
    block_info<MAXD, 5>(tl, K, d, q, bi, st);
    uint32_t* col = tbl + lane * kGroupWords;
    col[0] = (uint32_t)bi.CB;
    col[1] = (uint32_t)(bi.CB >> 32);
    col[2] = (uint32_t)bi.LBp;
    col[3] = (uint32_t)(bi.LBp >> 32);
    col[5] = __builtin_bit_cast(uint32_t, bi.fQ);
    col[6] = (uint32_t)st.base[0];
    col[7] = (uint32_t)(st.base[0] >> 32);
    const SufRow r0 = rows[st.mode[0]];
    col[8] = __builtin_bit_cast(uint32_t, r0.re);
    col[9] = __builtin_bit_cast(uint32_t, r0.im);
    // slots 1..5: the other active streams in ascending position, then slot 0's base, no edge
#pragma unroll
    for (int t = 1; t < 6; ++t) {
      col[10 + 4 * (t - 1)] = (uint32_t)st.base[0];
      col[11 + 4 * (t - 1)] = (uint32_t)(st.base[0] >> 32);
      col[12 + 4 * (t - 1)] = 0u;
      col[13 + 4 * (t - 1)] = 0u;
    }
    uint32_t nact = 1;
#pragma unroll
    for (int s = 1; s < MAXD; ++s)
      if (st.active[s]) {
        const SufRow rw = rows[st.mode[s]];
        col[10 + 4 * (nact - 1)] = (uint32_t)st.base[s];
        col[11 + 4 * (nact - 1)] = (uint32_t)(st.base[s] >> 32);
        col[12 + 4 * (nact - 1)] = __builtin_bit_cast(uint32_t, rw.re);
        col[13 + 4 * (nact - 1)] = __builtin_bit_cast(uint32_t, rw.im);
        ++nact;
      }
    col[4] = (uint32_t)bi.lo | ((uint32_t)bi.skip0 << 8) | ((uint32_t)bi.lastrun << 16) | (nact << 24);
    col[30] = bi.toff;
    col[31] = bi.L;
  }
  wave_sync();
  // shared by the group: the parent block R (slot 0's base), the tuple range, lo, skip0
  const uint64_t base0 = (uint64_t)rfl(tbl[6]) | ((uint64_t)rfl(tbl[7]) << 32);
  const uint32_t toff = rfl(tbl[30]);
  const uint32_t w60 = rfl(tbl[4]);
  const int lo = (int)(w60 & 0xffu), skip0 = (int)((w60 >> 8) & 1u);
  const float2* sp0 = par + base0;
  const uint64_t* tb = tuples5 + toff;
  const float2* const dsp[1] = {sp0};
  const float2 dcoef[1] = {make_float2(0.f, 0.f)};
  const ChunkCtx<1, 5> cx{dsp, dcoef, par, tb, nullptr, rows, rn, (uint32_t)cg_t[4 * stride], u_e, lane, 1.f, 0, skip0, lo,
                          1, a.scale, 0u, 1};
  uint32_t uc = u_b;
  // the shared tuple words and first-stream pair are loaded one step ahead
  TuplePair<5> tpn{};
  float4 v0n{};
  if (uc + 128u <= u_e) {
    tpn = load_tuple_pair(cx, uc);
    v0n = load_pair(sp0, (uc + 2u * (uint32_t)lane) * 8u);
  }
  for (; uc + 128u <= u_e; uc += 128u) {
    const uint32_t u = uc + 2u * (uint32_t)lane;
    const TuplePair<5> tp = tpn;
    const float4 v0 = v0n;
    if (uc + 256u <= u_e) {
      tpn = load_tuple_pair(cx, uc + 128u);
      v0n = load_pair(sp0, (u + 128u) * 8u);
    }
    float2 lc[2][5];
    uint32_t off[2][5];
    suffix_edges(cx, tp.a, lc[0], off[0]);
    suffix_edges(cx, tp.b, lc[1], off[1]);
    const uint32_t code[2] = {tuple_code<5>(tp.a), tuple_code<5>(tp.b)};
    const bool lead[2] = {skip0 && (int)((uint32_t)tp.a & 0xffu) == lo, skip0 && (int)((uint32_t)tp.b & 0xffu) == lo};
    float4 v0lo{};
    if (COMP && a.par_lo) v0lo = load_pair(sp0 + a.par_lo, u * 8u);
    for (int g = 0; g < G; ++g) {
      uint4 bw[8];
      group_words(tbl, g, bw);
      group_dispatch<MAXD>(bw[1].x >> 24, [&](auto ns) {
        group_block_pairs<decltype(ns)::value, KIND, COMP>(bw, par, v0, lc, off, u, code, lead, rn, a.scale, out, v0lo,
                                                            COMP ? a.par_lo : 0, COMP ? a.out_lo : 0);
      });
    }
  }
  for (; uc < u_e; uc += 64u) {
    const uint32_t ur = uc + (uint32_t)lane;
    const bool in = ur < u_e;
    const uint32_t u = in ? ur : u_e - 1u;
    const uint64_t w = tb[u];
    const float2 v0 = at_bytes(sp0, u * 8u);
    float2 lc[5];
    uint32_t off[5];
    suffix_edges(cx, w, lc, off);
    const uint32_t code = tuple_code<5>(w);
    const bool lead = skip0 && (int)((uint32_t)w & 0xffu) == lo;
    float2 v0lo{};
    if (COMP && a.par_lo) v0lo = at_bytes(sp0 + a.par_lo, u * 8u);
    for (int g = 0; g < G; ++g) {
      uint4 bw[8];
      group_words(tbl, g, bw);
      group_dispatch<MAXD>(bw[1].x >> 24, [&](auto ns) {
        group_block_one<decltype(ns)::value, KIND, COMP>(bw, par, v0, lc, off, u, in, code, lead, rn, a.scale, out, v0lo,
                                                          COMP ? a.par_lo : 0, COMP ? a.out_lo : 0);
      });
    }
  }
}

// ------------------------------------------------------------------------------------------
// Uneven batches: one launch per photon layer K covers every instance of every plan in a group
// (plans of different n, m, input and mask). The work is cut into work-group items, each inside one
// (plan, instance): plan g owns items [wg_begin, wg_begin + B * wgs_per_inst). The grid is
// persistent: XCD x walks the contiguous item range [x T/8, (x+1) T/8) with its work groups in
// turn, so each XCD's L2 serves neighbouring items.
// ------------------------------------------------------------------------------------------

constexpr int kMaxGroup = 8;

struct GroupItem {
  const uint64_t* cg;          // plan tables (device)
  const int* lo;
  const TupleWord* tuples;
  const float2* U;             // [B][m][m]
  const float2* par;           // [B][S_par]
  void* out;                   // [B][S]: float32 probabilities (kind 1) or complex64 (kind 0)
  uint64_t S_par, S, S_pre;
  uint64_t per_wave;           // block kernel: children per wave; simple kernel: states per thread
  float scale;
  unsigned flags;
  int me, multiset, m, col, kind;
  uint32_t wg_begin, wgs_per_inst;
  int64_t par_lo, out_lo;      // compensated sums: element offsets of the low-order parts (0: none)
};

struct GroupLaunch {
  GroupItem it[kMaxGroup];
  int count, K;
  uint32_t total;              // items of the launch
  uint32_t table_off;          // block kernel: byte offset of the per-wave block tables in dynamic LDS
};

// the plan of work-group index `item` (plans in ascending wg_begin)
__device__ __forceinline__ int group_index(const GroupLaunch& gl, uint32_t item) {
  int gi = 0;
#pragma unroll
  for (int i = 1; i < kMaxGroup; ++i)
    if (i < gl.count && item >= gl.it[i].wg_begin) gi = i;
  return (int)rfl((uint32_t)gi);
}

// persistent XCD-contiguous walk over [0, total): calls f(item) for this work group's items
template <class F>
__device__ __forceinline__ void for_each_item(uint32_t total, F&& f) {
  const uint32_t nx = gridDim.x >> 3;  // work groups per XCD (the grid is a multiple of 8)
  const uint32_t per_xcd = (total + 7u) >> 3;
  const uint32_t xcd = blockIdx.x & 7u;
  for (uint32_t t = blockIdx.x >> 3; t < per_xcd; t += nx) {
    const uint32_t item = xcd * per_xcd + t;
    if (item >= total) break;
    f(item);
  }
}

template <bool COMP, int MAXD>
__global__ __launch_bounds__(kBlockThreads) __attribute__((amdgpu_waves_per_eu(COMP || MAXD > 8 ? 3 : MAXD <= 5 ? 5 : 4))) void
k_layer_block_group(GroupLaunch gl) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ SufRow rows[kSufRows];
  __shared__ float rn[kNormRows * kNormCodes];
  const int K = gl.K;
  fill_norm_rows<kSuf>(rn);
  const int wv = (int)rfl(threadIdx.x >> 6);
  const uint64_t* cur = nullptr;  // plan whose tables are in LDS
  for_each_item(gl.total, [&](uint32_t item) {
    const GroupItem& it = gl.it[group_index(gl, item)];
    const uint32_t rel = item - it.wg_begin;
    const uint32_t b = rel / it.wgs_per_inst, w = rel - b * it.wgs_per_inst;
    const int stride = it.me + 1;
    const int ncg = (K + 1) * stride;
    uint64_t* cg_l = reinterpret_cast<uint64_t*>(smem);
    int* lo_l = reinterpret_cast<int*>(cg_l + ncg);
    __syncthreads();  // every wave is done with the previous item's LDS
    if (it.cg != cur) {
      for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg_l[i] = it.cg[i];
      for (int i = threadIdx.x; i < it.me; i += blockDim.x) lo_l[i] = it.lo[i];
      cur = it.cg;
    }
    fill_suffix_rows<kSuf>(rows, it.U + (size_t)b * it.m * it.m + it.col, it.m, it.cg, stride, it.me);
    __syncthreads();
    const BlockTables tg{it.cg, it.lo, it.tuples, it.me, stride, it.multiset};
    const BlockTables tl{cg_l, lo_l, it.tuples, it.me, stride, it.multiset};
    BlockLayer a;
    a.K = K; a.m = it.m; a.col = it.col; a.U = it.U; a.S_par = it.S_par; a.S = it.S; a.S_pre = it.S_pre;
    a.scale = it.scale; a.flags = it.flags; a.per_wave = it.per_wave; a.xcd_chunk = 1;
    a.par_lo = it.par_lo; a.out_lo = it.out_lo;
    const uint64_t g = uniform_u64(((uint64_t)w * kWaves + wv) * it.per_wave);
    const float2* par = it.par + (size_t)b * it.S_par;
    char* out = reinterpret_cast<char*>(it.out) + (size_t)b * it.S * (it.kind == 1 ? 4 : 8);
    uint32_t* tbl = reinterpret_cast<uint32_t*>(smem + gl.table_off) + (size_t)wv * kBlockBatch * block_words(MAXD);
    block_wave<3, COMP, MAXD, kSuf>(tg, tl, par, out, rows, rn, tbl, a, g, it.kind);
  });
}

}  // namespace slos
