// SLOS strong simulation on MI355X (gfx950): plan, C ABI and kernels.
//
// Forward path (replaces FockSLOSKernel._amplitudes / probs, photonic_engines/slos/kernel.py:201-283):
// one launch per photon layer, pull form — each child state sums its parents, whose lex ranks
// are computed from the child's own sorted positions with the counting table (space.hpp), so no
// state or edge table exists anywhere. Every child is written exactly once; there are no atomics.
// The last layer's launch squares, normalises and writes float32 probabilities (or scaled
// complex64 amplitudes) directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "../../include/slos_mi355x.h"
#include "space.hpp"
#include "layer_block.hpp"
#include "descend_block.hpp"
#include "layer_f64.hpp"
#include "pull_index.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                            \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess)                                                        \
      return fail(SLOS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kMaxPhotons = 32;

}  // namespace

struct DevTables {
  int device = -1;
  uint64_t* cg = nullptr;
  int* lo = nullptr;
  int* modes = nullptr;                    // the plan's input modes (spot permanents)
  slos::TupleWord* tuples = nullptr;       // block-kernel suffix-tuple words (may be absent)
  uint64_t* tuples5 = nullptr;             // the same for 5-position suffixes (may be absent)
  uint4* items[33] = {};                   // group-order work items of layer K (may be absent)
  uint32_t n_items[33] = {};
};

struct slos_plan {
  int n = 0, m = 0, mask = 0, me = 0, multiset = 1, shift = 0;
  std::vector<int> modes;
  std::vector<uint64_t> cg;  // (n + 2) x (me + 1)
  std::vector<int> lo;
  double s_fact = 1.0;       // prod_i s_i!
  std::vector<slos::TupleWord> tuples;   // suffix-tuple words of the block kernel (empty: unused)
  std::vector<uint64_t> tuples5;         // 5-position suffix-tuple words (empty: unused)
  std::vector<uint32_t> items[33];       // group-order work items of layer K, 4 words each (empty: unused)
  std::vector<DevTables> dev;
  std::mutex mu;
  slos::SpaceView host_view() const { return slos::SpaceView{cg.data(), lo.data(), me, me + 1}; }
  uint64_t size(int k) const { return cg[(size_t)k * (me + 1)]; }
};

namespace {

// A roctx range over a scope (rocprofv3 --marker-trace shows it; without a profiler it is a no-op
// call): the forward, each of its layers and each descent step are named ranges, so a trace ties
// every kernel to its layer.
struct RoctxRange {
  explicit RoctxRange(const char* name) { roctxRangePushA(name); }
  ~RoctxRange() { roctxRangePop(); }
  RoctxRange(const RoctxRange&) = delete;
  RoctxRange& operator=(const RoctxRange&) = delete;
};

// LDS one work group may use on the current GPU (hipDeviceAttributeMaxSharedMemoryPerBlock; 160 KB
// on gfx950), read once per device. Should the query fail, 64 KB (the LDS of every earlier CDNA
// part) is assumed, so launches only get more conservative.
size_t lds_limit() {
  static std::mutex mu;
  static std::vector<std::pair<int, size_t>> seen;
  int d = 0;
  if (hipGetDevice(&d) != hipSuccess) return 64 * 1024;
  std::lock_guard<std::mutex> g(mu);
  for (const auto& e : seen)
    if (e.first == d) return e.second;
  int v = 0;
  size_t lim = 64 * 1024;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, d) == hipSuccess && v > 0) lim = (size_t)v;
  seen.emplace_back(d, lim);
  return lim;
}

int device_tables(const slos_plan* cp, DevTables* out) {
  slos_plan* p = const_cast<slos_plan*>(cp);
  int d = 0;
  HIP_TRY(hipGetDevice(&d));
  std::lock_guard<std::mutex> g(p->mu);
  for (auto& t : p->dev)
    if (t.device == d) {
      *out = t;
      return SLOS_OK;
    }
  DevTables t;
  t.device = d;
  HIP_TRY(hipMalloc(&t.cg, p->cg.size() * sizeof(uint64_t)));
  HIP_TRY(hipMalloc(&t.lo, std::max<size_t>(1, p->lo.size()) * sizeof(int)));
  HIP_TRY(hipMemcpy(t.cg, p->cg.data(), p->cg.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(t.lo, p->lo.data(), p->lo.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&t.modes, p->modes.size() * sizeof(int)));
  HIP_TRY(hipMemcpy(t.modes, p->modes.data(), p->modes.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!p->tuples.empty()) {
    HIP_TRY(hipMalloc(&t.tuples, p->tuples.size() * sizeof(slos::TupleWord)));
    HIP_TRY(hipMemcpy(t.tuples, p->tuples.data(), p->tuples.size() * sizeof(slos::TupleWord),
                      hipMemcpyHostToDevice));
  }
  if (!p->tuples5.empty()) {
    HIP_TRY(hipMalloc(&t.tuples5, p->tuples5.size() * sizeof(uint64_t)));
    HIP_TRY(hipMemcpy(t.tuples5, p->tuples5.data(), p->tuples5.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  for (int K = 0; K <= p->n; ++K)
    if (!p->items[K].empty()) {
      HIP_TRY(hipMalloc(&t.items[K], p->items[K].size() * sizeof(uint32_t)));
      HIP_TRY(hipMemcpy(t.items[K], p->items[K].data(), p->items[K].size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      t.n_items[K] = (uint32_t)(p->items[K].size() / 4);
    }
  p->dev.push_back(t);
  *out = t;
  return SLOS_OK;
}

// ------------------------------------------------------------------------------------------
// thread-per-state layer kernel: unrank, parent ranks by prefix/suffix sums, gather-MAC
// ------------------------------------------------------------------------------------------

struct Acc {
  float re = 0.f, im = 0.f, cre = 0.f, cim = 0.f;
};

// (compensated arithmetic: no contraction, see bk_two_sum)
__device__ __forceinline__ void two_sum(float& s, float& c, float x) {
#pragma clang fp contract(off)
  const float t = s + x;
  const float bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

// acc += u * v, plain or compensated (product errors by fma, sums by two-sum)
template <bool COMP>
__device__ __forceinline__ void cmac(Acc& a, float2 u, float2 v) {
  if constexpr (!COMP) {
    a.re = fmaf(u.x, v.x, a.re);
    a.re = fmaf(-u.y, v.y, a.re);
    a.im = fmaf(u.x, v.y, a.im);
    a.im = fmaf(u.y, v.x, a.im);
  } else {
#pragma clang fp contract(off)
    float p, e;
    p = u.x * v.x; e = fmaf(u.x, v.x, -p); two_sum(a.re, a.cre, p); a.cre += e;
    p = -u.y * v.y; e = fmaf(-u.y, v.y, -p); two_sum(a.re, a.cre, p); a.cre += e;
    p = u.x * v.y; e = fmaf(u.x, v.y, -p); two_sum(a.im, a.cim, p); a.cim += e;
    p = u.y * v.x; e = fmaf(u.y, v.x, -p); two_sum(a.im, a.cim, p); a.cim += e;
  }
}

struct LayerArgs {
  const uint64_t* cg;
  const int* lo;
  int me, multiset, K, m, col;
  const float2* U;        // [B][m][m]
  const int32_t* cols;    // unused by the simple kernel
  const float2* par;      // [B][S_par]
  uint64_t S_par;
  void* out;              // [B][S]
  uint64_t S;
  int64_t B;
  float scale;
  unsigned flags;
  int64_t par_lo, out_lo;  // compensated sums: element offsets of the low-order parts (0: none)
};

// KIND 0: intermediate layer (complex64), 1: probabilities (float32), 2: scaled amplitudes
// (K <= KM: the per-state position arrays stay in registers)
template <int KIND, bool COMP, int KM>
__global__ __launch_bounds__(256) void k_layer_simple(LayerArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = a.me + 1;
  const int ncg = (a.K + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  float2* ucol = reinterpret_cast<float2*>(smem + ((ncg * 8 + a.me * 4 + 15) & ~15));
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = a.cg[i];
  for (int i = threadIdx.x; i < a.me; i += blockDim.x) lo[i] = a.lo[i];
  const slos::SpaceView sv{cg, lo, a.me, stride};
  const int K = a.K;
  for (int64_t b = blockIdx.y; b < a.B; b += gridDim.y) {
    __syncthreads();
    for (int j = threadIdx.x; j < a.m; j += blockDim.x)
      ucol[j] = a.U[(size_t)b * a.m * a.m + (size_t)j * a.m + a.col];
    __syncthreads();
    const float2* pb = a.par + (size_t)b * a.S_par;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.S;
         g += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      slos::unrank_fixed<KM>(sv, g, K, p);
      Acc acc;
      slos::parents_fixed<KM>(sv, a.multiset, K, p,
                              [&](uint64_t pr, int mode, int) { cmac<COMP>(acc, ucol[mode], pb[pr]); });
      if constexpr (COMP) {
        if (a.par_lo)  // the parents' low-order parts, into the correction terms
          slos::parents_fixed<KM>(sv, a.multiset, K, p, [&](uint64_t pr, int mode, int) {
            const float2 u = ucol[mode], w = pb[a.par_lo + pr];
            acc.cre = fmaf(u.x, w.x, acc.cre);
            acc.cre = fmaf(-u.y, w.y, acc.cre);
            acc.cim = fmaf(u.x, w.y, acc.cim);
            acc.cim = fmaf(u.y, w.x, acc.cim);
          });
      }
      float re = COMP ? acc.re + acc.cre : acc.re, im = COMP ? acc.im + acc.cim : acc.im;
      if constexpr (KIND == 0) {
        reinterpret_cast<float2*>(a.out)[(size_t)b * a.S + g] = make_float2(re, im);
        if constexpr (COMP) {
          if (a.out_lo)  // the rounding error of the stored value (fast two-sum)
            reinterpret_cast<float2*>(a.out)[a.out_lo + (size_t)b * a.S + g] =
                make_float2(acc.cre - (re - acc.re), acc.cim - (im - acc.im));
        }
      } else {
        const float f = slos::run_factorial_fixed<KM>(p, K, a.multiset);
        re *= a.scale;
        im *= a.scale;
        if constexpr (KIND == 1) {
          reinterpret_cast<float*>(a.out)[(size_t)b * a.S + g] = slos::prob_readout(re, im, f);
        } else {
          const float s = (a.flags & SLOS_FLAG_IMMEDIATE) ? sqrtf(f) : 1.f;
          reinterpret_cast<float2*>(a.out)[(size_t)b * a.S + g] = make_float2(re * s, im * s);
        }
      }
    }
  }
}

__global__ void k_fill_ones(float2* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_float2(1.f, 0.f);
}

// the output of layer 0, the vacuum (one state a batch row): probability |scale|^2, amplitude scale
__global__ void k_vacuum_out(void* out, int64_t n, int probs, float scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (probs) static_cast<float*>(out)[i] = scale * scale;
    else static_cast<float2*>(out)[i] = make_float2(scale, 0.f);
  }
}

// ------------------------------------------------------------------------------------------
// FP64 mass: per-block partial sums, then one block per batch row
// ------------------------------------------------------------------------------------------

constexpr int kMassBlocks = 1024;

__device__ double block_sum(double v) {
  __shared__ double sh[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_mass_partial(const float* p, uint64_t count, double* part) {
  const int64_t b = blockIdx.y;
  const float* row = p + (size_t)b * count;
  double s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    s += (double)row[i];
  s = block_sum(s);
  if (threadIdx.x == 0) part[(size_t)b * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mass_final(const double* part, int nb, double* out) {
  double s = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[(size_t)blockIdx.x * nb + i];
  s = block_sum(s);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// ------------------------------------------------------------------------------------------
// Group sums (GroupedSLOSKernel): a keyed reduction without atomics. The states are listed in group
// order (`order`), cut into chunks (4096 entries balance the waves); one wave sums a chunk's
// probabilities in FP64 (lane-strided, then a fixed butterfly), and one thread per group sums its
// chunks' partials in ascending order. Every addition has a fixed place, so the sums are
// reproducible run to run. T: float32 or float64 probabilities (and sums).
// ------------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_group_partial(const T* probs, uint64_t count, const int64_t* order,
                                                       const int64_t* chunks, uint32_t n_chunks, double* part) {
  const uint32_t c = blockIdx.x * 4u + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= n_chunks) return;
  const int64_t b = blockIdx.y;
  const T* pb = probs + (size_t)b * count;
  const int64_t i0 = chunks[2 * (size_t)c], i1 = chunks[2 * (size_t)c + 1];
  double s = 0.0;
  for (int64_t i = i0 + lane; i < i1; i += 64) s += (double)pb[order[i]];
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1) s += __shfl_xor(s, w, 64);
  if (lane == 0) part[(size_t)b * n_chunks + c] = s;
}

template <class T>
__global__ __launch_bounds__(256) void k_group_final(const double* part, uint32_t n_chunks, const uint32_t* first,
                                                     int n_groups, T* out) {
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (g >= n_groups) return;
  const int64_t b = blockIdx.y;
  double s = 0.0;
  for (uint32_t c = first[g]; c < first[g + 1]; ++c) s += part[(size_t)b * n_chunks + c];
  out[(size_t)b * n_groups + g] = (T)s;
}

// ------------------------------------------------------------------------------------------
// FP64 Ryser permanents for spot checks (one thread per output state)
// ------------------------------------------------------------------------------------------

__global__ __launch_bounds__(64) void k_spot_ryser(const uint64_t* cg_g, const int* lo_g, int me, int n, int m,
                                                   const double2* U, const int* modes, double s_fact,
                                                   int multiset, const uint64_t* idx, int64_t count,
                                                   double* out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const slos::SpaceView sv{cg_g, lo_g, me, me + 1};
  int p[kMaxPhotons];
  sv.unrank(idx[q], n, p);
  double tf = 1.0, run = 1.0;
  if (multiset)
    for (int i = 1; i < n; ++i) {
      run = (p[i] == p[i - 1]) ? run + 1.0 : 1.0;
      tf *= run;
    }
  double rre[kMaxPhotons], rim[kMaxPhotons];
  for (int i = 0; i < n; ++i) rre[i] = rim[i] = 0.0;
  double tre = 0.0, tim = 0.0;
  uint64_t prev = 0;
  for (uint64_t k = 1; k < (1ull << n); ++k) {
    const uint64_t gray = k ^ (k >> 1);
    const uint64_t diff = gray ^ prev;
    prev = gray;
    const int c = __builtin_ctzll(diff);
    const double sg = (gray & diff) ? 1.0 : -1.0;
    double pre = 1.0, pim = 0.0;
    for (int i = 0; i < n; ++i) {
      const double2 u = U[(size_t)p[i] * m + modes[c]];
      rre[i] += sg * u.x;
      rim[i] += sg * u.y;
      const double nr = pre * rre[i] - pim * rim[i];
      pim = pre * rim[i] + pim * rre[i];
      pre = nr;
    }
    const int pc = __builtin_popcountll(gray);
    if ((n - pc) & 1) { tre -= pre; tim -= pim; } else { tre += pre; tim += pim; }
  }
  out[q] = (tre * tre + tim * tim) / (s_fact * tf);
}


// ------------------------------------------------------------------------------------------
// uniform loss: q_{k-1}(t) = sum_j q_k(t + e_j) (t_j + 1) / k  (pull over children, no atomics)
// ------------------------------------------------------------------------------------------
// (child ranks: slos::children_fixed in pull_index.hpp; k - 1 <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_descend(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int k,
                                                 const float* __restrict__ qk, uint64_t Sk,
                                                 float* __restrict__ qout, uint64_t Sp, int64_t B, float scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (k + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  __syncthreads();
  const slos::SpaceView sv{cg, lo, me, stride};
  const float invk = 1.f / (float)k;
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    const float* qb = qk + (size_t)b * Sk;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < Sp; t += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      slos::unrank_fixed<KM>(sv, t, k - 1, p);
      float acc = 0.f;
      slos::children_fixed<KM>(sv, multiset, k, p, [&](uint64_t c, int, int cnt) { acc = fmaf(qb[c], (float)cnt, acc); });
      qout[(size_t)b * Sp + t] = (acc * invk) * scale;
    }
  }
}

// probability-level raise (a photon that interferes with nothing, partial distinguishability):
// q_k(c) = sum_{j distinct in c} w[b][j] q_{k-1}(c - e_j), w = |U[:, col]|^2 (pull over parents,
// ascending photon position, no atomics; k <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_raise(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int k, int m,
                                               const float* __restrict__ w, const float* __restrict__ qp, uint64_t Sp,
                                               float* __restrict__ qout, uint64_t Sk, int64_t B) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (k + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  float* wl = reinterpret_cast<float*>(lo + me);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  const slos::SpaceView sv{cg, lo, me, stride};
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    __syncthreads();
    for (int j = threadIdx.x; j < m; j += blockDim.x) wl[j] = w[(size_t)b * m + j];
    __syncthreads();
    const float* qb = qp + (size_t)b * Sp;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < Sk; c += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      slos::unrank_fixed<KM>(sv, c, k, p);
      float acc = 0.f;
      slos::parents_fixed<KM>(sv, multiset, k, p, [&](uint64_t pr, int mode, int) { acc = fmaf(wl[mode], qb[pr], acc); });
      qout[(size_t)b * Sk + c] = acc;
    }
  }
}

// adjoint of the descent (gradient of q_{k-1} w.r.t. q_k): g_k[c] = sum_{j distinct in c}
// g_{k-1}[c - e_j] * count_j(c) / k  — the forward pull structure with real weights (k <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_ascend(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int k,
                                                const float* __restrict__ gp, uint64_t Sp,
                                                float* __restrict__ gk, uint64_t Sk, int64_t B) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (k + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  __syncthreads();
  const slos::SpaceView sv{cg, lo, me, stride};
  const int K = k;
  const float invk = 1.f / (float)k;
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    const float* pb = gp + (size_t)b * Sp;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < Sk; g += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      slos::unrank_fixed<KM>(sv, g, K, p);
      float acc = 0.f;
      slos::parents_fixed<KM>(sv, multiset, K, p, [&](uint64_t pr, int, int cnt) { acc += pb[pr] * (float)cnt; });
      gk[(size_t)b * Sk + g] = acc * invk;
    }
  }
}


// ------------------------------------------------------------------------------------------
// adjoint (reverse-mode) kernels; torch's convention for complex leaves: grad z = dL/dRe z + i dL/dIm z
// ------------------------------------------------------------------------------------------

// readout: p = |a s|^2 f  =>  g_a = 2 s^2 f grad_p a. grad_p is read first: where it is zero the
// gradient is zero, and neither the amplitude nor the state's factor (its unrank) is needed. A loss
// over a few sampled outputs leaves grad_p zero almost everywhere, so the pass then reads grad_p and
// writes g only; elsewhere it computes the same products as without the test.
__global__ __launch_bounds__(256) void k_readout_adjoint(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int K,
                                                         const float2* __restrict__ a, const float* __restrict__ gp,
                                                         float2* __restrict__ g, uint64_t S, int64_t B, float s2) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (K + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  __syncthreads();
  const slos::SpaceView sv{cg, lo, me, stride};
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < S; c += (uint64_t)gridDim.x * blockDim.x) {
    bool have_f = false;
    float f = 0.f;
    for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
      const float gv = gp[(size_t)b * S + c];
      float2 r = make_float2(0.f, 0.f);
      if (gv != 0.f) {
        if (!have_f) {
          int p[kMaxPhotons];
          sv.unrank(c, K, p);
          f = slos::run_factorial(p, K, multiset);
          have_f = true;
        }
        const float2 v = a[(size_t)b * S + c];
        const float w = 2.f * s2 * f * gv;
        r = make_float2(w * v.x, w * v.y);
      }
      g[(size_t)b * S + c] = r;
    }
  }
}

// The same adjoint from layer K - 1 instead of layer K: where grad_p is nonzero the state's raw
// amplitude is pulled from its parents in `par` (the parent sum of k_layer_simple, plain sums);
// elsewhere the gradient is zero. For a grad_p that is zero almost everywhere this replaces the
// recompute of layer K and the read of its amplitudes by one pass that reads grad_p and writes g.
template <int KM>
__global__ __launch_bounds__(256) void k_readout_adjoint_pull(const uint64_t* cg_g, const int* lo_g, int me, int multiset,
                                                              int K, int m, int col, const float2* __restrict__ U,
                                                              const float2* __restrict__ par, uint64_t S_par,
                                                              const float* __restrict__ gp, float2* __restrict__ g,
                                                              uint64_t S, int64_t B, float s2) {
  // the counting tables are read from global memory: only the rare nonzero entries unrank, so
  // staging them in LDS per work group would cost more than it saves
  const slos::SpaceView sv{cg_g, lo_g, me, me + 1};
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < S; c += (uint64_t)gridDim.x * blockDim.x) {
    bool have_p = false;
    int p[KM];
    float f = 0.f;
    for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
      const float gv = gp[(size_t)b * S + c];
      float2 r = make_float2(0.f, 0.f);
      if (gv != 0.f) {
        if (!have_p) {
          slos::unrank_fixed<KM>(sv, c, K, p);
          f = slos::run_factorial_fixed<KM>(p, K, multiset);
          have_p = true;
        }
        const float2* ub = U + (size_t)b * m * m;
        const float2* pb = par + (size_t)b * S_par;
        Acc acc;
        slos::parents_fixed<KM>(sv, multiset, K, p,
                                [&](uint64_t pr, int mode, int) { cmac<false>(acc, ub[(size_t)mode * m + col], pb[pr]); });
        const float w = 2.f * s2 * f * gv;
        r = make_float2(w * acc.re, w * acc.im);
      }
      g[(size_t)b * S + c] = r;
    }
  }
}

// g_{k-1}[t] = sum_{children c = t + e_j} conj(U[b, j, col]) g_k[c]   (pull over children)
// and, per thread, gU_j += conj(psi_{k-1}[t]) g_k[t + e_j] in a private LDS row (no atomics);
// each block then reduces its rows to partial[b][block][j] (FP64 stage in k_reduce_gu)
// (k - 1 <= KM)
constexpr int kAdjThreads = 64;
template <int KM>
__global__ __launch_bounds__(kAdjThreads) void k_adjoint_layer(const uint64_t* cg_g, const int* lo_g, int me, int multiset,
                                                             int k, int m, int col, const float2* __restrict__ U,
                                                             const float2* __restrict__ psi, const float2* __restrict__ gk,
                                                             uint64_t Sk, float2* __restrict__ gout, uint64_t Sp,
                                                             double2* __restrict__ partial, int want_psi_grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (k + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  float2* ucol = reinterpret_cast<float2*>(smem + ((ncg * 8 + me * 4 + 15) & ~15));
  float2* slots = ucol + m;  // [kAdjThreads][me]
  const int64_t b = blockIdx.y;
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  for (int j = threadIdx.x; j < m; j += blockDim.x) ucol[j] = U[(size_t)b * m * m + (size_t)j * m + col];
  float2* my = slots + (size_t)threadIdx.x * me;
  for (int j = 0; j < me; ++j) my[j] = make_float2(0.f, 0.f);
  __syncthreads();
  const slos::SpaceView sv{cg, lo, me, stride};
  const float2* gb = gk + (size_t)b * Sk;
  const float2* pb = psi + (size_t)b * Sp;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < Sp; t += (uint64_t)gridDim.x * blockDim.x) {
    const float2 ps = pb[t];
    int p[KM];
    slos::unrank_fixed<KM>(sv, t, k - 1, p);
    float gr = 0.f, gi = 0.f;
    slos::children_fixed<KM>(sv, multiset, k, p, [&](uint64_t c, int j, int) {
      const float2 gv = gb[c];
      const float2 u = ucol[j];
      gr += u.x * gv.x + u.y * gv.y;   // conj(u) * gv
      gi += u.x * gv.y - u.y * gv.x;
      my[j].x += ps.x * gv.x + ps.y * gv.y;   // conj(psi) * gv
      my[j].y += ps.x * gv.y - ps.y * gv.x;
    });
    if (want_psi_grad) gout[(size_t)b * Sp + t] = make_float2(gr, gi);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < me; j += blockDim.x) {
    double sr = 0, si = 0;
    for (int r = 0; r < kAdjThreads; ++r) {
      sr += slots[(size_t)r * me + j].x;
      si += slots[(size_t)r * me + j].y;
    }
    partial[((size_t)b * gridDim.x + blockIdx.x) * me + j] = make_double2(sr, si);
  }
}

// gU[b][j][col] += sum over blocks of partial[b][block][j]
__global__ __launch_bounds__(64) void k_reduce_gu(const double2* __restrict__ partial, int nblk, int me, int m, int col,
                                                  float2* __restrict__ gU, int accumulate) {
  const int64_t b = blockIdx.y;
  const int j = blockIdx.x;
  double sr = 0, si = 0;
  for (int r = threadIdx.x; r < nblk; r += blockDim.x) {
    const double2 v = partial[((size_t)b * nblk + r) * me + j];
    sr += v.x;
    si += v.y;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_down(sr, o, 64);
    si += __shfl_down(si, o, 64);
  }
  if (threadIdx.x == 0) {
    float2* dst = gU + (size_t)b * m * m + (size_t)j * m + col;
    if (accumulate) *dst = make_float2(dst->x + (float)sr, dst->y + (float)si);
    else *dst = make_float2((float)sr, (float)si);
  }
}

// the stream-slot count instantiated for prefix width d (exact where the big layers live)
int block_maxd(int d) { return d <= 4 ? 4 : d <= 8 ? d : 16; }

// dynamic LDS of the single-plan block kernel for layer K with suffix length suf, or 0 if it
// does not fit the GPU's LDS
size_t block_lds_fit(int K, int me, int suf) {
  const size_t sh = slos::block_dyn_lds(K, me, block_maxd(K - suf));
  return sh + slos::kBlockStaticLds <= lds_limit() ? sh : 0;
}

// suffix length of the single-plan block kernel for layer K: 5-position suffixes where their table
// exists and the prefix is 1..8 positions long (5.4x fewer, larger blocks than with 4 at
// n = 11, m = 32: less per-block set-up and fewer part-filled sweep steps), else 4
int block_suffix(const slos_plan* p, const DevTables& dt, int K) {
  return (dt.tuples5 && K - 5 >= 1 && K - 5 <= 8 && block_lds_fit(K, p->me, 5)) ? 5 : slos::kSuf;
}

template <int KIND, bool COMP>
void launch_block_kind(int maxd, int suf, dim3 grid, size_t sh, hipStream_t st, const DevTables& dt,
                       const slos_plan* p, const float2* par, void* out, const slos::BlockLayer& L) {
#define SLOS_BLK(D, S, T)                                                                                        \
  hipLaunchKernelGGL((slos::k_layer_block<KIND, COMP, D, S>), grid, dim3(slos::kBlockThreads), sh, st, dt.cg, dt.lo, \
                     T, p->me, p->me + 1, p->multiset, par, out, L)
  // one stream slot per prefix position: exact widths where the big layers live (d = K - suf)
  if (suf == 5) {
    if (maxd <= 4) SLOS_BLK(4, 5, dt.tuples5);
    else if (maxd == 5) SLOS_BLK(5, 5, dt.tuples5);
    else if (maxd == 6) SLOS_BLK(6, 5, dt.tuples5);
    else if (maxd == 7) SLOS_BLK(7, 5, dt.tuples5);
    else SLOS_BLK(8, 5, dt.tuples5);
  } else {
    if (maxd <= 4) SLOS_BLK(4, 4, dt.tuples);
    else if (maxd == 5) SLOS_BLK(5, 4, dt.tuples);
    else if (maxd == 6) SLOS_BLK(6, 4, dt.tuples);
    else if (maxd == 7) SLOS_BLK(7, 4, dt.tuples);
    else if (maxd == 8) SLOS_BLK(8, 4, dt.tuples);
    else SLOS_BLK(16, 4, dt.tuples);
  }
#undef SLOS_BLK
}

int launch_block(const slos_plan* p, const DevTables& dt, const LayerArgs& a, int kind, bool comp, hipStream_t st) {
  slos::BlockLayer L;
  L.K = a.K; L.m = a.m; L.col = a.col; L.U = a.U; L.S_par = a.S_par; L.S = a.S;
  L.par_lo = a.par_lo; L.out_lo = a.out_lo;
  const int suf = block_suffix(p, dt, a.K);
  L.S_pre = p->size(a.K - suf);
  L.scale = a.scale; L.flags = a.flags;
  // children per wave: enough waves to fill the chip several times, but long ranges at scale
  uint64_t pw = (a.S / 16384 + 63) & ~63ull;
  // (2048 at most for the complex64 intermediate layers, 4096 for the read-out layer: measured best
  // at n = 11, m = 32)
  L.per_wave = std::min<uint64_t>(std::max<uint64_t>(pw, 256), kind == 0 ? 2048 : 4096);
  // XCD chunks of about 2^20 children (64 work groups at 4096 children per wave)
  L.xcd_chunk = (uint32_t)std::max<uint64_t>(1, (1ull << 20) / ((uint64_t)slos::kWaves * L.per_wave));
  const uint64_t waves = (a.S + L.per_wave - 1) / L.per_wave;
  const uint64_t wgs = (waves + slos::kWaves - 1) / slos::kWaves;
  if (wgs > 0x7fffffffull) return fail(SLOS_ERANGE, "layer too large for one launch");
  if (a.B > 65535) return fail(SLOS_ERANGE, "batch > 65535");
  // layers in group order where their work items exist (plain or compensated sums; probabilities or
  // complex64)
  const bool grp_last = kind == 1;
  if (suf == 5 && kind <= 1 && dt.items[a.K]) {
    const uint4* gitems = dt.items[a.K];
    const uint32_t gn = dt.n_items[a.K];
    const size_t shg = slos::group_dyn_lds(a.K, p->me);
    const uint64_t gwgs = ((uint64_t)gn + slos::kWaves - 1) / slos::kWaves;
    if (shg + slos::kBlockStaticLds <= lds_limit() && gwgs <= 0x7fffffffull) {
      L.xcd_chunk = (uint32_t)std::max<uint64_t>(1, (1ull << 20) / ((uint64_t)slos::kWaves * slos::kGroupItemChildren));
      const dim3 ggrid((unsigned)gwgs, (unsigned)a.B);
      const int maxd = a.K - suf;
#define SLOS_GRP2(D, C)                                                                                                \
  if (grp_last)                                                                                                        \
    hipLaunchKernelGGL((slos::k_layer_group<D, 1, C>), ggrid, dim3(slos::kBlockThreads), shg, st, dt.cg, dt.lo,          \
                       dt.tuples5, p->me, p->me + 1, a.par, a.out, L, gitems, gn);                                      \
  else                                                                                                                 \
    hipLaunchKernelGGL((slos::k_layer_group<D, 0, C>), ggrid, dim3(slos::kBlockThreads), shg, st, dt.cg, dt.lo,          \
                       dt.tuples5, p->me, p->me + 1, a.par, a.out, L, gitems, gn)
#define SLOS_GRP(D)      \
  if (comp) {            \
    SLOS_GRP2(D, true);  \
  } else {               \
    SLOS_GRP2(D, false); \
  }
      if (maxd <= 2) { SLOS_GRP(2); }
      else if (maxd == 3) { SLOS_GRP(3); }
      else if (maxd == 4) { SLOS_GRP(4); }
      else if (maxd == 5) { SLOS_GRP(5); }
      else { SLOS_GRP(6); }
#undef SLOS_GRP
#undef SLOS_GRP2
      HIP_TRY(hipGetLastError());
      return SLOS_OK;
    }
  }
  const size_t sh = block_lds_fit(a.K, p->me, suf);
  if (!sh) return fail(SLOS_ERANGE, "block kernel LDS exceeds the GPU's LDS");
  dim3 grid((unsigned)wgs, (unsigned)a.B);
  const int maxd = a.K - suf;
  if (kind == 0) comp ? launch_block_kind<0, true>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L)
                      : launch_block_kind<0, false>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L);
  else if (kind == 1) comp ? launch_block_kind<1, true>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L)
                           : launch_block_kind<1, false>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L);
  else comp ? launch_block_kind<2, true>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L)
            : launch_block_kind<2, false>(maxd, suf, grid, sh, st, dt, p, a.par, a.out, L);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

// calls f with the smallest photon bound KM (8, 16 or 32) of the thread-per-state kernels that
// holds `photons` (their per-state arrays are sized by it and stay in registers)
template <class F>
void with_km(int photons, F&& f) {
  if (photons <= 8) f(std::integral_constant<int, 8>{});
  else if (photons <= 16) f(std::integral_constant<int, 16>{});
  else f(std::integral_constant<int, 32>{});
}

// par_hilo / out_hilo (compensated sums only): the parent layer / the complex64 output layer is a
// hi/lo pair [2][B][S], the low-order parts after the high ones
int launch_layer(const slos_plan* p, const DevTables& dt, int K, int col, const float2* U, int64_t B,
                 const float2* par, uint64_t S_par, void* out, uint64_t S, int kind, float scale,
                 unsigned flags, hipStream_t st, bool par_hilo = false, bool out_hilo = false) {
  LayerArgs a;
  a.cg = dt.cg; a.lo = dt.lo; a.me = p->me; a.multiset = p->multiset; a.K = K; a.m = p->m; a.col = col;
  a.U = U; a.cols = nullptr; a.par = par; a.S_par = S_par; a.out = out; a.S = S; a.B = B;
  a.scale = scale; a.flags = flags;
  const bool comp = flags & SLOS_FLAG_COMPENSATED;
  a.par_lo = (comp && par_hilo) ? B * (int64_t)S_par : 0;
  a.out_lo = (comp && out_hilo && kind == 0) ? B * (int64_t)S : 0;
  if (!(flags & SLOS_FLAG_SIMPLE) && dt.tuples && K > slos::kSuf && K - slos::kSuf <= 16 &&
      block_lds_fit(K, p->me, slos::kSuf))
    return launch_block(p, dt, a, kind, comp, st);
  const size_t sh = (((size_t)(K + 1) * (p->me + 1) * 8 + p->me * 4 + 15) & ~(size_t)15) + (size_t)p->m * 8;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  const uint64_t blocks = std::min<uint64_t>((S + 255) / 256, 1u << 20);
  dim3 grid((unsigned)std::max<uint64_t>(blocks, 1), (unsigned)std::min<int64_t>(B, 65535));
  with_km(K, [&](auto km) {
    constexpr int KM = decltype(km)::value;
    if (kind == 0 && !comp) hipLaunchKernelGGL((k_layer_simple<0, false, KM>), grid, dim3(256), sh, st, a);
    else if (kind == 0) hipLaunchKernelGGL((k_layer_simple<0, true, KM>), grid, dim3(256), sh, st, a);
    else if (kind == 1 && !comp) hipLaunchKernelGGL((k_layer_simple<1, false, KM>), grid, dim3(256), sh, st, a);
    else if (kind == 1) hipLaunchKernelGGL((k_layer_simple<1, true, KM>), grid, dim3(256), sh, st, a);
    else if (!comp) hipLaunchKernelGGL((k_layer_simple<2, false, KM>), grid, dim3(256), sh, st, a);
    else hipLaunchKernelGGL((k_layer_simple<2, true, KM>), grid, dim3(256), sh, st, a);
  });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

// one complex128 layer (thread-per-state, FP64): kind 0 raw complex128, 1 float64 probabilities,
// 2 scaled complex128 amplitudes (SLOS_FLAG_IMMEDIATE: times sqrt(prod t!))
int launch_layer_f64(const slos_plan* p, const DevTables& dt, int K, int col, const void* U, int64_t B, const void* par,
                     uint64_t S_par, void* out, uint64_t S, int kind, double scale, unsigned flags, hipStream_t st) {
  slos::LayerArgsF64 a;
  a.cg = dt.cg; a.lo = dt.lo; a.me = p->me; a.multiset = p->multiset; a.K = K; a.m = p->m; a.col = col;
  a.U = static_cast<const double2*>(U); a.par = static_cast<const double2*>(par); a.S_par = S_par;
  a.out = out; a.S = S; a.B = B; a.scale = scale;
  a.immediate = (flags & SLOS_FLAG_IMMEDIATE) ? 1u : 0u;
  const size_t sh = slos::space_lds_f64(K, p->me) + (size_t)p->m * 16;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  const uint64_t blocks = std::min<uint64_t>((S + 255) / 256, 1u << 20);
  dim3 grid((unsigned)std::max<uint64_t>(blocks, 1), (unsigned)std::min<int64_t>(B, 65535));
#define SLOS_F64L(KM)                                                                         \
  if (kind == 0) hipLaunchKernelGGL((slos::k_layer_f64<0, KM>), grid, dim3(256), sh, st, a);  \
  else if (kind == 1) hipLaunchKernelGGL((slos::k_layer_f64<1, KM>), grid, dim3(256), sh, st, a); \
  else hipLaunchKernelGGL((slos::k_layer_f64<2, KM>), grid, dim3(256), sh, st, a)
  if (K <= 8) { SLOS_F64L(8); }
  else if (K <= 16) { SLOS_F64L(16); }
  else { SLOS_F64L(32); }
#undef SLOS_F64L
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

bool mul_overflow(uint64_t a, uint64_t b, uint64_t* r) { return __builtin_mul_overflow(a, b, r); }

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct WsLayout {
  size_t ones = 0, odd = 0, even = 0, total = 0;
};

// hilo: the intermediate layers are hi/lo pairs [2][B][S] (compensated sums carry their rounding
// errors from layer to layer), twice the bytes; f64: complex128 layers (16 bytes a state)
int forward_layout(const slos_plan* p, int64_t B, int layer, WsLayout* w, bool hilo = false, bool f64 = false) {
  uint64_t maxodd = 0, maxeven = 0;
  for (int k = 1; k < layer; ++k) (k & 1 ? maxodd : maxeven) = std::max(k & 1 ? maxodd : maxeven, p->size(k));
  uint64_t a = 0, b = 0;
  const uint64_t elem = f64 ? 16 : 8;
  const uint64_t per = (uint64_t)B * elem * (hilo ? 2 : 1);
  if (mul_overflow(maxodd, per, &a) || mul_overflow(maxeven, per, &b))
    return fail(SLOS_ERANGE, "workspace size overflows");
  w->ones = 0;
  w->odd = align256((size_t)B * elem);
  w->even = w->odd + align256(a);
  w->total = w->even + align256(b);
  return SLOS_OK;
}

// ------------------------------------------------------------------------------------------
// strong-then-sample: two-level inverse CDF over float32 probabilities in FP64, every addition in
// a fixed order (chunk = 64 entries summed in index order, block = 64 chunks summed in chunk
// order, block prefix = sequential inside segments of kSampleSeg blocks plus the sequential
// prefix of the segment totals), then the per-shot search. Replaces the float64 cumsum +
// searchsorted of imperfections.py:221-237 (above 2^24 categories) without an [S] float64 copy.
// ------------------------------------------------------------------------------------------
constexpr int kSampleChunk = 64;
constexpr int kSampleBlock = kSampleChunk * 64;   // 4096 entries
constexpr int kSampleSeg = 1024;                   // blocks per prefix segment

// One wave per block of 4096 entries; lane c loads its own chunk c (64 consecutive entries, 16 bytes
// at a time: the two cache lines of a chunk are read by one lane in consecutive loads) and sums it
// in index order (FP64, from 0.0; entries past S count as 0), then every lane sums the 64 chunk
// sums in chunk order. No LDS, so many waves stay resident to cover the stream.
constexpr int kSumWaves = 4;
__global__ __launch_bounds__(64 * kSumWaves) void k_sample_block_sums(const float* __restrict__ p, uint64_t S, uint64_t nb,
                                                                      double* __restrict__ bsum,
                                                                      double* __restrict__ csum) {
  const int lane = threadIdx.x & 63;
  const uint64_t b = (uint64_t)blockIdx.x * kSumWaves + (uint64_t)(threadIdx.x >> 6);
  if (b >= nb) return;
  const uint64_t i0 = b * kSampleBlock + (uint64_t)lane * kSampleChunk;
  double cs = 0.0;
  if (i0 + kSampleChunk <= S) {
    float4 v[kSampleChunk / 4];
#pragma unroll
    for (int j = 0; j < kSampleChunk / 4; ++j) __builtin_memcpy(&v[j], p + i0 + 4 * j, 16);
#pragma unroll
    for (int j = 0; j < kSampleChunk / 4; ++j) {
      cs += (double)v[j].x;
      cs += (double)v[j].y;
      cs += (double)v[j].z;
      cs += (double)v[j].w;
    }
  } else {
    for (int k = 0; k < kSampleChunk; ++k) cs += (double)(i0 + k < S ? p[i0 + k] : 0.f);
  }
  csum[b * 64 + (uint64_t)lane] = cs;
  double tot = 0.0;
  for (int c = 0; c < 64; ++c) tot += __shfl(cs, c, 64);
  if (lane == 0) bsum[b] = tot;
}

// Block prefix over segments of kSampleSeg blocks: one work group per segment stages its block sums
// in LDS (coalesced), one thread takes their inclusive prefix in order (from 0.0) and the group
// writes it back (coalesced); segtot[seg] = the segment's total.
__global__ __launch_bounds__(256) void k_sample_seg_prefix(const double* __restrict__ bsum, uint64_t nb,
                                                           double* __restrict__ cdf, double* __restrict__ segtot) {
  __shared__ double seg[kSampleSeg];
  const uint64_t a = (uint64_t)blockIdx.x * kSampleSeg;
  const int len = (int)(nb - a < (uint64_t)kSampleSeg ? nb - a : (uint64_t)kSampleSeg);
  for (int i = threadIdx.x; i < len; i += blockDim.x) seg[i] = bsum[a + i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (int i = 0; i < len; ++i) {
      acc += seg[i];
      seg[i] = acc;
    }
    segtot[blockIdx.x] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < len; i += blockDim.x) cdf[a + i] = seg[i];
}

// exclusive prefix of the segment totals, in order (one thread; at most 1024 segments)
__global__ __launch_bounds__(64) void k_sample_seg_offsets(double* __restrict__ segtot, int nseg) {
  if (threadIdx.x != 0) return;
  double run = 0.0;
  for (int i = 0; i < nseg; ++i) {
    const double v = segtot[i];
    segtot[i] = run;
    run += v;
  }
}

// cdf = segment offset + the segment's own prefix
__global__ __launch_bounds__(256) void k_sample_add_offsets(double* __restrict__ cdf, uint64_t nb,
                                                            const double* __restrict__ segoff) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * blockDim.x)
    cdf[i] = segoff[i / kSampleSeg] + cdf[i];
}

// one thread per shot: u = uniform * total; block = first cdf >= u (binary search), then the chunk
// and the entry by in-order prefix sums of the block pass's chunk sums and of the chunk's entries
__global__ __launch_bounds__(256) void k_sample_search(const float* __restrict__ p, uint64_t S, uint64_t nb,
                                                       const double* __restrict__ cdf, const double* __restrict__ csum,
                                                       const double* __restrict__ uni, int64_t shots,
                                                       uint64_t* __restrict__ idx) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= shots) return;
  const double total = cdf[nb - 1];
  const double u = uni[s] * total;
  uint64_t lo = 0, hi = nb - 1;  // first b with cdf[b] >= u, else nb - 1
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (cdf[mid] >= u) hi = mid;
    else lo = mid + 1;
  }
  const uint64_t b = lo;
  double r = u - (b > 0 ? cdf[b - 1] : 0.0);
  const double* cb = csum + b * 64;
  int c = 63;
  double pre = 0.0;
  for (int q = 0; q < 64; ++q) {
    const double v = cb[q];
    if (pre + v >= r) {
      c = q;
      break;
    }
    pre += v;
  }
  r -= pre;
  const uint64_t first = b * kSampleBlock + (uint64_t)c * kSampleChunk;
  int k = 63;
  double acc = 0.0;
  for (int q = 0; q < kSampleChunk; ++q) {
    acc += (double)(first + q < S ? p[first + q] : 0.f);
    if (acc >= r) {
      k = q;
      break;
    }
  }
  uint64_t out = first + (uint64_t)k;
  if (out >= S) out = S - 1;
  idx[s] = out;
}

// lex index -> occupation row (int32 [m]) of layer k, one thread per index
// (an index >= S_k gives a row of -1)
__global__ __launch_bounds__(256) void k_unrank_occ(const uint64_t* cg_g, const int* lo_g, int me, int k, int m,
                                                    const uint64_t* __restrict__ idx, int64_t count,
                                                    int32_t* __restrict__ occ) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int stride = me + 1;
  const int ncg = (k + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  __syncthreads();
  const slos::SpaceView sv{cg, lo, me, stride};
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < count; s += (int64_t)gridDim.x * blockDim.x) {
    int32_t* row = occ + (size_t)s * m;
    const uint64_t g = idx[s];
    if (g >= cg[(size_t)k * stride]) {
      for (int j = 0; j < m; ++j) row[j] = -1;
      continue;
    }
    int pos[kMaxPhotons];
    sv.unrank(g, k, pos);
    for (int j = 0; j < m; ++j) row[j] = 0;
    for (int i = 0; i < k; ++i) row[pos[i]] += 1;
  }
}

// ------------------------------------------------------------------------------------------
// uneven batches: a group of plans, one launch per photon layer (see layer_block.hpp)
// ------------------------------------------------------------------------------------------

// thread-per-state form of a group layer: item = (plan, instance, 256 * spt consecutive states)
// (K <= KM)
template <bool COMP, int KM>
__global__ __launch_bounds__(256) void k_layer_simple_group(slos::GroupLaunch gl) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int K = gl.K;
  const uint64_t* cur = nullptr;
  slos::for_each_item(gl.total, [&](uint32_t item) {
    const slos::GroupItem& it = gl.it[slos::group_index(gl, item)];
    const uint32_t rel = item - it.wg_begin;
    const uint32_t b = rel / it.wgs_per_inst, w = rel - b * it.wgs_per_inst;
    const int stride = it.me + 1;
    const int ncg = (K + 1) * stride;
    uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
    int* lo = reinterpret_cast<int*>(cg + ncg);
    float2* ucol = reinterpret_cast<float2*>(smem + ((ncg * 8 + it.me * 4 + 15) & ~15));
    __syncthreads();  // the previous item is done with LDS
    if (it.cg != cur) {
      for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = it.cg[i];
      for (int i = threadIdx.x; i < it.me; i += blockDim.x) lo[i] = it.lo[i];
      cur = it.cg;
    }
    for (int j = threadIdx.x; j < it.m; j += blockDim.x)
      ucol[j] = it.U[(size_t)b * it.m * it.m + (size_t)j * it.m + it.col];
    __syncthreads();
    const slos::SpaceView sv{cg, lo, it.me, stride};
    const float2* pb = it.par + (size_t)b * it.S_par;
    const uint32_t spt = (uint32_t)it.per_wave;
    const uint64_t g0 = (uint64_t)w * 256u * spt + threadIdx.x;
    for (uint32_t i = 0; i < spt; ++i) {
      const uint64_t g = g0 + 256ull * i;
      if (g >= it.S) break;
      int pos[KM];
      slos::unrank_fixed<KM>(sv, g, K, pos);
      Acc acc;
      slos::parents_fixed<KM>(sv, it.multiset, K, pos,
                              [&](uint64_t pr, int mode, int) { cmac<COMP>(acc, ucol[mode], pb[pr]); });
      if constexpr (COMP) {
        if (it.par_lo)  // the parents' low-order parts, into the correction terms
          slos::parents_fixed<KM>(sv, it.multiset, K, pos, [&](uint64_t pr, int mode, int) {
            const float2 u = ucol[mode], v = pb[it.par_lo + pr];
            acc.cre = fmaf(u.x, v.x, acc.cre);
            acc.cre = fmaf(-u.y, v.y, acc.cre);
            acc.cim = fmaf(u.x, v.y, acc.cim);
            acc.cim = fmaf(u.y, v.x, acc.cim);
          });
      }
      float re = COMP ? acc.re + acc.cre : acc.re, im = COMP ? acc.im + acc.cim : acc.im;
      if (it.kind == 1) {
        const float f = slos::run_factorial_fixed<KM>(pos, K, it.multiset);
        re *= it.scale;
        im *= it.scale;
        reinterpret_cast<float*>(it.out)[(size_t)b * it.S + g] = slos::prob_readout(re, im, f);
      } else {
        reinterpret_cast<float2*>(it.out)[(size_t)b * it.S + g] = make_float2(re, im);
        if constexpr (COMP) {
          if (it.out_lo)
            reinterpret_cast<float2*>(it.out)[it.out_lo + (size_t)b * it.S + g] =
                make_float2(acc.cre - (re - acc.re), acc.cim - (im - acc.im));
        }
      }
    }
  });
}

size_t simple_lds_bytes(int K, int me, int m) {
  return (((size_t)(K + 1) * (me + 1) * 8 + (size_t)me * 4 + 15) & ~(size_t)15) + (size_t)m * 8;
}

}  // namespace

// a group of plans run together (uneven batches)
struct slos_group {
  struct Cfg {
    int block = -1;         // 1 block kernel, 0 thread-per-state, 2 one single-plan launch per plan, -1 default
    uint64_t per_wave = 0;  // children per wave (block) / states per thread (simple); 0: default
    uint32_t grid = 0;      // work groups of the persistent launch; 0: one per item
  };
  std::vector<const slos_plan*> plans;
  int nmax = 0;
  std::vector<Cfg> cfg;     // [nmax + 1]
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int cap_device = -1;
};

namespace {

struct GroupArgs {
  std::vector<const float2*> U;
  std::vector<int64_t> B;
  std::vector<float*> out;
  std::vector<float> scale;
  unsigned flags = 0;
  unsigned char* ws = nullptr;
  size_t ws_bytes = 0;
};

struct GroupWs {
  std::vector<size_t> ones, buf0, buf1;
  size_t total = 0;
};

// hilo (compensated sums): every plan's intermediate layers are hi/lo pairs [2][B][S]
int group_layout(const slos_group* G, const int64_t* B, GroupWs* w, bool hilo) {
  size_t off = 0;
  for (size_t g = 0; g < G->plans.size(); ++g) {
    const slos_plan* p = G->plans[g];
    if (B[g] < 1 || B[g] > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
    uint64_t maxS = 0, bytes = 0;
    for (int k = 1; k < p->n; ++k) maxS = std::max(maxS, p->size(k));
    if (mul_overflow(maxS, (uint64_t)B[g] * 8 * (hilo ? 2 : 1), &bytes))
      return fail(SLOS_ERANGE, "workspace size overflows");
    w->ones.push_back(off);
    off += align256((size_t)B[g] * 8);
    w->buf0.push_back(off);
    off += align256(bytes);
    w->buf1.push_back(off);
    off += align256(bytes);
  }
  w->total = off;
  return SLOS_OK;
}

int group_args(const slos_group* G, const void* const* U, const int64_t* B, float* const* out, const double* scale,
               unsigned flags, void* ws, size_t ws_bytes, GroupArgs* a, GroupWs* w) {
  if (!G || !U || !B || !out || !scale) return fail(SLOS_EINVAL, "NULL argument");
  int rc = group_layout(G, B, w, (flags & SLOS_FLAG_COMPENSATED) != 0);
  if (rc) return rc;
  if (!ws || ws_bytes < w->total)
    return fail(SLOS_ENOMEM, "workspace too small: need " + std::to_string(w->total) + " bytes");
  for (size_t g = 0; g < G->plans.size(); ++g) {
    if (!U[g] || !out[g]) return fail(SLOS_EINVAL, "NULL unitary or output");
    a->U.push_back(static_cast<const float2*>(U[g]));
    a->B.push_back(B[g]);
    a->out.push_back(out[g]);
    a->scale.push_back((float)scale[g]);
  }
  a->flags = flags;
  a->ws = static_cast<unsigned char*>(ws);
  a->ws_bytes = ws_bytes;
  return SLOS_OK;
}

bool group_can_block(const slos_group* G, const std::vector<DevTables>& dts, int K, unsigned flags) {
  if ((flags & SLOS_FLAG_SIMPLE) || K <= slos::kSuf || K - slos::kSuf > 16) return false;
  size_t off = 0;
  for (size_t g = 0; g < G->plans.size(); ++g)
    if (G->plans[g]->n >= K) {
      if (!dts[g].tuples) return false;
      off = std::max(off, slos::block_table_offset(K, G->plans[g]->me));
    }
  const size_t sh = off + (size_t)slos::kWaves * slos::kBlockBatch * slos::block_words(block_maxd(K - slos::kSuf)) * 4;
  return sh + slos::kBlockStaticLds <= lds_limit();
}

template <bool COMP>
void launch_block_group(int maxd, uint32_t grid, size_t sh, hipStream_t st, const slos::GroupLaunch& gl) {
#define SLOS_BLKG(D) hipLaunchKernelGGL((slos::k_layer_block_group<COMP, D>), dim3(grid), dim3(slos::kBlockThreads), sh, st, gl)
  if (maxd <= 4) SLOS_BLKG(4);
  else if (maxd == 5) SLOS_BLKG(5);
  else if (maxd == 6) SLOS_BLKG(6);
  else if (maxd == 7) SLOS_BLKG(7);
  else if (maxd == 8) SLOS_BLKG(8);
  else SLOS_BLKG(16);
#undef SLOS_BLKG
}

// layer K of every plan with n >= K as one launch per plan, through the single-plan path (the
// 5-position-suffix and group-order kernels where they apply): the same values as the group kernel
int group_layer_per_plan(const slos_group* G, int K, const GroupArgs& a, const GroupWs& w,
                         const std::vector<DevTables>& dts, hipStream_t st) {
  const bool hilo = (a.flags & SLOS_FLAG_COMPENSATED) != 0;
  for (size_t g = 0; g < G->plans.size(); ++g) {
    const slos_plan* p = G->plans[g];
    if (p->n < K) continue;
    float2* bufs[2] = {reinterpret_cast<float2*>(a.ws + w.buf0[g]), reinterpret_cast<float2*>(a.ws + w.buf1[g])};
    const float2* par = K == 1 ? reinterpret_cast<const float2*>(a.ws + w.ones[g]) : bufs[(K - 1) & 1];
    const bool last = K == p->n;
    void* out = last ? static_cast<void*>(a.out[g]) : static_cast<void*>(bufs[K & 1]);
    const int rc = launch_layer(p, dts[g], K, p->modes[K - 1], a.U[g], a.B[g], par, p->size(K - 1), out, p->size(K),
                                last ? 1 : 0, last ? a.scale[g] : 1.f, a.flags & ~SLOS_FLAG_HILO, st, hilo && K >= 2,
                                hilo && !last);
    if (rc) return rc;
  }
  return SLOS_OK;
}

// layer K of every plan with n >= K, in one launch (or one per plan: Cfg.block 2, and by default
// where a single plan reaches layer K)
int group_layer(const slos_group* G, int K, const GroupArgs& a, const GroupWs& w, const std::vector<DevTables>& dts,
                const slos_group::Cfg& c, hipStream_t st) {
  int active = 0;
  for (size_t g = 0; g < G->plans.size(); ++g) active += G->plans[g]->n >= K;
  if (c.block == 2 || (c.block == -1 && active == 1)) return group_layer_per_plan(G, K, a, w, dts, st);
  slos::GroupLaunch gl{};
  gl.K = K;
  const bool block = group_can_block(G, dts, K, a.flags) && c.block != 0;
  uint64_t children = 0;
  for (size_t g = 0; g < G->plans.size(); ++g)
    if (G->plans[g]->n >= K) children += G->plans[g]->size(K) * (uint64_t)a.B[g];
  uint64_t pw_default = std::min<uint64_t>(std::max<uint64_t>((children / 16384 + 63) & ~63ull, 256), 4096);
  size_t sh = 0;
  uint64_t total = 0;
  int cnt = 0;
  for (size_t g = 0; g < G->plans.size(); ++g) {
    const slos_plan* p = G->plans[g];
    if (p->n < K) continue;
    slos::GroupItem& it = gl.it[cnt++];
    const DevTables& dt = dts[g];
    it.cg = dt.cg;
    it.lo = dt.lo;
    it.tuples = dt.tuples;
    it.U = a.U[g];
    const float2* bufs[2] = {reinterpret_cast<const float2*>(a.ws + w.buf0[g]),
                             reinterpret_cast<const float2*>(a.ws + w.buf1[g])};
    it.par = K == 1 ? reinterpret_cast<const float2*>(a.ws + w.ones[g]) : bufs[(K - 1) & 1];
    const bool last = K == p->n;
    it.out = last ? static_cast<void*>(a.out[g]) : const_cast<float2*>(bufs[K & 1]);
    it.S_par = p->size(K - 1);
    it.S = p->size(K);
    it.S_pre = block ? p->size(K - slos::kSuf) : 0;
    it.kind = last ? 1 : 0;
    it.scale = last ? a.scale[g] : 1.f;
    it.flags = a.flags;
    // compensated sums carry hi/lo layers (the vacuum row of layer 0 has no low part)
    const bool hilo = (a.flags & SLOS_FLAG_COMPENSATED) != 0;
    it.par_lo = (hilo && K >= 2) ? a.B[g] * (int64_t)it.S_par : 0;
    it.out_lo = (hilo && !last) ? a.B[g] * (int64_t)it.S : 0;
    it.me = p->me;
    it.multiset = p->multiset;
    it.m = p->m;
    it.col = p->modes[K - 1];
    uint64_t wgs;
    if (block) {
      it.per_wave = c.per_wave ? c.per_wave : pw_default;
      wgs = ((it.S + it.per_wave - 1) / it.per_wave + slos::kWaves - 1) / slos::kWaves;
      sh = std::max(sh, slos::block_table_offset(K, p->me));
    } else {
      it.per_wave = c.per_wave ? c.per_wave : 1;
      wgs = (it.S + 256 * it.per_wave - 1) / (256 * it.per_wave);
      sh = std::max(sh, simple_lds_bytes(K, p->me, p->m));
    }
    it.wg_begin = (uint32_t)total;
    it.wgs_per_inst = (uint32_t)wgs;
    total += wgs * (uint64_t)a.B[g];
    if (total > 0x7fffffffull) return fail(SLOS_ERANGE, "group layer too large for one launch");
  }
  if (cnt == 0) return SLOS_OK;
  gl.count = cnt;
  gl.total = (uint32_t)total;
  if (block) {
    gl.table_off = (uint32_t)sh;
    sh += (size_t)slos::kWaves * slos::kBlockBatch * slos::block_words(block_maxd(K - slos::kSuf)) * 4;
  }
  if (sh + (block ? slos::kBlockStaticLds : 0) > lds_limit()) return fail(SLOS_ERANGE, "group layer LDS exceeds the GPU's LDS");
  const uint32_t all = (uint32_t)(8 * ((total + 7) / 8));
  const uint32_t grid = c.grid ? std::min<uint32_t>(std::max<uint32_t>(c.grid & ~7u, 8), all) : all;
  const bool comp = a.flags & SLOS_FLAG_COMPENSATED;
  if (block) comp ? launch_block_group<true>(K - slos::kSuf, grid, sh, st, gl)
                  : launch_block_group<false>(K - slos::kSuf, grid, sh, st, gl);
  else
    with_km(K, [&](auto km) {
      constexpr int KM = decltype(km)::value;
      if (comp) hipLaunchKernelGGL((k_layer_simple_group<true, KM>), dim3(grid), dim3(256), sh, st, gl);
      else hipLaunchKernelGGL((k_layer_simple_group<false, KM>), dim3(grid), dim3(256), sh, st, gl);
    });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int group_tables(const slos_group* G, std::vector<DevTables>* dts) {
  dts->resize(G->plans.size());
  for (size_t g = 0; g < G->plans.size(); ++g) {
    int rc = device_tables(G->plans[g], &(*dts)[g]);
    if (rc) return rc;
  }
  return SLOS_OK;
}

int group_forward(const slos_group* G, const GroupArgs& a, const GroupWs& w, const std::vector<DevTables>& dts,
                  hipStream_t st) {
  for (size_t g = 0; g < G->plans.size(); ++g) {
    hipLaunchKernelGGL(k_fill_ones, dim3((unsigned)std::min<int64_t>((a.B[g] + 255) / 256, 4096)), dim3(256), 0, st,
                       reinterpret_cast<float2*>(a.ws + w.ones[g]), a.B[g]);
    HIP_TRY(hipGetLastError());
  }
  for (int K = 1; K <= G->nmax; ++K) {
    int rc = group_layer(G, K, a, w, dts, G->cfg[K], st);
    if (rc) return rc;
  }
  return SLOS_OK;
}

void group_drop_graph(slos_group* G) {
  if (G->exec) (void)hipGraphExecDestroy(G->exec);
  if (G->graph) (void)hipGraphDestroy(G->graph);
  G->exec = nullptr;
  G->graph = nullptr;
}

}  // namespace

void build_tuples(slos_plan* p) {
  const auto sv = p->host_view();
  const uint64_t N = p->size(slos::kSuf);
  p->tuples.resize(N);
  for (uint64_t u = 0; u < N; ++u) p->tuples[u] = slos::make_tuple_word(sv, p->multiset, u);
  // 5-position suffixes: positions < 64 keep every rank term inside its packed field
  // (cg[4][1] < 2^24, cg[3][1] and cg[2][1] < 2^16, cg[1][0] < 2^8)
  if (p->n > 5 && p->size(5) <= (4u << 20)) {
    const uint64_t N5 = p->size(5);
    p->tuples5.resize(N5);
    for (uint64_t u = 0; u < N5; ++u) p->tuples5[u] = slos::make_tuple_word<5>(sv, p->multiset, u);
    // the last layer in group order: multiset spaces, prefixes of 2..6 positions
    // (item ranks are 32-bit; at most about 32M items a layer)
    for (int K = 7; K <= p->n && K <= 11; ++K)
      if (p->multiset && p->size(K - 6) <= 0xffffffffull &&
          p->size(K) / slos::kGroupItemChildren + p->size(K - 6) <= (32u << 20))
        slos::group_items(sv, K, p->items[K]);
  }
}

// the keyed group sums of slos_group_sums / slos_group_sums_f64 (k_group_partial, k_group_final)
namespace {

template <class T>
int group_sums(const T* probs, int64_t batch, uint64_t count, const int64_t* order, const int64_t* chunks,
               uint32_t n_chunks, const uint32_t* first, int n_groups, T* out, void* ws, size_t ws_bytes,
               slos_stream_t stream) {
  if (!probs || !order || !first || !out || batch < 1 || batch > 65535 || n_groups < 1 || (n_chunks && !chunks))
    return fail(SLOS_EINVAL, "bad argument");
  if (n_chunks && (!ws || ws_bytes < (size_t)batch * n_chunks * sizeof(double)))
    return fail(SLOS_ENOMEM, "workspace too small");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  if (n_chunks) {
    hipLaunchKernelGGL(k_group_partial<T>, dim3((n_chunks + 3) / 4, (unsigned)batch), dim3(256), 0, st, probs, count,
                       order, chunks, n_chunks, part);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_group_final<T>, dim3((unsigned)((n_groups + 255) / 256), (unsigned)batch), dim3(256), 0, st,
                     part, n_chunks, first, n_groups, out);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

}  // namespace

extern "C" {

const char* slos_version(void) { return "slos_mi355x 0.1 (gfx950)"; }

const char* slos_last_error(void) { return g_err.c_str(); }

int slos_plan_create(int n, int m, int mask, const int32_t* input_modes, slos_plan** out) {
  if (!out) return fail(SLOS_EINVAL, "out is NULL");
  *out = nullptr;
  if (n < 0) return fail(SLOS_EINVAL, "n must be >= 0");
  if (n > kMaxPhotons) return fail(SLOS_ERANGE, "at most 32 photons");
  if (m < 1 || m > 4096) return fail(SLOS_EINVAL, "m must be in [1, 4096]");
  if (!input_modes) return fail(SLOS_EINVAL, "input_modes is NULL");
  auto* p = new slos_plan();
  p->n = n; p->m = m; p->mask = mask;
  switch (mask) {
    case SLOS_MASK_FOCK: p->me = m; p->multiset = 1; break;
    case SLOS_MASK_UNBUNCHED: p->me = m; p->multiset = 0; p->shift = 0; break;
    case SLOS_MASK_DUAL_RAIL: p->me = m; p->multiset = 0; p->shift = 1; break;
    case SLOS_MASK_SPARSE_1IN4: p->me = m; p->multiset = 0; p->shift = 2; break;
    case SLOS_MASK_TAIL_EMPTY: p->me = m - 2; p->multiset = 1; break;
    default: delete p; return fail(SLOS_EINVAL, "unknown mask id " + std::to_string(mask));
  }
  if (p->me < 1) { delete p; return fail(SLOS_EINVAL, "empty output space (too few modes for the mask)"); }
  std::vector<int> cnt(m, 0);
  for (int k = 0; k < n; ++k) {
    if (input_modes[k] < 0 || input_modes[k] >= m) { delete p; return fail(SLOS_EINVAL, "input mode out of range"); }
    p->modes.push_back(input_modes[k]);
    cnt[input_modes[k]]++;
  }
  for (int c : cnt) for (int i = 2; i <= c; ++i) p->s_fact *= i;
  p->lo.resize(p->me);
  for (int x = 0; x < p->me; ++x) {
    const int v = p->multiset ? x : (((x >> p->shift) + 1) << p->shift);
    p->lo[x] = std::min(v, p->me);
  }
  const int st = p->me + 1;
  p->cg.assign((size_t)(n + 2) * st, 0);
  for (int x = 0; x <= p->me; ++x) p->cg[x] = 1;
  for (int r = 1; r <= n + 1; ++r)
    for (int x = p->me - 1; x >= 0; --x) {
      uint64_t v;
      if (__builtin_add_overflow(p->cg[(size_t)r * st + x + 1], p->cg[(size_t)(r - 1) * st + p->lo[x]], &v)) {
        if (r <= n) { delete p; return fail(SLOS_ERANGE, "layer size exceeds 64 bits"); }
        v = UINT64_MAX;  // row n+1 is only a guard
      }
      p->cg[(size_t)r * st + x] = v;
    }
  // the block kernel's tuple words hold positions < 64 (6 bits and a mark bit per byte)
  if (n > slos::kSuf && p->m <= slos::kBlockModes && p->size(slos::kSuf) <= (4u << 20)) build_tuples(p);
  if (p->size(n) == 0) {
    delete p;
    return fail(SLOS_EINVAL, "empty output space for n=" + std::to_string(n) + " m=" + std::to_string(m));
  }
  *out = p;
  return SLOS_OK;
}

int slos_plan_destroy(slos_plan* p) {
  if (!p) return SLOS_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& t : p->dev) {
    (void)hipSetDevice(t.device);
    (void)hipFree(t.cg);
    (void)hipFree(t.lo);
    (void)hipFree(t.modes);
    if (t.tuples) (void)hipFree(t.tuples);
    if (t.tuples5) (void)hipFree(t.tuples5);
    for (auto* it : t.items)
      if (it) (void)hipFree(it);
  }
  (void)hipSetDevice(cur);
  delete p;
  return SLOS_OK;
}

int slos_layer_size(const slos_plan* p, int k, uint64_t* size) {
  if (!p || !size) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 0 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  *size = p->size(k);
  return SLOS_OK;
}

int slos_rank(const slos_plan* p, int k, const int32_t* pos, int64_t count, uint64_t* idx) {
  if (!p || (!pos && count) || (!idx && count)) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 0 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  const auto sv = p->host_view();
  for (int64_t q = 0; q < count; ++q) {
    const int32_t* t = pos + q * k;
    int l = 0, run = 0;
    for (int i = 0; i < k; ++i) {
      if (t[i] < l || t[i] >= p->me) return fail(SLOS_EINVAL, "positions do not describe an admissible sorted state");
      run = (i > 0 && t[i] == t[i - 1]) ? run + 1 : 1;
      if (sv.C(k - i - 1, p->lo[t[i]]) == 0) return fail(SLOS_EINVAL, "state has no admissible completion");
      l = p->lo[t[i]];
    }
    (void)run;
    idx[q] = sv.rank(reinterpret_cast<const int*>(t), k);
  }
  return SLOS_OK;
}

int slos_unrank(const slos_plan* p, int k, const uint64_t* idx, int64_t count, int32_t* pos) {
  if (!p || (!pos && count) || (!idx && count)) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 0 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  const auto sv = p->host_view();
  const uint64_t S = p->size(k);
  for (int64_t q = 0; q < count; ++q) {
    if (idx[q] >= S) return fail(SLOS_ERANGE, "index >= layer size");
    sv.unrank(idx[q], k, reinterpret_cast<int*>(pos + q * k));
  }
  return SLOS_OK;
}

int slos_prob_norms(const slos_plan* p, int k, uint64_t first, int64_t count, double* out) {
  if (!p || (!out && count)) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 0 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  const uint64_t S = p->size(k);
  if (count < 0 || first > S || (uint64_t)count > S - first) return fail(SLOS_ERANGE, "range outside the layer");
  const auto sv = p->host_view();
  int t[kMaxPhotons];
  for (int64_t q = 0; q < count; ++q) {
    sv.unrank(first + q, k, t);
    double f = 1, run = 1;
    if (p->multiset)
      for (int i = 1; i < k; ++i) {
        run = (t[i] == t[i - 1]) ? run + 1 : 1;
        f *= run;
      }
    out[q] = f;
  }
  return SLOS_OK;
}

int slos_forward_workspace(const slos_plan* p, int64_t batch, int layer, unsigned flags, size_t* bytes) {
  if (!p || !bytes) return fail(SLOS_EINVAL, "NULL argument");
  if (batch < 1) return fail(SLOS_EINVAL, "batch must be >= 1");
  if (layer < 0 || layer > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if ((flags & SLOS_FLAG_F64) && (flags & (SLOS_FLAG_COMPENSATED | SLOS_FLAG_HILO)))
    return fail(SLOS_EINVAL, "SLOS_FLAG_F64 computes in FP64: no compensated sums");
  WsLayout w;
  int rc = forward_layout(p, batch, layer, &w, (flags & SLOS_FLAG_COMPENSATED) != 0, (flags & SLOS_FLAG_F64) != 0);
  if (rc) return rc;
  *bytes = w.total;
  return SLOS_OK;
}

namespace {

// argument checks shared by slos_forward and slos_forward_noalloc: the input columns and the
// workspace layout
int forward_check(const slos_plan* p, const void* U, int64_t batch, const int32_t* modes, int layer, int out_kind,
                  void* out, unsigned flags, int* cols, WsLayout* w) {
  if (!p || !U || !out) return fail(SLOS_EINVAL, "NULL argument");
  if (batch < 1) return fail(SLOS_EINVAL, "batch must be >= 1");
  if (layer < 0 || layer > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (out_kind != SLOS_OUT_PROBS && out_kind != SLOS_OUT_AMPS) return fail(SLOS_EINVAL, "unknown output kind");
  for (int k = 0; k < layer; ++k) {
    cols[k] = modes ? modes[k] : p->modes[k];
    if (cols[k] < 0 || cols[k] >= p->m) return fail(SLOS_EINVAL, "input mode out of range");
  }
  if ((flags & SLOS_FLAG_F64) && (flags & (SLOS_FLAG_COMPENSATED | SLOS_FLAG_HILO)))
    return fail(SLOS_EINVAL, "SLOS_FLAG_F64 computes in FP64: no compensated sums");
  return forward_layout(p, batch, layer, w, (flags & SLOS_FLAG_COMPENSATED) != 0, (flags & SLOS_FLAG_F64) != 0);
}

// the complex128 forward (SLOS_FLAG_F64): the same layer walk as forward_run, FP64 kernels
int forward_run_f64(const slos_plan* p, const DevTables& dt, const void* U, int64_t batch, const int* cols, int layer,
                    int out_kind, void* out, double scale, unsigned flags, const WsLayout& w, void* ws,
                    hipStream_t st) {
  auto* base = static_cast<unsigned char*>(ws);
  double2* ones = reinterpret_cast<double2*>(base + w.ones);
  void* bufs[2] = {base + w.even, base + w.odd};
  const dim3 g1((unsigned)std::min<int64_t>((batch + 255) / 256, 4096));
  if (layer == 0) {  // no photon: the vacuum
    hipLaunchKernelGGL(slos::k_vacuum_out_f64, g1, dim3(256), 0, st, out, batch, out_kind == SLOS_OUT_PROBS ? 1 : 0,
                       scale);
    HIP_TRY(hipGetLastError());
    return SLOS_OK;
  }
  hipLaunchKernelGGL(slos::k_fill_ones_f64, g1, dim3(256), 0, st, ones, batch);
  HIP_TRY(hipGetLastError());
  const void* par = ones;
  uint64_t S_par = 1;
  const RoctxRange whole("slos_forward");
  for (int k = 1; k <= layer; ++k) {
    const uint64_t S = p->size(k);
    const bool last = (k == layer);
    void* dst = last ? out : bufs[k & 1];
    const int kind = last ? (out_kind == SLOS_OUT_PROBS ? 1 : 2) : 0;
    char name[32];
    std::snprintf(name, sizeof name, "slos layer %d", k);
    const RoctxRange range(name);
    const int rc = launch_layer_f64(p, dt, k, cols[k - 1], U, batch, par, S_par, dst, S, kind, scale, flags, st);
    if (rc) return rc;
    par = dst;
    S_par = S;
  }
  return SLOS_OK;
}

// the layer launches of a forward whose workspace `ws` follows layout `w` (no allocation)
int forward_run(const slos_plan* p, const DevTables& dt, const void* U, int64_t batch, const int* cols, int layer,
                int out_kind, void* out, double scale, unsigned flags, const WsLayout& w, void* ws, hipStream_t st) {
  if (flags & SLOS_FLAG_F64) return forward_run_f64(p, dt, U, batch, cols, layer, out_kind, out, scale, flags, w, ws, st);
  int rc = SLOS_OK;
  auto* base = static_cast<unsigned char*>(ws);
  float2* ones = reinterpret_cast<float2*>(base + w.ones);
  float2* bufs[2] = {reinterpret_cast<float2*>(base + w.even), reinterpret_cast<float2*>(base + w.odd)};
  hipLaunchKernelGGL(k_fill_ones, dim3((unsigned)std::min<int64_t>((batch + 255) / 256, 4096)), dim3(256), 0, st,
                     ones, batch);
  HIP_TRY(hipGetLastError());
  if (layer == 0) {  // no photon: the vacuum
    hipLaunchKernelGGL(k_vacuum_out, dim3((unsigned)std::min<int64_t>((batch + 255) / 256, 4096)), dim3(256), 0, st, out,
                       batch, out_kind == SLOS_OUT_PROBS ? 1 : 0, (float)scale);
    HIP_TRY(hipGetLastError());
    return SLOS_OK;
  }
  const float2* Uf = static_cast<const float2*>(U);
  const float2* par = ones;
  uint64_t S_par = 1;
  // compensated sums: every intermediate layer is a hi/lo pair [2][B][S] (the buffers are sized for it)
  const bool hilo = (flags & SLOS_FLAG_COMPENSATED) != 0;
  const RoctxRange whole("slos_forward");
  for (int k = 1; k <= layer; ++k) {
    const uint64_t S = p->size(k);
    const bool last = (k == layer);
    void* dst = last ? out : bufs[k & 1];
    const int kind = last ? (out_kind == SLOS_OUT_PROBS ? 1 : 2) : 0;
    char name[32];
    std::snprintf(name, sizeof name, "slos layer %d", k);
    const RoctxRange range(name);
    rc = launch_layer(p, dt, k, cols[k - 1], Uf, batch, par, S_par, dst, S, kind, (float)scale, flags & ~SLOS_FLAG_HILO,
                      st, hilo && k >= 2, hilo && !last);
    if (rc) return rc;
    par = static_cast<const float2*>(dst);
    S_par = S;
  }
  return SLOS_OK;
}

}  // namespace

// Scratch of slos_forward_noalloc: the vacuum row and the two ping-pong layers, one copy per GPU
// (held until the process ends). Host code never passes the symbol itself to a kernel: its device
// address comes from hipGetSymbolAddress, once per device.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct ScratchSlot {
  int device;
  void* ptr;
};
std::mutex g_scratch_mu;
std::vector<ScratchSlot> g_scratch;

// device address of the current GPU's scratch array (looked up on the first call on each GPU)
int scratch_address(void** out) {
  int d = 0;
  HIP_TRY(hipGetDevice(&d));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  for (const auto& s : g_scratch)
    if (s.device == d) {
      *out = s.ptr;
      return SLOS_OK;
    }
  void* ptr = nullptr;
  HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(SLOS_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
  g_scratch.push_back(ScratchSlot{d, ptr});
  *out = ptr;
  return SLOS_OK;
}

// the plan's tables on the current GPU if they are already resident (no allocation)
int resident_tables(const slos_plan* cp, DevTables* out) {
  slos_plan* p = const_cast<slos_plan*>(cp);
  int d = 0;
  HIP_TRY(hipGetDevice(&d));
  std::lock_guard<std::mutex> g(p->mu);
  for (auto& t : p->dev)
    if (t.device == d) {
      *out = t;
      return SLOS_OK;
    }
  return fail(SLOS_EINVAL, "the plan has no tables on this GPU yet: run slos_forward (or slos_layer_step) with "
                           "this plan once on it first");
}

}  // namespace

int slos_forward(const slos_plan* p, const void* U, int64_t batch, const int32_t* modes, int layer, int out_kind,
                 void* out, double scale, unsigned flags, void* ws, size_t ws_bytes, slos_stream_t stream) {
  int cols[kMaxPhotons] = {};  // layer <= n <= 32 (checked before cols is written)
  WsLayout w;
  int rc = forward_check(p, U, batch, modes, layer, out_kind, out, flags, cols, &w);
  if (rc) return rc;
  if (!ws || ws_bytes < w.total)
    return fail(SLOS_ENOMEM, "workspace too small: need " + std::to_string(w.total) + " bytes");
  DevTables dt;
  rc = device_tables(p, &dt);
  if (rc) return rc;
  return forward_run(p, dt, U, batch, cols, layer, out_kind, out, scale, flags, w, ws,
                     reinterpret_cast<hipStream_t>(stream));
}

int slos_forward_noalloc(const slos_plan* p, const void* U, int64_t batch, const int32_t* modes, int layer,
                         int out_kind, void* out, double scale, unsigned flags, slos_stream_t stream) {
  int cols[kMaxPhotons] = {};  // layer <= n <= 32 (checked before cols is written)
  WsLayout w;
  int rc = forward_check(p, U, batch, modes, layer, out_kind, out, flags, cols, &w);
  if (rc) return rc;
  if (w.total > kNoAllocScratchBytes)
    return fail(SLOS_ERANGE, "slos_forward_noalloc: the intermediate layers need " + std::to_string(w.total) +
                                 " bytes, above its fixed scratch of " + std::to_string(kNoAllocScratchBytes) +
                                 " bytes; use slos_forward with a workspace");
  DevTables dt;
  rc = resident_tables(p, &dt);
  if (rc) return rc;
  void* ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  return forward_run(p, dt, U, batch, cols, layer, out_kind, out, scale, flags, w, ws,
                     reinterpret_cast<hipStream_t>(stream));
}

int slos_layer_step(const slos_plan* p, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                    int out_kind, void* out, double scale, unsigned flags, slos_stream_t stream) {
  if (!p || !U || !psi_prev || !out) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (col < 0 || col >= p->m) return fail(SLOS_EINVAL, "input mode out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  if (out_kind < 0 || out_kind > 2) return fail(SLOS_EINVAL, "unknown output kind");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const int kind = out_kind == 2 ? 0 : (out_kind == SLOS_OUT_PROBS ? 1 : 2);
  if (flags & SLOS_FLAG_F64) {
    if (flags & (SLOS_FLAG_COMPENSATED | SLOS_FLAG_HILO))
      return fail(SLOS_EINVAL, "SLOS_FLAG_F64 computes in FP64: no compensated sums");
    return launch_layer_f64(p, dt, k, col, U, batch, psi_prev, p->size(k - 1), out, p->size(k), kind, scale, flags,
                            reinterpret_cast<hipStream_t>(stream));
  }
  const bool hilo = (flags & SLOS_FLAG_HILO) && (flags & SLOS_FLAG_COMPENSATED);
  if ((flags & SLOS_FLAG_HILO) && !(flags & SLOS_FLAG_COMPENSATED))
    return fail(SLOS_EINVAL, "SLOS_FLAG_HILO needs SLOS_FLAG_COMPENSATED");
  return launch_layer(p, dt, k, col, static_cast<const float2*>(U), batch, static_cast<const float2*>(psi_prev),
                      p->size(k - 1), out, p->size(k), kind, (float)scale, flags & ~SLOS_FLAG_HILO,
                      reinterpret_cast<hipStream_t>(stream), hilo, hilo && kind == 0);
}

int slos_readout_adjoint(const slos_plan* p, int64_t batch, const void* amps, const float* grad_p, double scale,
                         void* grad_amps, slos_stream_t stream) {
  if (!p || !amps || !grad_p || !grad_amps) return fail(SLOS_EINVAL, "NULL argument");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const uint64_t S = p->size(p->n);
  const size_t sh = (size_t)(p->n + 1) * (p->me + 1) * 8 + (size_t)p->me * 4;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((S + 255) / 256, 1u << 20)),
            (unsigned)std::min<int64_t>(batch, 16));
  hipLaunchKernelGGL(k_readout_adjoint, grid, dim3(256), sh, reinterpret_cast<hipStream_t>(stream), dt.cg, dt.lo,
                     p->me, p->multiset, p->n, static_cast<const float2*>(amps), grad_p,
                     static_cast<float2*>(grad_amps), S, batch, (float)(scale * scale));
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_readout_adjoint_pull(const slos_plan* p, const void* U, int64_t batch, const void* psi_prev,
                              const float* grad_p, double scale, void* grad_amps, slos_stream_t stream) {
  if (!p || !U || !psi_prev || !grad_p || !grad_amps) return fail(SLOS_EINVAL, "NULL argument");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  if (p->n < 1) return fail(SLOS_EINVAL, "no photon layer");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const int K = p->n;
  const uint64_t S = p->size(K);
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((S + 255) / 256, 1u << 20)),
            (unsigned)std::min<int64_t>(batch, 16));
  with_km(K, [&](auto km) {
    hipLaunchKernelGGL((k_readout_adjoint_pull<decltype(km)::value>), grid, dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), dt.cg, dt.lo, p->me, p->multiset, K, p->m,
                       p->modes[K - 1], static_cast<const float2*>(U), static_cast<const float2*>(psi_prev),
                       p->size(K - 1), grad_p, static_cast<float2*>(grad_amps), S, batch, (float)(scale * scale));
  });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

// work groups of the thread-per-state adjoint (grid-stride over the parents; one FP64 partial per
// work group and mode)
static int adjoint_blocks(uint64_t Sp) { return (int)std::max<uint64_t>(1, std::min<uint64_t>((Sp + 63) / 64, 1u << 16)); }

namespace {

// parents per XCD chunk of the block adjoint (see slos::xcd_chunk_index)
constexpr uint64_t kDescChunkParents = 1ull << 18;

// How one adjoint layer runs: the block kernel (multiset spaces whose parent layer has a 1..8
// position prefix before its 4-position suffix, up to 32 modes) or the thread-per-state kernel;
// `nblk` work groups, each writing one FP64 partial per mode.
struct AdjLaunch {
  bool block = false;
  int maxd = 0;
  uint64_t per_wave = 0;
  uint32_t xcd_chunk = 1;
  size_t lds = 0;
  uint64_t nblk = 0;
};

int adjoint_launch(const slos_plan* p, int k, AdjLaunch* L) {
  const uint64_t Sp = p->size(k - 1);
  const int dp = k - 1 - slos::kSuf;
  if (p->multiset && !p->tuples.empty() && dp >= 1 && dp <= 8 && p->me <= 32) {
    const int maxd = dp <= 4 ? 4 : dp <= 6 ? 6 : 8;
    const size_t shb = slos::desc_dyn_lds(k, p->me, maxd);
    if (shb + sizeof(double) * 2 * slos::kWaves * 32 <= lds_limit()) {
      // parents per wave: at most about 2^16 work groups (FP64 partials per mode), at least 256
      // (at n = 8, m = 24 the top layer took 0.46 ms with at least 2048 a wave, 0.25 ms so)
      uint64_t pw = (Sp / ((uint64_t)slos::kWaves << 16) + 63) & ~63ull;
      pw = std::max<uint64_t>(pw, 256);
      const uint64_t wgs = ((Sp + pw - 1) / pw + slos::kWaves - 1) / slos::kWaves;
      if (wgs > 0x7fffffffull) return fail(SLOS_ERANGE, "layer too large for one launch");
      L->block = true;
      L->maxd = maxd;
      L->per_wave = pw;
      L->xcd_chunk = (uint32_t)std::max<uint64_t>(1, kDescChunkParents / ((uint64_t)slos::kWaves * pw));
      L->lds = shb;
      L->nblk = wgs;
      return SLOS_OK;
    }
  }
  const size_t sh = (((size_t)(k + 1) * (p->me + 1) * 8 + p->me * 4 + 15) & ~(size_t)15) + (size_t)p->m * 8 +
                    (size_t)kAdjThreads * p->me * 8;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "adjoint LDS exceeds the GPU's LDS (too many modes)");
  L->block = false;
  L->lds = sh;
  L->nblk = adjoint_blocks(Sp);
  return SLOS_OK;
}

}  // namespace

int slos_adjoint_workspace(const slos_plan* p, int64_t batch, int k, size_t* bytes) {
  if (!p || !bytes) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n || batch < 1) return fail(SLOS_EINVAL, "bad argument");
  AdjLaunch L;
  int rc = adjoint_launch(p, k, &L);
  if (rc) return rc;
  *bytes = (size_t)batch * L.nblk * p->me * sizeof(double2);
  return SLOS_OK;
}

int slos_adjoint_layer(const slos_plan* p, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                       const void* grad_k, void* grad_prev, void* grad_U, int accumulate, void* ws, size_t ws_bytes,
                       slos_stream_t stream) {
  if (!p || !U || !psi_prev || !grad_k || !grad_U) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (col < 0 || col >= p->m) return fail(SLOS_EINVAL, "input mode out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  AdjLaunch L;
  int rc = adjoint_launch(p, k, &L);
  if (rc) return rc;
  if (!ws || ws_bytes < (size_t)batch * L.nblk * p->me * sizeof(double2)) return fail(SLOS_ENOMEM, "workspace too small");
  DevTables dt;
  rc = device_tables(p, &dt);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  const auto* Uf = static_cast<const float2*>(U);
  const auto* psi = static_cast<const float2*>(psi_prev);
  const auto* gk = static_cast<const float2*>(grad_k);
  auto* gout = static_cast<float2*>(grad_prev);
  auto* part = static_cast<double2*>(ws);
  if (L.block && dt.tuples) {
    const dim3 grid((unsigned)L.nblk, (unsigned)batch);
    const uint64_t S_pre = p->size(k - 1 - slos::kSuf);
#define SLOS_ADJ(D, E)                                                                                                 \
  hipLaunchKernelGGL((slos::k_adjoint_block<D, E>), grid, dim3(slos::kBlockThreads), L.lds, st, dt.cg, dt.lo, dt.tuples, \
                     p->me, p->me + 1, k, p->m, col, Uf, psi, gk, Sk, gout, Sp, S_pre, L.per_wave, L.xcd_chunk, part)
#define SLOS_ADJ_E(D)                   \
  if (p->me <= 16) SLOS_ADJ(D, 16);     \
  else SLOS_ADJ(D, 32)
    if (L.maxd == 4) { SLOS_ADJ_E(4); }
    else if (L.maxd == 6) { SLOS_ADJ_E(6); }
    else { SLOS_ADJ_E(8); }
#undef SLOS_ADJ_E
#undef SLOS_ADJ
  } else {
    with_km(k - 1, [&](auto km) {
      hipLaunchKernelGGL((k_adjoint_layer<decltype(km)::value>), dim3((unsigned)L.nblk, (unsigned)batch),
                         dim3(kAdjThreads), L.lds, st, dt.cg, dt.lo, p->me, p->multiset, k, p->m, col, Uf, psi, gk, Sk,
                         gout, Sp, part, grad_prev != nullptr);
    });
  }
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_gu, dim3((unsigned)p->me, (unsigned)batch), dim3(64), 0, st,
                     static_cast<const double2*>(ws), (int)L.nblk, p->me, p->m, col, static_cast<float2*>(grad_U), accumulate);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_loss_descend(const slos_plan* p, int k, const float* q_k, float* q_km1, int64_t batch, unsigned flags,
                      double scale, slos_stream_t stream) {
  if (!p || !q_k || !q_km1) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  const int dp = k - 1 - slos::kSuf;
  char name[32];
  std::snprintf(name, sizeof name, "slos descent %d", k);
  const RoctxRange range(name);
  if (!(flags & SLOS_FLAG_SIMPLE) && p->multiset && dt.tuples && dp >= 1 && dp <= 8) {
    const int maxd = dp <= 4 ? 4 : dp <= 6 ? 6 : 8;
    const size_t shb = slos::desc_dyn_lds(k, p->me, maxd);
    if (shb <= lds_limit()) {
      // parents per wave: 1024 on layers of 2^27 parents or more, 512 below, fewer on small layers;
      // XCD chunks of 2^16 parents (measured at n = 11, m = 32: the k = 11, 10, 9 steps 34.4, 9.0,
      // 2.5 ms with 4096 per wave and chunks of 2^18, 32.6, 8.1, 2.0 ms so)
      uint64_t pw = (Sp / 16384 + 63) & ~63ull;
      pw = std::min<uint64_t>(std::max<uint64_t>(pw, 256), Sp >= (1ull << 27) ? 1024 : 512);
      const uint64_t wgs = ((Sp + pw - 1) / pw + slos::kWaves - 1) / slos::kWaves;
      if (wgs > 0x7fffffffull) return fail(SLOS_ERANGE, "layer too large for one launch");
      const dim3 grid((unsigned)wgs, (unsigned)batch);
      const uint32_t xc = (uint32_t)std::max<uint64_t>(1, (1ull << 16) / ((uint64_t)slos::kWaves * pw));
      const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
      const uint64_t S_pre = p->size(dp);
#define SLOS_DESC(D, E)                                                                                                  \
  hipLaunchKernelGGL((slos::k_descend_block<D, E>), grid, dim3(slos::kBlockThreads), shb, st, dt.cg, dt.lo, dt.tuples, \
                     p->me, p->me + 1, k, q_k, Sk, q_km1, Sp, S_pre, pw, (float)scale, xc)
#define SLOS_DESC_E(D)          \
  if (p->me <= 16) SLOS_DESC(D, 16); \
  else if (p->me <= 32) SLOS_DESC(D, 32); \
  else SLOS_DESC(D, 64)
      if (maxd == 4) { SLOS_DESC_E(4); }
      else if (maxd == 6) { SLOS_DESC_E(6); }
      else { SLOS_DESC_E(8); }
#undef SLOS_DESC_E
#undef SLOS_DESC
      HIP_TRY(hipGetLastError());
      return SLOS_OK;
    }
  }
  const size_t sh = (size_t)(k + 1) * (p->me + 1) * 8 + (size_t)p->me * 4;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((Sp + 255) / 256, 1u << 20)), (unsigned)batch);
  with_km(k - 1, [&](auto km) {
    hipLaunchKernelGGL((k_descend<decltype(km)::value>), grid, dim3(256), sh, reinterpret_cast<hipStream_t>(stream),
                       dt.cg, dt.lo, p->me, p->multiset, k, q_k, Sk, q_km1, Sp, batch, (float)scale);
  });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_raise(const slos_plan* p, int k, const float* w, const float* q_km1, float* q_k, int64_t batch,
               slos_stream_t stream) {
  if (!p || !w || !q_km1 || !q_k) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  const size_t sh = (size_t)(k + 1) * (p->me + 1) * 8 + (size_t)p->me * 4 + (size_t)p->m * 4;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((Sk + 255) / 256, 1u << 20)), (unsigned)batch);
  with_km(k, [&](auto km) {
    hipLaunchKernelGGL((k_raise<decltype(km)::value>), grid, dim3(256), sh, reinterpret_cast<hipStream_t>(stream),
                       dt.cg, dt.lo, p->me, p->multiset, k, p->m, w, q_km1, Sp, q_k, Sk, batch);
  });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_loss_ascend(const slos_plan* p, int k, const float* g_km1, float* g_k, int64_t batch, slos_stream_t stream) {
  if (!p || !g_km1 || !g_k) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  const size_t sh = (size_t)(k + 1) * (p->me + 1) * 8 + (size_t)p->me * 4;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((Sk + 255) / 256, 1u << 20)), (unsigned)batch);
  with_km(k, [&](auto km) {
    hipLaunchKernelGGL((k_ascend<decltype(km)::value>), grid, dim3(256), sh, reinterpret_cast<hipStream_t>(stream),
                       dt.cg, dt.lo, p->me, p->multiset, k, g_km1, Sp, g_k, Sk, batch);
  });
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_mass_workspace(int64_t batch, uint64_t count, size_t* bytes) {
  (void)count;
  if (!bytes || batch < 1) return fail(SLOS_EINVAL, "bad argument");
  *bytes = (size_t)batch * kMassBlocks * sizeof(double);
  return SLOS_OK;
}

int slos_total_mass(const float* probs, int64_t batch, uint64_t count, double* out, void* ws, size_t ws_bytes,
                    slos_stream_t stream) {
  if (!probs || !out || batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "bad argument");
  if (!ws || ws_bytes < (size_t)batch * kMassBlocks * sizeof(double)) return fail(SLOS_ENOMEM, "workspace too small");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_mass_partial, dim3(kMassBlocks, (unsigned)batch), dim3(256), 0, st, probs, count, part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_mass_final, dim3((unsigned)batch), dim3(256), 0, st, part, kMassBlocks, out);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}


int slos_group_sums(const float* probs, int64_t batch, uint64_t count, const int64_t* order, const int64_t* chunks,
                    uint32_t n_chunks, const uint32_t* first, int n_groups, float* out, void* ws, size_t ws_bytes,
                    slos_stream_t stream) {
  return group_sums(probs, batch, count, order, chunks, n_chunks, first, n_groups, out, ws, ws_bytes, stream);
}

int slos_group_sums_f64(const double* probs, int64_t batch, uint64_t count, const int64_t* order,
                        const int64_t* chunks, uint32_t n_chunks, const uint32_t* first, int n_groups, double* out,
                        void* ws, size_t ws_bytes, slos_stream_t stream) {
  return group_sums(probs, batch, count, order, chunks, n_chunks, first, n_groups, out, ws, ws_bytes, stream);
}

int slos_spot_probs(const slos_plan* p, const void* U, const uint64_t* idx, int64_t count, double* out,
                    slos_stream_t stream) {
  if (!p || !U || (!idx && count) || (!out && count)) return fail(SLOS_EINVAL, "NULL argument");
  if (p->n > 24) return fail(SLOS_ERANGE, "spot permanents support n <= 24");
  if (count == 0) return SLOS_OK;
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_spot_ryser, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, st, dt.cg, dt.lo, p->me, p->n,
                     p->m, static_cast<const double2*>(U), dt.modes, p->s_fact, p->multiset, idx, count, out);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_sample_workspace(uint64_t count, size_t* bytes) {
  if (!bytes) return fail(SLOS_EINVAL, "NULL argument");
  if (count < 1) return fail(SLOS_EINVAL, "count must be >= 1");
  const uint64_t nb = (count + kSampleBlock - 1) / kSampleBlock;
  if (nb > (uint64_t)kSampleSeg * 1024) return fail(SLOS_ERANGE, "too many categories for the sampler");
  // block sums, the cdf, the chunk sums and the segment totals
  *bytes = 2 * align256(nb * sizeof(double)) + align256(nb * 64 * sizeof(double)) + align256(1024 * sizeof(double));
  return SLOS_OK;
}

int slos_sample(const float* probs, uint64_t count, const double* uniforms, int64_t shots, uint64_t* idx, void* ws,
                size_t ws_bytes, slos_stream_t stream) {
  if (shots < 0) return fail(SLOS_EINVAL, "shots must be >= 0");
  if (!probs || (shots > 0 && (!uniforms || !idx))) return fail(SLOS_EINVAL, "NULL argument");
  size_t need = 0;
  int rc = slos_sample_workspace(count, &need);
  if (rc) return rc;
  if (!ws || ws_bytes < need) return fail(SLOS_ENOMEM, "workspace too small: need " + std::to_string(need) + " bytes");
  if (shots == 0) return SLOS_OK;
  const uint64_t nb = (count + kSampleBlock - 1) / kSampleBlock;
  double* bsum = static_cast<double*>(ws);
  double* cdf = reinterpret_cast<double*>(static_cast<unsigned char*>(ws) + align256(nb * sizeof(double)));
  double* csum = reinterpret_cast<double*>(static_cast<unsigned char*>(ws) + 2 * align256(nb * sizeof(double)));
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_sample_block_sums, dim3((unsigned)((nb + kSumWaves - 1) / kSumWaves)), dim3(64 * kSumWaves), 0, st,
                     probs, count, nb, bsum, csum);
  HIP_TRY(hipGetLastError());
  double* segtot = reinterpret_cast<double*>(static_cast<unsigned char*>(ws) + 2 * align256(nb * sizeof(double)) +
                                             align256(nb * 64 * sizeof(double)));
  const uint64_t nseg = (nb + kSampleSeg - 1) / kSampleSeg;  // <= 1024 (slos_sample_workspace)
  hipLaunchKernelGGL(k_sample_seg_prefix, dim3((unsigned)nseg), dim3(256), 0, st, bsum, nb, cdf, segtot);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_sample_seg_offsets, dim3(1), dim3(64), 0, st, segtot, (int)nseg);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_sample_add_offsets, dim3((unsigned)std::min<uint64_t>((nb + 255) / 256, 4096)), dim3(256), 0, st,
                     cdf, nb, segtot);
  HIP_TRY(hipGetLastError());
  const uint64_t g = ((uint64_t)shots + 255) / 256;
  if (g > 0x7fffffffull) return fail(SLOS_ERANGE, "too many shots for one launch");
  hipLaunchKernelGGL(k_sample_search, dim3((unsigned)g), dim3(256), 0, st, probs, count, nb, cdf, csum, uniforms, shots,
                     idx);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_unrank_device(const slos_plan* p, int k, const uint64_t* idx, int64_t count, int32_t* occ,
                       slos_stream_t stream) {
  if (!p || !idx || !occ) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 0 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (count < 0) return fail(SLOS_EINVAL, "count must be >= 0");
  if (count == 0) return SLOS_OK;
  DevTables dt;
  int rc = device_tables(p, &dt);
  if (rc) return rc;
  const size_t sh = (size_t)(k + 1) * (p->me + 1) * 8 + (size_t)p->me * 4;
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  const uint64_t blocks = std::min<uint64_t>(((uint64_t)count + 255) / 256, 1u << 16);
  hipLaunchKernelGGL(k_unrank_occ, dim3((unsigned)blocks), dim3(256), sh, reinterpret_cast<hipStream_t>(stream), dt.cg,
                     dt.lo, p->me, k, p->m, idx, count, occ);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_group_create(int count, const slos_plan* const* plans, slos_group** out) {
  if (!out || !plans) return fail(SLOS_EINVAL, "NULL argument");
  *out = nullptr;
  if (count < 1 || count > slos::kMaxGroup)
    return fail(SLOS_EINVAL, "a group holds 1 to " + std::to_string(slos::kMaxGroup) + " plans");
  auto* G = new slos_group();
  for (int g = 0; g < count; ++g) {
    const slos_plan* p = plans[g];
    if (!p) {
      delete G;
      return fail(SLOS_EINVAL, "NULL plan");
    }
    if (simple_lds_bytes(p->n, p->me, p->m) > lds_limit()) {
      delete G;
      return fail(SLOS_ERANGE, "plan too large for a group launch (LDS)");
    }
    G->plans.push_back(p);
    G->nmax = std::max(G->nmax, p->n);
  }
  G->cfg.assign(G->nmax + 1, slos_group::Cfg());
  *out = G;
  return SLOS_OK;
}

int slos_group_destroy(slos_group* G) {
  if (!G) return SLOS_OK;
  group_drop_graph(G);
  if (G->cap) (void)hipStreamDestroy(G->cap);
  delete G;
  return SLOS_OK;
}

int slos_group_workspace(const slos_group* G, const int64_t* batch, unsigned flags, size_t* bytes) {
  if (!G || !batch || !bytes) return fail(SLOS_EINVAL, "NULL argument");
  GroupWs w;
  int rc = group_layout(G, batch, &w, (flags & SLOS_FLAG_COMPENSATED) != 0);
  if (rc) return rc;
  *bytes = w.total;
  return SLOS_OK;
}

int slos_group_forward(slos_group* G, const void* const* U, const int64_t* batch, float* const* out,
                       const double* scale, unsigned flags, void* ws, size_t ws_bytes, slos_stream_t stream) {
  GroupArgs a;
  GroupWs w;
  int rc = group_args(G, U, batch, out, scale, flags, ws, ws_bytes, &a, &w);
  if (rc) return rc;
  std::vector<DevTables> dts;
  rc = group_tables(G, &dts);
  if (rc) return rc;
  return group_forward(G, a, w, dts, reinterpret_cast<hipStream_t>(stream));
}

int slos_group_tune(slos_group* G, const void* const* U, const int64_t* batch, float* const* out,
                    const double* scale, unsigned flags, void* ws, size_t ws_bytes, int reps, slos_stream_t stream) {
  GroupArgs a;
  GroupWs w;
  int rc = group_args(G, U, batch, out, scale, flags, ws, ws_bytes, &a, &w);
  if (rc) return rc;
  if (reps < 1) return fail(SLOS_EINVAL, "reps must be >= 1");
  std::vector<DevTables> dts;
  rc = group_tables(G, &dts);
  if (rc) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  int dev = 0, ncu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  group_drop_graph(G);
  std::fill(G->cfg.begin(), G->cfg.end(), slos_group::Cfg());
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // layers in order: each candidate rewrites layer K from the (already final) layer K - 1
  rc = group_forward(G, a, w, dts, st);
  for (int K = 1; K <= G->nmax && rc == SLOS_OK; ++K) {
    std::vector<slos_group::Cfg> cand;
    const uint32_t grids[4] = {0, (uint32_t)ncu * 2, (uint32_t)ncu * 4, (uint32_t)ncu * 8};
    if (group_can_block(G, dts, K, flags))
      for (uint64_t pw : {256, 512, 1024, 2048, 4096})
        for (uint32_t gr : grids) cand.push_back({1, pw, gr});
    for (uint64_t spt : {1, 2, 4})
      for (uint32_t gr : grids) cand.push_back({0, spt, gr});
    cand.push_back({2, 0, 0});  // one single-plan launch per plan
    float best = 1e30f;
    slos_group::Cfg pick;
    for (const auto& c : cand) {
      rc = group_layer(G, K, a, w, dts, c, st);  // warm-up
      if (rc) break;
      if ((rc = (hipEventRecord(e0, st) == hipSuccess ? SLOS_OK : fail(SLOS_EHIP, "hipEventRecord")))) break;
      for (int r = 0; r < reps && rc == SLOS_OK; ++r) rc = group_layer(G, K, a, w, dts, c, st);
      if (rc) break;
      float ms = 0.f;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(SLOS_EHIP, "event timing failed");
        break;
      }
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    G->cfg[K] = pick;
    // leave layer K final under the chosen launch (all candidates compute the same values)
    if (rc == SLOS_OK) rc = group_layer(G, K, a, w, dts, pick, st);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int slos_group_config(const slos_group* G, int K, int* block, uint64_t* per_wave, uint32_t* grid) {
  if (!G || !block || !per_wave || !grid) return fail(SLOS_EINVAL, "NULL argument");
  if (K < 1 || K > G->nmax) return fail(SLOS_EINVAL, "layer out of range");
  *block = G->cfg[K].block;
  *per_wave = G->cfg[K].per_wave;
  *grid = G->cfg[K].grid;
  return SLOS_OK;
}

int slos_group_capture(slos_group* G, const void* const* U, const int64_t* batch, float* const* out,
                       const double* scale, unsigned flags, void* ws, size_t ws_bytes) {
  GroupArgs a;
  GroupWs w;
  int rc = group_args(G, U, batch, out, scale, flags, ws, ws_bytes, &a, &w);
  if (rc) return rc;
  std::vector<DevTables> dts;
  rc = group_tables(G, &dts);  // device tables exist before capture: nothing is allocated inside it
  if (rc) return rc;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (G->cap && G->cap_device != dev) {
    (void)hipStreamDestroy(G->cap);
    G->cap = nullptr;
  }
  if (!G->cap) {
    HIP_TRY(hipStreamCreateWithFlags(&G->cap, hipStreamNonBlocking));
    G->cap_device = dev;
  }
  group_drop_graph(G);
  HIP_TRY(hipStreamBeginCapture(G->cap, hipStreamCaptureModeThreadLocal));
  rc = group_forward(G, a, w, dts, G->cap);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(G->cap, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ec != hipSuccess) return fail(SLOS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  G->graph = graph;
  HIP_TRY(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  return SLOS_OK;
}

int slos_group_replay(slos_group* G, slos_stream_t stream) {
  if (!G) return fail(SLOS_EINVAL, "NULL argument");
  if (!G->exec) return fail(SLOS_EINVAL, "no captured forward: call slos_group_capture first");
  HIP_TRY(hipGraphLaunch(G->exec, reinterpret_cast<hipStream_t>(stream)));
  return SLOS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// complex128 / float64 twins of the entry points that take typed pointers or no flags
// (FockSLOSKernel(dtype=torch.complex128)); thread-per-state FP64 kernels of layer_f64.hpp
// ------------------------------------------------------------------------------------------

namespace {

// plan tables on the device, the batch range and the layer range shared by the twins below
int f64_common(const slos_plan* p, int k, int64_t batch, DevTables* dt) {
  if (k < 1 || k > p->n) return fail(SLOS_EINVAL, "layer out of range");
  if (batch < 1 || batch > 65535) return fail(SLOS_EINVAL, "batch must be in [1, 65535]");
  return device_tables(p, dt);
}

// work groups and dynamic LDS of k_adjoint_layer_f64 for layer k
int adjoint_launch_f64(const slos_plan* p, int k, size_t* lds, uint64_t* nblk) {
  *lds = slos::space_lds_f64(k, p->me) + (size_t)p->m * 16 + (size_t)slos::kAdjThreadsF64 * p->me * 16;
  if (*lds > lds_limit()) return fail(SLOS_ERANGE, "adjoint LDS exceeds the GPU's LDS (too many modes)");
  *nblk = (uint64_t)adjoint_blocks(p->size(k - 1));
  return SLOS_OK;
}

}  // namespace

int slos_loss_descend_f64(const slos_plan* p, int k, const double* q_k, double* q_km1, int64_t batch, double scale,
                          slos_stream_t stream) {
  if (!p || !q_k || !q_km1) return fail(SLOS_EINVAL, "NULL argument");
  DevTables dt;
  int rc = f64_common(p, k, batch, &dt);
  if (rc) return rc;
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  const size_t sh = slos::space_lds_f64(k, p->me);
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  char name[32];
  std::snprintf(name, sizeof name, "slos descent %d", k);
  const RoctxRange range(name);
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((Sp + 255) / 256, 1u << 20)), (unsigned)batch);
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define SLOS_F64D(KM) \
  hipLaunchKernelGGL((slos::k_descend_f64<KM>), grid, dim3(256), sh, st, dt.cg, dt.lo, p->me, p->multiset, k, q_k, Sk, \
                     q_km1, Sp, batch, scale)
  if (k - 1 <= 8) { SLOS_F64D(8); }
  else if (k - 1 <= 16) { SLOS_F64D(16); }
  else { SLOS_F64D(32); }
#undef SLOS_F64D
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_loss_ascend_f64(const slos_plan* p, int k, const double* g_km1, double* g_k, int64_t batch,
                         slos_stream_t stream) {
  if (!p || !g_km1 || !g_k) return fail(SLOS_EINVAL, "NULL argument");
  DevTables dt;
  int rc = f64_common(p, k, batch, &dt);
  if (rc) return rc;
  const uint64_t Sp = p->size(k - 1), Sk = p->size(k);
  const size_t sh = slos::space_lds_f64(k, p->me);
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((Sk + 255) / 256, 1u << 20)), (unsigned)batch);
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define SLOS_F64A(KM)                                                                                                \
  hipLaunchKernelGGL((slos::k_ascend_f64<KM>), grid, dim3(256), sh, st, dt.cg, dt.lo, p->me, p->multiset, k, g_km1, Sp, \
                     g_k, Sk, batch)
  if (k <= 8) { SLOS_F64A(8); }
  else if (k <= 16) { SLOS_F64A(16); }
  else { SLOS_F64A(32); }
#undef SLOS_F64A
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_readout_adjoint_f64(const slos_plan* p, int64_t batch, const void* amps, const double* grad_p, double scale,
                             void* grad_amps, slos_stream_t stream) {
  if (!p || !amps || !grad_p || !grad_amps) return fail(SLOS_EINVAL, "NULL argument");
  if (p->n < 1) return fail(SLOS_EINVAL, "no photon: nothing depends on U");
  DevTables dt;
  int rc = f64_common(p, p->n, batch, &dt);
  if (rc) return rc;
  const uint64_t S = p->size(p->n);
  const size_t sh = slos::space_lds_f64(p->n, p->me);
  if (sh > lds_limit()) return fail(SLOS_ERANGE, "counting table does not fit LDS");
  dim3 grid((unsigned)std::max<uint64_t>(1, std::min<uint64_t>((S + 255) / 256, 1u << 20)),
            (unsigned)std::min<int64_t>(batch, 16));
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define SLOS_F64R(KM)                                                                                                 \
  hipLaunchKernelGGL((slos::k_readout_adjoint_f64<KM>), grid, dim3(256), sh, st, dt.cg, dt.lo, p->me, p->multiset, p->n, \
                     static_cast<const double2*>(amps), grad_p, static_cast<double2*>(grad_amps), S, batch, scale * scale)
  if (p->n <= 8) { SLOS_F64R(8); }
  else if (p->n <= 16) { SLOS_F64R(16); }
  else { SLOS_F64R(32); }
#undef SLOS_F64R
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}

int slos_adjoint_workspace_f64(const slos_plan* p, int64_t batch, int k, size_t* bytes) {
  if (!p || !bytes) return fail(SLOS_EINVAL, "NULL argument");
  if (k < 1 || k > p->n || batch < 1) return fail(SLOS_EINVAL, "bad argument");
  size_t lds = 0;
  uint64_t nblk = 0;
  int rc = adjoint_launch_f64(p, k, &lds, &nblk);
  if (rc) return rc;
  *bytes = (size_t)batch * nblk * p->me * sizeof(double2);
  return SLOS_OK;
}

int slos_adjoint_layer_f64(const slos_plan* p, const void* U, int64_t batch, int k, int col, const void* psi_prev,
                           const void* grad_k, void* grad_prev, void* grad_U, int accumulate, void* ws,
                           size_t ws_bytes, slos_stream_t stream) {
  if (!p || !U || !psi_prev || !grad_k || !grad_U) return fail(SLOS_EINVAL, "NULL argument");
  if (col < 0 || col >= p->m) return fail(SLOS_EINVAL, "input mode out of range");
  DevTables dt;
  int rc = f64_common(p, k, batch, &dt);
  if (rc) return rc;
  size_t lds = 0;
  uint64_t nblk = 0;
  rc = adjoint_launch_f64(p, k, &lds, &nblk);
  if (rc) return rc;
  if (!ws || ws_bytes < (size_t)batch * nblk * p->me * sizeof(double2)) return fail(SLOS_ENOMEM, "workspace too small");
  const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
#define SLOS_F64J(KM)                                                                                              \
  hipLaunchKernelGGL((slos::k_adjoint_layer_f64<KM>), dim3((unsigned)nblk, (unsigned)batch), dim3(slos::kAdjThreadsF64), \
                     lds, st, dt.cg, dt.lo, p->me, p->multiset, k, p->m, col, static_cast<const double2*>(U),         \
                     static_cast<const double2*>(psi_prev), static_cast<const double2*>(grad_k), p->size(k),          \
                     static_cast<double2*>(grad_prev), p->size(k - 1), static_cast<double2*>(ws), grad_prev != nullptr)
  if (k - 1 <= 8) { SLOS_F64J(8); }
  else if (k - 1 <= 16) { SLOS_F64J(16); }
  else { SLOS_F64J(32); }
#undef SLOS_F64J
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(slos::k_reduce_gu_f64, dim3((unsigned)p->me, (unsigned)batch), dim3(64), 0, st,
                     static_cast<const double2*>(ws), (int)nblk, p->me, p->m, col, static_cast<double2*>(grad_U),
                     accumulate);
  HIP_TRY(hipGetLastError());
  return SLOS_OK;
}
