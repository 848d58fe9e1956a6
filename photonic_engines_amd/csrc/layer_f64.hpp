// complex128 / float64 forms of the thread-per-state kernels: FockSLOSKernel(dtype=torch.complex128)
// (photonic_engines/slos/kernel.py:117-118 computes every layer in complex128 and the probabilities
// in float64). The same pull recurrence, readout, loss descent and adjoint as the complex64 kernels
// of slos.hip, one thread per state, with the parents and children of pull_index.hpp (the forms
// for at most KM photons, whose per-state arrays stay in registers); every sum in FP64, every state
// written exactly once, no atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pull_index.hpp"
#include "space.hpp"

namespace slos {

// prod_j t_j! of a sorted K-tuple (K <= KM) in FP64 (exact up to 18! per mode, far beyond the
// float form)
template <int KM>
SLOS_HD double run_factorial_f64(const int (&p)[KM], int K, int multiset) {
  double f = 1.0, run = 1.0;
  if (multiset) {
#pragma unroll
    for (int i = 1; i < KM; ++i)
      if (i < K) {
        run = (p[i] == p[i - 1]) ? run + 1.0 : 1.0;
        f *= run;
      }
  }
  return f;
}

struct LayerArgsF64 {
  const uint64_t* cg;
  const int* lo;
  int me, multiset, K, m, col;
  const double2* U;    // [B][m][m]
  const double2* par;  // [B][S_par]
  uint64_t S_par;
  void* out;           // [B][S]: complex128 (kinds 0, 2) or float64 (kind 1)
  uint64_t S;
  int64_t B;
  double scale;
  unsigned immediate;  // kind 2: amplitudes times sqrt(prod t!)
};

// the counting table and lo of layer K into LDS (the first bytes of dynamic LDS)
__device__ __forceinline__ SpaceView load_space_f64(unsigned char* smem, const uint64_t* cg_g, const int* lo_g, int me,
                                                    int K) {
  const int stride = me + 1;
  const int ncg = (K + 1) * stride;
  uint64_t* cg = reinterpret_cast<uint64_t*>(smem);
  int* lo = reinterpret_cast<int*>(cg + ncg);
  for (int i = threadIdx.x; i < ncg; i += blockDim.x) cg[i] = cg_g[i];
  for (int i = threadIdx.x; i < me; i += blockDim.x) lo[i] = lo_g[i];
  return SpaceView{cg, lo, me, stride};
}

// byte offset of the first free 16-byte slot of dynamic LDS after the table of load_space_f64
__host__ __device__ inline size_t space_lds_f64(int K, int me) {
  return ((size_t)(K + 1) * (me + 1) * 8 + (size_t)me * 4 + 15) & ~(size_t)15;
}

// layer K from layer K-1 (K <= KM): KIND 0 raw complex128, 1 float64 probabilities
// |amp s|^2 prod t!, 2 scaled complex128 amplitudes
template <int KIND, int KM>
__global__ __launch_bounds__(256) void k_layer_f64(LayerArgsF64 a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SpaceView sv = load_space_f64(smem, a.cg, a.lo, a.me, a.K);
  double2* ucol = reinterpret_cast<double2*>(smem + space_lds_f64(a.K, a.me));
  const int K = a.K;
  for (int64_t b = blockIdx.y; b < a.B; b += gridDim.y) {
    __syncthreads();
    for (int j = threadIdx.x; j < a.m; j += blockDim.x) ucol[j] = a.U[(size_t)b * a.m * a.m + (size_t)j * a.m + a.col];
    __syncthreads();
    const double2* pb = a.par + (size_t)b * a.S_par;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.S; g += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      unrank_fixed<KM>(sv, g, K, p);
      double re = 0.0, im = 0.0;
      parents_fixed<KM>(sv, a.multiset, K, p, [&](uint64_t pr, int mode, int) {
        const double2 u = ucol[mode], v = pb[pr];
        re = fma(u.x, v.x, re);
        re = fma(-u.y, v.y, re);
        im = fma(u.x, v.y, im);
        im = fma(u.y, v.x, im);
      });
      if constexpr (KIND == 0) {
        static_cast<double2*>(a.out)[(size_t)b * a.S + g] = make_double2(re, im);
      } else {
        const double f = run_factorial_f64<KM>(p, K, a.multiset);
        re *= a.scale;
        im *= a.scale;
        if constexpr (KIND == 1) {
          static_cast<double*>(a.out)[(size_t)b * a.S + g] = (re * re + im * im) * f;
        } else {
          const double s = a.immediate ? sqrt(f) : 1.0;
          static_cast<double2*>(a.out)[(size_t)b * a.S + g] = make_double2(re * s, im * s);
        }
      }
    }
  }
}

__global__ void k_fill_ones_f64(double2* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_double2(1.0, 0.0);
}

// the vacuum's output (one state a batch row): probability scale^2, amplitude scale
__global__ void k_vacuum_out_f64(void* out, int64_t n, int probs, double scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    if (probs) static_cast<double*>(out)[i] = scale * scale;
    else static_cast<double2*>(out)[i] = make_double2(scale, 0.0);
  }
}

// uniform loss: q_{k-1}(t) = scale * (sum_j q_k(t + e_j) (t_j + 1)) / k (children in ascending
// mode; k - 1 <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_descend_f64(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int k,
                                                     const double* __restrict__ qk, uint64_t Sk,
                                                     double* __restrict__ qout, uint64_t Sp, int64_t B, double scale) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SpaceView sv = load_space_f64(smem, cg_g, lo_g, me, k);
  __syncthreads();
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    const double* qb = qk + (size_t)b * Sk;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < Sp; t += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      unrank_fixed<KM>(sv, t, k - 1, p);
      double acc = 0.0;
      children_fixed<KM>(sv, multiset, k, p, [&](uint64_t c, int, int cnt) { acc = fma(qb[c], (double)cnt, acc); });
      qout[(size_t)b * Sp + t] = (acc / (double)k) * scale;
    }
  }
}

// adjoint of the descent: g_k[c] = sum_{j distinct in c} g_{k-1}[c - e_j] count_j(c) / k (k <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_ascend_f64(const uint64_t* cg_g, const int* lo_g, int me, int multiset, int k,
                                                    const double* __restrict__ gp, uint64_t Sp,
                                                    double* __restrict__ gk, uint64_t Sk, int64_t B) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SpaceView sv = load_space_f64(smem, cg_g, lo_g, me, k);
  __syncthreads();
  for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
    const double* pb = gp + (size_t)b * Sp;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < Sk; g += (uint64_t)gridDim.x * blockDim.x) {
      int p[KM];
      unrank_fixed<KM>(sv, g, k, p);
      double acc = 0.0;
      parents_fixed<KM>(sv, multiset, k, p, [&](uint64_t pr, int, int cnt) { acc = fma(pb[pr], (double)cnt, acc); });
      gk[(size_t)b * Sk + g] = acc / (double)k;
    }
  }
}

// readout: p = |a s|^2 f  =>  g_a = 2 s^2 f grad_p a (K <= KM)
template <int KM>
__global__ __launch_bounds__(256) void k_readout_adjoint_f64(const uint64_t* cg_g, const int* lo_g, int me, int multiset,
                                                             int K, const double2* __restrict__ a,
                                                             const double* __restrict__ gp, double2* __restrict__ g,
                                                             uint64_t S, int64_t B, double s2) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SpaceView sv = load_space_f64(smem, cg_g, lo_g, me, K);
  __syncthreads();
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < S; c += (uint64_t)gridDim.x * blockDim.x) {
    bool have_f = false;
    double f = 0.0;
    for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
      // zero grad_p: zero gradient, without the amplitude or the state's factor
      const double gv = gp[(size_t)b * S + c];
      double2 r = make_double2(0.0, 0.0);
      if (gv != 0.0) {
        if (!have_f) {
          int p[KM];
          unrank_fixed<KM>(sv, c, K, p);
          f = run_factorial_f64<KM>(p, K, multiset);
          have_f = true;
        }
        const double2 v = a[(size_t)b * S + c];
        const double w = 2.0 * s2 * f * gv;
        r = make_double2(w * v.x, w * v.y);
      }
      g[(size_t)b * S + c] = r;
    }
  }
}

// g_{k-1}[t] = sum_{children c = t + e_j} conj(U[b, j, col]) g_k[c] (pull over children), and per
// thread gU_j += conj(psi_{k-1}[t]) g_k[t + e_j] in a private LDS row; each work group then sums its
// rows in thread order to partial[b][block][j] (k - 1 <= KM)
constexpr int kAdjThreadsF64 = 64;
template <int KM>
__global__ __launch_bounds__(kAdjThreadsF64) void k_adjoint_layer_f64(const uint64_t* cg_g, const int* lo_g, int me,
                                                                       int multiset, int k, int m, int col,
                                                                       const double2* __restrict__ U,
                                                                       const double2* __restrict__ psi,
                                                                       const double2* __restrict__ gk, uint64_t Sk,
                                                                       double2* __restrict__ gout, uint64_t Sp,
                                                                       double2* __restrict__ partial, int want_psi_grad) {
  extern __shared__ __align__(16) unsigned char smem[];
  const SpaceView sv = load_space_f64(smem, cg_g, lo_g, me, k);
  double2* ucol = reinterpret_cast<double2*>(smem + space_lds_f64(k, me));
  double2* slots = ucol + m;  // [kAdjThreadsF64][me]
  const int64_t b = blockIdx.y;
  for (int j = threadIdx.x; j < m; j += blockDim.x) ucol[j] = U[(size_t)b * m * m + (size_t)j * m + col];
  double2* my = slots + (size_t)threadIdx.x * me;
  for (int j = 0; j < me; ++j) my[j] = make_double2(0.0, 0.0);
  __syncthreads();
  const double2* gb = gk + (size_t)b * Sk;
  const double2* pb = psi + (size_t)b * Sp;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < Sp; t += (uint64_t)gridDim.x * blockDim.x) {
    const double2 ps = pb[t];
    int p[KM];
    unrank_fixed<KM>(sv, t, k - 1, p);
    double gr = 0.0, gi = 0.0;
    children_fixed<KM>(sv, multiset, k, p, [&](uint64_t c, int j, int) {
      const double2 gv = gb[c];
      const double2 u = ucol[j];
      gr += u.x * gv.x + u.y * gv.y;  // conj(u) * gv
      gi += u.x * gv.y - u.y * gv.x;
      my[j].x += ps.x * gv.x + ps.y * gv.y;  // conj(psi) * gv
      my[j].y += ps.x * gv.y - ps.y * gv.x;
    });
    if (want_psi_grad) gout[(size_t)b * Sp + t] = make_double2(gr, gi);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < me; j += blockDim.x) {
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < kAdjThreadsF64; ++r) {
      sr += slots[(size_t)r * me + j].x;
      si += slots[(size_t)r * me + j].y;
    }
    partial[((size_t)b * gridDim.x + blockIdx.x) * me + j] = make_double2(sr, si);
  }
}

// gU[b][j][col] (+)= sum over work groups of partial[b][block][j] (complex128 gU)
__global__ __launch_bounds__(64) void k_reduce_gu_f64(const double2* __restrict__ partial, int nblk, int me, int m,
                                                      int col, double2* __restrict__ gU, int accumulate) {
  const int64_t b = blockIdx.y;
  const int j = blockIdx.x;
  double sr = 0.0, si = 0.0;
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
    double2* dst = gU + (size_t)b * m * m + (size_t)j * m + col;
    if (accumulate) *dst = make_double2(dst->x + sr, dst->y + si);
    else *dst = make_double2(sr, si);
  }
}

}  // namespace slos
