// Per-state edge enumeration on the lex-ranked layers (host and device), for the thread-per-state
// kernels: the forward layer fallback, the loss descent and its adjoint, the gradient kernels.
//
// for_each_parent(K, g): the parents of child g of layer K, one per distinct mode (the raising
//   edges of photonic_engines/slos/basis.py:160-193), with the rank formula of space.hpp.
// for_each_child(k, t): the children of state t of layer k-1 (every admissible inserted mode j,
//   the transpose of the same edges). Inserting j at position i (i = number of parent photons
//   <= j) gives, with the parent's own terms T_l and PB_i = sum_{l<i} (cg[k-l][lo_l] -
//   cg[k-l][p_l]):
//     i < P: rank = PB_i + cg[k-i][lo_i] - cg[k-i-1][p_i] + sum_{l>i} T_l - cg[k-i][j+1]
//     i = P: rank = PB_P + cg[1][lo_P] - cg[1][j]
//   because cg[r][j] = cg[r][j+1] + cg[r-1][lo[j]]. Segment i holds j in [lo_i, p_i) with
//   lo[j] <= p_i (P = k-1 parent photons, p_P treated as me).
// The callback receives (rank, mode, copies of the mode in the child).
#pragma once
#include <stdint.h>

#include "space.hpp"

namespace slos {

constexpr int kMaxK = 32;

template <class F>
SLOS_HD void for_each_parent(const SpaceView& sv, int multiset, int K, uint64_t g, F&& f) {
  int p[kMaxK];
  int lop[kMaxK + 1];
  uint64_t suf[kMaxK + 2];
  sv.unrank(g, K, p);
  lop[0] = 0;
  for (int i = 0; i < K; ++i) lop[i + 1] = sv.lo[p[i]];
  suf[K] = 0;
  suf[K + 1] = 0;
  for (int i = K - 1; i >= 0; --i) suf[i] = suf[i + 1] + (sv.C(K - i, lop[i]) - sv.C(K - i, p[i]));
  uint64_t pre = 0;
  for (int l = 0; l < K; ++l) {
    if (!multiset || l == 0 || p[l] != p[l - 1]) {
      int cnt = 1;
      if (multiset)
        while (l + cnt < K && p[l + cnt] == p[l]) ++cnt;
      const uint64_t pr =
          (l == K - 1) ? pre : pre + (sv.C(K - l - 1, lop[l]) - sv.C(K - l - 1, p[l + 1])) + suf[l + 2];
      f(pr, p[l], cnt);
    }
    if (l < K - 1) pre += sv.C(K - l - 1, lop[l]) - sv.C(K - l - 1, p[l]);
  }
}

// children of state t of layer k-1 in layer k
template <class F>
SLOS_HD void for_each_child(const SpaceView& sv, int multiset, int k, uint64_t t, F&& f) {
  const int P = k - 1;
  int p[kMaxK];
  uint64_t SP[kMaxK + 1];
  if (P > 0) sv.unrank(t, P, p);
  {
    int lov[kMaxK + 1];
    lov[0] = 0;
    for (int i = 0; i < P; ++i) lov[i + 1] = sv.lo[p[i]];
    SP[P] = 0;
    for (int i = P - 1; i >= 0; --i) SP[i] = SP[i + 1] + (sv.C(P - i, lov[i]) - sv.C(P - i, p[i]));
  }
  uint64_t PB = 0;
  int lo_i = 0, run = 0;
  for (int i = 0; i <= P; ++i) {
    if (i < P) {
      const uint64_t base = PB + sv.C(k - i, lo_i) - sv.C(k - i - 1, p[i]) + SP[i + 1];
      for (int j = lo_i; j < p[i]; ++j) {
        if (sv.lo[j] > p[i]) break;
        const int cnt = (multiset && i > 0 && j == p[i - 1]) ? run + 1 : 1;
        f(base - sv.C(k - i, j + 1), j, cnt);
      }
      PB += sv.C(k - i, lo_i) - sv.C(k - i, p[i]);
      run = (i > 0 && p[i] == p[i - 1]) ? run + 1 : 1;
      lo_i = sv.lo[p[i]];
    } else {
      const uint64_t base = PB + sv.C(1, lo_i);
      for (int j = lo_i; j < sv.me; ++j) {
        const int cnt = (multiset && i > 0 && j == p[i - 1]) ? run + 1 : 1;
        f(base - sv.C(1, j), j, cnt);
      }
    }
  }
}

// prod_j t_j! of a sorted tuple
SLOS_HD float run_factorial(const int* p, int K, int multiset) {
  float f = 1.f, run = 1.f;
  if (multiset)
    for (int i = 1; i < K; ++i) {
      run = (p[i] == p[i - 1]) ? run + 1.f : 1.f;
      f *= run;
    }
  return f;
}

// ------------------------------------------------------------------------------------------
// The same enumerations for at most KM photons, every per-state array indexed by compile-time
// positions (loops over 0..KM-1, fully unrolled, with the runtime count as a guard), so that the
// arrays live in registers instead of per-thread scratch memory. Same rank terms, same callbacks in
// the same order as for_each_parent / for_each_child: bit-identical sums.
// ------------------------------------------------------------------------------------------

// (an index kept inside [0, N) for code that the runtime guards never reach)
template <int N>
SLOS_HD constexpr int in_range(int i) { return i < 0 ? 0 : i >= N ? N - 1 : i; }

// sorted positions of lex index g of layer K (K <= KM) into p[0..K)
template <int KM>
SLOS_HD void unrank_fixed(const SpaceView& sv, uint64_t g, int K, int (&p)[KM]) {
  int l = 0;
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    p[i] = 0;
    if (i < K) {
      const int rem = K - i;
      const uint64_t base = sv.C(rem, l);
      int x = l;
      while (base - sv.C(rem, x + 1) <= g) ++x;
      g -= base - sv.C(rem, x);
      p[i] = x;
      l = sv.lo[x];
    }
  }
}

// the parents of the K-tuple p (K <= KM), as for_each_parent
template <int KM, class F>
SLOS_HD void parents_fixed(const SpaceView& sv, int multiset, int K, const int (&p)[KM], F&& f) {
  int lop[KM + 1];
  uint64_t suf[KM + 2];
  int run[KM];
  lop[0] = 0;
#pragma unroll
  for (int i = 0; i < KM; ++i) lop[i + 1] = i < K ? sv.lo[p[i]] : 0;
  suf[KM] = 0;
  suf[KM + 1] = 0;
#pragma unroll
  for (int i = KM - 1; i >= 0; --i) {
    suf[i] = i < K ? suf[i + 1] + (sv.C(K - i, lop[i]) - sv.C(K - i, p[i])) : 0;
    // copies of p[i] from i on (the count of a mode's first copy)
    const bool same = i + 1 < K && p[in_range<KM>(i + 1)] == p[i];
    run[i] = same ? run[in_range<KM>(i + 1)] + 1 : 1;
  }
  uint64_t pre = 0;
#pragma unroll
  for (int l = 0; l < KM; ++l) {
    if (l < K) {
      if (!multiset || l == 0 || p[l] != p[in_range<KM>(l - 1)]) {
        const int cnt = multiset ? run[l] : 1;
        const uint64_t pr =
            (l == K - 1) ? pre
                         : pre + (sv.C(K - l - 1, lop[l]) - sv.C(K - l - 1, p[in_range<KM>(l + 1)])) + suf[l + 2];
        f(pr, p[l], cnt);
      }
      if (l < K - 1) pre += sv.C(K - l - 1, lop[l]) - sv.C(K - l - 1, p[l]);
    }
  }
}

// the children in layer k of the (k-1)-tuple p (k - 1 <= KM), as for_each_child
template <int KM, class F>
SLOS_HD void children_fixed(const SpaceView& sv, int multiset, int k, const int (&p)[KM], F&& f) {
  const int P = k - 1;
  uint64_t SP[KM + 1];
  SP[KM] = 0;
#pragma unroll
  for (int i = KM - 1; i >= 0; --i) {
    const int lov = i == 0 ? 0 : sv.lo[p[in_range<KM>(i - 1)]];
    SP[i] = i < P ? SP[i + 1] + (sv.C(P - i, lov) - sv.C(P - i, p[i])) : 0;
  }
  uint64_t PB = 0;
  int lo_i = 0, run = 0;
#pragma unroll
  for (int i = 0; i <= KM; ++i) {
    const int pi = p[in_range<KM>(i)], pprev = p[in_range<KM>(i - 1)];
    if (i < P) {
      const uint64_t base = PB + sv.C(k - i, lo_i) - sv.C(k - i - 1, pi) + SP[in_range<KM + 1>(i + 1)];
      for (int j = lo_i; j < pi; ++j) {
        if (sv.lo[j] > pi) break;
        const int cnt = (multiset && i > 0 && j == pprev) ? run + 1 : 1;
        f(base - sv.C(k - i, j + 1), j, cnt);
      }
      PB += sv.C(k - i, lo_i) - sv.C(k - i, pi);
      run = (i > 0 && pi == pprev) ? run + 1 : 1;
      lo_i = sv.lo[pi];
    } else if (i == P) {
      const uint64_t base = PB + sv.C(1, lo_i);
      for (int j = lo_i; j < sv.me; ++j) {
        const int cnt = (multiset && i > 0 && j == pprev) ? run + 1 : 1;
        f(base - sv.C(1, j), j, cnt);
      }
    }
  }
}

// prod_j t_j! of the K-tuple p (K <= KM), as run_factorial
template <int KM>
SLOS_HD float run_factorial_fixed(const int (&p)[KM], int K, int multiset) {
  float f = 1.f, run = 1.f;
  if (multiset) {
#pragma unroll
    for (int i = 1; i < KM; ++i)
      if (i < K) {
        run = (p[i] == p[i - 1]) ? run + 1.f : 1.f;
        f *= run;
      }
  }
  return f;
}

}  // namespace slos
