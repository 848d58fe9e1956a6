// Counting tables and lex rank / unrank of the mask-pruned Fock layers (host and device).
//
// A k-photon state is the sorted tuple of its photon positions p_1 <= ... <= p_k. Layer order is
// the lexicographic order of these tuples (photonic_engines/slos/basis.py:120-148: children are
// every parent followed by every admissible position a >= last, parent order first).
//
// Two tables describe every mask:
//   lo[x]     smallest position admissible after a photon at x: x for the multiset masks
//             (fock, tail_empty), the first position of the next block for the set masks
//             (unbunched: block 1, dual_rail: 2, sparse_1in4: 4). Positions are < me, where
//             me = m - 2 for tail_empty (its last two modes have cap 0, basis.py:46-47) and m else.
//   cg[r][x]  number of admissible sorted r-tuples whose first element is >= x (x in [0, me]).
//             cg[0][x] = 1, cg[r][me] = 0 (r > 0), cg[r][x] = cg[r][x+1] + cg[r-1][lo[x]].
// Layer size S_k = cg[k][0]. With these, for a K-tuple (lo_0 = 0, lo_i = lo[p_i]):
//   rank(p)  = sum_i T_i,          T_i = cg[K-i+1][lo_{i-1}] - cg[K-i+1][p_i]      (1-based i)
// and the parent that drops element l (1-based) has rank
//   sum_{i<l} A_i + M_l + sum_{i>=l+2} T_i   (l < K),   sum_{i<K} A_i   (l = K)
//   A_i = cg[K-i][lo_{i-1}] - cg[K-i][p_i],   M_l = cg[K-l][lo_{l-1}] - cg[K-l][p_{l+1}]
// because dropping p_l leaves the terms after p_{l+1} unchanged and shifts the earlier ones by one
// photon. For the multiset masks only the first copy of a repeated mode gives an edge (one edge
// per distinct mode, basis.py:166-189).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SLOS_HD __host__ __device__ __forceinline__
#else
#define SLOS_HD inline
#endif

namespace slos {

struct SpaceView {
  const uint64_t* cg;  // (rows) x (me + 1), row-major
  const int* lo;       // me
  int me;              // number of usable positions
  int stride;          // me + 1
  SLOS_HD uint64_t C(int r, int x) const { return cg[r * stride + x]; }

  // lex rank of a sorted K-tuple
  SLOS_HD uint64_t rank(const int* p, int K) const {
    uint64_t r = 0;
    int l = 0;
    for (int i = 0; i < K; ++i) {
      r += C(K - i, l) - C(K - i, p[i]);
      l = lo[p[i]];
    }
    return r;
  }

  // inverse of rank: g must be < C(K, 0)
  SLOS_HD void unrank(uint64_t g, int K, int* p) const {
    int l = 0;
    for (int i = 0; i < K; ++i) {
      const int rem = K - i;
      const uint64_t base = C(rem, l);
      int x = l;
      while (base - C(rem, x + 1) <= g) ++x;
      g -= base - C(rem, x);
      p[i] = x;
      l = lo[x];
    }
  }
};

}  // namespace slos
