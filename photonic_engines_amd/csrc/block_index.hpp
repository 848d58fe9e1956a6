// Index arithmetic of the block-structured layer kernel, shared by the device kernel and host code
// (the plan builds the suffix-tuple table with it; the CPU tests execute it unchanged).
//
// A child state of layer K is c = (Q, tau): Q = its first d = K - SUF sorted positions, tau = the
// last SUF positions, all >= lo = lo[last(Q)] (lo = 0 when d = 0). The children of one Q are
// contiguous (the "block" of Q, size cg[SUF][lo]), and tau runs over the SUF-tuples >= lo in
// lex order, which is the tail of the lex order of all SUF-tuples of the space. So with u the
// position of c inside its block:
//  * dropping the prefix photon q_l (first copy of its mode) gives the parent (Q - q_l, tau); for
//    a fixed Q these parents form an aligned contiguous run of layer K-1: parent = base_l + u;
//  * dropping the suffix photon x_j gives (Q, tau - x_j) inside the block of Q in layer K-1
//    (suffix length SUF-1): parent = LBp + lr_j, lr_j = global (SUF-1)-tuple index of tau - x_j.
//    With cg[r][x] - cg[r-1][lo[x]] = cg[r][x+1] (the counting recurrence), the global index of a
//    3-tuple (a, b, c) is cg[3][0] - cg[3][a+1] - cg[2][b+1] - cg[1][c]: three per-position terms.
// Ranks use the counting table of space.hpp: with lo_0 = 0, lo_{i+1} = lo[q_i],
//   T_i = cg[K-i][lo_i] - cg[K-i][q_i]        (child terms, photon count K)
//   A_i = cg[K-1-i][lo_i] - cg[K-1-i][q_i]    (same position, one photon fewer)
//   M_l = cg[K-1-l][lo_l] - cg[K-1-l][q_{l+1}] (q_{l+1} moved into position l)
//   block start CB = sum_{i<d} T_i, local parent base LB = sum_{i<d} A_i,
//   prefix parent block of Q - q_l: sum_{i<l} A_i + M_l + sum_{l+2<=i<d} T_i (l < d-1),
//                                  sum_{i<d-1} A_i (l = d-1).
#pragma once
#include <stdint.h>

#include "space.hpp"

namespace slos {

constexpr int kSuf = 4;
constexpr int kNoEdge = 255;

// Suffix-tuple word (4 B) of a SUF-tuple tau, in lex order over all SUF-tuples: byte j is
// x_j | mark_j << 6, where mark_j = 1 when x_j repeats x_{j-1} (a multiset's later copy of a mode,
// which is not an edge). Positions are < 64 (kBlockModes, checked at plan creation). A byte is
// also the row of the per-launch suffix table (128 rows): rows x hold U[x, col] and the rank terms
// of position x, rows x | 64 the same rank terms with a zero coefficient.
typedef uint32_t TupleWord;
constexpr int kBlockModes = 64;
constexpr uint32_t kTupleMark = 64;
constexpr int kSufRows = 128;
constexpr int kNoEdgeRow = 127;  // a zero-coefficient row: inactive prefix-stream slots

// the marks of x1, x2, x3 as a 3-bit code (bit j-1: x_j repeats x_{j-1})
SLOS_HD uint32_t tuple_code(TupleWord w) { return ((w >> 14) & 1u) | ((w >> 21) & 2u) | ((w >> 28) & 4u); }
// copies of x0 in tau
SLOS_HD int code_lead(uint32_t c) { return 1 + (int)(c & 1u) + ((c & 3u) == 3u) + (c == 7u); }
// prod of the run factorials of tau
SLOS_HD int code_fct(uint32_t c) {
  int f = 1, run = 1;
  for (int j = 0; j < kSuf - 1; ++j) {
    run = ((c >> j) & 1u) ? run + 1 : 1;
    f *= run;
  }
  return f;
}

// Global (SUF-1)-tuple indices of tau without x_j, j = 0..3, from the per-position terms of tau's
// real positions: a3[j] = cg[3][x_j + 1], b2[j] = cg[2][x_j + 1], b1[j] = cg[1][x_j], C0 = cg[3][0].
SLOS_HD void suffix_parent_ranks(uint32_t C0, const uint32_t (&a3)[kSuf], const uint32_t (&b2)[kSuf],
                                 const uint32_t (&b1)[kSuf], uint32_t (&lr)[kSuf]) {
  const uint32_t s23 = b2[2] + b1[3];
  const uint32_t t0 = C0 - a3[0];
  const uint32_t t01 = t0 - b2[1];
  lr[0] = C0 - a3[1] - s23;  // (x1, x2, x3)
  lr[1] = t0 - s23;          // (x0, x2, x3)
  lr[2] = t01 - b1[3];       // (x0, x1, x3)
  lr[3] = t01 - b1[2];       // (x0, x1, x2)
}

struct BlockTables {
  const uint64_t* cg;        // (n + 2) x (me + 1)
  const int* lo;             // me
  const TupleWord* tuples;   // cg[SUF][0] entries
  int me, stride, multiset;
};

struct BlockInfo {
  uint64_t CB;      // index of the block's first child in layer K
  uint64_t LBp;     // local parent base minus the (SUF-1)-tuple tail offset of lo
  uint32_t L;       // block size
  uint32_t toff;    // tuple-table index of the block's first tau
  int lo;           // smallest admissible suffix position
  int skip0;        // multiset and d >= 1: a tau starting at lo continues last(Q)'s run
  int lastrun;      // copies of last(Q) in Q (multiset)
  float fQ;         // prod of run factorials of Q
};

// Prefix streams, one slot per prefix position l (active when q_l is the first copy of its mode),
// so that every array index is a compile-time constant after unrolling and the whole per-wave
// state stays in registers.
template <int MAXD>
struct Streams {
  uint64_t base[MAXD];
  int mode[MAXD];
  bool active[MAXD];
};

template <int MAXD>
SLOS_HD void block_info(const BlockTables& t, int K, int d, const int* q, BlockInfo& bi, Streams<MAXD>& st) {
  const uint64_t* cg = t.cg;
  const int W = t.stride;
#define SLOS_C(r, x) cg[(r) * W + (x)]
  int lov[MAXD + 1];
  lov[0] = 0;
  int lo = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i) {
    lov[i + 1] = (i < d) ? t.lo[q[i]] : 0;
    if (i < d) lo = lov[i + 1];
  }
  uint64_t CB = 0, LB = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i)
    if (i < d) {
      CB += SLOS_C(K - i, lov[i]) - SLOS_C(K - i, q[i]);
      LB += SLOS_C(K - 1 - i, lov[i]) - SLOS_C(K - 1 - i, q[i]);
    }
  bi.CB = CB;
  bi.L = (uint32_t)SLOS_C(kSuf, lo);
  bi.toff = (uint32_t)(SLOS_C(kSuf, 0) - SLOS_C(kSuf, lo));
  bi.LBp = LB - (SLOS_C(kSuf - 1, 0) - SLOS_C(kSuf - 1, lo));
  bi.lo = lo;
  bi.skip0 = t.multiset && d >= 1;
  float fQ = 1.f, run = 1.f;
  if (t.multiset) {
#pragma unroll
    for (int i = 1; i < MAXD; ++i)
      if (i < d) {
        run = (q[i] == q[i - 1]) ? run + 1.f : 1.f;
        fQ *= run;
      }
  }
  bi.fQ = fQ;
  bi.lastrun = d >= 1 ? (int)run : 0;
  uint64_t sufT[MAXD + 2];
  sufT[MAXD] = 0;
  sufT[MAXD + 1] = 0;
#pragma unroll
  for (int i = MAXD - 1; i >= 0; --i)
    sufT[i] = sufT[i + 1] + ((i < d) ? (SLOS_C(K - i, lov[i]) - SLOS_C(K - i, q[i])) : 0);
  uint64_t preA = 0;
  const uint64_t CS_lo = SLOS_C(kSuf, lo);
#pragma unroll
  for (int l = 0; l < MAXD; ++l) {
    st.active[l] = false;
    st.base[l] = 0;
    st.mode[l] = 0;
    if (l < d) {
      const bool first = !t.multiset || l == 0 || q[l] != q[l - 1];
      if (first) {
        uint64_t pr;
        int lop;
        if (l == d - 1) {
          pr = preA;
          lop = lov[l];
        } else {
          pr = preA + (SLOS_C(K - 1 - l, lov[l]) - SLOS_C(K - 1 - l, q[l + 1])) + sufT[l + 2];
          lop = lo;
        }
        st.base[l] = pr + (SLOS_C(kSuf, lop) - CS_lo);
        st.mode[l] = q[l];
        st.active[l] = true;
      }
      preA += SLOS_C(K - 1 - l, lov[l]) - SLOS_C(K - 1 - l, q[l]);
    }
  }
#undef SLOS_C
}

// lex-next prefix Q (length d) whose block in layer K is non-empty; false after the last one.
// Unrolled over positions so that q stays in registers: find the last position c that can be
// raised (to the smallest admissible x > q_c leaving a completion), then refill c+1.. minimally.
template <int MAXD>
SLOS_HD bool next_prefix(const BlockTables& t, int K, int d, int* q) {
  const int W = t.stride;
  int c = -1, cx = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i)
    if (i < d) {
      int x = q[i] + 1;
      while (x < t.me && t.cg[(K - 1 - i) * W + t.lo[x]] == 0) ++x;
      if (x < t.me) {
        c = i;
        cx = x;
      }
    }
  if (c < 0) return false;
  int l = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i)
    if (i < d && i >= c) {
      int v = cx;
      if (i > c) {
        v = l;
        while (t.cg[(K - 1 - i) * W + t.lo[v]] == 0) ++v;
      }
      q[i] = v;
      l = t.lo[v];
    }
  return true;
}

// lex rank of the prefix Q among all d-tuples of the space (layer d)
template <int MAXD>
SLOS_HD uint64_t prefix_rank(const BlockTables& t, int d, const int* q) {
  const int W = t.stride;
  uint64_t r = 0;
  int l = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i)
    if (i < d) {
      r += t.cg[(d - i) * W + l] - t.cg[(d - i) * W + q[i]];
      l = t.lo[q[i]];
    }
  return r;
}

// prefix Q of the child with lex index g in layer K; returns g's position u inside Q's block
template <int MAXD>
SLOS_HD uint64_t unrank_prefix(const BlockTables& t, int K, int d, uint64_t g, int* q) {
  const int W = t.stride;
  int l = 0;
#pragma unroll
  for (int i = 0; i < MAXD; ++i)
    if (i < d) {
      const int r = K - i;
      const uint64_t base = t.cg[r * W + l];
      int x = l;
      while (base - t.cg[r * W + x + 1] <= g) ++x;
      g -= base - t.cg[r * W + x];
      q[i] = x;
      l = t.lo[x];
    }
  return g;
}

// the tuple word of one SUF-tuple (host side, at plan creation)
inline TupleWord make_tuple_word(const SpaceView& sv, int multiset, uint64_t u) {
  int x[kSuf];
  sv.unrank(u, kSuf, x);
  TupleWord w = 0;
  for (int j = 0; j < kSuf; ++j) {
    const bool mark = multiset && j > 0 && x[j] == x[j - 1];
    w |= ((uint32_t)x[j] | (mark ? kTupleMark : 0u)) << (8 * j);
  }
  return w;
}

static_assert(kSuf <= 4, "lead_binom covers runs of up to 4 copies");

// binom(r + lead, lead) for lead <= kSuf, exact in integers (each step divides exactly)
SLOS_HD float lead_binom(int r, int lead) {
  uint32_t p = (uint32_t)r + 1u;
  const uint32_t p2 = p * ((uint32_t)r + 2u) / 2u;
  const uint32_t p3 = p2 * ((uint32_t)r + 3u) / 3u;
  const uint32_t p4 = p3 * ((uint32_t)r + 4u) / 4u;
  p = lead >= 4 ? p4 : lead == 3 ? p3 : lead == 2 ? p2 : lead == 1 ? p : 1u;
  return (float)p;
}

// prod t_j! of the child (Q, tau) from the block's run data and the tuple word
SLOS_HD float child_norm(const BlockInfo& bi, TupleWord w, int multiset) {
  if (!multiset) return 1.f;
  const uint32_t c = tuple_code(w);
  float f = bi.fQ * (float)code_fct(c);
  // tau's leading run continues the last run of Q: (r + c)! = r! c! binom(r + c, c)
  if (bi.skip0 && (int)(w & 0xffu) == bi.lo) f *= lead_binom(bi.lastrun, code_lead(c));
  return f;
}

}  // namespace slos
