// Index arithmetic of the block-structured layer kernel, shared by the device kernel and host code
// (the plan builds the suffix-tuple table with it; the CPU tests execute it unchanged).
//
// A child state of layer K is c = (Q, tau): Q = its first d = K - SUF sorted positions, tau = the
// last SUF positions (SUF = 4, or 5 for the big layers), all >= lo = lo[last(Q)] (lo = 0 when d = 0). The children of one Q are
// contiguous (the "block" of Q, size cg[SUF][lo]), and tau runs over the SUF-tuples >= lo in
// lex order, which is the tail of the lex order of all SUF-tuples of the space. So with u the
// position of c inside its block:
//  * dropping the prefix photon q_l (first copy of its mode) gives the parent (Q - q_l, tau); for
//    a fixed Q these parents form an aligned contiguous run of layer K-1: parent = base_l + u;
//  * dropping the suffix photon x_j gives (Q, tau - x_j) inside the block of Q in layer K-1
//    (suffix length SUF-1): parent = LBp + lr_j, lr_j = global (SUF-1)-tuple index of tau - x_j.
//    With cg[r][x] - cg[r-1][lo[x]] = cg[r][x+1] (the counting recurrence), the global index of a
//    3-tuple (a, b, c) is cg[3][0] - cg[3][a+1] - cg[2][b+1] - cg[1][c]: three per-position terms
//    (a 4-tuple (a, b, c, e): cg[4][0] - cg[4][a+1] - cg[3][b+1] - cg[2][c+1] - cg[1][e]).
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

constexpr int kSuf = 4;     // suffix length of the loss descent and the uneven-batch launches
constexpr int kSufMax = 5;  // the forward layer kernel also runs with 5-position suffixes
constexpr int kNoEdge = 255;

// Suffix-tuple word of a SUF-tuple tau, in lex order over all SUF-tuples: byte j is
// x_j | mark_j << 6, where mark_j = 1 when x_j repeats x_{j-1} (a multiset's later copy of a mode,
// which is not an edge); 4 B for SUF = 4, 8 B (bytes 5..7 zero) for SUF = 5. Positions are < 64
// (kBlockModes, checked at plan creation). A byte is also the row of the per-launch suffix table
// (128 rows): rows x hold U[x, col] and the rank terms of position x, rows x | 64 the same rank
// terms with a zero coefficient.
template <int SUF>
struct SufWord;
template <>
struct SufWord<4> {
  typedef uint32_t type;
};
template <>
struct SufWord<5> {
  typedef uint64_t type;
};
template <int SUF>
using TupleWordT = typename SufWord<SUF>::type;
typedef TupleWordT<4> TupleWord;
constexpr int kBlockModes = 64;
constexpr uint32_t kTupleMark = 64;
constexpr int kSufRows = 128;
constexpr int kNoEdgeRow = 127;  // a zero-coefficient row: inactive prefix-stream slots

// byte j of a tuple word
template <class W>
SLOS_HD uint32_t tw_byte(W w, int j) {
  return (uint32_t)(w >> (8 * j)) & 0xffu;
}
// the marks of x_1 .. x_{SUF-1} as a (SUF-1)-bit code (bit j-1: x_j repeats x_{j-1})
template <int SUF = kSuf>
SLOS_HD uint32_t tuple_code(TupleWordT<SUF> w) {
  uint32_t c = 0;
#pragma unroll
  for (int j = 1; j < SUF; ++j) c |= ((tw_byte(w, j) >> 6) & 1u) << (j - 1);
  return c;
}
// copies of x0 in tau
template <int SUF = kSuf>
SLOS_HD int code_lead(uint32_t c) {
  int lead = 1;
  bool run = true;
#pragma unroll
  for (int j = 0; j < SUF - 1; ++j) {
    run = run && ((c >> j) & 1u);
    lead += run ? 1 : 0;
  }
  return lead;
}
// prod of the run factorials of tau
template <int SUF = kSuf>
SLOS_HD int code_fct(uint32_t c) {
  int f = 1, run = 1;
  for (int j = 0; j < SUF - 1; ++j) {
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

// The same for SUF = 5: global 4-tuple indices of tau without x_j, j = 0..4, from
// t4[j] = cg[4][x_j + 1], t3[j] = cg[3][x_j + 1], t2[j] = cg[2][x_j + 1], t1[j] = cg[1][x_j], C0 = cg[4][0].
SLOS_HD void suffix_parent_ranks5(uint32_t C0, const uint32_t (&t4)[5], const uint32_t (&t3)[5], const uint32_t (&t2)[5],
                                  const uint32_t (&t1)[5], uint32_t (&lr)[5]) {
  const uint32_t s34 = t2[3] + t1[4];
  const uint32_t c0 = C0 - t4[0];
  const uint32_t c01 = c0 - t3[1];
  const uint32_t c012 = c01 - t2[2];
  lr[0] = C0 - t4[1] - t3[2] - s34;  // (x1, x2, x3, x4)
  lr[1] = c0 - t3[2] - s34;          // (x0, x2, x3, x4)
  lr[2] = c01 - s34;                 // (x0, x1, x3, x4)
  lr[3] = c012 - t1[4];              // (x0, x1, x2, x4)
  lr[4] = c012 - t1[3];              // (x0, x1, x2, x3)
}

struct BlockTables {
  const uint64_t* cg;        // (n + 2) x (me + 1)
  const int* lo;             // me
  const void* tuples;        // cg[SUF][0] tuple words of the launch's SUF
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

template <int MAXD, int SUF = kSuf>
SLOS_HD void block_info(const BlockTables& t, int K, int d, const int* q, BlockInfo& bi, Streams<MAXD>& st) {
  constexpr int kSuf = SUF;  // the suffix length of this launch (shadows the default)
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
template <int SUF = kSuf>
inline TupleWordT<SUF> make_tuple_word(const SpaceView& sv, int multiset, uint64_t u) {
  int x[SUF];
  sv.unrank(u, SUF, x);
  TupleWordT<SUF> w = 0;
  for (int j = 0; j < SUF; ++j) {
    const bool mark = multiset && j > 0 && x[j] == x[j - 1];
    w |= (TupleWordT<SUF>)((uint32_t)x[j] | (mark ? kTupleMark : 0u)) << (8 * j);
  }
  return w;
}

// ---- the last layer in group order (5-position suffixes, multiset spaces) ----------------------
// A block Q = (a, R) of the last layer drops its first photon into the parent block R of layer
// K-1 (same tuple offsets u, since lo[last(Q)] = lo[last(R)]). The blocks (a, R), a = 0..R_0,
// form the group of R: they share that parent block, their tuple words and their local-parent
// ranks. A work item is (rank of R among the (d-1)-prefixes, u_begin, u_end): a range of suffix
// positions of every block of the group, at most about 2048 children in all (measured best at
// n = 11, m = 32: 27.5 ms for the last layer against 28.2 ms with 4096, 28.9 ms with 1024).
constexpr uint32_t kGroupItemChildren = 2048;
SLOS_HD uint32_t group_item_width(int G) {
  const uint32_t w = (kGroupItemChildren / (uint32_t)G) & ~127u;
  return w < 128u ? 128u : w;
}

// the work items of layer K (prefix length d = K - 5 >= 2), 4 words each (r, u_b, u_e, 0)
template <class Vec>
inline void group_items(const SpaceView& sv, int K, Vec& out) {
  const int d = K - 5;
  const uint64_t nR = sv.C(d - 1, 0);
  int R[64];
  for (uint64_t r = 0; r < nR; ++r) {
    sv.unrank(r, d - 1, R);
    const uint32_t L = (uint32_t)sv.C(5, sv.lo[R[d - 2]]);
    const uint32_t w = group_item_width(R[0] + 1);
    for (uint32_t ub = 0; ub < L; ub += w) {
      out.push_back((uint32_t)r);
      out.push_back(ub);
      out.push_back(L - ub < w ? L : ub + w);
      out.push_back(0u);
    }
  }
}

// ---- loss descent on the block structure (pull over children, multiset spaces: lo[x] = x) -----
// A parent of layer D = K - 1 is (P, s): P its first dp = D - SUF positions, s = (s0..s3) its last
// four, u its offset in the block of P. Inserting mode a gives the child (K photons):
//  * a in segment i < dp, i.e. a in [lo_i, p_i) (lo_0 = 0, lo_i = p_{i-1}): the child lies in block
//    P + a of layer K at the same offset u, child = B_i - cg[K-i][a+1] + u, with
//    B_i = sum_{l<i} (cg[K-l][lo_l] - cg[K-l][p_l]) + cg[K-i][lo_i] - cg[D-i][p_i]
//          + sum_{i<l<dp} (cg[D-l][lo_l] - cg[D-l][p_l]);
//  * last(P) <= a <= s0: child = CB0 + cg[5][last(P)+1] - cg[5][a+1] + u (block (P, a)),
//    CB0 = sum_{l<dp} (cg[K-l][lo_l] - cg[K-l][p_l]) the first child of block (P, last(P));
//  * a > s0: child = CB0 + cg[5][last(P)] - cg[5][s0+1] - G(sorted(s1, s2, s3, a)) (block (P, s0)),
//    G(r) = cg[4][r0+1] + cg[3][r1+1] + cg[2][r2+1] + cg[1][r3].
// (the blocks (P, x), x >= last(P), are consecutive in layer K, cg[5][x] - cg[4][x] = cg[5][x+1]).
// The copies of a in the child, the weight's numerator, are 1 + its copies in P and in s.
template <int MAXD>
struct DescBlock {
  uint64_t CBp;      // first parent of the block in layer D
  uint64_t CB0;      // first child of block (P, last(P)) in layer K
  uint64_t B[MAXD];  // segment bases
  uint32_t L;        // block size
  uint32_t toff;     // tuple-word index of the block's first suffix
  int lastP, lastrun;
  int p[MAXD];       // the prefix positions
  int run[MAXD];     // run[i], i >= 1: copies of p_{i-1} in P
};

template <int MAXD>
SLOS_HD void descend_block_info(const BlockTables& t, int K, int dp, const int* q, DescBlock<MAXD>& db) {
  const uint64_t* cg = t.cg;
  const int W = t.stride;
  const int D = K - 1;
#define SLOS_C(r, x) cg[(r) * W + (x)]
  uint64_t TD[MAXD];  // parent terms cg[D-l][lo_l] - cg[D-l][p_l]
  uint64_t CBp = 0, CB0 = 0;
  int lo = 0, run = 0, prev = -1;
#pragma unroll
  for (int l = 0; l < MAXD; ++l) {
    TD[l] = 0;
    db.p[l] = 0;
    db.run[l] = 0;
    db.B[l] = 0;
    if (l < dp) {
      db.p[l] = q[l];
      db.run[l] = run;                    // copies of p_{l-1} in P so far (segment l's first mode)
      TD[l] = SLOS_C(D - l, lo) - SLOS_C(D - l, q[l]);
      CBp += TD[l];
      CB0 += SLOS_C(K - l, lo) - SLOS_C(K - l, q[l]);
      run = (q[l] == prev) ? run + 1 : 1;
      prev = q[l];
      lo = q[l];
    }
  }
  // segment bases, from the back: SP = sum_{i<l<dp} TD[l], PB = sum_{l<i} child terms
  uint64_t SP = 0;
#pragma unroll
  for (int i = MAXD - 1; i >= 0; --i)
    if (i < dp) {
      uint64_t PB = 0;
      int lo_l = 0;
#pragma unroll
      for (int l = 0; l < MAXD; ++l)
        if (l < i) {
          PB += SLOS_C(K - l, lo_l) - SLOS_C(K - l, q[l]);
          lo_l = q[l];
        }
      db.B[i] = PB + SLOS_C(K - i, lo_l) - SLOS_C(D - i, q[i]) + SP;
      SP += TD[i];
    }
  db.CBp = CBp;
  db.CB0 = CB0;
  db.lastP = q[dp - 1];
  db.lastrun = run;
  db.L = (uint32_t)SLOS_C(kSuf, db.lastP);
  db.toff = (uint32_t)(SLOS_C(kSuf, 0) - SLOS_C(kSuf, db.lastP));
#undef SLOS_C
}

// Per-parent terms of the children with a > s0: child - CB0 = Zl - Wp[p] - A_p(a), p = number of
// s1..s3 below a, A_0..A_3(a) = cg[4][a+1], cg[3][a+1], cg[2][a+1], cg[1][a].
struct DescLane {
  uint32_t Zl;
  uint32_t Wp[4];
};

SLOS_HD void desc_lane_init(const uint64_t* cg, int W, int lastP, const int (&s)[kSuf], DescLane& dl) {
#define SLOS_C(r, x) (uint32_t)cg[(r) * W + (x)]
  dl.Zl = SLOS_C(5, lastP) - SLOS_C(5, s[0] + 1);
  const uint32_t c4s1 = SLOS_C(4, s[1] + 1), c3s1 = SLOS_C(3, s[1] + 1), c3s2 = SLOS_C(3, s[2] + 1);
  const uint32_t c2s2 = SLOS_C(2, s[2] + 1), c2s3 = SLOS_C(2, s[3] + 1), c1s3 = SLOS_C(1, s[3]);
  dl.Wp[0] = c3s1 + c2s2 + c1s3;   // (a, s1, s2, s3)
  dl.Wp[1] = c4s1 + c2s2 + c1s3;   // (s1, a, s2, s3)
  dl.Wp[2] = c4s1 + c3s2 + c1s3;   // (s1, s2, a, s3)
  dl.Wp[3] = c4s1 + c3s2 + c2s3;   // (s1, s2, s3, a)
#undef SLOS_C
}

// offset from CB0 of the child with mode a > s0 (block (P, s0)); A = (cg[4][a+1], cg[3][a+1],
// cg[2][a+1], cg[1][a]) are the same for every parent
SLOS_HD uint32_t desc_local_offset(const DescLane& dl, const int (&s)[kSuf], int a, uint32_t A0, uint32_t A1, uint32_t A2,
                                   uint32_t A3) {
  // values, not array elements, in the selects (no select of addresses, nothing kept in memory)
  const int p = (s[1] < a) + (s[2] < a) + (s[3] < a);
  const uint32_t w0 = dl.Wp[0], w1 = dl.Wp[1], w2 = dl.Wp[2], w3 = dl.Wp[3];
  const uint32_t wsel = p == 0 ? w0 : p == 1 ? w1 : p == 2 ? w2 : w3;
  const uint32_t asel = p == 0 ? A0 : p == 1 ? A1 : p == 2 ? A2 : A3;
  return dl.Zl - wsel - asel;
}

// offset from CB0 of the child with mode a >= last(P): the stream of block (P, a) at u while
// a <= s0 (S5a = cg[5][last(P)+1] - cg[5][a+1], the same for every parent), else block (P, s0)
SLOS_HD uint32_t desc_tail_offset(const DescLane& dl, const int (&s)[kSuf], uint32_t u, int a, uint32_t S5a,
                                  const uint32_t (&A)[4]) {
  const uint32_t a0 = A[0], a1 = A[1], a2 = A[2], a3 = A[3];
  return a <= s[0] ? S5a + u : desc_local_offset(dl, s, a, a0, a1, a2, a3);
}

// copies of mode a >= last(P) in the child of parent (P, s)
SLOS_HD int desc_tail_copies(const int (&s)[kSuf], int a, int lastP, int lastrun) {
  return 1 + (a == lastP ? lastrun : 0) + (s[0] == a) + (s[1] == a) + (s[2] == a) + (s[3] == a);
}

static_assert(kSufMax <= 5, "lead_binom covers runs of up to 5 copies");

// binom(r + lead, lead) for lead <= 5, exact in integers (each step divides exactly)
SLOS_HD float lead_binom(int r, int lead) {
  uint32_t p = (uint32_t)r + 1u;
  const uint32_t p2 = p * ((uint32_t)r + 2u) / 2u;
  const uint32_t p3 = p2 * ((uint32_t)r + 3u) / 3u;
  const uint32_t p4 = p3 * ((uint32_t)r + 4u) / 4u;
  const uint32_t p5 = p4 * ((uint32_t)r + 5u) / 5u;
  p = lead >= 5 ? p5 : lead == 4 ? p4 : lead == 3 ? p3 : lead == 2 ? p2 : lead == 1 ? p : 1u;
  return (float)p;
}

// prod t_j! of the child (Q, tau) from the block's run data and the tuple word
template <int SUF = kSuf>
SLOS_HD float child_norm(const BlockInfo& bi, TupleWordT<SUF> w, int multiset) {
  if (!multiset) return 1.f;
  const uint32_t c = tuple_code<SUF>(w);
  float f = bi.fQ * (float)code_fct<SUF>(c);
  // tau's leading run continues the last run of Q: (r + c)! = r! c! binom(r + c, c)
  if (bi.skip0 && (int)tw_byte(w, 0) == bi.lo) f *= lead_binom(bi.lastrun, code_lead<SUF>(c));
  return f;
}

}  // namespace slos
