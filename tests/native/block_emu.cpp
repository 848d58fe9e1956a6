// Host execution of the block layer kernel's index code (photonic_engines_amd/csrc/block_index.hpp)
// with the kernel's own wave / block / 64-lane loop structure, in complex128. Test infrastructure:
// it lets the CPU test suite check the GPU kernel's decomposition (block starts, prefix-stream
// bases, suffix-tuple table, norms, wave boundaries) against the oracle without a GPU.
#include <cmath>
#include <complex>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../photonic_engines_amd/csrc/block_index.hpp"
#include "../../photonic_engines_amd/csrc/pull_index.hpp"

using cd = std::complex<double>;

namespace {

struct Space {
  int me = 0, multiset = 1;
  std::vector<uint64_t> cg;
  std::vector<int> lo;
  std::vector<slos::TupleWord> tuples;
  std::vector<uint64_t> tuples5;  // 5-position suffix words (layers K >= 6)
};

bool build(int n, int m, int mask, Space& s) {
  int shift = 0;
  switch (mask) {
    case 0: s.me = m; s.multiset = 1; break;
    case 1: s.me = m; s.multiset = 0; break;
    case 2: s.me = m; s.multiset = 0; shift = 1; break;
    case 3: s.me = m; s.multiset = 0; shift = 2; break;
    case 4: s.me = m - 2; s.multiset = 1; break;
    default: return false;
  }
  s.lo.resize(s.me);
  for (int x = 0; x < s.me; ++x) {
    int v = s.multiset ? x : (((x >> shift) + 1) << shift);
    s.lo[x] = v > s.me ? s.me : v;
  }
  const int W = s.me + 1;
  s.cg.assign((size_t)(n + 2) * W, 0);
  for (int x = 0; x <= s.me; ++x) s.cg[x] = 1;
  for (int r = 1; r <= n + 1; ++r)
    for (int x = s.me - 1; x >= 0; --x) s.cg[(size_t)r * W + x] = s.cg[(size_t)r * W + x + 1] + s.cg[(size_t)(r - 1) * W + s.lo[x]];
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  if (n > slos::kSuf) {
    const uint64_t N = s.cg[(size_t)slos::kSuf * W];
    s.tuples.resize(N);
    for (uint64_t u = 0; u < N; ++u) s.tuples[u] = slos::make_tuple_word(sv, s.multiset, u);
  }
  if (n > 5) {
    const uint64_t N5 = s.cg[(size_t)5 * W];
    s.tuples5.resize(N5);
    for (uint64_t u = 0; u < N5; ++u) s.tuples5[u] = slos::make_tuple_word<5>(sv, s.multiset, u);
  }
  return true;
}

// thread-per-state formula (the simple kernel's slos::for_each_parent) for layers K <= SUF
void layer_simple(const Space& s, int K, const cd* ucol, const std::vector<cd>& par, std::vector<cd>& out) {
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
  const uint64_t S = sv.C(K, 0);
  out.assign(S, 0);
  for (uint64_t g = 0; g < S; ++g) {
    cd acc = 0;
    slos::for_each_parent(sv, s.multiset, K, g, [&](uint64_t pr, int mode, int) { acc += ucol[mode] * par[pr]; });
    out[g] = acc;
  }
}

// The device's suffix-table row x | mark << 6: the U coefficient (zero for marked rows) and the
// per-position rank terms packed as the kernel packs them (SufRow in layer_block.hpp), unpacked as
// it unpacks them: t[i] = cg[SUF-1-i][x+1] for i < SUF-2, t[SUF-2] = cg[1][x].
template <int SUF>
struct Row {
  cd c;
  uint32_t t[SUF - 1];
};
template <int SUF>
Row<SUF> suffix_row(const Space& s, const std::vector<cd>& ucol, int i) {
  const int W = s.me + 1;
  const int x = i & (slos::kBlockModes - 1);
  Row<SUF> r{};
  if (i < slos::kBlockModes && x < (int)ucol.size()) r.c = ucol[x];
  uint32_t t0 = 0, t1 = 0;
  if (x < s.me) {
    if constexpr (SUF == 4) {
      t0 = (uint32_t)s.cg[(size_t)3 * W + x + 1];
      t1 = (uint32_t)s.cg[(size_t)2 * W + x + 1] | ((uint32_t)s.cg[(size_t)W + x] << 16);
    } else {
      t0 = (uint32_t)s.cg[(size_t)4 * W + x + 1] | ((uint32_t)s.cg[(size_t)W + x] << 24);
      t1 = (uint32_t)s.cg[(size_t)3 * W + x + 1] | ((uint32_t)s.cg[(size_t)2 * W + x + 1] << 16);
    }
  }
  if constexpr (SUF == 4) {
    r.t[0] = t0;
    r.t[1] = t1 & 0xffffu;
    r.t[2] = t1 >> 16;
  } else {
    r.t[0] = t0 & 0xffffffu;
    r.t[1] = t1 & 0xffffu;
    r.t[2] = t1 >> 16;
    r.t[3] = t0 >> 24;
  }
  return r;
}

// local-parent ranks of one suffix word, through the kernel's functions
template <int SUF>
void local_ranks(uint32_t C0, const Row<SUF>* rw, uint32_t (&lr)[SUF]) {
  if constexpr (SUF == 4) {
    uint32_t a3[4], b2[4], b1[4];
    for (int j = 0; j < 4; ++j) {
      a3[j] = rw[j].t[0];
      b2[j] = rw[j].t[1];
      b1[j] = rw[j].t[2];
    }
    slos::suffix_parent_ranks(C0, a3, b2, b1, lr);
  } else {
    uint32_t t4[5], t3[5], t2[5], t1[5];
    for (int j = 0; j < 5; ++j) {
      t4[j] = rw[j].t[0];
      t3[j] = rw[j].t[1];
      t2[j] = rw[j].t[2];
      t1[j] = rw[j].t[3];
    }
    slos::suffix_parent_ranks5(C0, t4, t3, t2, t1, lr);
  }
}

// the block kernel's wave loop: batches of 64 consecutive prefixes (one per lane), then the blocks
// one at a time, 64 children per step; writes[c] counts how often child c is written
template <int MAXD, int SUF>
void layer_block(const Space& s, int K, const std::vector<cd>& ucol_in, const std::vector<cd>& par, std::vector<cd>& out,
                 std::vector<double>* norms, uint64_t per_wave, std::vector<int>& writes) {
  const void* tw = SUF == 4 ? (const void*)s.tuples.data() : (const void*)s.tuples5.data();
  slos::BlockTables t{s.cg.data(), s.lo.data(), tw, s.me, s.me + 1, s.multiset};
  const auto* words = static_cast<const slos::TupleWordT<SUF>*>(tw);
  const int W = s.me + 1;
  const uint64_t S = s.cg[(size_t)K * W];
  const int d = K - SUF;
  const uint64_t S_pre = s.cg[(size_t)d * W];
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  std::vector<Row<SUF>> rows(slos::kSufRows);
  for (int i = 0; i < slos::kSufRows; ++i) rows[i] = suffix_row<SUF>(s, ucol_in, i);
  const uint32_t C0 = (uint32_t)s.cg[(size_t)(SUF - 1) * W];
  out.assign(S, 0);
  writes.assign(S, 0);
  if (norms) norms->assign(S, 0.0);
  const uint64_t waves = (S + per_wave - 1) / per_wave;
  for (uint64_t w = 0; w < waves; ++w) {
    uint64_t g = w * per_wave;
    const uint64_t g_end = (g + per_wave < S) ? g + per_wave : S;
    int q0[MAXD];
    uint64_t u0 = slos::unrank_prefix<MAXD>(t, K, d, g, q0);
    uint64_t r = slos::prefix_rank<MAXD>(t, d, q0);
    while (g < g_end) {
      if (r >= S_pre) break;
      const int nb = (int)((S_pre - r < 64) ? S_pre - r : 64);
      for (int j = 0; j < nb && g < g_end; ++j) {
        int q[MAXD];
        sv.unrank(r + j, d, q);
        slos::BlockInfo bi;
        slos::Streams<MAXD> st;
        slos::block_info<MAXD, SUF>(t, K, d, q, bi, st);
        if (bi.L == 0) continue;
        const uint64_t avail = g_end - g;
        const uint32_t u_stop = (uint32_t)(((uint64_t)bi.L - u0 < avail) ? (uint64_t)bi.L : u0 + avail);
        for (uint32_t uc = (uint32_t)u0; uc < u_stop; uc += 64)
          for (int lane = 0; lane < 64; ++lane) {
            const uint32_t u = uc + lane;
            if (u >= u_stop) continue;
            cd acc = 0;
            for (int si = 0; si < MAXD; ++si)
              if (st.active[si]) acc += rows.at(st.mode[si]).c * par.at(st.base[si] + u);
            const slos::TupleWordT<SUF> w = words[bi.toff + u];
            Row<SUF> rw[SUF];
            uint32_t lr[SUF];
            cd coef[SUF];
            for (int jj = 0; jj < SUF; ++jj) {
              rw[jj] = rows.at(slos::tw_byte(w, jj));
              coef[jj] = rw[jj].c;
            }
            // tau's first copy of lo continuing the prefix's last run is no edge (edge 0 only)
            if (bi.skip0 && (int)slos::tw_byte(w, 0) == bi.lo) coef[0] = cd(0, 0);
            local_ranks<SUF>(C0, rw, lr);
            for (int jj = 0; jj < SUF; ++jj) acc += coef[jj] * par.at(bi.LBp + lr[jj]);
            const uint64_t c = bi.CB + u;
            out.at(c) = acc;
            writes.at(c)++;
            if (norms) (*norms)[c] = slos::child_norm<SUF>(bi, w, t.multiset);
          }
        g += u_stop - u0;
        u0 = 0;
      }
      r += 64;
    }
  }
}

// The last layer in group order (the device's k_layer_group): the work items of
// slos::group_items, and per item the blocks (a, R), a = 0..R_0, from slos::block_info. Checks the
// sharing the kernel relies on (every block of a group has the same first stream base, block size,
// tuple offset, lo and skip0) and returns false if it fails.
template <int MAXD>
bool layer_group(const Space& s, int K, const std::vector<cd>& ucol_in, const std::vector<cd>& par, std::vector<cd>& out,
                 std::vector<double>* norms, std::vector<int>& writes) {
  slos::BlockTables t{s.cg.data(), s.lo.data(), s.tuples5.data(), s.me, s.me + 1, s.multiset};
  const int W = s.me + 1;
  const uint64_t S = s.cg[(size_t)K * W];
  const int d = K - 5;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  std::vector<Row<5>> rows(slos::kSufRows);
  for (int i = 0; i < slos::kSufRows; ++i) rows[i] = suffix_row<5>(s, ucol_in, i);
  const uint32_t C0 = (uint32_t)s.cg[(size_t)4 * W];
  std::vector<uint32_t> items;
  slos::group_items(sv, K, items);
  out.assign(S, 0);
  writes.assign(S, 0);
  if (norms) norms->assign(S, 0.0);
  for (size_t it = 0; it < items.size(); it += 4) {
    int R[64];
    sv.unrank(items[it], d - 1, R);
    const uint32_t ub = items[it + 1], ue = items[it + 2];
    const int G = R[0] + 1;
    slos::BlockInfo b0{};
    uint64_t base0 = 0;
    for (int a = 0; a < G; ++a) {
      int q[MAXD];
      q[0] = a;
      for (int i = 1; i < MAXD; ++i) q[i] = i < d ? R[i - 1] : 0;
      slos::BlockInfo bi;
      slos::Streams<MAXD> st;
      slos::block_info<MAXD, 5>(t, K, d, q, bi, st);
      if (!st.active[0]) return false;
      if (a == 0) {
        b0 = bi;
        base0 = st.base[0];
      } else if (st.base[0] != base0 || bi.L != b0.L || bi.toff != b0.toff || bi.lo != b0.lo || bi.skip0 != b0.skip0) {
        return false;
      }
      if (ue > bi.L || ub >= ue) return false;
      for (uint32_t u = ub; u < ue; ++u) {
        cd acc = rows.at(st.mode[0]).c * par.at(base0 + u);
        for (int si = 1; si < MAXD; ++si)
          if (st.active[si]) acc += rows.at(st.mode[si]).c * par.at(st.base[si] + u);
        const slos::TupleWordT<5> w = s.tuples5.at(b0.toff + u);
        Row<5> rw[5];
        uint32_t lr[5];
        cd coef[5];
        for (int jj = 0; jj < 5; ++jj) {
          rw[jj] = rows.at(slos::tw_byte(w, jj));
          coef[jj] = rw[jj].c;
        }
        if (b0.skip0 && (int)slos::tw_byte(w, 0) == b0.lo) coef[0] = cd(0, 0);
        local_ranks<5>(C0, rw, lr);
        for (int jj = 0; jj < 5; ++jj) acc += coef[jj] * par.at(bi.LBp + lr[jj]);
        const uint64_t c = bi.CB + u;
        out.at(c) = acc;
        writes.at(c)++;
        if (norms) (*norms)[c] = slos::child_norm<5>(bi, w, t.multiset);
      }
    }
  }
  return true;
}

}  // namespace

extern "C" {

// Forward pass from the vacuum to layer n; probs[S_n] = |amp * in_norm|^2 * prod t_j!. suf 5 runs
// the layers with K - 5 in [1, 8] with 5-position suffixes (the device's choice where its table
// exists), suf 4 every block layer with 4, suf 6 as 5 but the layers of a multiset space that
// have a prefix of 2..6 positions (K = 7..11) in group order. Returns 0 on success, 2 if some child was not written
// exactly once, 3 if a group's blocks do not share what the group kernel shares, 1 on bad input.
int emu_forward_suf(int n, int m, int mask, const int* modes, const double* U, double in_norm, uint64_t per_wave,
                    int suf, double* probs) {
  Space s;
  if (!build(n, m, mask, s) || n > 32) return 1;
  const cd* Uc = reinterpret_cast<const cd*>(U);
  std::vector<cd> par(1, cd(1, 0)), out, ucol(m);
  std::vector<double> norms;
  std::vector<int> writes;
  for (int K = 1; K <= n; ++K) {
    for (int j = 0; j < m; ++j) ucol[j] = Uc[(size_t)j * m + modes[K - 1]];
    const bool last = K == n;
    if (K <= slos::kSuf) {
      layer_simple(s, K, ucol.data(), par, out);
      if (last) {
        slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
        norms.assign(out.size(), 1.0);
        int p[64];
        for (uint64_t g = 0; g < out.size(); ++g) {
          sv.unrank(g, K, p);
          norms[g] = slos::run_factorial(p, K, s.multiset);
        }
      }
    } else if (suf == 6 && s.multiset && K - 5 >= 2 && K - 5 <= 6) {
      const int d = K - 5;
      auto* nm = last ? &norms : nullptr;
      bool ok;
      if (d == 2) ok = layer_group<2>(s, K, ucol, par, out, nm, writes);
      else if (d == 3) ok = layer_group<3>(s, K, ucol, par, out, nm, writes);
      else if (d == 4) ok = layer_group<4>(s, K, ucol, par, out, nm, writes);
      else if (d == 5) ok = layer_group<5>(s, K, ucol, par, out, nm, writes);
      else ok = layer_group<6>(s, K, ucol, par, out, nm, writes);
      if (!ok) return 3;
      for (int w : writes)
        if (w != 1) return 2;
    } else if (suf >= 5 && K - 5 >= 1 && K - 5 <= 8) {
      const int d = K - 5;
      auto* nm = last ? &norms : nullptr;
      if (d <= 4) layer_block<4, 5>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 5) layer_block<5, 5>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 6) layer_block<6, 5>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 7) layer_block<7, 5>(s, K, ucol, par, out, nm, per_wave, writes);
      else layer_block<8, 5>(s, K, ucol, par, out, nm, per_wave, writes);
      for (int w : writes)
        if (w != 1) return 2;
    } else {
      const int d = K - slos::kSuf;
      // the stream widths the device launcher instantiates: exact for d = 5..8
      auto* nm = last ? &norms : nullptr;
      if (d <= 4) layer_block<4, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 5) layer_block<5, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 6) layer_block<6, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 7) layer_block<7, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      else if (d == 8) layer_block<8, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      else layer_block<16, 4>(s, K, ucol, par, out, nm, per_wave, writes);
      for (int w : writes)
        if (w != 1) return 2;
    }
    par.swap(out);
  }
  for (uint64_t g = 0; g < par.size(); ++g) probs[g] = std::norm(par[g] * in_norm) * norms[g];
  return 0;
}

int emu_forward(int n, int m, int mask, const int* modes, const double* U, double in_norm, uint64_t per_wave,
                double* probs) {
  return emu_forward_suf(n, m, mask, modes, U, in_norm, per_wave, 4, probs);
}

// q_{k-1}(t) = sum_j q_k(t + e_j) count / k via slos::for_each_child (the descent kernel's code)
int emu_descend(int n, int m, int mask, int k, const double* qk, double* out) {
  Space s;
  if (!build(n, m, mask, s) || k < 1 || k > n) return 1;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
  const uint64_t Sp = sv.C(k - 1, 0);
  for (uint64_t t = 0; t < Sp; ++t) {
    double acc = 0;
    slos::for_each_child(sv, s.multiset, k, t, [&](uint64_t c, int, int cnt) { acc += qk[c] * cnt; });
    out[t] = acc / k;
  }
  return 0;
}

// g_k(c) = sum_{j distinct in c} g_{k-1}(c - e_j) count / k via slos::for_each_parent
int emu_ascend(int n, int m, int mask, int k, const double* gp, double* gk) {
  Space s;
  if (!build(n, m, mask, s) || k < 1 || k > n) return 1;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
  const uint64_t Sk = sv.C(k, 0);
  for (uint64_t c = 0; c < Sk; ++c) {
    double acc = 0;
    slos::for_each_parent(sv, s.multiset, k, c, [&](uint64_t p, int, int cnt) { acc += gp[p] * cnt; });
    gk[c] = acc / k;
  }
  return 0;
}

// edges k-1 -> k enumerated from the parent side: (parent, child, mode, count) rows; returns count
int64_t emu_child_edges(int n, int m, int mask, int k, int64_t cap, int64_t* out) {
  Space s;
  if (!build(n, m, mask, s) || k < 1 || k > n) return -1;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
  const uint64_t Sp = sv.C(k - 1, 0);
  int64_t e = 0;
  for (uint64_t t = 0; t < Sp; ++t)
    slos::for_each_child(sv, s.multiset, k, t, [&](uint64_t c, int j, int cnt) {
      if (e < cap) { out[4 * e] = (int64_t)t; out[4 * e + 1] = (int64_t)c; out[4 * e + 2] = j; out[4 * e + 3] = cnt; }
      ++e;
    });
  return e;
}

}  // extern "C"

extern "C" {
// every child rank that slos::for_each_child reports for parents [first, first + count) of layer
// k-1 equals the rank of the child's own sorted positions and is < S_k; returns the number of
// mismatches (0: all correct), -1 on bad input
}  // extern "C"

// Every SUF-tuple of the space: the tuple word's positions and marks, the suffix-table rows (packed
// and unpacked as the kernel does), the SUF local-parent ranks from the per-position terms
// (slos::suffix_parent_ranks / suffix_parent_ranks5), and the run data from the mark code against
// direct computation (unrank, rank of tau minus one position, runs). Returns the number of
// mismatches, or -1 for bad input.
template <int SUF>
int64_t check_tuple_words(int m, int mask) {
  Space s;
  if (!build(SUF + 1, m, mask, s)) return -1;
  const int W = s.me + 1;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  const uint32_t C0 = (uint32_t)s.cg[(size_t)(SUF - 1) * W];
  const std::vector<cd> ucol(m, cd(1, 0));
  const uint64_t N = s.cg[(size_t)SUF * W];
  int64_t bad = 0;
  for (uint64_t u = 0; u < N; ++u) {
    int x[SUF];
    sv.unrank(u, SUF, x);
    const slos::TupleWordT<SUF> w = SUF == 4 ? (slos::TupleWordT<SUF>)s.tuples[u] : (slos::TupleWordT<SUF>)s.tuples5[u];
    Row<SUF> rw[SUF];
    uint32_t lr[SUF];
    for (int j = 0; j < SUF; ++j) {
      const uint32_t byte = slos::tw_byte(w, j);
      const int pos = (int)(byte & (slos::kBlockModes - 1));
      const bool mark = (byte & slos::kTupleMark) != 0;
      if (pos != x[j] || mark != (s.multiset && j > 0 && x[j] == x[j - 1]) || byte >= (uint32_t)slos::kSufRows) ++bad;
      rw[j] = suffix_row<SUF>(s, ucol, (int)byte);
      // a marked row has a zero coefficient and the same rank terms
      if ((std::abs(rw[j].c) != 0) == mark) ++bad;
    }
    if (SUF == 5 && (slos::tw_byte(w, 5) | slos::tw_byte(w, 6) | slos::tw_byte(w, 7))) ++bad;
    local_ranks<SUF>(C0, rw, lr);
    for (int j = 0; j < SUF; ++j) {
      int r[SUF - 1], c = 0;
      for (int i = 0; i < SUF; ++i)
        if (i != j) r[c++] = x[i];
      if ((uint64_t)lr[j] != sv.rank(r, SUF - 1)) ++bad;
    }
    int lead = 1, fct = 1, run = 1;
    if (s.multiset) {
      while (lead < SUF && x[lead] == x[0]) ++lead;
      for (int j = 1; j < SUF; ++j) {
        run = (x[j] == x[j - 1]) ? run + 1 : 1;
        fct *= run;
      }
    }
    const uint32_t code = slos::tuple_code<SUF>(w);
    if (s.multiset && (slos::code_lead<SUF>(code) != lead || slos::code_fct<SUF>(code) != fct)) ++bad;
    if (!s.multiset && code != 0) ++bad;
  }
  return bad;
}

extern "C" {

int64_t emu_check_tuple_words(int m, int mask) { return check_tuple_words<4>(m, mask); }
int64_t emu_check_tuple_words5(int m, int mask) { return check_tuple_words<5>(m, mask); }

// The block descent's child of every parent in [first, first + count) of layer k-1 and every mode
// (slos::descend_block_info, desc_lane_init, desc_tail_offset, desc_tail_copies, as the device
// kernel uses them) against slos::for_each_child: same child index, same copies, same order.
// Multiset spaces with k-1 >= SUF+1 only. Returns the number of mismatches, or -1 for bad input.
int64_t emu_check_descend_block(int n, int m, int mask, int k, uint64_t first, uint64_t count) {
  constexpr int MAXD = 16;
  Space s;
  if (!build(n, m, mask, s) || !s.multiset || k - 1 < slos::kSuf + 1 || k > n || k - 1 - slos::kSuf > MAXD) return -1;
  const int W = s.me + 1, D = k - 1, dp = D - slos::kSuf;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  slos::BlockTables t{s.cg.data(), s.lo.data(), s.tuples.data(), s.me, W, s.multiset};
  const uint64_t Sp = s.cg[(size_t)D * W], Sk = s.cg[(size_t)k * W];
  int64_t bad = 0;
  for (uint64_t g = first; g < first + count && g < Sp; ++g) {
    int q[MAXD];
    const uint64_t u = slos::unrank_prefix<MAXD>(t, D, dp, g, q);
    slos::DescBlock<MAXD> db;
    slos::descend_block_info<MAXD>(t, k, dp, q, db);
    if (db.CBp + u != g || u >= db.L) { ++bad; continue; }
    const slos::TupleWord w = s.tuples[db.toff + u];
    int sx[slos::kSuf];
    for (int j = 0; j < slos::kSuf; ++j) sx[j] = (int)((w >> (8 * j)) & 63u);
    slos::DescLane dl;
    slos::desc_lane_init(s.cg.data(), W, db.lastP, sx, dl);
    std::vector<std::pair<uint64_t, int>> mine, ref;
    int lo_i = 0;
    for (int i = 0; i < dp; ++i) {
      for (int a = lo_i; a < db.p[i]; ++a)
        mine.push_back({db.B[i] - s.cg[(size_t)(k - i) * W + a + 1] + u, (i > 0 && a == lo_i) ? db.run[i] + 1 : 1});
      lo_i = db.p[i];
    }
    const uint32_t c5l = (uint32_t)s.cg[(size_t)5 * W + db.lastP + 1];
    for (int a = db.lastP; a < s.me; ++a) {
      const uint32_t A[4] = {(uint32_t)s.cg[(size_t)4 * W + a + 1], (uint32_t)s.cg[(size_t)3 * W + a + 1],
                             (uint32_t)s.cg[(size_t)2 * W + a + 1], (uint32_t)s.cg[(size_t)W + a]};
      const uint32_t S5a = c5l - (uint32_t)s.cg[(size_t)5 * W + a + 1];
      mine.push_back({db.CB0 + slos::desc_tail_offset(dl, sx, (uint32_t)u, a, S5a, A),
                      slos::desc_tail_copies(sx, a, db.lastP, db.lastrun)});
    }
    slos::for_each_child(sv, s.multiset, k, g, [&](uint64_t r, int, int cnt) { ref.push_back({r, cnt}); });
    if (mine != ref) ++bad;
    for (auto& e : mine)
      if (e.first >= Sk) ++bad;
  }
  return bad;
}

// The block descent on a whole layer, in the device kernel's order of operations (float32 sums of
// q * copies in ascending mode order, then times 1/k): out[t] for every parent t of layer k-1.
int emu_descend_block(int n, int m, int mask, int k, const float* qk, float* out) {
  constexpr int MAXD = 16;
  Space s;
  if (!build(n, m, mask, s) || !s.multiset || k - 1 < slos::kSuf + 1 || k > n) return 1;
  const int W = s.me + 1, D = k - 1, dp = D - slos::kSuf;
  slos::BlockTables t{s.cg.data(), s.lo.data(), s.tuples.data(), s.me, W, s.multiset};
  const uint64_t Sp = s.cg[(size_t)D * W], Sk = s.cg[(size_t)k * W];
  const float invk = 1.f / (float)k;
  for (uint64_t g = 0; g < Sp; ++g) {
    int q[MAXD];
    const uint64_t u = slos::unrank_prefix<MAXD>(t, D, dp, g, q);
    slos::DescBlock<MAXD> db;
    slos::descend_block_info<MAXD>(t, k, dp, q, db);
    const slos::TupleWord w = s.tuples[db.toff + u];
    int sx[slos::kSuf];
    for (int j = 0; j < slos::kSuf; ++j) sx[j] = (int)((w >> (8 * j)) & 63u);
    slos::DescLane dl;
    slos::desc_lane_init(s.cg.data(), W, db.lastP, sx, dl);
    float acc = 0.f;
    int lo_i = 0;
    for (int i = 0; i < dp; ++i) {
      for (int a = lo_i; a < db.p[i]; ++a) {
        const uint64_t c = db.B[i] - s.cg[(size_t)(k - i) * W + a + 1] + u;
        if (c >= Sk) return 2;
        acc = std::fma(qk[c], (float)((i > 0 && a == lo_i) ? db.run[i] + 1 : 1), acc);
      }
      lo_i = db.p[i];
    }
    const uint32_t c5l = (uint32_t)s.cg[(size_t)5 * W + db.lastP + 1];
    for (int a = db.lastP; a < s.me; ++a) {
      const uint32_t A[4] = {(uint32_t)s.cg[(size_t)4 * W + a + 1], (uint32_t)s.cg[(size_t)3 * W + a + 1],
                             (uint32_t)s.cg[(size_t)2 * W + a + 1], (uint32_t)s.cg[(size_t)W + a]};
      const uint64_t c = db.CB0 + slos::desc_tail_offset(dl, sx, (uint32_t)u, a, c5l - (uint32_t)s.cg[(size_t)5 * W + a + 1], A);
      if (c >= Sk) return 2;
      acc = std::fma(qk[c], (float)slos::desc_tail_copies(sx, a, db.lastP, db.lastrun), acc);
    }
    out[g] = acc * invk;
  }
  return 0;
}

int64_t emu_check_children(int n, int m, int mask, int k, uint64_t first, uint64_t count) {
  Space s;
  if (!build(n, m, mask, s) || k < 1 || k > n) return -1;
  const int W = s.me + 1;
  slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, W};
  const uint64_t Sp = s.cg[(size_t)(k - 1) * W], Sk = s.cg[(size_t)k * W];
  int64_t bad = 0;
  for (uint64_t t = first; t < first + count && t < Sp; ++t) {
    int p[64], c[64];
    if (k > 1) sv.unrank(t, k - 1, p);
    slos::for_each_child(sv, s.multiset, k, t, [&](uint64_t r, int j, int) {
      int q = 0, placed = 0;
      for (int i = 0; i < k - 1; ++i) {
        if (!placed && j < p[i]) { c[q++] = j; placed = 1; }
        c[q++] = p[i];
      }
      if (!placed) c[q++] = j;
      if (r >= Sk || sv.rank(c, k) != r) ++bad;
    });
  }
  return bad;
}
}
