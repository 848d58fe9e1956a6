// CPU pull-form restatement of the SLOS layer recurrence — TEST INFRASTRUCTURE ONLY (the oracle).
//
// Restates photonic_engines/slos/kernel.py:201-283 (recurrence, deferred normalisation, readout)
// without stored edge tables: every child computes its parents' lex ranks from its own sorted
// photon positions (the parent rule is basis.py:166-189: remove one photon of a distinct mode).
// Order: the lex order of sorted position tuples produced by basis.py:120-148.
// Also: Ryser permanents (reference.py:24-36) with Gray-code row-sum updates, batched spot checks.
//
// Counting table: cg[r][x] = number of admissible sorted r-tuples whose first element is >= x
// (x in [0, me]); lo[x] = smallest admissible position after a photon at x (x itself for the
// multiset masks fock / tail_empty, the start of the next block for unbunched / dual_rail /
// sparse_1in4). Built by the recursion cg[r][x] = cg[r][x+1] + cg[r-1][lo[x]].
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <omp.h>

namespace {

struct Space {
  int n = 0, m = 0, me = 0, multiset = 1;
  std::vector<uint64_t> cg;  // (n+2) x (me+1)
  std::vector<int> lo;       // me
  uint64_t C(int r, int x) const { return cg[(size_t)r * (me + 1) + x]; }
};

// mask ids: 0 fock, 1 unbunched, 2 dual_rail, 3 sparse_1in4, 4 tail_empty
bool build_space(int n, int m, int mask, Space& s) {
  s.n = n; s.m = m;
  int shift = 0;
  switch (mask) {
    case 0: s.me = m; s.multiset = 1; break;
    case 1: s.me = m; s.multiset = 0; shift = 0; break;
    case 2: s.me = m; s.multiset = 0; shift = 1; break;
    case 3: s.me = m; s.multiset = 0; shift = 2; break;
    case 4: s.me = m - 2; s.multiset = 1; break;
    default: return false;
  }
  if (s.me <= 0) return false;
  s.lo.resize(s.me);
  for (int x = 0; x < s.me; ++x) {
    int v = s.multiset ? x : (((x >> shift) + 1) << shift);
    s.lo[x] = v > s.me ? s.me : v;
  }
  s.cg.assign((size_t)(n + 2) * (s.me + 1), 0);
  for (int x = 0; x <= s.me; ++x) s.cg[x] = 1;
  for (int r = 1; r <= n + 1; ++r)
    for (int x = s.me - 1; x >= 0; --x)
      s.cg[(size_t)r * (s.me + 1) + x] = s.C(r, x + 1) + s.C(r - 1, s.lo[x]);
  return true;
}

uint64_t rank_of(const Space& s, const int* p, int K) {
  uint64_t r = 0;
  int lo = 0;
  for (int i = 0; i < K; ++i) {
    r += s.C(K - i, lo) - s.C(K - i, p[i]);
    lo = s.lo[p[i]];
  }
  return r;
}

void unrank_to(const Space& s, uint64_t g, int K, int* p) {
  int lo = 0;
  for (int i = 0; i < K; ++i) {
    int rem = K - i;
    uint64_t base = s.C(rem, lo);
    int x = lo;
    // first element x: states with first element in [lo, x) number base - C(rem, x)
    while (base - s.C(rem, x + 1) <= g) ++x;
    g -= base - s.C(rem, x);
    p[i] = x;
    lo = s.lo[x];
  }
}

// next tuple in lex order; returns false after the last one
bool successor(const Space& s, int* p, int K) {
  for (int i = K - 1; i >= 0; --i) {
    int rem = K - i;
    for (int x = p[i] + 1; x < s.me; ++x) {
      if (s.C(rem - 1, s.lo[x]) == 0) continue;   // no completion starts at x
      p[i] = x;
      int lo = s.lo[x];
      for (int j = i + 1; j < K; ++j) {
        int y = lo;
        while (s.C(K - j - 1, s.lo[y]) == 0) ++y;
        p[j] = y;
        lo = s.lo[y];
      }
      return true;
    }
  }
  return false;
}

template <class T>
void layer_pull(const Space& s, int K, const std::complex<T>* ucol, const std::complex<T>* par,
                std::complex<T>* chi, uint64_t S) {
  const uint64_t chunk = 4096;
  const uint64_t nchunks = (S + chunk - 1) / chunk;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t c = 0; c < (int64_t)nchunks; ++c) {
    int p[64];
    uint64_t g0 = (uint64_t)c * chunk, g1 = std::min<uint64_t>(S, g0 + chunk);
    unrank_to(s, g0, K, p);
    uint64_t Tt[65], A[65], M[65], suf[66];
    int lo[65];
    for (uint64_t g = g0; g < g1; ++g) {
      lo[0] = 0;
      for (int i = 0; i < K; ++i) lo[i + 1] = s.lo[p[i]];
      for (int i = 0; i < K; ++i) {
        Tt[i] = s.C(K - i, lo[i]) - s.C(K - i, p[i]);
        A[i] = (i < K - 1) ? s.C(K - i - 1, lo[i]) - s.C(K - i - 1, p[i]) : 0;
        M[i] = (i < K - 1) ? s.C(K - i - 1, lo[i]) - s.C(K - i - 1, p[i + 1]) : 0;
      }
      suf[K] = 0; suf[K + 1] = 0;
      for (int i = K - 1; i >= 0; --i) suf[i] = suf[i + 1] + Tt[i];
      std::complex<T> acc(0, 0);
      uint64_t pre = 0;
      for (int l = 0; l < K; ++l) {
        bool first = !s.multiset || l == 0 || p[l] != p[l - 1];
        if (first) {
          uint64_t pr = (l == K - 1) ? pre : pre + M[l] + suf[l + 2];
          acc += ucol[p[l]] * par[pr];
        }
        pre += A[l];
      }
      chi[g] = acc;
      successor(s, p, K);
    }
  }
}

template <class T>
int forward_impl(int n, int m, int mask, const int* modes, const double* U /*[m][m][2]*/,
                 double in_norm, std::complex<T>* amps_out, double* probs_out) {
  Space s;
  if (!build_space(n, m, mask, s)) return 1;
  std::vector<std::complex<T>> a(1, std::complex<T>(1, 0)), b;
  std::vector<std::complex<T>> ucol(m);
  for (int k = 1; k <= n; ++k) {
    uint64_t S = s.C(k, 0);
    int col = modes[k - 1];
    for (int j = 0; j < m; ++j)
      ucol[j] = std::complex<T>((T)U[2 * ((size_t)j * m + col)], (T)U[2 * ((size_t)j * m + col) + 1]);
    b.assign(S, std::complex<T>(0, 0));
    layer_pull<T>(s, k, ucol.data(), a.data(), b.data(), S);
    a.swap(b);
  }
  uint64_t S = s.C(n, 0);
  if (amps_out)
    for (uint64_t g = 0; g < S; ++g) amps_out[g] = a[g] * (T)in_norm;
  if (probs_out) {
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)((S + 4095) / 4096); ++c) {
      int p[64];
      uint64_t g0 = (uint64_t)c * 4096, g1 = std::min<uint64_t>(S, g0 + 4096);
      unrank_to(s, g0, n, p);
      for (uint64_t g = g0; g < g1; ++g) {
        double f = 1, run = 1;
        for (int i = 1; i < n; ++i) {
          run = (p[i] == p[i - 1]) ? run + 1 : 1;
          f *= run;
        }
        std::complex<T> v = a[g] * (T)in_norm;
        T pr = v.real() * v.real() + v.imag() * v.imag();
        probs_out[g] = (double)(s.multiset ? pr * (T)f : pr);
        successor(s, p, n);
      }
    }
  }
  return 0;
}

std::complex<double> ryser(int n, const std::complex<double>* A /*row-major n x n*/) {
  if (n == 0) return 1.0;
  std::vector<std::complex<double>> rs(n, 0.0);
  std::complex<double> total = 0.0;
  uint64_t gray_prev = 0;
  for (uint64_t k = 1; k < (1ull << n); ++k) {
    uint64_t gray = k ^ (k >> 1);
    uint64_t diff = gray ^ gray_prev;
    int col = __builtin_ctzll(diff);
    double sg = (gray & diff) ? 1.0 : -1.0;
    for (int i = 0; i < n; ++i) rs[i] += sg * A[(size_t)i * n + col];
    gray_prev = gray;
    std::complex<double> prod = 1.0;
    for (int i = 0; i < n; ++i) prod *= rs[i];
    int pc = __builtin_popcountll(gray);
    total += ((n - pc) % 2 == 0) ? prod : -prod;
  }
  return total;
}

// timed sample of the forward pass: layers 1..n-1 in full, then the first `last_count` states of
// layer n (probabilities discarded after a checksum); t[0] = prefix seconds, t[1] = last-layer seconds
template <class T>
int sample_impl(int n, int m, int mask, const int* modes, const double* U, uint64_t last_count, double* t,
                double* checksum) {
  Space s;
  if (!build_space(n, m, mask, s)) return 1;
  std::vector<std::complex<T>> a(1, std::complex<T>(1, 0)), b;
  std::vector<std::complex<T>> ucol(m);
  auto load_col = [&](int k) {
    const int col = modes[k - 1];
    for (int j = 0; j < m; ++j)
      ucol[j] = std::complex<T>((T)U[2 * ((size_t)j * m + col)], (T)U[2 * ((size_t)j * m + col) + 1]);
  };
  double t0 = omp_get_wtime();
  for (int k = 1; k < n; ++k) {
    load_col(k);
    b.assign(s.C(k, 0), std::complex<T>(0, 0));
    layer_pull<T>(s, k, ucol.data(), a.data(), b.data(), s.C(k, 0));
    a.swap(b);
  }
  double t1 = omp_get_wtime();
  load_col(n);
  const uint64_t cnt = std::min<uint64_t>(last_count, s.C(n, 0));
  b.assign(cnt, std::complex<T>(0, 0));
  layer_pull<T>(s, n, ucol.data(), a.data(), b.data(), cnt);
  double t2 = omp_get_wtime();
  double cs = 0;
  for (uint64_t g = 0; g < cnt; ++g) cs += (double)std::norm(b[g]);
  t[0] = t1 - t0;
  t[1] = t2 - t1;
  *checksum = cs;
  return 0;
}

}  // namespace

extern "C" {

int oracle_forward_sample(int n, int m, int mask, const int* modes, const double* U, int prec, uint64_t last_count,
                          int threads, double* t /*[2]*/, double* checksum) {
  if (threads > 0) omp_set_num_threads(threads);
  if (prec == 64) return sample_impl<float>(n, m, mask, modes, U, last_count, t, checksum);
  return sample_impl<double>(n, m, mask, modes, U, last_count, t, checksum);
}

int oracle_layer_size(int n, int m, int mask, int k, uint64_t* out) {
  Space s;
  if (!build_space(n, m, mask, s) || k < 0 || k > n) return 1;
  *out = s.C(k, 0);
  return 0;
}

int oracle_rank(int n, int m, int mask, int k, const int* pos, int64_t count, uint64_t* out) {
  Space s;
  if (!build_space(n, m, mask, s)) return 1;
  for (int64_t i = 0; i < count; ++i) out[i] = rank_of(s, pos + i * k, k);
  return 0;
}

int oracle_unrank(int n, int m, int mask, int k, const uint64_t* idx, int64_t count, int* pos) {
  Space s;
  if (!build_space(n, m, mask, s)) return 1;
  for (int64_t i = 0; i < count; ++i) unrank_to(s, idx[i], k, pos + i * k);
  return 0;
}

// probs (double, [S]) of the complex128 or complex64 pull recurrence; prec 64 or 128
int oracle_forward_probs(int n, int m, int mask, const int* modes, const double* U, double in_norm,
                         int prec, double* probs, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  if (prec == 64) return forward_impl<float>(n, m, mask, modes, U, in_norm, nullptr, probs);
  return forward_impl<double>(n, m, mask, modes, U, in_norm, nullptr, probs);
}

int oracle_forward_amps(int n, int m, int mask, const int* modes, const double* U, double in_norm,
                        double* amps /*[S][2]*/, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
  return forward_impl<double>(n, m, mask, modes, U, in_norm,
                              reinterpret_cast<std::complex<double>*>(amps), nullptr);
}

int oracle_permanent(int n, const double* A /*[n][n][2]*/, double* out /*[2]*/) {
  std::complex<double> v = ryser(n, reinterpret_cast<const std::complex<double>*>(A));
  out[0] = v.real(); out[1] = v.imag();
  return 0;
}

// p(t) = |per(U[t rows][:, s cols])|^2 / (prod s! prod t!) for output lex indices idx[]
int oracle_spot_probs(int n, int m, int mask, const int* modes, const double* U, double s_fact,
                      const uint64_t* idx, int64_t count, double* out, int threads) {
  Space s;
  if (!build_space(n, m, mask, s)) return 1;
  if (threads > 0) omp_set_num_threads(threads);
  const std::complex<double>* Uc = reinterpret_cast<const std::complex<double>*>(U);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < count; ++q) {
    int p[64];
    unrank_to(s, idx[q], n, p);
    std::vector<std::complex<double>> A((size_t)n * n);
    double tf = 1, run = 1;
    for (int i = 0; i < n; ++i) {
      if (i > 0) { run = (p[i] == p[i - 1]) ? run + 1 : 1; tf *= run; }
      for (int c = 0; c < n; ++c) A[(size_t)i * n + c] = Uc[(size_t)p[i] * m + modes[c]];
    }
    std::complex<double> v = ryser(n, A.data());
    out[q] = std::norm(v) / (s_fact * tf);
  }
  return 0;
}

}  // extern "C"
