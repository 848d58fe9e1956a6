// Standalone driver of the CPU emulation (block_emu.cpp, which compiles the product's index headers)
// built with -fsanitize=address,undefined: the forward decompositions (4- and 5-position suffix
// blocks, group order), the descent's child indices, and the sampling kernels' index logic, each
// at the sizes the GPU tests use. Test infrastructure; exits non-zero on the first failure.
//
// The sampling kernels (k_sample_block_sums, k_sample_seg_prefix, k_sample_seg_offsets,
// k_sample_add_offsets, k_sample_search, k_unrank_occ in
// photonic_engines_amd/csrc/slos.hip) are restated here lane by lane, with every global-memory
// index checked against its buffer's size, at the launch shapes slos_sample and
// slos_unrank_device use.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "block_emu.cpp"

namespace {

int g_fail = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what.c_str());
    ++g_fail;
  }
}

template <class T>
struct Checked {  // a device buffer: every access bounds-checked
  std::vector<T> v;
  const char* name;
  explicit Checked(size_t n, const char* nm) : v(n), name(nm) {}
  T& operator[](uint64_t i) {
    if (i >= v.size()) throw std::out_of_range(std::string(name) + " index " + std::to_string(i));
    return v[i];
  }
};

std::vector<std::complex<double>> haar_like(int m, unsigned seed) {
  // any unitary-ish dense matrix works for the index checks: Gaussian entries
  std::mt19937_64 r(seed);
  std::normal_distribution<double> nd(0.0, 1.0 / std::sqrt(2.0 * m));
  std::vector<std::complex<double>> U((size_t)m * m);
  for (auto& z : U) z = {nd(r), nd(r)};
  return U;
}

// forward at (n, m, fock, first-n modes): every block decomposition against the 4-suffix one
void check_forward(int n, int m) {
  std::vector<int> modes(n);
  for (int i = 0; i < n; ++i) modes[i] = i;
  const auto U = haar_like(m, 1000u * n + m);
  const uint64_t S = [&] {
    Space s;
    build(n, m, 0, s);
    slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
    return sv.C(n, 0);
  }();
  std::vector<double> ref(S), out(S);
  int rc = emu_forward_suf(n, m, 0, modes.data(), reinterpret_cast<const double*>(U.data()), 1.0, 256, 4, ref.data());
  expect(rc == 0, "forward suf 4 n=" + std::to_string(n) + " m=" + std::to_string(m));
  for (int suf : {5, 6})
    for (uint64_t pw : {64ull, 1000ull, 4096ull}) {
      rc = emu_forward_suf(n, m, 0, modes.data(), reinterpret_cast<const double*>(U.data()), 1.0, pw, suf, out.data());
      expect(rc == 0, "forward suf " + std::to_string(suf) + " rc " + std::to_string(rc));
      double worst = 0;
      for (uint64_t g = 0; g < S; ++g) worst = std::max(worst, std::abs(out[g] - ref[g]) / (1e-30 + std::abs(ref[g])));
      expect(worst < 1e-9, "forward suf " + std::to_string(suf) + " differs from suf 4 by " + std::to_string(worst));
    }
}

// slos_sample's kernels over p[S] for `shots` uniforms, restated lane by lane
std::vector<uint64_t> sample_emulated(const std::vector<float>& pin, const std::vector<double>& uni) {
  constexpr int kChunk = 64, kBlock = kChunk * 64, kSeg = 1024, kSumWaves = 4;
  const uint64_t S = pin.size();
  const uint64_t nb = (S + kBlock - 1) / kBlock;
  if (nb > (uint64_t)kSeg * 1024) throw std::runtime_error("too many categories");
  Checked<float> p(S, "probs");
  for (uint64_t i = 0; i < S; ++i) p[i] = pin[i];
  Checked<double> bsum(nb, "bsum"), cdf(nb, "cdf"), csum(nb * 64, "csum"), segtot(1024, "segtot");
  // k_sample_block_sums: grid ceil(nb / kSumWaves) x (64 kSumWaves) threads, one wave per block,
  // lane c sums chunk c
  const uint64_t grid1 = (nb + kSumWaves - 1) / kSumWaves;
  for (uint64_t bx = 0; bx < grid1; ++bx)
    for (int w = 0; w < kSumWaves; ++w) {
      const uint64_t b = bx * kSumWaves + (uint64_t)w;
      if (b >= nb) continue;
      double cs[64];
      for (int lane = 0; lane < 64; ++lane) {
        const uint64_t i0 = b * kBlock + (uint64_t)lane * kChunk;
        double acc = 0.0;
        if (i0 + kChunk <= S) {
          for (int j = 0; j < kChunk / 4; ++j)
            for (int e = 0; e < 4; ++e) acc += (double)p[i0 + 4 * j + e];
        } else {
          for (int k = 0; k < kChunk; ++k) acc += (double)(i0 + k < S ? p[i0 + k] : 0.f);
        }
        cs[lane] = acc;
        csum[b * 64 + (uint64_t)lane] = acc;
      }
      double tot = 0.0;
      for (int c = 0; c < 64; ++c) tot += cs[c];
      bsum[b] = tot;
    }
  // k_sample_seg_prefix: one work group per segment of kSeg blocks
  const uint64_t nseg = (nb + kSeg - 1) / kSeg;
  for (uint64_t g = 0; g < nseg; ++g) {
    const uint64_t a = g * kSeg;
    const int len = (int)(nb - a < (uint64_t)kSeg ? nb - a : (uint64_t)kSeg);
    std::vector<double> seg((size_t)kSeg, 0.0);
    for (int i = 0; i < len; ++i) seg.at((size_t)i) = bsum[a + (uint64_t)i];
    double acc = 0.0;
    for (int i = 0; i < len; ++i) {
      acc += seg.at((size_t)i);
      seg.at((size_t)i) = acc;
    }
    segtot[g] = acc;
    for (int i = 0; i < len; ++i) cdf[a + (uint64_t)i] = seg.at((size_t)i);
  }
  // k_sample_seg_offsets: one thread, in order
  double run = 0.0;
  for (uint64_t i = 0; i < nseg; ++i) {
    const double v = segtot[i];
    segtot[i] = run;
    run += v;
  }
  // k_sample_add_offsets: grid-stride over the blocks
  for (uint64_t i = 0; i < nb; ++i) cdf[i] = segtot[i / kSeg] + cdf[i];
  // k_sample_search: ceil(shots / 256) work groups of 256 threads, one thread per shot
  const int64_t shots = (int64_t)uni.size();
  Checked<double> u_d(uni.size(), "uniforms");
  for (size_t i = 0; i < uni.size(); ++i) u_d[i] = uni[i];
  Checked<uint64_t> idx(uni.size(), "idx");
  const int64_t grid3 = (shots + 255) / 256;
  for (int64_t bx = 0; bx < grid3; ++bx)
    for (int t = 0; t < 256; ++t) {
      const int64_t s = bx * 256 + t;
      if (s >= shots) continue;
      const double total = cdf[nb - 1];
      const double u = u_d[(uint64_t)s] * total;
      uint64_t lo = 0, hi = nb - 1;
      while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] >= u) hi = mid;
        else lo = mid + 1;
      }
      const uint64_t b = lo;
      double r = u - (b > 0 ? cdf[b - 1] : 0.0);
      int c = 63;
      double pre = 0.0;
      for (int q = 0; q < 64; ++q) {
        const double v = csum[b * 64 + (uint64_t)q];
        if (pre + v >= r) {
          c = q;
          break;
        }
        pre += v;
      }
      r -= pre;
      const uint64_t first = b * kBlock + (uint64_t)c * kChunk;
      int k = 63;
      double acc = 0.0;
      for (int q = 0; q < kChunk; ++q) {
        acc += (double)(first + (uint64_t)q < S ? p[first + (uint64_t)q] : 0.f);
        if (acc >= r) {
          k = q;
          break;
        }
      }
      uint64_t o = first + (uint64_t)k;
      if (o >= S) o = S - 1;
      idx[(uint64_t)s] = o;
    }
  return idx.v;
}

void check_sampling(uint64_t S, int64_t shots, unsigned seed, bool zero_runs) {
  std::mt19937_64 r(seed);
  std::uniform_real_distribution<double> ud(0.0, 1.0);
  std::vector<float> p(S);
  for (auto& x : p) {
    const double v = ud(r);
    x = ud(r) < 0.25 ? 0.f : (float)(v * v * v);
  }
  if (zero_runs && S > 10 * 4096)
    for (uint64_t i = 2 * 4096; i < 7 * 4096; ++i) p[i] = 0.f;
  p[S - 1] = 1e-3f;
  std::vector<double> uni((size_t)shots);
  for (auto& u : uni) u = ud(r);
  try {
    const auto idx = sample_emulated(p, uni);
    uint64_t bad = 0;
    for (auto i : idx) bad += (i >= S || p[i] <= 0.f);
    expect(bad == 0, "sampling S=" + std::to_string(S) + ": " + std::to_string(bad) + " draws of zero weight");
  } catch (const std::exception& e) {
    expect(false, std::string("sampling S=") + std::to_string(S) + ": " + e.what());
  }
}

// k_unrank_occ at slos_unrank_device's launch shape: grid min(ceil(count/256), 65536) x 256,
// grid-stride loop; rows of indices >= S_k are -1
void check_unrank(int n, int m, int k, const std::vector<uint64_t>& idx) {
  Space s;
  build(n, m, 0, s);
  const int stride = s.me + 1;
  Checked<uint64_t> cg_g(s.cg.size(), "cg");
  for (size_t i = 0; i < s.cg.size(); ++i) cg_g[i] = s.cg[i];
  const int ncg = (k + 1) * stride;
  std::vector<uint64_t> cg_l((size_t)ncg);
  for (int i = 0; i < ncg; ++i) cg_l[(size_t)i] = cg_g[(uint64_t)i];  // the LDS copy
  const slos::SpaceView sv{cg_l.data(), s.lo.data(), s.me, stride};
  const int64_t count = (int64_t)idx.size();
  Checked<uint64_t> id(idx.size(), "idx");
  for (size_t i = 0; i < idx.size(); ++i) id[i] = idx[i];
  Checked<int32_t> occ((size_t)count * m, "occ");
  const uint64_t blocks = std::min<uint64_t>(((uint64_t)count + 255) / 256, 1u << 16);
  for (uint64_t bx = 0; bx < blocks; ++bx)
    for (int tx = 0; tx < 256; ++tx)
      for (int64_t q = (int64_t)(bx * 256 + tx); q < count; q += (int64_t)(blocks * 256)) {
        const uint64_t g = id[(uint64_t)q];
        const uint64_t row = (uint64_t)q * (uint64_t)m;
        if (g >= cg_l[(size_t)k * stride]) {
          for (int j = 0; j < m; ++j) occ[row + j] = -1;
          continue;
        }
        int pos[32];
        sv.unrank(g, k, pos);
        for (int j = 0; j < m; ++j) occ[row + j] = 0;
        for (int i = 0; i < k; ++i) occ[row + (uint64_t)pos[i]] += 1;
      }
  const uint64_t Sk = sv.C(k, 0);
  for (int64_t q = 0; q < count; ++q) {
    int tot = 0;
    for (int j = 0; j < m; ++j) tot += occ[(uint64_t)q * m + j];
    expect(idx[(size_t)q] >= Sk ? tot == -m : tot == k, "unrank row sum");
  }
}

}  // namespace

int main() {
  try {
    // forward decompositions, including group order (n >= 7) and the C3 / test shapes
    for (auto nm : {std::pair<int, int>{6, 10}, {7, 9}, {8, 12}, {9, 11}, {10, 12}, {8, 20}})
      check_forward(nm.first, nm.second);
    // descent child indices at n = 10, m = 20 (every block step), start / middle / end
    for (int k = 6; k <= 10; ++k) {
      Space s;
      build(10, 20, 0, s);
      slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
      const uint64_t Sp = sv.C(k - 1, 0);
      for (uint64_t first : {uint64_t(0), Sp / 2, Sp > 50000 ? Sp - 50000 : uint64_t(0)})
        expect(emu_check_descend_block(10, 20, 0, k, first, 50000) == 0, "descent indices k=" + std::to_string(k));
    }
    // suffix-tuple words of the sizes in use
    for (int m : {12, 20, 24, 32}) {
      expect(emu_check_tuple_words(m, 0) == 0, "tuple words m=" + std::to_string(m));
      expect(emu_check_tuple_words5(m, 0) == 0, "tuple words5 m=" + std::to_string(m));
    }
    // the sampler at the sizes of the GPU tests: 1, 4095, 4096, 12.6M, and the 30,045,015 entries of
    // a lossy n = 10, m = 20 distribution over layers 0..10
    check_sampling(1, 1000, 1, false);
    check_sampling(4095, 20000, 2, false);
    check_sampling(4096, 20000, 3, false);
    check_sampling(3ull * 4096 * 1024 + 777, 20000, 4, true);
    check_sampling(30045015ull, 100000, 5, false);
    // device unrank of every layer of n = 10, m = 20, in and out of range
    std::mt19937_64 r(7);
    for (int k = 0; k <= 10; ++k) {
      Space s;
      build(10, 20, 0, s);
      slos::SpaceView sv{s.cg.data(), s.lo.data(), s.me, s.me + 1};
      const uint64_t Sk = sv.C(k, 0);
      std::vector<uint64_t> idx(20000);
      for (auto& v : idx) v = r() % (Sk + 3);
      check_unrank(10, 20, k, idx);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FAIL: exception %s\n", e.what());
    return 1;
  }
  if (g_fail) return 1;
  std::printf("emu_sanitize: all checks passed\n");
  return 0;
}
