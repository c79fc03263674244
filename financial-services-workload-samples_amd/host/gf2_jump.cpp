// GF(2) polynomial arithmetic for jump-ahead (host side): Berlekamp-Massey
// for the minimal polynomials, carry-less multiplication (PCLMULQDQ where
// the host has it), Barrett reduction modulo phi, and x^J mod phi.
#include "gf2_jump.hpp"

#include <immintrin.h>

#include <algorithm>
#include <cstring>
#include <mutex>

namespace fsmc {
namespace gf2 {
namespace {

using Words = std::vector<uint64_t>;

inline bool bit(const Words& a, long long i) { return (a[(size_t)(i >> 6)] >> (i & 63)) & 1u; }
inline void flip(Words& a, long long i) { a[(size_t)(i >> 6)] ^= 1ull << (i & 63); }

// 64 x 64 -> 128 carry-less product.
__attribute__((target("pclmul"))) inline void clmul_hw(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  const __m128i p = _mm_clmulepi64_si128(_mm_set_epi64x(0, (long long)a), _mm_set_epi64x(0, (long long)b), 0);
  lo = (uint64_t)_mm_cvtsi128_si64(p);
  hi = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(p, p));
}
inline void clmul_sw(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  uint64_t l = 0, h = 0;
  for (int i = 0; i < 64; ++i)
    if ((b >> i) & 1u) {
      l ^= a << i;
      h ^= i ? a >> (64 - i) : 0;
    }
  lo = l;
  hi = h;
}
bool has_pclmul() {
  static const bool v = __builtin_cpu_supports("pclmul");
  return v;
}

// r = a * b (words), r sized a.size() + b.size().
__attribute__((target("pclmul"))) void mul_hw(const uint64_t* a, size_t na, const uint64_t* b, size_t nb,
                                              uint64_t* r) {
  std::memset(r, 0, sizeof(uint64_t) * (na + nb));
  for (size_t i = 0; i < na; ++i) {
    if (!a[i]) continue;
    const __m128i ai = _mm_set_epi64x(0, (long long)a[i]);
    for (size_t j = 0; j < nb; ++j) {
      const __m128i p = _mm_clmulepi64_si128(ai, _mm_set_epi64x(0, (long long)b[j]), 0);
      r[i + j] ^= (uint64_t)_mm_cvtsi128_si64(p);
      r[i + j + 1] ^= (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(p, p));
    }
  }
}
void mul(const Words& a, const Words& b, Words& r) {
  r.assign(a.size() + b.size(), 0);
  if (has_pclmul()) {
    mul_hw(a.data(), a.size(), b.data(), b.size(), r.data());
    return;
  }
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < b.size(); ++j) {
      uint64_t lo, hi;
      clmul_sw(a[i], b[j], lo, hi);
      r[i + j] ^= lo;
      r[i + j + 1] ^= hi;
    }
  (void)clmul_hw;
}

// a >> s (bits), keeping `words` words.
Words shr(const Words& a, long long s, size_t words) {
  Words r(words, 0);
  const size_t q = (size_t)(s >> 6);
  const int b = (int)(s & 63);
  for (size_t i = 0; i < words; ++i) {
    const size_t k = i + q;
    uint64_t v = k < a.size() ? a[k] >> b : 0;
    if (b && k + 1 < a.size()) v |= a[k + 1] << (64 - b);
    r[i] = v;
  }
  return r;
}

// The generator's minimal polynomial and its Barrett constant
// mu = floor(x^(2n) / phi), n = the degree of phi.
struct Modulus {
  long long n;  // degree of phi
  Words phi;    // n + 1 coefficients in kWords + 1 words
  Words mu;     // floor(x^(2n) / phi): n + 1 coefficients
};
inline uint64_t top_mask(long long n) { return (n % 64) ? ((1ull << (n % 64)) - 1) : ~0ull; }

// Berlekamp-Massey over GF(2) on the bit sequence s[0..N): the shortest
// connection polynomial C (C_0 = 1) with s_n = sum_{i=1..L} C_i s_{n-i}.
Words berlekamp_massey(const std::vector<uint8_t>& s, int& L_out) {
  const long long N = (long long)s.size();
  const size_t W = (size_t)(N + 64) / 64 + 1;
  Words C(W, 0), B(W, 0);
  C[0] = B[0] = 1;
  // the sequence reversed, as bits: r[j] = s[N - 1 - j], so that
  // s_{n-i} = r[N - 1 - n + i] and the discrepancy is a dot product of C with
  // r shifted by N - 1 - n
  Words r(W + 1, 0);
  for (long long j = 0; j < N; ++j)
    if (s[(size_t)(N - 1 - j)]) flip(r, j);
  long long L = 0, m = 1;
  for (long long n = 0; n < N; ++n) {
    const long long off = N - 1 - n;
    const size_t q = (size_t)(off >> 6);
    const int b = (int)(off & 63);
    uint64_t acc = 0;
    const size_t cw = (size_t)(L >> 6) + 1;
    for (size_t i = 0; i < cw; ++i) {
      const size_t k = i + q;
      uint64_t v = k < r.size() ? r[k] >> b : 0;
      if (b && k + 1 < r.size()) v |= r[k + 1] << (64 - b);
      acc ^= C[i] & v;
    }
    // bits of C above L are zero, so whole words are safe
    if (!(__builtin_popcountll(acc) & 1)) {
      ++m;
      continue;
    }
    const Words T = (2 * L <= n) ? C : Words();
    // C ^= B << m
    const size_t qm = (size_t)(m >> 6);
    const int bm = (int)(m & 63);
    for (size_t i = W; i-- > 0;) {
      if (i < qm) break;
      uint64_t v = B[i - qm] << bm;
      if (bm && i > qm) v |= B[i - qm - 1] >> (64 - bm);
      C[i] ^= v;
    }
    if (2 * L <= n) {
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      ++m;
    }
  }
  L_out = (int)L;
  return C;
}

std::mutex g_mu;
Modulus* g_mod[2] = {nullptr, nullptr};
bool g_tried[2] = {false, false};

Modulus* build_modulus(Gen g) {
  // low bits of the raw word sequence of seed 4357 (any certified state)
  const long long nbits = 2LL * 19968 + 128;
  std::vector<uint8_t> s((size_t)nbits);
  if (g == kMT19937) {
    const std::vector<uint32_t> x = mt19937_raw(4357u, nbits);
    for (long long i = 0; i < nbits; ++i) s[(size_t)i] = x[(size_t)i] & 1u;
  } else {
    const std::vector<uint32_t> x = sfmt19937_raw(4357u, 4 * nbits);
    for (long long i = 0; i < nbits; ++i) s[(size_t)i] = x[(size_t)(4 * i)] & 1u;  // lane 0 of each 128-bit word
  }
  int L = 0;
  const Words C = berlekamp_massey(s, L);
  if (L != degree(g)) return nullptr;
  Modulus* M = new Modulus();
  const long long n = L;
  M->n = n;
  M->phi.assign(kWords + 1, 0);
  for (long long i = 0; i <= n; ++i)  // phi_i = C_{L - i}: the reciprocal of the connection polynomial
    if (bit(C, n - i)) flip(M->phi, i);
  // mu = floor(x^(2n) / phi) by long division
  Words rem(2 * kWords + 2, 0);
  flip(rem, 2 * n);
  M->mu.assign(kWords + 1, 0);
  for (long long d = 2 * n; d >= n; --d) {
    if (!bit(rem, d)) continue;
    flip(M->mu, d - n);
    // rem ^= phi << (d - n), word by word
    const long long sh = d - n;
    const size_t q = (size_t)(sh >> 6);
    const int bs = (int)(sh & 63);
    for (size_t i = 0; i < M->phi.size(); ++i) {
      const uint64_t v = M->phi[i];
      if (!v) continue;
      rem[i + q] ^= v << bs;
      if (bs && i + q + 1 < rem.size()) rem[i + q + 1] ^= v >> (64 - bs);
    }
  }
  return M;
}

// nullptr if Berlekamp-Massey did not find degree(g) (a broken generator copy)
const Modulus* modulus(Gen g) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_tried[g]) {
    g_mod[g] = build_modulus(g);
    g_tried[g] = true;
  }
  return g_mod[g];
}

// a mod phi for deg a < 2n (Barrett: q = ((a >> n) mu) >> n, r = a - q phi).
Words reduce(const Modulus& M, const Words& a) {
  const long long n = M.n;
  const Words hi = shr(a, n, kWords + 1);
  Words t;
  mul(hi, M.mu, t);
  const Words q = shr(t, n, kWords + 1);
  Words qp;
  mul(q, M.phi, qp);
  Words r(kWords, 0);
  for (size_t i = 0; i < (size_t)kWords; ++i) r[i] = (i < a.size() ? a[i] : 0) ^ (i < qp.size() ? qp[i] : 0);
  // bits at and above n are zero by construction; clear the padding of the top word
  r[kWords - 1] &= top_mask(n);
  return r;
}

}  // namespace

int degree(Gen g) { return g == kMT19937 ? 19937 : 19968; }

const std::vector<uint64_t>& minpoly(Gen g) {
  static const std::vector<uint64_t> empty;
  const Modulus* m = modulus(g);
  return m ? m->phi : empty;
}

std::vector<uint64_t> mulmod(Gen g, const std::vector<uint64_t>& a, const std::vector<uint64_t>& b) {
  const Modulus* M = modulus(g);
  if (!M) return {};
  Words p;
  mul(a, b, p);
  return reduce(*M, p);
}

std::vector<uint64_t> xpow(Gen g, unsigned long long J) {
  const Modulus* Mp = modulus(g);
  if (!Mp) return {};
  const Modulus& M = *Mp;
  Words r(kWords, 0);
  r[0] = 1;
  int top = 63;
  while (top >= 0 && !((J >> top) & 1ull)) --top;
  for (int k = top; k >= 0; --k) {
    // r = r^2 mod phi: squaring spreads the bits (no cross terms over GF(2))
    Words sq(2 * kWords, 0);
    for (size_t i = 0; i < (size_t)kWords; ++i) {
      uint64_t lo = 0, hi = 0;
      const uint64_t v = r[i];
      for (int b = 0; b < 32; ++b) {
        lo |= ((v >> b) & 1ull) << (2 * b);
        hi |= ((v >> (b + 32)) & 1ull) << (2 * b);
      }
      sq[2 * i] = lo;
      sq[2 * i + 1] = hi;
    }
    r = reduce(M, sq);
    if ((J >> k) & 1ull) {  // r = r x mod phi
      uint64_t carry = 0;
      for (size_t i = 0; i < (size_t)kWords; ++i) {
        const uint64_t v = r[i];
        r[i] = (v << 1) | carry;
        carry = v >> 63;
      }
      const long long n = M.n;
      if (n % 64 ? bit(r, n) : carry != 0) {  // the coefficient of x^n (in word 312 when n = 19968)
        for (size_t i = 0; i < (size_t)kWords; ++i) r[i] ^= M.phi[i];
        r[kWords - 1] &= top_mask(n);
      }
    }
  }
  return r;
}

// ---- host reference generators ------------------------------------------------
std::vector<uint32_t> mt19937_raw(uint32_t seed, long long n) {
  constexpr int N = 624, M = 397;
  uint32_t mt[N];
  mt[0] = 19650218u;
  for (int i = 1; i < N; ++i) mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (uint32_t)i;
  int i = 1;
  for (int k = N; k > 0; --k) {  // init_by_array with the one-word key {seed}
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1664525u)) + seed;
    if (++i >= N) {
      mt[0] = mt[N - 1];
      i = 1;
    }
  }
  for (int k = N - 1; k > 0; --k) {
    mt[i] = (mt[i] ^ ((mt[i - 1] ^ (mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
    if (++i >= N) {
      mt[0] = mt[N - 1];
      i = 1;
    }
  }
  mt[0] = 0x80000000u;
  std::vector<uint32_t> out((size_t)n);
  for (long long t = 0; t < n; ++t) {
    const int k = (int)(t % N);
    const uint32_t y = (mt[k] & 0x80000000u) | (mt[(k + 1) % N] & 0x7fffffffu);
    mt[k] = mt[(k + M) % N] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    out[(size_t)t] = mt[k];
  }
  return out;
}

std::vector<uint32_t> sfmt19937_raw(uint32_t seed, long long n) {
  constexpr int N32 = 624, Q = 156, lag = 11, mid = (N32 - lag) / 2;
  constexpr uint32_t msk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};
  constexpr uint32_t parity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};
  uint32_t s[N32];
  for (int i = 0; i < N32; ++i) s[i] = 0x8b8b8b8bu;
  int i = 0;
  for (int k = 0; k < N32; ++k) {  // init_by_array({seed}): the key pass
    const uint32_t x = s[i] ^ s[(i + mid) % N32] ^ s[(i + N32 - 1) % N32];
    uint32_t r = (x ^ (x >> 27)) * 1664525u;
    s[(i + mid) % N32] += r;
    r += (k == 0 ? 1u : 0u) + (k == 1 ? seed : 0u) + (uint32_t)i;
    s[(i + mid + lag) % N32] += r;
    s[i] = r;
    i = (i + 1) % N32;
  }
  for (int k = 0; k < N32; ++k) {  // the mixing pass
    const uint32_t x = s[i] + s[(i + mid) % N32] + s[(i + N32 - 1) % N32];
    uint32_t r = (x ^ (x >> 27)) * 1566083941u;
    s[(i + mid) % N32] ^= r;
    r -= (uint32_t)i;
    s[(i + mid + lag) % N32] ^= r;
    s[i] = r;
    i = (i + 1) % N32;
  }
  uint32_t inner = 0;
  for (int k = 0; k < 4; ++k) inner ^= s[k] & parity[k];
  for (int sh = 16; sh > 0; sh >>= 1) inner ^= inner >> sh;
  if (!(inner & 1u)) s[0] ^= 1u;  // period certification: the first parity bit
  auto lsh8 = [](const uint32_t* in, uint32_t* o) {  // 128-bit shift left by one byte
    const uint64_t th = ((uint64_t)in[3] << 32) | in[2], tl = ((uint64_t)in[1] << 32) | in[0];
    const uint64_t oh = (th << 8) | (tl >> 56), ol = tl << 8;
    o[0] = (uint32_t)ol;
    o[1] = (uint32_t)(ol >> 32);
    o[2] = (uint32_t)oh;
    o[3] = (uint32_t)(oh >> 32);
  };
  auto rsh8 = [](const uint32_t* in, uint32_t* o) {  // 128-bit shift right by one byte
    const uint64_t th = ((uint64_t)in[3] << 32) | in[2], tl = ((uint64_t)in[1] << 32) | in[0];
    const uint64_t ol = (tl >> 8) | (th << 56), oh = th >> 8;
    o[0] = (uint32_t)ol;
    o[1] = (uint32_t)(ol >> 32);
    o[2] = (uint32_t)oh;
    o[3] = (uint32_t)(oh >> 32);
  };
  std::vector<uint32_t> out((size_t)n);
  long long t = 0;
  while (t < n) {
    for (int q = 0; q < Q; ++q) {
      uint32_t* a = s + 4 * q;
      const uint32_t* b = s + 4 * ((q + 122) % Q);
      const uint32_t* c = s + 4 * ((q + Q - 2) % Q);
      const uint32_t* d = s + 4 * ((q + Q - 1) % Q);
      uint32_t al[4], cr[4];
      lsh8(a, al);
      rsh8(c, cr);
      for (int j = 0; j < 4; ++j) a[j] = a[j] ^ al[j] ^ ((b[j] >> 11) & msk[j]) ^ cr[j] ^ (d[j] << 18);
      for (int j = 0; j < 4 && t < n; ++j) out[(size_t)t++] = a[j];
    }
  }
  return out;
}

}  // namespace gf2
}  // namespace fsmc

#include "../../include/fsmc.h"

extern "C" int fsmc_jump_poly(int generator, unsigned long long steps, uint64_t* coeffs, int n_words) {
  using namespace fsmc::gf2;
  if ((generator != kMT19937 && generator != kSFMT19937) || coeffs == nullptr || n_words < kWords)
    return FSMC_ERR_INVALID_ARG;
  const std::vector<uint64_t> g = xpow((Gen)generator, steps);
  if (g.size() != (size_t)kWords) return FSMC_ERR_DEVICE;  // the minimal polynomial was not found
  std::copy(g.begin(), g.end(), coeffs);
  return degree((Gen)generator);
}
