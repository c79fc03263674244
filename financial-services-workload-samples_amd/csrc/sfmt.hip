// MKL-compatible SFMT19937 normals on the GPU: the generator the
// MonteCarloEuropeanOptions sample draws from (MonteCarloInsideBlockingDP.cpp:
// 229, vslNewStream(VSL_BRNG_SFMT19937, 777 + thread), and :249,
// vsRngGaussian(VSL_RNG_METHOD_GAUSSIAN_ICDF, ..., MuByT, VBySqrtT)).
//
// SFMT19937 (Saito and Matsumoto, "SIMD-oriented Fast Mersenne Twister",
// MCQMC 2006): 156 128-bit words w, regenerated in place in order
//   w[i] = w[i] ^ (w[i] <<_128 8) ^ ((w[i + 122] >>_32 11) & MSK)
//          ^ (w[i - 2] >>_128 8) ^ (w[i - 1] <<_32 18),
// where <<_128 / >>_128 shift the whole 128-bit word by 8 bits and <<_32 /
// >>_32 shift each 32-bit lane; indices are cyclic, so w[i - 1] and w[i - 2]
// are the words just written.  The outputs are the 624 32-bit lanes in order.
// MKL seeds a single-seed stream with init_by_array({seed}) followed by the
// period certification, and its ICDF method reads each output b as a signed
// integer: u = int32(b) 2^-32 + 1/2 (MKL 2021.4's first 2048 outputs and
// float normals for seed 777 pin both).  The inverse CDF is AS241 in FP64,
// rounded to float once with the affine map mean + stddev z.
//
// One stream is sequential: a workgroup per stream (or per substream, after
// a skip-ahead) seeds the state in LDS (one lane); one wave runs the 156-word
// recursion in registers (one 32-bit column per lane of its first quad)
// while the other turns the previous block into normals.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../host/gf2_jump.hpp"
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int kWords = 624;  // 32-bit words of the state
constexpr uint32_t kMsk[4] = {0xdfffffefu, 0xddfecb7fu, 0xbffaffffu, 0xbffffff6u};
constexpr uint32_t kParity[4] = {0x00000001u, 0x00000000u, 0x00000000u, 0x13c9e684u};

// init_by_array({seed}) of the SFMT reference, then the period check, into
// the LDS state s (one thread).
__device__ void sfmt_seed(uint32_t key, uint32_t* s) {
  constexpr int lag = 11, mid = (kWords - lag) / 2;
  for (int i = 0; i < kWords; ++i) s[i] = 0x8b8b8b8bu;
  int i = 0;
  for (int k = 0; k < kWords; ++k) {  // the key pass: length 1, then indices
    const uint32_t x = s[i] ^ s[(i + mid) % kWords] ^ s[(i + kWords - 1) % kWords];
    uint32_t r = (x ^ (x >> 27)) * 1664525u;
    s[(i + mid) % kWords] += r;
    r += (k == 0 ? 1u : 0u) + (k == 1 ? key : 0u) + (uint32_t)i;
    s[(i + mid + lag) % kWords] += r;
    s[i] = r;
    i = (i + 1) % kWords;
  }
  for (int k = 0; k < kWords; ++k) {  // the mixing pass
    const uint32_t x = s[i] + s[(i + mid) % kWords] + s[(i + kWords - 1) % kWords];
    uint32_t r = (x ^ (x >> 27)) * 1566083941u;
    s[(i + mid) % kWords] ^= r;
    r -= (uint32_t)i;
    s[(i + mid + lag) % kWords] ^= r;
    s[i] = r;
    i = (i + 1) % kWords;
  }
  uint32_t inner = 0;
  for (int k = 0; k < 4; ++k) inner ^= s[k] & kParity[k];
  for (int sh = 16; sh > 0; sh >>= 1) inner ^= inner >> sh;
  if (!(inner & 1u)) s[0] ^= 1u;  // the first parity bit (bit 0 of word 0)
}

constexpr int kQ = kWords / 4;  // 128-bit words of the state

// One regeneration of the 156-word state held in registers, in place, on
// lanes 0-3 of a wave (lane j holds column j of every word; lanes 4-63 hold
// zeros and are inactive).  The 128-bit byte shifts take the neighbouring
// column from the adjacent lane by a row shift (DPP row_shr:1 / row_shl:1
// with zero fill: lane 0 has no lower neighbour, lane 3's upper one is
// inactive and zero), and the funnel shift joins the two columns:
// (w <<_128 8) of column j = alignbit(w_j, w_{j-1}, 24),
// (w >>_128 8) of column j = alignbit(w_{j+1}, w_j, 8).
__device__ __forceinline__ void sfmt_regen(uint32_t (&st)[kQ], uint32_t msk) {
#pragma unroll
  for (int q = 0; q < kQ; ++q) {  // the recursion, word by word, in place
    const uint32_t a = st[q], bb = st[(q + 122) % kQ], c = st[(q + kQ - 2) % kQ], d = st[(q + kQ - 1) % kQ];
    const uint32_t a_lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)a, 0x111, 0xF, 0xF, true);  // row_shr:1
    const uint32_t c_hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)c, 0x101, 0xF, 0xF, true);  // row_shl:1
    const uint32_t a_sl = __builtin_amdgcn_alignbit(a, a_lo, 24);
    const uint32_t c_sr = __builtin_amdgcn_alignbit(c_hi, c, 8);
    // d is the word just written: its term goes in last, so the chain from
    // one word to the next is a shift and an XOR
    uint32_t rest = (a ^ a_sl) ^ (c_sr ^ ((bb >> 11) & msk));
    asm("" : "+v"(rest));  // keeps the compiler from folding d's term in first
    st[q] = rest ^ (d << 18);
  }
}

__device__ __forceinline__ uint32_t column_mask(int j) {
  return j == 0 ? kMsk[0] : j == 1 ? kMsk[1] : j == 2 ? kMsk[2] : kMsk[3];
}

// Outputs [k L, min(n_per_stream, (k + 1) L)) of stream t (blockIdx.y) into
// out[t * n_per_stream + i], k = blockIdx.x; with mean and stddev the value
// of global index g is mean[g / seg] + stddev[g / seg] z.  The substream
// starts from start[(t gridDim.x + k)] (624 words: the state the block at
// output k L is regenerated from, L a multiple of 624), or, without start
// (one substream), from the seeded state of seed0 + t.  Wave 0 is the
// producer: lanes 0-3 keep the state in registers and write block b to LDS
// buffer b & 1, while wave 1 maps block b - 1 from the other buffer
// through the inverse CDF; one barrier per block separates the two.
constexpr int kGaussThreads = 128;
__global__ __launch_bounds__(kGaussThreads) void sfmt_gaussian_kernel(uint32_t seed0, long long n_per_stream,
                                                                      long long L, const uint32_t* start, long long seg,
                                                                      const double* mean, const double* stddev,
                                                                      float* out) {
  __shared__ uint32_t w[2][kWords];
  const int t = threadIdx.x, j = t & 3;
  const bool producer = t < 64;  // wave-uniform
  const int k = blockIdx.x, stream = blockIdx.y;
  const long long first = (long long)k * L;
  const long long m = first + L < n_per_stream ? L : n_per_stream - first;  // this substream's outputs
  const size_t base = (size_t)stream * (size_t)n_per_stream + (size_t)first;
  const long long n_blocks = (m + kWords - 1) / kWords;
  if (start) {
    const uint32_t* src = start + ((size_t)stream * gridDim.x + k) * kWords;
    for (int i = t; i < kWords; i += kGaussThreads) w[1][i] = src[i];
  } else if (t == 0) {
    sfmt_seed(seed0 + stream, w[1]);
  }
  __syncthreads();
  // The two waves run separate loops with the same number of barriers, so
  // that the state's registers are not live in the consumer's code (which
  // keeps the kernel within 256 VGPRs: two waves per SIMD).
  if (producer) {
    uint32_t st[kQ];
    const uint32_t msk = column_mask(j);
#pragma unroll
    for (int q = 0; q < kQ; ++q) st[q] = t < 4 ? w[1][4 * q + t] : 0u;
    __syncthreads();  // the start buffer is read before block 1 reuses it
    __builtin_amdgcn_s_setprio(3);  // the recursion is the critical path: it issues first
    for (long long b = 0; b <= n_blocks; ++b) {
      if (t < 4 && b < n_blocks) {  // lanes 0-3 run the recursion
        sfmt_regen(st, msk);
        uint32_t* dst = w[b & 1];
#pragma unroll
        for (int q = 0; q < kQ; ++q) dst[4 * q + t] = st[q];
      }
      __syncthreads();
    }
  } else {
    __syncthreads();  // pairs with the producer's first barrier
    for (long long b = 0; b <= n_blocks; ++b) {
      if (b >= 1) {
        const long long done = (b - 1) * kWords;
        const uint32_t* src = w[(b - 1) & 1];
        const int mm = (int)std::min<long long>(kWords, m - done);
        // segment of the block's first output and its offset in it (uniform);
        // with seg >= 624 the block spans at most two segments, whose means
        // and deviations are loaded once
        const size_t g0 = base + (size_t)done;
        const size_t kk0 = g0 / (size_t)seg, r0 = g0 - kk0 * (size_t)seg;
        const bool wide = seg >= kWords;
        const bool two = wide && r0 + (size_t)mm > (size_t)seg;
        double m0 = 0.0, s0 = 1.0, m1 = 0.0, s1 = 1.0;
        if (mean && wide) {
          m0 = mean[kk0];
          s0 = stddev[kk0];
          if (two) {
            m1 = mean[kk0 + 1];
            s1 = stddev[kk0 + 1];
          }
        }
        for (int i = t - 64; i < mm; i += kGaussThreads - 64) {
          // int32 = -2^31 gives u = 0, where the inverse CDF diverges: 2^-33 instead
          const double u = fmax((double)(int32_t)src[i] * 0x1.0p-32 + 0.5, 0x1.0p-33);
          const double z = icdf(u);
          const size_t g = g0 + (size_t)i;
          if (mean) {
            if (wide) {
              const bool hi = r0 + (size_t)i >= (size_t)seg;
              out[g] = (float)((hi ? m1 : m0) + (hi ? s1 : s0) * z);
            } else {
              const size_t kk = g / (size_t)seg;
              out[g] = (float)(mean[kk] + stddev[kk] * z);
            }
          } else {
            out[g] = (float)z;
          }
        }
      }
      __syncthreads();
    }
  }
}

// ---- skip-ahead ------------------------------------------------------------
// The 128-bit words Y_s of a stream (Y_0 .. Y_155 the seeded state, then the
// regenerated blocks) obey a linear recursion whose step has characteristic
// polynomial phi of degree 19968; the state J words later is, word by word,
// W[w] = XOR over the set coefficients i of x^J mod phi of Y_{i + w}
// (tests/test_jump_poly.py checks the identity on the oracle's stream).
// A call of K = B0 B1 substreams of L outputs builds their start states in
// at most two rounds of independent jumps: round 0 jumps the seeded state
// (substream 0's) by m B1 L / 4 words into substream m B1, m = 1 .. B0 - 1;
// round 1 jumps each substream m B1 by j L / 4 words into m B1 + j,
// j = 1 .. B1 - 1 (one round, B1 = 1, for K <= 32).  Every jump of a round
// runs in its own workgroup, so a call waits on two jump latencies, not
// log2 K of them.

__global__ __launch_bounds__(64) void sfmt_seed_kernel(uint32_t seed0, int K, uint32_t* states) {
  __shared__ uint32_t w[kWords];
  if (threadIdx.x == 0) sfmt_seed(seed0 + blockIdx.x, w);
  __syncthreads();
  uint32_t* dst = states + (size_t)blockIdx.x * K * kWords;
  for (int i = threadIdx.x; i < kWords; i += 64) dst[i] = w[i];
}

constexpr int kJumpThreads = 128;
constexpr int kRingWords = 2048;                                 // LDS ring of 128-bit words (32 KB)
constexpr int kPiece = kRingWords - 2 * kQ;                      // 1736 coefficient values per piece
constexpr int kPieces = (gf2::kMaxDeg + kPiece - 1) / kPiece;  // 12 pieces cover x^0 .. x^19967

// One jump per workgroup (x, stream): in round 0 substream 0's state by
// list x into substream (x + 1) B1; in round 1 substream m B1's state by
// list B0 - 1 + j - 1 into substream m B1 + j (m = x / (B1 - 1),
// j = 1 + x % (B1 - 1)).  List l's sorted coefficients (values < 19968) in
// pieces: piece p (values in [p kPiece, (p + 1) kPiece)) is
// coef[poff[l (kPieces + 1) + p] .. poff[l (kPieces + 1) + p + 1]).  Lanes
// 0-3 of wave 0 regenerate the sequence from the state into the ring; thread
// t of the two waves accumulates the window's 32-bit words t + 128 h,
// h = 0 .. 4 (word c / 4, lane c % 4).  Two waves of at most 256 VGPRs
// each let four workgroups share a CU.  A piece's values are short enough that Y_{i .. i + 155} of
// all of them is in the ring at once; its coefficients are first copied to
// LDS (coalesced) and then read eight at a time by every thread (one
// broadcast 16-byte read), so the loop waits on no global memory per
// coefficient.
__global__ __launch_bounds__(kJumpThreads) void sfmt_jump_kernel(uint32_t* states, int K, int round, int B0, int B1,
                                                                 const uint16_t* coef, const int* poff_all) {
  __shared__ uint32_t ring[kRingWords * 4];
  __shared__ __align__(16) uint16_t C[kPiece + 8];
  const int t = threadIdx.x;
  const int x = (int)blockIdx.x;
  int from, to, list;
  if (round == 0) {
    from = 0;
    to = (x + 1) * B1;
    list = x;
  } else {
    const int m = x / (B1 - 1), j = 1 + x % (B1 - 1);
    from = m * B1;
    to = from + j;
    list = B0 - 1 + j - 1;
  }
  if (to >= K) return;  // uniform per workgroup, before any barrier
  const int* poff = poff_all + (size_t)list * (kPieces + 1);
  const uint32_t* src = states + ((size_t)blockIdx.y * K + from) * kWords;
  uint32_t* dst = states + ((size_t)blockIdx.y * K + to) * kWords;
  for (int i = t; i < kWords; i += kJumpThreads) ring[i] = src[i];  // Y_0 .. Y_155 at ring words 0 .. 155
  __syncthreads();
  const bool producer = t < 64;  // wave-uniform
  const int j = t & 3;
  uint32_t st[kQ];
  const uint32_t msk = column_mask(j);
  if (producer) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) st[q] = t < 4 ? ring[4 * q + t] : 0u;
  }
  // this thread's window words c = t + 128 h: ring word offset c / 4, lane c % 4 (= t % 4)
  constexpr int kAcc = (kWords + kJumpThreads - 1) / kJumpThreads;  // 5
  const bool has_last = t + (kAcc - 1) * kJumpThreads < kWords;     // h = 4 exists for t < 112
  uint32_t acc[kAcc];
#pragma unroll
  for (int h = 0; h < kAcc; ++h) acc[h] = 0u;
  long long G = kQ;  // Y_s is in the ring for s < G
  for (int p = 0; p < kPieces; ++p) {
    const int q0 = poff[p], nq = poff[p + 1] - q0;  // uniform
    if (nq == 0) continue;
    __syncthreads();  // the previous piece's reads of the ring and of C
    for (int i = t; i < nq; i += kJumpThreads) C[i] = coef[q0 + i];
    // Y_s for s < (p + 1) kPiece + 155 is needed; the oldest word still read is p kPiece
    while (G < (long long)(p + 1) * kPiece + kQ) {
      if (t < 4) {  // the producer: lanes 0-3 of wave 0
        sfmt_regen(st, msk);
        const int g = (int)(G & (kRingWords - 1));
        if (g + kQ <= kRingWords) {  // uniform: the block does not wrap, constant offsets
          uint32_t* r = ring + 4 * g + t;
#pragma unroll
          for (int qq = 0; qq < kQ; ++qq) r[4 * qq] = st[qq];
        } else {
#pragma unroll
          for (int qq = 0; qq < kQ; ++qq) ring[(((g + qq) & (kRingWords - 1)) << 2) + t] = st[qq];
        }
      }
      G += kQ;
    }
    __syncthreads();  // the ring and C are ready
    // ring index of window word c for coefficient i: ((i + c / 4) mod 2048) * 4 + c % 4
    auto at = [&](uint32_t i, int h) {
      return ring[(((i + (uint32_t)((t + h * kJumpThreads) >> 2)) & (kRingWords - 1)) << 2) + j];
    };
    const uint4* C8 = reinterpret_cast<const uint4*>(C);
    int q = 0;
    for (; q + 8 <= nq; q += 8) {
      const uint4 v = C8[q >> 3];
      const uint32_t c[8] = {v.x & 0xffffu, v.x >> 16, v.y & 0xffffu, v.y >> 16,
                             v.z & 0xffffu, v.z >> 16, v.w & 0xffffu, v.w >> 16};
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma unroll
        for (int h = 0; h < kAcc - 1; ++h) acc[h] ^= at(c[u], h);
        if (has_last) acc[kAcc - 1] ^= at(c[u], kAcc - 1);
      }
    }
    for (; q < nq; ++q) {
      const uint32_t c = C[q];
#pragma unroll
      for (int h = 0; h < kAcc - 1; ++h) acc[h] ^= at(c, h);
      if (has_last) acc[kAcc - 1] ^= at(c, kAcc - 1);
    }
  }
#pragma unroll
  for (int h = 0; h < kAcc - 1; ++h) dst[t + h * kJumpThreads] = acc[h];
  if (has_last) dst[t + (kAcc - 1) * kJumpThreads] = acc[kAcc - 1];
}

int sfmt_validate(int n_streams, long long n_per_stream, long long seg, const double* mean, const double* stddev,
                  const float* out) {
  FSMC_REQUIRE(n_streams >= 1 && n_streams <= 65535, "n_streams %d outside [1, 65535]", n_streams);
  FSMC_REQUIRE(n_per_stream >= 0 && n_per_stream <= (1LL << 40) / n_streams, "n_per_stream %lld out of range",
               n_per_stream);
  FSMC_REQUIRE((mean == nullptr) == (stddev == nullptr), "mean and stddev must both be given or both be NULL");
  FSMC_REQUIRE(mean == nullptr || seg >= 1, "seg %lld must be >= 1", seg);
  FSMC_REQUIRE(n_per_stream == 0 || out != nullptr, "NULL output");
  return FSMC_OK;
}

// Substreams of a call: K (a power of two) substreams of L outputs (a
// multiple of 624) per stream, enough workgroups to fill the GPU, each at
// least kMinPer outputs; K = 1 for short streams.
constexpr long long kMinPer = 624LL * 32;
constexpr int kMaxSub = 1024;
struct SkipPlan {
  int K = 1, B0 = 1, B1 = 1;  // K = B0 B1 substreams
  long long L = 0;
  int lists() const { return B0 - 1 + B1 - 1; }
};
SkipPlan skip_plan(int n_streams, long long n_per_stream) {
  SkipPlan p;
  p.L = n_per_stream;
  int K = 1, lg = 0;
  while (2 * K <= kMaxSub && 2 * K * n_streams <= 2048 && 2 * K * kMinPer <= n_per_stream) {
    K *= 2;
    ++lg;
  }
  if (K == 1) return p;
  p.K = K;
  p.B1 = lg <= 5 ? 1 : 1 << (lg / 2);
  p.B0 = K / p.B1;
  p.L = ((n_per_stream + K - 1) / K + kWords - 1) / kWords * kWords;
  return p;
}

// The jump lists of a plan (round 0's x^(m B1 L / 4), m = 1 .. B0 - 1, then
// round 1's x^(j L / 4), j = 1 .. B1 - 1, mod phi): sorted coefficients
// concatenated, with the absolute offsets of every list's pieces; cached
// per plan for the process.
struct SkipLists {
  std::vector<uint16_t> coef;
  std::vector<int> poff;  // lists() (kPieces + 1) entries
};
std::mutex g_skip_mu;
std::map<std::tuple<long long, int, int>, SkipLists>& skip_cache() {
  static std::map<std::tuple<long long, int, int>, SkipLists> c;
  return c;
}
void append_list(const std::vector<uint64_t>& poly, SkipLists& out) {
  int piece = 0;
  for (int w = 0; w < gf2::kWords; ++w)
    for (uint64_t v = poly[w]; v; v &= v - 1) {
      const int c = 64 * w + __builtin_ctzll(v);
      while (piece <= c / kPiece) {
        out.poff.push_back((int)out.coef.size());
        ++piece;
      }
      out.coef.push_back((uint16_t)c);
    }
  while (piece <= kPieces) {
    out.poff.push_back((int)out.coef.size());
    ++piece;
  }
}
int skip_lists(const SkipPlan& p, const SkipLists** out) {
  std::lock_guard<std::mutex> lock(g_skip_mu);
  const auto key = std::make_tuple(p.L, p.B0, p.B1);
  auto it = skip_cache().find(key);
  if (it == skip_cache().end()) {
    SkipLists sl;
    const std::vector<uint64_t> step = gf2::xpow(gf2::kSFMT19937, (unsigned long long)(p.L / 4));
    FSMC_REQUIRE(step.size() == (size_t)gf2::kWords, "SFMT19937 jump polynomial unavailable");
    std::vector<uint64_t> big = step;  // x^(B1 L / 4)
    for (int j = 1; j < p.B1; ++j) big = gf2::mulmod(gf2::kSFMT19937, big, step);
    std::vector<uint64_t> q = big;
    for (int m = 1; m < p.B0; ++m) {
      if (m > 1) q = gf2::mulmod(gf2::kSFMT19937, q, big);
      append_list(q, sl);
    }
    q = step;
    for (int j = 1; j < p.B1; ++j) {
      if (j > 1) q = gf2::mulmod(gf2::kSFMT19937, q, step);
      append_list(q, sl);
    }
    it = skip_cache().emplace(key, std::move(sl)).first;
  }
  *out = &it->second;
  return FSMC_OK;
}

// Workspace: the K start states per stream, then the lists' coefficients
// (at most 19968 each) and their piece offsets.
size_t skip_workspace(const SkipPlan& p, int n_streams) {
  if (p.K == 1) return 0;
  return align_up((size_t)n_streams * p.K * kWords * sizeof(uint32_t), 256) +
         align_up((size_t)p.lists() * gf2::kMaxDeg * sizeof(uint16_t), 256) +
         align_up((size_t)p.lists() * (kPieces + 1) * sizeof(int), 256);
}

int sfmt_enqueue(uint32_t seed0, int n_streams, long long n_per_stream, long long seg, const double* d_mean,
                 const double* d_stddev, float* d_out, void* d_ws, size_t ws_bytes, hipStream_t s) {
  if (n_per_stream == 0) return FSMC_OK;
  SkipPlan p;
  if (d_ws) {
    p = skip_plan(n_streams, n_per_stream);
  } else {
    p.L = n_per_stream;
  }
  const long long sg = d_mean ? seg : 1LL;
  if (p.K == 1) {
    hipLaunchKernelGGL(sfmt_gaussian_kernel, dim3(1, n_streams), dim3(kGaussThreads), 0, s, seed0, n_per_stream,
                       n_per_stream, (const uint32_t*)nullptr, sg, d_mean, d_stddev, d_out);
    FSMC_HIP(hipGetLastError());
    return FSMC_OK;
  }
  if (ws_bytes < skip_workspace(p, n_streams)) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, skip_workspace(p, n_streams));
    return FSMC_ERR_WORKSPACE;
  }
  const SkipLists* sl = nullptr;
  FSMC_TRY(skip_lists(p, &sl));
  uint32_t* states = (uint32_t*)d_ws;
  char* cbase = (char*)d_ws + align_up((size_t)n_streams * p.K * kWords * sizeof(uint32_t), 256);
  char* obase = cbase + align_up((size_t)p.lists() * gf2::kMaxDeg * sizeof(uint16_t), 256);
  FSMC_HIP(hipMemcpyAsync(cbase, sl->coef.data(), sl->coef.size() * sizeof(uint16_t), hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(obase, sl->poff.data(), sl->poff.size() * sizeof(int), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(sfmt_seed_kernel, dim3(n_streams), dim3(64), 0, s, seed0, p.K, states);
  hipLaunchKernelGGL(sfmt_jump_kernel, dim3(p.B0 - 1, n_streams), dim3(kJumpThreads), 0, s, states, p.K, 0, p.B0,
                     p.B1, (const uint16_t*)cbase, (const int*)obase);
  if (p.B1 > 1)
    hipLaunchKernelGGL(sfmt_jump_kernel, dim3(p.B0 * (p.B1 - 1), n_streams), dim3(kJumpThreads), 0, s, states, p.K,
                       1, p.B0, p.B1, (const uint16_t*)cbase, (const int*)obase);
  hipLaunchKernelGGL(sfmt_gaussian_kernel, dim3(p.K, n_streams), dim3(kGaussThreads), 0, s, seed0, n_per_stream, p.L,
                     (const uint32_t*)states, sg, d_mean, d_stddev, d_out);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

}  // namespace
}  // namespace fsmc

using namespace fsmc;

extern "C" {

int fsmc_sfmt19937_gaussian_async(uint32_t seed0, int n_streams, long long n_per_stream, long long seg,
                                  const double* d_mean, const double* d_stddev, float* d_out, void* stream) {
  FSMC_TRY(sfmt_validate(n_streams, n_per_stream, seg, d_mean, d_stddev, d_out));
  return sfmt_enqueue(seed0, n_streams, n_per_stream, seg, d_mean, d_stddev, d_out, nullptr, 0,
                      (hipStream_t)stream);
}

size_t fsmc_sfmt19937_workspace(int n_streams, long long n_per_stream) {
  if (n_streams < 1 || n_streams > 65535 || n_per_stream < 0) return 0;
  return skip_workspace(skip_plan(n_streams, n_per_stream), n_streams);
}

int fsmc_sfmt19937_gaussian_skip_async(uint32_t seed0, int n_streams, long long n_per_stream, long long seg,
                                       const double* d_mean, const double* d_stddev, float* d_out, void* d_workspace,
                                       size_t workspace_bytes, void* stream) {
  FSMC_TRY(sfmt_validate(n_streams, n_per_stream, seg, d_mean, d_stddev, d_out));
  FSMC_REQUIRE(d_workspace != nullptr || fsmc_sfmt19937_workspace(n_streams, n_per_stream) == 0,
               "NULL workspace");
  return sfmt_enqueue(seed0, n_streams, n_per_stream, seg, d_mean, d_stddev, d_out, d_workspace, workspace_bytes,
                      (hipStream_t)stream);
}

int fsmc_sfmt19937_gaussian(fsmc_ctx* ctx, uint32_t seed, long long n, double mean, double stddev, float* out) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(sfmt_validate(1, n, 1, nullptr, nullptr, out));
  if (n == 0) return FSMC_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t ws = fsmc_sfmt19937_workspace(1, n);
  FSMC_TRY(ctx_prepare(ctx, 2 * sizeof(double), 0, (size_t)n * sizeof(float), ws));
  double* d_ms = (double*)ctx->inputs.ptr;
  const double ms[2] = {mean, stddev};
  hipStream_t s = ctx->stream;
  FSMC_HIP(hipMemcpyAsync(d_ms, ms, sizeof(ms), hipMemcpyHostToDevice, s));
  FSMC_TRY(sfmt_enqueue(seed, 1, n, n, d_ms, d_ms + 1, (float*)ctx->outputs.ptr, ws ? ctx->workspace.ptr : nullptr,
                        ws, s));
  FSMC_HIP(hipMemcpyAsync(out, ctx->outputs.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"
