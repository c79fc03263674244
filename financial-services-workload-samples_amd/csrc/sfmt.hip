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
// One stream is sequential: a two-wave workgroup per stream seeds the state
// in LDS (one lane); one wave runs the 156-word recursion in registers (one
// 32-bit column per lane) while the other turns the previous block into
// normals.
#include <algorithm>
#include <map>
#include <mutex>
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

// Value of lane (lane & ~3) + (lane & 3) + shift within each group of four
// lanes, by a quad permute (kCtrl: quad_perm selectors).
template <int kCtrl>
__device__ __forceinline__ uint32_t quad_mov(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, 0xF, 0xF, false);
}

constexpr int kQ = kWords / 4;  // 128-bit words of the state

// One regeneration of the 156-word state held in registers, in place: this
// lane holds column j = lane & 3 of every word (every quad of the wave runs
// the same recursion; the 128-bit byte shifts take the neighbouring column
// from the adjacent lane by a quad permute).
__device__ __forceinline__ void sfmt_regen(uint32_t (&st)[kQ], uint32_t msk, int j) {
#pragma unroll
  for (int q = 0; q < kQ; ++q) {  // the recursion, word by word, in place
    const uint32_t a = st[q], bb = st[(q + 122) % kQ], c = st[(q + kQ - 2) % kQ], d = st[(q + kQ - 1) % kQ];
    const uint32_t a_lo = quad_mov<0x90>(a);  // column j - 1 (quad_perm 0,0,1,2)
    const uint32_t c_hi = quad_mov<0xF9>(c);  // column j + 1 (quad_perm 1,2,3,3)
    const uint32_t a_sl = (a << 8) | (j > 0 ? a_lo >> 24 : 0u);
    const uint32_t c_sr = (c >> 8) | (j < 3 ? c_hi << 24 : 0u);
    st[q] = a ^ a_sl ^ ((bb >> 11) & msk) ^ c_sr ^ (d << 18);
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
// (one substream), from the seeded state of seed0 + t.  Two waves,
// producer and consumer: wave 0 keeps the state in registers and writes
// block b to LDS buffer b & 1, while wave 1 maps block b - 1 from the other
// buffer through the inverse CDF; one barrier per block separates the two.
__global__ __launch_bounds__(128) void sfmt_gaussian_kernel(uint32_t seed0, long long n_per_stream, long long L,
                                                            const uint32_t* start, long long seg, const double* mean,
                                                            const double* stddev, float* out) {
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
    for (int i = t; i < kWords; i += 128) w[1][i] = src[i];
  } else if (t == 0) {
    sfmt_seed(seed0 + stream, w[1]);
  }
  __syncthreads();
  uint32_t st[kQ];
  const uint32_t msk = column_mask(j);
  if (producer) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) st[q] = w[1][4 * q + j];
  }
  __syncthreads();  // the start buffer is read before block 1 reuses it
  for (long long b = 0; b <= n_blocks; ++b) {
    if (producer && b < n_blocks) {
      sfmt_regen(st, msk, j);
      if (t < 4) {
        uint32_t* dst = w[b & 1];
#pragma unroll
        for (int q = 0; q < kQ; ++q) dst[4 * q + t] = st[q];
      }
    }
    if (!producer && b >= 1) {
      const long long done = (b - 1) * kWords;
      const uint32_t* src = w[(b - 1) & 1];
      const int mm = (int)std::min<long long>(kWords, m - done);
      for (int i = t - 64; i < mm; i += 64) {
        // int32 = -2^31 gives u = 0, where the inverse CDF diverges: 2^-33 instead
        const double u = fmax((double)(int32_t)src[i] * 0x1.0p-32 + 0.5, 0x1.0p-33);
        const double z = icdf(u);
        const size_t g = base + (size_t)(done + i);
        if (mean) {
          const size_t kk = g / (size_t)seg;
          out[g] = (float)(mean[kk] + stddev[kk] * z);
        } else {
          out[g] = (float)z;
        }
      }
    }
    __syncthreads();
  }
}

// ---- skip-ahead ------------------------------------------------------------
// The 128-bit words Y_s of a stream (Y_0 .. Y_155 the seeded state, then the
// regenerated blocks) obey a linear recursion whose step has characteristic
// polynomial phi of degree 19968; the state J words later is, word by word,
// W[w] = XOR over the set coefficients i of x^J mod phi of Y_{i + w}
// (tests/test_jump_poly.py checks the identity on the oracle's stream).
// A call of K substreams of L outputs builds their start states by doubling:
// the seeded state is substream 0's, and round r jumps substreams
// [0, 2^r) by 2^r L / 4 words into [2^r, 2^(r+1)).

__global__ __launch_bounds__(64) void sfmt_seed_kernel(uint32_t seed0, int K, uint32_t* states) {
  __shared__ uint32_t w[kWords];
  if (threadIdx.x == 0) sfmt_seed(seed0 + blockIdx.x, w);
  __syncthreads();
  uint32_t* dst = states + (size_t)blockIdx.x * K * kWords;
  for (int i = threadIdx.x; i < kWords; i += 64) dst[i] = w[i];
}

constexpr int kJumpThreads = 256;
constexpr int kRingWords = 2048;  // LDS ring of 128-bit words (32 KB)

// Round r: workgroup (jj, stream) jumps substream jj's state by the
// polynomial coef[0 .. ncoef) (sorted coefficients, values < 19968) into
// substream 2^r + jj.  Wave 0 regenerates the sequence from the state into
// the ring; thread t accumulates the window's 32-bit words t, t + 256,
// t + 512 (word c / 4, lane c % 4), taking the coefficients in ascending
// pieces short enough that Y_{i .. i + 155} is still in the ring.
__global__ __launch_bounds__(kJumpThreads) void sfmt_jump_kernel(uint32_t* states, int K, int r,
                                                                 const uint16_t* coef, int ncoef) {
  __shared__ uint32_t ring[kRingWords * 4];
  const int t = threadIdx.x;
  const int k = (1 << r) + (int)blockIdx.x;
  if (k >= K) return;  // uniform per workgroup, before any barrier
  const uint32_t* src = states + ((size_t)blockIdx.y * K + blockIdx.x) * kWords;
  uint32_t* dst = states + ((size_t)blockIdx.y * K + k) * kWords;
  for (int i = t; i < kWords; i += kJumpThreads) ring[i] = src[i];  // Y_0 .. Y_155 at ring words 0 .. 155
  __syncthreads();
  const bool producer = t < 64;  // wave-uniform
  const int j = t & 3;
  uint32_t st[kQ];
  const uint32_t msk = column_mask(j);
  if (producer) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) st[q] = ring[4 * q + j];
  }
  uint32_t acc[3] = {0u, 0u, 0u};
  long long G = kQ;  // Y_s is in the ring for s < G
  int q = 0;
  while (q < ncoef) {
    const int a = coef[q];
    const int lim = a + (kRingWords - 2 * kQ);  // the coefficients of this piece are < lim
    while (G < (long long)lim + kQ) {
      __syncthreads();  // the previous piece's reads of the ring before it moves on
      if (producer) {
        sfmt_regen(st, msk, j);
        if (t < 4) {
#pragma unroll
          for (int qq = 0; qq < kQ; ++qq) ring[(((G + qq) & (kRingWords - 1)) << 2) + t] = st[qq];
        }
      }
      __syncthreads();
      G += kQ;
    }
    int qq = q;
    for (; qq < ncoef && (int)coef[qq] < lim; ++qq) {
      const int i = coef[qq];
#pragma unroll
      for (int h = 0; h < 3; ++h) {
        const int c = t + h * kJumpThreads;
        if (c < kWords) acc[h] ^= ring[(((i + (c >> 2)) & (kRingWords - 1)) << 2) + (c & 3)];
      }
    }
    q = qq;  // the same for every thread
  }
#pragma unroll
  for (int h = 0; h < 3; ++h) {
    const int c = t + h * kJumpThreads;
    if (c < kWords) dst[c] = acc[h];
  }
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
  int K = 1, rounds = 0;
  long long L = 0;
};
SkipPlan skip_plan(int n_streams, long long n_per_stream) {
  SkipPlan p;
  p.L = n_per_stream;
  int K = 1;
  while (2 * K <= kMaxSub && 2 * K * n_streams <= 2048 && 2 * K * kMinPer <= n_per_stream) K *= 2;
  if (K == 1) return p;
  p.K = K;
  p.L = ((n_per_stream + K - 1) / K + kWords - 1) / kWords * kWords;
  while ((1 << p.rounds) < K) ++p.rounds;
  return p;
}

// Sorted coefficients of x^(2^r L / 4) mod phi, r < rounds, cached per
// (L, rounds) for the process.
std::mutex g_skip_mu;
std::map<std::pair<long long, int>, std::vector<std::vector<uint16_t>>>& skip_cache() {
  static std::map<std::pair<long long, int>, std::vector<std::vector<uint16_t>>> c;
  return c;
}
int skip_lists(const SkipPlan& p, const std::vector<std::vector<uint16_t>>** out) {
  std::lock_guard<std::mutex> lock(g_skip_mu);
  const auto key = std::make_pair(p.L, p.rounds);
  auto it = skip_cache().find(key);
  if (it == skip_cache().end()) {
    std::vector<std::vector<uint16_t>> lists;
    std::vector<uint64_t> poly = gf2::xpow(gf2::kSFMT19937, (unsigned long long)(p.L / 4));
    FSMC_REQUIRE(poly.size() == (size_t)gf2::kWords, "SFMT19937 jump polynomial unavailable");
    for (int r = 0; r < p.rounds; ++r) {
      if (r > 0) poly = gf2::mulmod(gf2::kSFMT19937, poly, poly);
      std::vector<uint16_t> c;
      for (int w = 0; w < gf2::kWords; ++w)
        for (uint64_t v = poly[w]; v; v &= v - 1) c.push_back((uint16_t)(64 * w + __builtin_ctzll(v)));
      lists.push_back(std::move(c));
    }
    it = skip_cache().emplace(key, std::move(lists)).first;
  }
  *out = &it->second;
  return FSMC_OK;
}

size_t skip_workspace(const SkipPlan& p, int n_streams) {
  if (p.K == 1) return 0;
  return align_up((size_t)n_streams * p.K * kWords * sizeof(uint32_t), 256) +
         (size_t)p.rounds * align_up((size_t)gf2::kWords * 64 * sizeof(uint16_t), 256);
}

int sfmt_enqueue(uint32_t seed0, int n_streams, long long n_per_stream, long long seg, const double* d_mean,
                 const double* d_stddev, float* d_out, void* d_ws, size_t ws_bytes, hipStream_t s) {
  if (n_per_stream == 0) return FSMC_OK;
  const SkipPlan p = d_ws ? skip_plan(n_streams, n_per_stream) : SkipPlan{1, 0, n_per_stream};
  const long long sg = d_mean ? seg : 1LL;
  if (p.K == 1) {
    hipLaunchKernelGGL(sfmt_gaussian_kernel, dim3(1, n_streams), dim3(128), 0, s, seed0, n_per_stream, n_per_stream,
                       (const uint32_t*)nullptr, sg, d_mean, d_stddev, d_out);
    FSMC_HIP(hipGetLastError());
    return FSMC_OK;
  }
  if (ws_bytes < skip_workspace(p, n_streams)) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, skip_workspace(p, n_streams));
    return FSMC_ERR_WORKSPACE;
  }
  const std::vector<std::vector<uint16_t>>* lists = nullptr;
  FSMC_TRY(skip_lists(p, &lists));
  uint32_t* states = (uint32_t*)d_ws;
  char* cbase = (char*)d_ws + align_up((size_t)n_streams * p.K * kWords * sizeof(uint32_t), 256);
  const size_t cstride = align_up((size_t)gf2::kWords * 64 * sizeof(uint16_t), 256);
  for (int r = 0; r < p.rounds; ++r)
    FSMC_HIP(hipMemcpyAsync(cbase + r * cstride, (*lists)[r].data(), (*lists)[r].size() * sizeof(uint16_t),
                            hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(sfmt_seed_kernel, dim3(n_streams), dim3(64), 0, s, seed0, p.K, states);
  for (int r = 0; r < p.rounds; ++r)
    hipLaunchKernelGGL(sfmt_jump_kernel, dim3(1u << r, n_streams), dim3(kJumpThreads), 0, s, states, p.K, r,
                       (const uint16_t*)(cbase + r * cstride), (int)(*lists)[r].size());
  hipLaunchKernelGGL(sfmt_gaussian_kernel, dim3(p.K, n_streams), dim3(128), 0, s, seed0, n_per_stream, p.L,
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
