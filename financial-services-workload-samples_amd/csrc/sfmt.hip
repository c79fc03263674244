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

// n_per_stream normals of stream t into out[t * n_per_stream + i]; with mean
// and stddev the value of global index g is mean[g / seg] + stddev[g / seg] z.
// Two waves per stream, producer and consumer: wave 0 keeps the state in
// registers, column j = lane & 3 of the 156 words in each lane (every quad
// runs the same recursion; the 128-bit byte shifts take the neighbouring
// column from the adjacent lane by a quad permute) and writes block b to LDS
// buffer b & 1, while wave 1 maps block b - 1 from the other buffer through
// the inverse CDF; one barrier per block separates the two.
__global__ __launch_bounds__(128) void sfmt_gaussian_kernel(uint32_t seed0, long long n_per_stream, long long seg,
                                                            const double* mean, const double* stddev, float* out) {
  constexpr int kQ = kWords / 4;  // 128-bit words
  __shared__ uint32_t w[2][kWords];
  const int t = threadIdx.x, j = t & 3;
  const bool producer = t < 64;  // wave-uniform
  const size_t base = (size_t)blockIdx.x * (size_t)n_per_stream;
  const long long n_blocks = (n_per_stream + kWords - 1) / kWords;
  if (t == 0) sfmt_seed(seed0 + blockIdx.x, w[1]);
  __syncthreads();
  uint32_t st[kQ];
  const uint32_t msk = j == 0 ? kMsk[0] : j == 1 ? kMsk[1] : j == 2 ? kMsk[2] : kMsk[3];
  if (producer) {
#pragma unroll
    for (int q = 0; q < kQ; ++q) st[q] = w[1][4 * q + j];
  }
  __syncthreads();  // the seed buffer is read before block 1 reuses it
  for (long long b = 0; b <= n_blocks; ++b) {
    if (producer && b < n_blocks) {
#pragma unroll
      for (int q = 0; q < kQ; ++q) {  // the recursion, word by word, in place
        const uint32_t a = st[q], bb = st[(q + 122) % kQ], c = st[(q + kQ - 2) % kQ], d = st[(q + kQ - 1) % kQ];
        const uint32_t a_lo = quad_mov<0x90>(a);  // column j - 1 (quad_perm 0,0,1,2)
        const uint32_t c_hi = quad_mov<0xF9>(c);  // column j + 1 (quad_perm 1,2,3,3)
        const uint32_t a_sl = (a << 8) | (j > 0 ? a_lo >> 24 : 0u);
        const uint32_t c_sr = (c >> 8) | (j < 3 ? c_hi << 24 : 0u);
        st[q] = a ^ a_sl ^ ((bb >> 11) & msk) ^ c_sr ^ (d << 18);
      }
      if (t < 4) {
        uint32_t* dst = w[b & 1];
#pragma unroll
        for (int q = 0; q < kQ; ++q) dst[4 * q + t] = st[q];
      }
    }
    if (!producer && b >= 1) {
      const long long done = (b - 1) * kWords;
      const uint32_t* src = w[(b - 1) & 1];
      const int m = (int)std::min<long long>(kWords, n_per_stream - done);
      for (int i = t - 64; i < m; i += 64) {
        // int32 = -2^31 gives u = 0, where the inverse CDF diverges: 2^-33 instead
        const double u = fmax((double)(int32_t)src[i] * 0x1.0p-32 + 0.5, 0x1.0p-33);
        const double z = icdf(u);
        const size_t g = base + (size_t)(done + i);
        if (mean) {
          const size_t k = g / (size_t)seg;
          out[g] = (float)(mean[k] + stddev[k] * z);
        } else {
          out[g] = (float)z;
        }
      }
    }
    __syncthreads();
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

}  // namespace
}  // namespace fsmc

using namespace fsmc;

extern "C" {

int fsmc_sfmt19937_gaussian_async(uint32_t seed0, int n_streams, long long n_per_stream, long long seg,
                                  const double* d_mean, const double* d_stddev, float* d_out, void* stream) {
  FSMC_TRY(sfmt_validate(n_streams, n_per_stream, seg, d_mean, d_stddev, d_out));
  if (n_per_stream == 0) return FSMC_OK;
  hipLaunchKernelGGL(sfmt_gaussian_kernel, dim3(n_streams), dim3(128), 0, (hipStream_t)stream, seed0, n_per_stream,
                     d_mean ? seg : 1LL, d_mean, d_stddev, d_out);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int fsmc_sfmt19937_gaussian(fsmc_ctx* ctx, uint32_t seed, long long n, double mean, double stddev, float* out) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(sfmt_validate(1, n, 1, nullptr, nullptr, out));
  if (n == 0) return FSMC_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  FSMC_TRY(ctx_prepare(ctx, 2 * sizeof(double), 0, (size_t)n * sizeof(float), 0));
  double* d_ms = (double*)ctx->inputs.ptr;
  const double ms[2] = {mean, stddev};
  hipStream_t s = ctx->stream;
  FSMC_HIP(hipMemcpyAsync(d_ms, ms, sizeof(ms), hipMemcpyHostToDevice, s));
  FSMC_TRY(fsmc_sfmt19937_gaussian_async(seed, 1, n, n, d_ms, d_ms + 1, (float*)ctx->outputs.ptr, s));
  FSMC_HIP(hipMemcpyAsync(out, ctx->outputs.ptr, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"
