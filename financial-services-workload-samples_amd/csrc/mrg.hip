// MKL-compatible MRG32k3a stream with inverse-CDF normals on the GPU: the
// device side of the generateRand seam (LiborMarketModel/mklRng.h:9,
// mklRng.cpp:22-32 and 44-139).  generateRand fills z[k], k < num, with the
// k-th normal of vslNewStream(VSL_BRNG_MRG32K3A, seed) under
// VSL_RNG_METHOD_GAUSSIAN_ICDF, whatever its thread count (each OpenMP thread
// skips ahead to its start).  Here every GPU thread owns a block of kBlockLen
// consecutive stream positions: it jumps its state there with precomputed
// powers of the one-step transition matrices, then runs the recurrences.
//
//   x_n = (1403580 x_{n-2} - 810728 x_{n-3}) mod m1,   m1 = 2^32 - 209
//   y_n = (527612 y_{n-1} - 1370589 y_{n-3}) mod m2,   m2 = 2^32 - 22853
//   u_n = ((x_n - y_n) mod m1) * (1 / m1)              (seed state x = (seed mod m1, 1, 1), y = (1, 1, 1))
//   z_n = mean + stddev * icdf(u_n)                    (Wichura's AS241, FP64)
//
// The uniforms equal MKL's bit for bit; MKL's own ICDF is not reproducible
// exactly (it differs from the exact inverse by up to ~7.3e-9 relative), so
// the normals agree with MKL's to that and with the exact inverse to ~1e-16.
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr uint64_t kM1 = 4294967087ULL, kM2 = 4294944443ULL;
constexpr uint64_t kC1 = (1ULL << 32) - kM1, kC2 = (1ULL << 32) - kM2;  // 209, 22853
constexpr int kBlockLen = 1024;  // stream positions per thread
constexpr int kThreads = 256;

// v mod (2^32 - c) for v < 2^64 by folding the high word twice.
__host__ __device__ constexpr uint64_t fold(uint64_t v, uint64_t c, uint64_t m) {
  v = (v >> 32) * c + (v & 0xffffffffULL);
  v = (v >> 32) * c + (v & 0xffffffffULL);
  while (v >= m) v -= m;
  return v;
}
__host__ __device__ constexpr uint64_t mulmod(uint64_t a, uint64_t b, uint64_t c, uint64_t m) {
  return fold(a * b, c, m);  // a, b < 2^32
}

struct Mat3 {
  uint64_t a[9];
};
struct JumpTable {
  Mat3 x[64], y[64];  // A1^(2^j), A2^(2^j)
};

__host__ __device__ constexpr Mat3 mat_mul(const Mat3& A, const Mat3& B, uint64_t c, uint64_t m) {
  Mat3 C{};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      uint64_t acc = 0;
      for (int k = 0; k < 3; ++k) acc = fold(acc + mulmod(A.a[3 * i + k], B.a[3 * k + j], c, m), c, m);
      C.a[3 * i + j] = acc;
    }
  return C;
}

constexpr JumpTable make_jump_table() {
  JumpTable t{};
  // one step maps (s0, s1, s2) = (x_{n-3}, x_{n-2}, x_{n-1}) to (x_{n-2}, x_{n-1}, x_n)
  t.x[0] = Mat3{{0, 1, 0, 0, 0, 1, kM1 - 810728ULL, 1403580ULL, 0}};
  t.y[0] = Mat3{{0, 1, 0, 0, 0, 1, kM2 - 1370589ULL, 0, 527612ULL}};
  for (int j = 1; j < 64; ++j) {
    t.x[j] = mat_mul(t.x[j - 1], t.x[j - 1], kC1, kM1);
    t.y[j] = mat_mul(t.y[j - 1], t.y[j - 1], kC2, kM2);
  }
  return t;
}

__constant__ JumpTable kJump = make_jump_table();
constexpr JumpTable kJumpHost = make_jump_table();

__host__ __device__ inline void mat_vec(const Mat3& A, uint64_t (&v)[3], uint64_t c, uint64_t m) {
  uint64_t t[3];
  for (int i = 0; i < 3; ++i) {
    uint64_t acc = 0;
    for (int k = 0; k < 3; ++k) acc = fold(acc + mulmod(A.a[3 * i + k], v[k], c, m), c, m);
    t[i] = acc;
  }
  v[0] = t[0];
  v[1] = t[1];
  v[2] = t[2];
}

struct MrgArgs {
  uint64_t x[3], y[3];  // state at the first requested position
  long long n;
  double mean, stddev;
  double* out;
};

// kNormal = false writes the uniforms themselves.
template <bool kNormal>
__global__ __launch_bounds__(kThreads) void mrg_kernel(MrgArgs a) {
  const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
  const long long first = t * kBlockLen;
  if (first >= a.n) return;
  uint64_t x[3] = {a.x[0], a.x[1], a.x[2]}, y[3] = {a.y[0], a.y[1], a.y[2]};
  // jump by first = t * kBlockLen positions
  for (int j = 0; j < 63; ++j) {
    if (((unsigned long long)first >> j) & 1ULL) {
      mat_vec(kJump.x[j], x, kC1, kM1);
      mat_vec(kJump.y[j], y, kC2, kM2);
    }
  }
  const long long last = min(a.n, first + kBlockLen);
  const double inv_m1 = 1.0 / 4294967087.0;
  for (long long i = first; i < last; ++i) {
    // -a x mod m = a (m - x): every product stays below 2^53
    const uint64_t xn = fold(1403580ULL * x[1] + 810728ULL * (kM1 - x[0]), kC1, kM1);
    const uint64_t yn = fold(527612ULL * y[2] + 1370589ULL * (kM2 - y[0]), kC2, kM2);
    x[0] = x[1];
    x[1] = x[2];
    x[2] = xn;
    y[0] = y[1];
    y[1] = y[2];
    y[2] = yn;
    const uint64_t z = xn >= yn ? xn - yn : xn + kM1 - yn;
    const double u = z ? (double)z * inv_m1 : 0.5 * inv_m1;  // z = 0 (p = 2^-32): half a step up
    a.out[i] = kNormal ? a.mean + a.stddev * icdf(u) : u;
  }
}

int mrg_validate(long long n, const void* out) {
  FSMC_REQUIRE(n >= 0, "n %lld < 0", n);
  FSMC_REQUIRE(n == 0 || out != nullptr, "NULL output");
  FSMC_REQUIRE(n <= (1LL << 40), "n %lld above 2^40", n);
  return FSMC_OK;
}

MrgArgs mrg_args(uint32_t seed, uint64_t offset, long long n, double mean, double stddev, double* out) {
  MrgArgs a;
  uint64_t x[3] = {seed % kM1, 1, 1}, y[3] = {1, 1, 1};
  for (int j = 0; j < 64; ++j) {
    if ((offset >> j) & 1ULL) {
      mat_vec(kJumpHost.x[j], x, kC1, kM1);
      mat_vec(kJumpHost.y[j], y, kC2, kM2);
    }
  }
  for (int i = 0; i < 3; ++i) {
    a.x[i] = x[i];
    a.y[i] = y[i];
  }
  a.n = n;
  a.mean = mean;
  a.stddev = stddev;
  a.out = out;
  return a;
}

int mrg_enqueue(bool normal, const MrgArgs& a, hipStream_t s) {
  if (a.n == 0) return FSMC_OK;
  const long long threads = (a.n + kBlockLen - 1) / kBlockLen;
  const dim3 grid((unsigned)((threads + kThreads - 1) / kThreads));
  if (normal)
    hipLaunchKernelGGL(mrg_kernel<true>, grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(mrg_kernel<false>, grid, dim3(kThreads), 0, s, a);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int mrg_sync(fsmc_ctx* ctx, bool normal, uint32_t seed, uint64_t offset, long long n, double mean, double stddev,
             double* host_out) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(mrg_validate(n, host_out));
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t bytes = (size_t)n * sizeof(double);
  FSMC_TRY(ctx_prepare(ctx, 0, bytes, 0, 0));
  hipStream_t s = ctx->stream;
  FSMC_TRY(mrg_enqueue(normal, mrg_args(seed, offset, n, mean, stddev, (double*)ctx->normals.ptr), s));
  if (n) FSMC_HIP(hipMemcpyAsync(host_out, ctx->normals.ptr, bytes, hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // namespace

int mrg_gaussian_enqueue(uint32_t seed, uint64_t offset, long long n, double mean, double stddev, double* d_out,
                         hipStream_t s) {
  FSMC_TRY(mrg_validate(n, d_out));
  return mrg_enqueue(true, mrg_args(seed, offset, n, mean, stddev, d_out), s);
}

}  // namespace fsmc

using namespace fsmc;

extern "C" {

int fsmc_mrg32k3a_uniform(fsmc_ctx* ctx, uint32_t seed, uint64_t offset, long long n, double* u) {
  return mrg_sync(ctx, false, seed, offset, n, 0.0, 1.0, u);
}

int fsmc_mrg32k3a_gaussian(fsmc_ctx* ctx, uint32_t seed, uint64_t offset, long long n, double mean, double stddev,
                           double* z) {
  return mrg_sync(ctx, true, seed, offset, n, mean, stddev, z);
}

int fsmc_mrg32k3a_gaussian_async(uint32_t seed, uint64_t offset, long long n, double mean, double stddev,
                                 double* d_z, void* stream) {
  FSMC_TRY(mrg_validate(n, d_z));
  return mrg_enqueue(true, mrg_args(seed, offset, n, mean, stddev, d_z), (hipStream_t)stream);
}

}  // extern "C"
