// Drop-in price_european (MonteCarloEuropeanOptions sample) on the engine's
// GBM kernels.  See include/fsmc_dropin.hpp for the contract.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "dropin_common.hpp"
#include "fsmc_dropin.hpp"

namespace fsmc_dropin {

EuropeanConfig& european_config() {
  static EuropeanConfig c;
  return c;
}

}  // namespace fsmc_dropin

namespace {

// Device buffers of one sample-stream pricing, released on every return.
struct EuropeanBuffers {
  double *in = nullptr, *out = nullptr;
  float* samples = nullptr;
  void* ws = nullptr;
  void* skip = nullptr;  // SFMT19937 skip-ahead workspace
  hipStream_t s = nullptr;
  ~EuropeanBuffers() {
    if (s) (void)hipStreamSynchronize(s);
    if (s) (void)hipStreamDestroy(s);
    for (void* p : {(void*)in, (void*)out, (void*)samples, ws, skip})
      if (p) (void)hipFree(p);
  }
};

bool hip_ok(hipError_t e, const char* where) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "fsmc: %s failed: %s\n", where, hipGetErrorString(e));
  return false;
}

bool fsmc_ok(int st, const char* where) {
  if (st == FSMC_OK) return true;
  fsmc_dropin::report(where, st);
  return false;
}

// The sample's own normals (MonteCarloInsideBlockingDP.cpp:229-249): thread t
// of `threads` prices options [t n / threads, (t + 1) n / threads) from the
// SFMT19937 stream seed + t, drawing each option's paths in order as
// s = MuByT + VBySqrtT z; then the GBM kernel in buffer mode on those samples.
bool price_sample_stream(const fsmc_gbm_params& p, int precision, unsigned long long seed, int threads,
                         const double* S, const double* X, const double* T, double* res, double* conf) {
  const int n = p.n_options;
  if (threads < 1 || n % threads != 0) {
    std::fprintf(stderr, "fsmc: price_european: %d options are not a multiple of %d sample threads\n", n, threads);
    return false;
  }
  if (!fsmc_dropin::context()) return false;
  // MonteCarloInsideBlockingDP.cpp:71-76, 239-243: log2-domain drift and scale per option
  const double l2e = 1.4426950408889634074;
  std::vector<double> host(5 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    host[i] = S[i];
    host[n + i] = X[i];
    host[2 * n + i] = T[i];
    host[3 * n + i] = l2e * (p.riskfree - 0.5 * p.volatility * p.volatility) * T[i];
    host[4 * n + i] = l2e * p.volatility * std::sqrt(T[i]);
  }
  const fsmc_mode m = {precision, FSMC_RNG_BUFFER, 0, 0, 0};
  const size_t ws_bytes = fsmc_gbm_european_workspace(&p, &m);
  if (ws_bytes == 0) return fsmc_ok(FSMC_ERR_INVALID_ARG, "fsmc_gbm_european_workspace");
  EuropeanBuffers b;
  const size_t z = (size_t)n * (size_t)p.n_paths;
  if (!hip_ok(hipMalloc((void**)&b.in, host.size() * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.out, 2 * (size_t)n * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.samples, z * sizeof(float)), "hipMalloc") ||
      !hip_ok(hipMalloc(&b.ws, ws_bytes), "hipMalloc") || !hip_ok(hipStreamCreate(&b.s), "hipStreamCreate") ||
      !hip_ok(hipMemcpyAsync(b.in, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, b.s),
              "hipMemcpyAsync"))
    return false;
  const double* d = b.in;
  // each sample thread's stream spread over substreams started by skip-ahead
  const size_t skip_bytes = fsmc_sfmt19937_workspace(threads, (long long)(z / threads));
  if (skip_bytes && !hip_ok(hipMalloc(&b.skip, skip_bytes), "hipMalloc")) return false;
  if (!fsmc_ok(fsmc_sfmt19937_gaussian_skip_async((uint32_t)seed, threads, (long long)(z / threads), p.n_paths,
                                                  d + 3 * n, d + 4 * n, b.samples, b.skip, skip_bytes, b.s),
               "fsmc_sfmt19937_gaussian_skip_async") ||
      !fsmc_ok(fsmc_gbm_european_async(&p, &m, d, d + n, d + 2 * n, b.samples, b.out, b.out + n, nullptr, b.ws,
                                       ws_bytes, b.s),
               "fsmc_gbm_european_async") ||
      !hip_ok(hipMemcpyAsync(res, b.out, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, b.s),
              "hipMemcpyAsync") ||
      !hip_ok(hipMemcpyAsync(conf, b.out + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, b.s),
              "hipMemcpyAsync") ||
      !hip_ok(hipStreamSynchronize(b.s), "hipStreamSynchronize"))
    return false;
  return true;
}

template <typename Real>
void european(const Real* S, const Real* X, const Real* T, Real r, Real sigma, int nOpt, int nPaths, int blockLen,
              Real* CallResult, Real* CallConfidence, unsigned long long seed) {
  for (int i = 0; i < nOpt; ++i) CallResult[i] = CallConfidence[i] = std::numeric_limits<Real>::quiet_NaN();
  // the sample's checks of RAND_BLOCK_LENGTH (MonteCarloInsideBlockingDP.cpp:165-174)
  if (nPaths < 16 || blockLen < 16 || blockLen % 16 != 0 || blockLen > nPaths || nPaths % blockLen != 0) {
    std::fprintf(stderr,
                 "fsmc: price_european: blockLen %d must be a multiple of 16, at least 16, at most and "
                 "dividing nPaths %d\n",
                 blockLen, nPaths);
    return;
  }
  std::vector<double> s(nOpt), x(nOpt), t(nOpt), res(nOpt), conf(nOpt);
  for (int i = 0; i < nOpt; ++i) {
    s[i] = (double)S[i];
    x[i] = (double)X[i];
    t[i] = (double)T[i];
  }
  const fsmc_gbm_params p = {nOpt, nPaths, (double)r, (double)sigma};
  const int precision = sizeof(Real) == 8 ? FSMC_FP64 : FSMC_FP32;
  const fsmc_dropin::EuropeanConfig& c = fsmc_dropin::european_config();
  if (c.sample_stream) {
    if (!price_sample_stream(p, precision, seed, c.threads, s.data(), x.data(), t.data(), res.data(),
                             conf.data()))
      return;
  } else {
    fsmc_ctx* ctx = fsmc_dropin::context();
    if (!ctx) return;
    const fsmc_mode m = {precision, FSMC_RNG_PHILOX, seed, 0, 0};
    const int st = fsmc_gbm_european(ctx, &p, &m, s.data(), x.data(), t.data(), nullptr, res.data(), conf.data(),
                                     nullptr);
    if (st != FSMC_OK) {
      fsmc_dropin::report("fsmc_gbm_european", st);
      return;
    }
  }
  for (int i = 0; i < nOpt; ++i) {
    CallResult[i] = (Real)res[i];
    CallConfidence[i] = (Real)conf[i];
  }
}

}  // namespace

void price_european(const double* S, const double* X, const double* T, double r, double sigma, int nOpt,
                    int nPaths, int blockLen, double* CallResult, double* CallConfidence, unsigned long long seed) {
  european(S, X, T, r, sigma, nOpt, nPaths, blockLen, CallResult, CallConfidence, seed);
}

void price_european(const float* S, const float* X, const float* T, float r, float sigma, int nOpt, int nPaths,
                    int blockLen, float* CallResult, float* CallConfidence, unsigned long long seed) {
  european(S, X, T, r, sigma, nOpt, nPaths, blockLen, CallResult, CallConfidence, seed);
}
