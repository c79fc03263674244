// MonteCarloInsideBlocking command line on the GPU engine, output-compatible
// with the sample (MonteCarloEuropeanOptions/MonteCarloInsideBlockingDP.cpp:
// argv at :141-177, banner at :180-185, results at :306-322) so that its
// MPTest scripts (which grep "Elapsed") parse GPU runs unchanged:
//
//   european_dp <nthreads> <noptions> <path_length> <path_block_length> [verbose]
//
// Options are drawn exactly as the sample draws them: "thread" t of
// <nthreads> takes noptions / nthreads options from glibc rand_r seeded with
// 777 + t, S ~ U[5, 50], X ~ U[10, 25], T ~ U[1, 5] (:45-48, :208-226), with
// r = 0.06 and sigma = 0.10 (:71-72).  All of them are priced in one call of
// the drop-in price_european with path_length paths each, on the sample's own
// normals: thread t's options draw in order from the SFMT19937 stream of seed
// 777 + t (:229-249), generated on the GPU (the block length is validated as
// in the sample; with inverse-CDF normals it does not change the stream).  As in the sample the batch is priced 4 times and the last pass
// is timed.  With verbose != 0 the Black-Scholes comparison of :275-316 is
// printed.  -DUSE_DOUBLE selects REAL = double and the FP64 engine path,
// otherwise REAL = float and FP32.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fsmc_dropin.hpp"

#ifdef USE_DOUBLE
typedef double tfloat;
static const char* kName = "Double";
#else
typedef float tfloat;
static const char* kName = "Single";
#endif

static const double kRiskfree = 0.06, kVolatility = 0.10;
static const int kIters = 4;

static void die(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  exit(EXIT_FAILURE);
}

// Size suffixes of the sample: k = 1024, K = 1000, m = 2^20, M = 1e6, g = 2^30, G = 1e9.
static long long suffixed(const char* s) {
  const size_t n = strlen(s);
  long long mul = 1;
  switch (n ? s[n - 1] : 0) {
    case 'k': mul = 1024; break;
    case 'K': mul = 1000; break;
    case 'm': mul = 1024 * 1024; break;
    case 'M': mul = 1000000; break;
    case 'g': mul = 1024LL * 1024 * 1024; break;
    case 'G': mul = 1000000000; break;
    default: break;
  }
  return mul * atoll(s);
}

static double uniform(double lo, double hi, unsigned* seed) {
  const double t = (double)rand_r(seed) / (double)RAND_MAX;
  return (1.0 - t) * lo + t * hi;
}

static double bs_call(double S, double X, double T, double r, double v) {
  const double sq = v * sqrt(T);
  const double d1 = (log(S / X) + (r + 0.5 * v * v) * T) / sq;
  const double d2 = d1 - sq;
  return S * 0.5 * erfc(-d1 / sqrt(2.0)) - X * exp(-r * T) * 0.5 * erfc(-d2 / sqrt(2.0));
}

int main(int argc, char** argv) {
  if (argc < 5)
    die("Usage: %s <nthreads> <noptions> <path_length> <path_block_length> <verbose(default=0)> \n", argv[0]);
  const long long nthreads = suffixed(argv[1]);
  if (nthreads < 1) die("NTHREADS < 1: (%lld)\n", nthreads);
  if (nthreads > 288) die("NTHREADS > MAX_THREADS: (%lld)\n", nthreads);
  if (nthreads % 4 != 0 && nthreads != 1) die("NTHREADS really ought to be a multiple of 4! (%lld)\n", nthreads);
  const long long opt_n = suffixed(argv[2]);
  const long long per_thread = opt_n / nthreads;
  if (per_thread < 1) die("OPTS_N should be at least 1! (%lld)\n", opt_n);
  const long long rand_n = suffixed(argv[3]);
  if (rand_n < 16) die("RAND_N should be at least 16 for SIMD's sake! (%lld)\n", rand_n);
  if (rand_n > 0x7fffffff) die("RAND_N above 2^31 - 1 (%lld)\n", rand_n);
  const long long block = suffixed(argv[4]);
  if (block < 16) die("RAND_BLOCK_LENGTH should be at least 16! (%lld)\n", block);
  if (block > rand_n) die("RAND_BLOCK_LENGTH should be no more than RAND_N(=%lld)! (%lld)\n", rand_n, block);
  if (block % 16 != 0) die("RAND_BLOCK_LENGTH should divisibly by 16! (/16 = %lf)\n", block / 16.0);
  if (rand_n % block != 0)
    die("RAND_BLOCK_LENGTH should evenly divide RAND_N! (RAND_N/BLOCK_LENGTH = %lf)\n", rand_n / (double)block);
  const int verbose = argc >= 6 ? (int)suffixed(argv[5]) : 0;

  printf("%s Precision Monte-Carlo European Option Pricing, version 1.2 (MI355X engine)\n\n", kName);
  printf("Path Length      = %lld\n", rand_n);
  printf("Number of Options= %lld\n", opt_n);
  printf("Block Size       = %lld\n", block);
  printf("Worker Threads   = %lld\n\n", nthreads);

  const size_t n = (size_t)(per_thread * nthreads);
  std::vector<tfloat> S(n), X(n), T(n), res(n), conf(n);
  for (long long t = 0; t < nthreads; ++t) {
    unsigned seed = 777u + (unsigned)t;
    for (long long i = 0; i < per_thread; ++i) {
      const size_t k = (size_t)(t * per_thread + i);
      S[k] = (tfloat)uniform(5.0, 50.0, &seed);
      X[k] = (tfloat)uniform(10.0, 25.0, &seed);
      T[k] = (tfloat)uniform(1.0, 5.0, &seed);
    }
  }
  // each sample thread's options draw from its own SFMT19937 stream (777 + thread)
  fsmc_dropin::european_config().threads = (int)nthreads;
  double elapsed = 0.0;
  for (int it = 0; it < kIters; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    price_european(S.data(), X.data(), T.data(), (tfloat)kRiskfree, (tfloat)kVolatility, (int)n, (int)rand_n,
                   (int)block, res.data(), conf.data(), 777);
    elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (res[0] != res[0]) die("price_european failed\n");
  }

  if (verbose) {
    double sum_delta = 0.0, sum_ref = 0.0, reserve = 0.0, max_delta = 0.0;
    for (size_t i = 0; i < n; ++i) {
      const double ref = bs_call(S[i], X[i], T[i], kRiskfree, kVolatility);
      const double delta = fabs(ref - res[i]);
      sum_delta += delta;
      sum_ref += fabs(ref);
      if (delta > 1e-6) reserve += conf[i] / delta;
      if (delta > max_delta) max_delta = delta;
    }
    reserve /= (double)opt_n;
    printf("L1_Norm          = %E\n", sum_delta / sum_ref);
    printf("Average RESERVE  = %f\n", reserve);
    printf("Max Error        = %E\n", max_delta);
    printf(reserve > 1.0 ? "Test passed\n" : "Test failed!\n");
  }
  printf("==========================================\n");
  printf("Time Elapsed = %lf\n", elapsed);
  printf("Opt/sec      = %lf\n", opt_n / elapsed);
  printf("==========================================\n");
  return 0;
}
