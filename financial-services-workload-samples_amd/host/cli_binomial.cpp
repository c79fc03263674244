// BinomialOptions command line on the GPU engine, printing the sample's
// report (BinomialOptions/binomial_cpu.cpp:41-45, 58-59, 105-106 and
// binomial_main.cpp:63-98) so that its MPTest getresults.sh (grep Completed,
// field 3) parses GPU runs unchanged:
//
//   binomial_dp [verbose]
//
// 48000 European calls on a 2048-step Cox-Ross-Rubinstein lattice
// (binomial.h:35-39) with the sample's float constants r = 0.06f and
// sigma = 0.10f promoted to double; options drawn as the sample draws them
// with the unseeded std::rand(): all S ~ U[5, 50], then all X ~ U[10, 25],
// then all T ~ U[1, 5] (binomial_main.cpp:112-114).  Verbose prints the
// sample's Black-Scholes check (L1 norm, TEST PASSED below 5e-4).
// -DUSE_DOUBLE selects the FP64 engine path, otherwise FP32.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fsmc.h"

#ifdef USE_DOUBLE
static const int kPrecision = FSMC_FP64;
static const char* kName = "Double";
#else
static const int kPrecision = FSMC_FP32;
static const char* kName = "Single";
#endif

int main(int argc, char** argv) {
  const int opt_n = 48000, steps = 2048;
  const double r = (double)0.06f, vol = (double)0.10f;
  const bool verbose = argc > 1 && atoi(argv[1]) != 0;
  auto get_rand = [](double lo, double hi) {
    const double t = (double)std::rand() / (double)RAND_MAX;
    return (1.0 - t) * lo + t * hi;
  };
  std::vector<double> S(opt_n), X(opt_n), T(opt_n), V(opt_n, vol), call(opt_n);
  for (auto& s : S) s = get_rand(5.0, 50.0);
  for (auto& x : X) x = get_rand(10.0, 25.0);
  for (auto& t : T) t = get_rand(1.0, 5.0);
  printf("%s Precision Binomial Option Pricing, version 1.1 (MI355X engine)\n", kName);
  printf("Input Dataset     = %d\n", opt_n);
  printf("\nUsing the GPU lattice kernel.\n");
  printf("Pricing %d Option with time step of %d.\n", opt_n, steps);
  fsmc_ctx* ctx = nullptr;
  if (fsmc_ctx_create(0, &ctx) != FSMC_OK) {
    fprintf(stderr, "fsmc_ctx_create: %s\n", fsmc_last_error());
    return 1;
  }
  fsmc_mode m = {kPrecision, FSMC_RNG_PHILOX, 0, 0, 0};
  const auto t0 = std::chrono::steady_clock::now();
  if (fsmc_lattice(ctx, opt_n, &m, S.data(), X.data(), T.data(), V.data(), r, steps, 0u, call.data()) != FSMC_OK) {
    fprintf(stderr, "fsmc_lattice: %s\n", fsmc_last_error());
    return 1;
  }
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fsmc_ctx_destroy(ctx);
  printf("Completed in %10.5f seconds. Options per second: %10.5f\n", sec, opt_n / sec);
  printf("Time Elapsed =  %10.5f seconds\n", sec);
  if (verbose) {
    printf("Creating the reference result...\n");
    printf("Benchmarking the reference result...\n");
    double sum_delta = 0.0, sum_ref = 0.0;
    for (int i = 0; i < opt_n; ++i) {
      const double sq = vol * sqrt(T[i]);
      const double d1 = (log(S[i] / X[i]) + (r + 0.5 * vol * vol) * T[i]) / sq;
      const double ref = S[i] * 0.5 * erfc(-d1 / sqrt(2.0)) - X[i] * exp(-r * T[i]) * 0.5 * erfc(-(d1 - sq) / sqrt(2.0));
      sum_delta += fabs(ref - call[i]);
      sum_ref += fabs(ref);
    }
    const double err = sum_ref > 1e-5 ? sum_delta / sum_ref : sum_delta / opt_n;
    printf(sum_ref > 1e-5 ? "L1 norm: %E\n" : "Avg. diff: %E\n", err);
    printf(err < 5e-4 ? "TEST PASSED\n" : "TEST FAILED\n");
  }
  return 0;
}
