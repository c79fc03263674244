// AsianOptionsPricing command line on the GPU engine, printing the sample's
// MonteCarloTimeStep report in its formats (AsianOptionsPricing/MCAsianOptions.cpp:
// :204-208, and the rate on stderr).  Its MPTest getresults.sh greps the rate
// line with three spaces after "-", which "%8.4f" produces only for rates
// below 10 options/s; GPU rates need the pattern without them (awk's $8 is
// still the rate).  No
// arguments, as the sample: 100 options, 131072 paths, 128 steps, r = 0.03,
// rho = -0.65 (:27-43), parameters drawn with the process's unseeded
// std::rand() in the sample's order S, X, T, v0, kappa, theta, xi per option
// (:246-261).  The sample's own arithmetic is kept (variance used as the
// volatility, running val2: the two Heston-Asian quirk flags); the normals are
// the engine's Philox stream generated inside the kernel, so the "RNG" time is
// zero.  Timing is the mean of 2 pricings, as the sample's NUM_LOOPS.
// -DUSE_DOUBLE selects the FP64 engine path, otherwise FP32.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fsmc.h"

#ifdef USE_DOUBLE
static const int kPrecision = FSMC_FP64;
#else
static const int kPrecision = FSMC_FP32;
#endif

int main() {
  const int opt_n = 100, rand_n = 131072, steps = 128, loops = 2;
  auto get_rand = [](double lo, double hi) {
    const double t = (double)std::rand() / (double)RAND_MAX;
    return (1.0 - t) * lo + t * hi;
  };
  std::vector<double> S(opt_n), X(opt_n), T(opt_n), v0(opt_n), kappa(opt_n), theta(opt_n), xi(opt_n);
  std::vector<double> res(opt_n), conf(opt_n);
  for (int i = 0; i < opt_n; ++i) {
    S[i] = get_rand(5.0, 50.0);
    X[i] = get_rand(10.0, 25.0);
    T[i] = get_rand(1.0, 5.0);
    v0[i] = get_rand(0.01, 0.015);
    kappa[i] = get_rand(6.0, 7.0);
    theta[i] = get_rand(0.02, 0.03);
    xi[i] = get_rand(0.45, 0.6);
  }
  printf("Pricing Asian Options under Heston Model ... \n");
  fsmc_ctx* ctx = nullptr;
  if (fsmc_ctx_create(0, &ctx) != FSMC_OK) {
    fprintf(stderr, "fsmc_ctx_create: %s\n", fsmc_last_error());
    return EXIT_FAILURE;
  }
  fsmc_heston_params p = {opt_n, rand_n, steps, 0.03, -0.65};
  fsmc_mode m = {kPrecision, FSMC_RNG_PHILOX, 42, 0, FSMC_QUIRK_ASIAN_VAR_AS_VOL | FSMC_QUIRK_ASIAN_RUNNING_VAL2};
  double total = 0.0;
  for (int l = 0; l < loops; ++l) {
    const auto t0 = std::chrono::steady_clock::now();
    if (fsmc_heston_asian(ctx, &p, &m, S.data(), X.data(), T.data(), v0.data(), kappa.data(), theta.data(),
                          xi.data(), nullptr, nullptr, res.data(), conf.data(), nullptr) != FSMC_OK) {
      fprintf(stderr, "fsmc_heston_asian: %s\n", fsmc_last_error());
      return EXIT_FAILURE;
    }
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  fsmc_ctx_destroy(ctx);
  const double time = total / loops;
  printf("Completed RNG %8.4f seconds \n", 0.0);
  printf("Completed MC by timestep in %8.4f seconds \n", time);
  printf("Computation rate MC by time Step - %8.4f.\n", opt_n / time);
  fprintf(stderr, "%8.4f\n", opt_n / time);
  bool finite = true;
  for (int i = 0; i < opt_n; ++i) finite = finite && res[i] == res[i] && res[i] >= 0.0;
  return finite ? EXIT_SUCCESS : EXIT_FAILURE;
}
