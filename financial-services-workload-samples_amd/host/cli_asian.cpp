// AsianOptionsPricing command line on the GPU engine: the sample's main
// (AsianOptionsPricing/MCAsianOptions.cpp:217-290) over the drop-in
// MonteCarloTimeStep, which prints the sample's report (:208-212, the rate
// also on stderr).  Its MPTest getresults.sh greps the rate line with three
// spaces after "-", which "%8.4f" produces only for rates below 10
// options/s; GPU rates need the pattern without them (awk's $8 is still the
// rate).  Defaults are the sample's: 100 options, 131072 paths, 128 steps,
// r = 0.03, rho = -0.65 (:23-43), parameters drawn with the process's
// unseeded std::rand() in the sample's order S, X, T, v0, kappa, theta, xi
// per option (:254-266), two pricings (NUM_LOOPS) on the sample's two
// MKL-compatible MT19937 streams (seeds 42 and 43) with the sample's
// arithmetic (variance as the volatility, running val2).
// Optional arguments (not the sample's): options paths steps [prices]; with
// "prices" each option's price and confidence are printed after the report.
// -DUSE_DOUBLE selects the FP64 engine path (the double overload), otherwise
// FP32 (the float overload).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fsmc_dropin.hpp"

#ifdef USE_DOUBLE
typedef double tfloat;
#else
typedef float tfloat;
#endif

int main(int argc, char** argv) {
  fsmc_dropin::AsianConfig& c = fsmc_dropin::asian_config();
  if (argc >= 4) {
    c.opt_n = atoi(argv[1]);
    c.rand_n = atoi(argv[2]);
    c.simsteps = atoi(argv[3]);
  }
  const bool print_prices = argc >= 5 && strcmp(argv[4], "prices") == 0;
  if (c.opt_n < 1 || c.rand_n < 64 || c.rand_n % 64 != 0 || c.simsteps < 1) {
    fprintf(stderr, "usage: %s [options paths steps [prices]] (paths a multiple of 64)\n", argv[0]);
    return EXIT_FAILURE;
  }
  const int n = c.opt_n;
  auto get_rand = [](double lo, double hi) {
    const double t = (double)std::rand() / (double)RAND_MAX;
    return (1.0 - t) * lo + t * hi;
  };
  std::vector<tfloat> res(n, 0), conf(n, -1), S(n), X(n), T(n), v0(n), kappa(n), theta(n), xi(n);
  for (int i = 0; i < n; ++i) {
    S[i] = (tfloat)get_rand(5.0, 50.0);
    X[i] = (tfloat)get_rand(10.0, 25.0);
    T[i] = (tfloat)get_rand(1.0, 5.0);
    v0[i] = (tfloat)get_rand(0.01, 0.015);
    kappa[i] = (tfloat)get_rand(6.0, 7.0);
    theta[i] = (tfloat)get_rand(0.02, 0.03);
    xi[i] = (tfloat)get_rand(0.45, 0.6);
  }
  printf("Pricing Asian Options under Heston Model ... \n");
  MonteCarloTimeStep(res.data(), conf.data(), S.data(), X.data(), T.data(), v0.data(), kappa.data(), theta.data(),
                     xi.data());
  bool finite = true;
  for (int i = 0; i < n; ++i) {
    finite = finite && res[i] == res[i] && res[i] >= 0;
    if (print_prices) printf("option %d price %.17g confidence %.17g\n", i, (double)res[i], (double)conf[i]);
  }
  return finite ? EXIT_SUCCESS : EXIT_FAILURE;
}
