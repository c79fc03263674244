// AsianOptionsPricing command line on the GPU engine, printing the sample's
// MonteCarloTimeStep report in its formats (AsianOptionsPricing/MCAsianOptions.cpp:
// :204-212, and the rate on stderr).  Its MPTest getresults.sh greps the rate
// line with three spaces after "-", which "%8.4f" produces only for rates
// below 10 options/s; GPU rates need the pattern without them (awk's $8 is
// still the rate).  Defaults are the sample's: 100 options, 131072 paths,
// 128 steps, r = 0.03, rho = -0.65 (:27-43), parameters drawn with the
// process's unseeded std::rand() in the sample's order S, X, T, v0, kappa,
// theta, xi per option (:246-261), two pricings (NUM_LOOPS).  The sample's
// arithmetic is kept (variance used as the volatility, running val2: the two
// Heston-Asian quirk flags), and so is its random stream: two MKL-compatible
// MT19937 streams, seeds 42 and 43, drawn as Box-Muller normals in the
// sample's order (:92-93,134-135: per option, per 64-path chunk, 64 x steps
// normals from each stream; the second pricing continues both streams).
// Per option the engine generates that option's normals on the GPU and
// prices it in buffer mode.  "Completed RNG" is the GPU time of the
// generation, as the sample times its vdRngGaussian calls.
// Optional arguments (not the sample's): options paths steps [prices]; with
// "prices" each option's price and confidence are printed after the report.
// -DUSE_DOUBLE selects the FP64 engine path, otherwise FP32.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "fsmc.h"

#ifdef USE_DOUBLE
static const int kPrecision = FSMC_FP64;
#else
static const int kPrecision = FSMC_FP32;
#endif

#define CHECK_FSMC(call)                                            \
  do {                                                              \
    if ((call) != FSMC_OK) {                                        \
      fprintf(stderr, "%s: %s\n", #call, fsmc_last_error());        \
      return EXIT_FAILURE;                                          \
    }                                                               \
  } while (0)
#define CHECK_HIP(call)                                                          \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));                 \
      return EXIT_FAILURE;                                                       \
    }                                                                            \
  } while (0)

int main(int argc, char** argv) {
  int opt_n = 100, rand_n = 131072, steps = 128;
  const int loops = 2;
  if (argc >= 4) {
    opt_n = atoi(argv[1]);
    rand_n = atoi(argv[2]);
    steps = atoi(argv[3]);
  }
  const bool print_prices = argc >= 5 && strcmp(argv[4], "prices") == 0;
  if (opt_n < 1 || rand_n < 64 || rand_n % 64 != 0 || steps < 1) {
    fprintf(stderr, "usage: %s [options paths steps [prices]] (paths a multiple of 64)\n", argv[0]);
    return EXIT_FAILURE;
  }
  auto get_rand = [](double lo, double hi) {
    const double t = (double)std::rand() / (double)RAND_MAX;
    return (1.0 - t) * lo + t * hi;
  };
  // SoA parameters, one array each: S, X, T, v0, kappa, theta, xi
  std::vector<double> prm(7 * (size_t)opt_n);
  for (int i = 0; i < opt_n; ++i) {
    prm[0 * opt_n + i] = get_rand(5.0, 50.0);
    prm[1 * opt_n + i] = get_rand(10.0, 25.0);
    prm[2 * opt_n + i] = get_rand(1.0, 5.0);
    prm[3 * opt_n + i] = get_rand(0.01, 0.015);
    prm[4 * opt_n + i] = get_rand(6.0, 7.0);
    prm[5 * opt_n + i] = get_rand(0.02, 0.03);
    prm[6 * opt_n + i] = get_rand(0.45, 0.6);
  }
  printf("Pricing Asian Options under Heston Model ... \n");
  fsmc_ctx* ctx = nullptr;
  CHECK_FSMC(fsmc_ctx_create(0, &ctx));
  fsmc_mt19937 *g1 = nullptr, *g2 = nullptr;
  CHECK_FSMC(fsmc_mt19937_create(ctx, 42, &g1));
  CHECK_FSMC(fsmc_mt19937_create(ctx, 43, &g2));
  const fsmc_heston_params p1 = {1, rand_n, steps, 0.03, -0.65};
  const fsmc_mode m = {kPrecision, FSMC_RNG_BUFFER, 0, 0, FSMC_QUIRK_ASIAN_VAR_AS_VOL | FSMC_QUIRK_ASIAN_RUNNING_VAL2};
  const long long per_option = (long long)rand_n * steps;
  const size_t ws_bytes = fsmc_heston_asian_workspace(&p1, &m);
  if (ws_bytes == 0) {
    fprintf(stderr, "fsmc_heston_asian_workspace: %s\n", fsmc_last_error());
    return EXIT_FAILURE;
  }
  double *d_prm, *d_z1, *d_z2, *d_res, *d_conf;
  void* d_ws;
  CHECK_HIP(hipMalloc((void**)&d_prm, prm.size() * sizeof(double)));
  CHECK_HIP(hipMalloc((void**)&d_z1, per_option * sizeof(double)));
  CHECK_HIP(hipMalloc((void**)&d_z2, per_option * sizeof(double)));
  CHECK_HIP(hipMalloc((void**)&d_res, opt_n * sizeof(double)));
  CHECK_HIP(hipMalloc((void**)&d_conf, opt_n * sizeof(double)));
  CHECK_HIP(hipMalloc(&d_ws, ws_bytes));
  CHECK_HIP(hipMemcpy(d_prm, prm.data(), prm.size() * sizeof(double), hipMemcpyHostToDevice));
  hipStream_t s;
  CHECK_HIP(hipStreamCreate(&s));
  std::vector<hipEvent_t> ev(2 * (size_t)opt_n);
  for (auto& e : ev) CHECK_HIP(hipEventCreate(&e));

  double total = 0.0, rng = 0.0;
  for (int l = 0; l < loops; ++l) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int o = 0; o < opt_n; ++o) {
      CHECK_HIP(hipEventRecord(ev[2 * o], s));
      CHECK_FSMC(fsmc_mt19937_gaussian_async(g1, per_option, 0.0, 1.0, d_z1, s));
      CHECK_FSMC(fsmc_mt19937_gaussian_async(g2, per_option, 0.0, 1.0, d_z2, s));
      CHECK_HIP(hipEventRecord(ev[2 * o + 1], s));
      const double* q = d_prm + o;
      CHECK_FSMC(fsmc_heston_asian_async(&p1, &m, q, q + opt_n, q + 2 * opt_n, q + 3 * opt_n, q + 4 * opt_n,
                                         q + 5 * opt_n, q + 6 * opt_n, d_z1, d_z2, d_res + o, d_conf + o, nullptr,
                                         d_ws, ws_bytes, s));
    }
    CHECK_HIP(hipStreamSynchronize(s));
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int o = 0; o < opt_n; ++o) {
      float ms = 0.0f;
      CHECK_HIP(hipEventElapsedTime(&ms, ev[2 * o], ev[2 * o + 1]));
      rng += 1e-3 * ms;
    }
  }
  std::vector<double> res(opt_n), conf(opt_n);
  CHECK_HIP(hipMemcpy(res.data(), d_res, opt_n * sizeof(double), hipMemcpyDeviceToHost));
  CHECK_HIP(hipMemcpy(conf.data(), d_conf, opt_n * sizeof(double), hipMemcpyDeviceToHost));
  const double time = total / loops;
  printf("Completed RNG %8.4f seconds \n", rng / loops);
  printf("Completed MC by timestep in %8.4f seconds \n", time);
  printf("Computation rate MC by time Step - %8.4f.\n", opt_n / time);
  fprintf(stderr, "%8.4f\n", opt_n / time);
  bool finite = true;
  for (int i = 0; i < opt_n; ++i) {
    finite = finite && res[i] == res[i] && res[i] >= 0.0;
    if (print_prices) printf("option %d price %.17g confidence %.17g\n", i, res[i], conf[i]);
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(s);
  (void)hipFree(d_prm);
  (void)hipFree(d_z1);
  (void)hipFree(d_z2);
  (void)hipFree(d_res);
  (void)hipFree(d_conf);
  (void)hipFree(d_ws);
  fsmc_mt19937_destroy(g1);
  fsmc_mt19937_destroy(g2);
  fsmc_ctx_destroy(ctx);
  return finite ? EXIT_SUCCESS : EXIT_FAILURE;
}
