// MonteCarloAmericanOptions command line on the GPU engine, output-compatible
// with the sample's report (MonteCarloAmericanOptions/main.cpp:43, :111 and
// :287-295) so that its MPTest scripts (which grep "Computation") parse GPU
// runs unchanged.  No arguments, as the sample: 524288 paths, 256 steps
// (utilities.h:24-25), spot 100, strike 110, vol 0.25, maturity 2
// (main.cpp:46-49).  Each loop runs the sample's algorithm
// (fsmc_american_lsmc_ref: path matrix, lagged quadratic regression with ridge
// 0.01, cofactor inverse) on the sample's own normals: MRG32k3a seed 12345,
// stream s-1 skipped ahead by 512 (s-1) and advancing by numPaths per loop, as
// the single-thread run draws them (main.cpp:93-106, utilities.cpp:52-72),
// generated on the GPU.  Everything runs as one stream-ordered sequence, so
// the three phase timers are not separated and print 0.  50 timed loops after
// one untimed, as the sample's progLoops.
// -DUSE_DOUBLE selects the FP64 engine path, otherwise FP32.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>

#include "fsmc.h"

#ifdef USE_DOUBLE
static const int kPrecision = FSMC_FP64;
#else
static const int kPrecision = FSMC_FP32;
#endif

int main() {
  const int num_paths = 131072 * 2 * 2, num_steps = 256, loops = 50;
  std::cout << "numPaths = " << num_paths << ", numSteps = " << num_steps << std::endl;
  std::cout << "Number of Loops = " << loops << std::endl;
  fsmc_ctx* ctx = nullptr;
  if (fsmc_ctx_create(0, &ctx) != FSMC_OK) {
    std::cerr << "fsmc_ctx_create: " << fsmc_last_error() << std::endl;
    return EXIT_FAILURE;
  }
  fsmc_lsmc_ref_params p = {num_paths, num_steps, 100.0, 110.0, 0.25, 2.0};
  fsmc_mode m = {kPrecision, FSMC_RNG_MRG32K3A, 12345, 0, 0};
  double american = 0.0, european = 0.0, compute_ms = 0.0;
  for (int l = 0; l <= loops; ++l) {
    m.stream = (uint32_t)l;  // loop l continues every stream by numPaths draws
    const auto t0 = std::chrono::steady_clock::now();
    if (fsmc_american_lsmc_ref(ctx, &p, &m, nullptr, &american, &european) != FSMC_OK) {
      std::cerr << "fsmc_american_lsmc_ref: " << fsmc_last_error() << std::endl;
      return EXIT_FAILURE;
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (l >= 1) compute_ms += ms;
  }
  fsmc_ctx_destroy(ctx);
  const double elapsed = compute_ms / loops;
  std::cout << "init asset time = " << 0 << std::endl;
  std::cout << "reginit Time = " << 0 << std::endl;
  std::cout << "apply Time   = " << 0 << std::endl;
  std::cout << "Computation time per loop: " << elapsed << " ms" << std::endl;
  std::cout << "Paths/sec: " << num_paths / elapsed << " Kpaths/sec" << std::endl;
  std::cout << std::setprecision(9) << "European Put Option Price: " << european << std::endl;
  std::cout << std::setprecision(9) << "American Put Option Price: " << american << std::endl;
  return american == american ? EXIT_SUCCESS : EXIT_FAILURE;
}
