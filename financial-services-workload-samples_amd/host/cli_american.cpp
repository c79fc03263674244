// MonteCarloAmericanOptions command line on the GPU engine, output-compatible
// with the sample's report (MonteCarloAmericanOptions/main.cpp:43, :111 and
// :287-295) so that its MPTest scripts (which grep "Computation") parse GPU
// runs unchanged.  No arguments, as the sample: 524288 paths, 256 steps
// (utilities.h:24-25), spot 100, strike 110, vol 0.25, maturity 2
// (main.cpp:46-49).  Each loop runs the sample's algorithm through the
// drop-in price_american_lsmc (path matrix, lagged quadratic regression with
// ridge 0.01, cofactor inverse) on the sample's own normals: MRG32k3a seed 12345,
// stream s-1 skipped ahead by 512 (s-1) and advancing by numPaths per loop, as
// the single-thread run draws them (main.cpp:93-106, utilities.cpp:52-72),
// generated on the GPU.  Everything runs as one stream-ordered sequence, so
// the three phase timers are not separated and print 0.  50 timed loops after
// one untimed, as the sample's progLoops.
// -DUSE_DOUBLE selects REAL = double and the FP64 engine path, otherwise
// REAL = float and FP32.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>

#include "fsmc_dropin.hpp"

#ifdef USE_DOUBLE
typedef double REAL;
#else
typedef float REAL;
#endif

int main() {
  const int num_paths = 131072 * 2 * 2, num_steps = 256, loops = 50;
  std::cout << "numPaths = " << num_paths << ", numSteps = " << num_steps << std::endl;
  std::cout << "Number of Loops = " << loops << std::endl;
  REAL american = 0, european = 0;
  double compute_ms = 0.0;
  for (int l = 0; l <= loops; ++l) {
    const auto t0 = std::chrono::steady_clock::now();
    // loop l continues every stream by numPaths draws
    price_american_lsmc(num_paths, num_steps, (REAL)100, (REAL)110, (REAL)0.25, (REAL)2, &american, &european,
                        (unsigned)l);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (american != american) return EXIT_FAILURE;
    if (l >= 1) compute_ms += ms;
  }
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
