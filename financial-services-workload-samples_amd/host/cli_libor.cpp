// liborSwaptionGreeks command line, output-compatible with the sample's main
// (LiborMarketModel/main.cpp:33-97) so that its MPTest scripts parse GPU runs
// unchanged: argv[1] = number of paths (default 131072); prints
// "Result: value=<v>,\tlambda=<l>", "Total time: <ms> ms" and "Paths/sec".
// Build-time choice of REAL as in the sample: -DUSE_DOUBLE.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <vector>

#include "fsmc_dropin.hpp"

#ifdef USE_DOUBLE
typedef double REAL;
#else
typedef float REAL;
#endif

int main(int argc, char** argv) {
  int numPath = 131072;
  if (argc > 1) numPath = std::atoi(argv[1]);
  std::cout << "Number of paths: " << numPath << std::endl;
  const int num = 40 * numPath;
  std::vector<REAL> z(num);
  generateRand(z.data(), num, REAL(0), REAL(1));
  std::cout << "Running... " << std::endl;
  REAL result = 0, Lb = 0;
  const auto t0 = std::chrono::steady_clock::now();
  liborSwaptionGreeks(z.data(), num, REAL(0), REAL(1), 1234, &result, &Lb, numPath);
  const double elapsed = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "Result: value=" << result << ",\tlambda=" << Lb << std::endl;
  std::cout << "Total time: " << elapsed << " ms" << std::endl;
  std::cout << "Paths/sec: " << numPath / elapsed << " KPaths/sec " << std::endl;
  return (result == result) ? EXIT_SUCCESS : EXIT_FAILURE;
}
