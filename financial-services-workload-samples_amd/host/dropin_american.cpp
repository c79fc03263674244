// Drop-in perform_sims, invert3_arr and price_american_lsmc
// (MonteCarloAmericanOptions sample) on the engine's LSMC kernels.  See
// include/fsmc_dropin.hpp for the contract.
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>

#include "dropin_common.hpp"
#include "fsmc_dropin.hpp"

namespace fsmc_dropin {

AmericanConfig& american_config() {
  static AmericanConfig c;
  return c;
}

}  // namespace fsmc_dropin

namespace {

template <typename Real>
void sims(Real** a, Real fwdFactor, Real volSqrtdt) {
  fsmc_dropin::AmericanConfig& c = fsmc_dropin::american_config();
  const int n = c.num_paths;
  if (!a || n < 1 || c.num_steps < 1) {
    std::fprintf(stderr, "fsmc: perform_sims: no path matrix or an empty shape (%d paths, %d steps)\n", n,
                 c.num_steps);
    return;
  }
  // the kernel starts every path at one spot, as main.cpp:84-87 fills a[0]
  const Real spot = a[0][0];
  for (int p = 1; p < n; ++p)
    if (!(a[0][p] == spot)) {
      std::fprintf(stderr, "fsmc: perform_sims: a[0][%d] = %g differs from a[0][0] = %g\n", p, (double)a[0][p],
                   (double)spot);
      return;
    }
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx) return;
  const fsmc_mode m = {sizeof(Real) == 8 ? FSMC_FP64 : FSMC_FP32, FSMC_RNG_MRG32K3A, c.seed, c.pricing, 0};
  const int st = fsmc_american_paths_ref(ctx, n, c.num_steps, (double)spot, (double)fwdFactor, (double)volSqrtdt, &m,
                                         nullptr, (void* const*)a);
  if (st != FSMC_OK) {
    fsmc_dropin::report("fsmc_american_paths_ref", st);
    return;
  }
  ++c.pricing;  // every stream has drawn numPaths more
}

// Cofactor inverse over the determinant (utilities.cpp:90-123).  The six
// determinant terms are formed in double, as det3 takes double arguments,
// and the sum is rounded to REAL before the test and the divisions, as
// determinant() returns REAL.
template <typename Real>
void invert3(Real a[][3], Real x[][3]) {
  const Real d = (Real)((double)x[0][0] * x[1][1] * x[2][2] + (double)x[1][0] * x[2][1] * x[0][2] +
                        (double)x[2][0] * x[0][1] * x[1][2] - (double)x[2][0] * x[1][1] * x[0][2] -
                        (double)x[1][0] * x[0][1] * x[2][2] - (double)x[0][0] * x[2][1] * x[1][2]);
  if (std::abs((double)d) < 1e-8) std::cout << "Trying to invert non invertible matrix" << std::endl;
  // a[i][j] = cofactor(j, i) / det
  for (int i = 0; i < 3; ++i) {
    const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; ++j) {
      const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      a[i][j] = (x[j1][i1] * x[j2][i2] - x[j2][i1] * x[j1][i2]) / d;
    }
  }
}

template <typename Real>
void american(int numPaths, int numSteps, Real spot, Real strike, Real vol, Real maturity, Real* am, Real* eu,
              unsigned pricing) {
  *am = *eu = std::numeric_limits<Real>::quiet_NaN();
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx) return;
  const fsmc_lsmc_ref_params p = {numPaths, numSteps, (double)spot, (double)strike, (double)vol, (double)maturity};
  const fsmc_mode m = {sizeof(Real) == 8 ? FSMC_FP64 : FSMC_FP32, FSMC_RNG_MRG32K3A,
                       fsmc_dropin::american_config().seed, pricing, 0};
  double a = 0.0, e = 0.0;
  const int st = fsmc_american_lsmc_ref(ctx, &p, &m, nullptr, &a, &e);
  if (st != FSMC_OK) {
    fsmc_dropin::report("fsmc_american_lsmc_ref", st);
    return;
  }
  *am = (Real)a;
  *eu = (Real)e;
}

}  // namespace

void perform_sims(double** a, void** streams, int nthreads, double fwdFactor, double volSqrtdt) {
  (void)streams;
  (void)nthreads;
  sims(a, fwdFactor, volSqrtdt);
}

void perform_sims(float** a, void** streams, int nthreads, float fwdFactor, float volSqrtdt) {
  (void)streams;
  (void)nthreads;
  sims(a, fwdFactor, volSqrtdt);
}

void invert3_arr(double a[][3], double x[][3], int numRows, int numCols) {
  (void)numRows;
  (void)numCols;
  invert3(a, x);
}

void invert3_arr(float a[][3], float x[][3], int numRows, int numCols) {
  (void)numRows;
  (void)numCols;
  invert3(a, x);
}

void price_american_lsmc(int numPaths, int numSteps, double spot, double strike, double vol, double maturity,
                         double* american_price, double* european_price, unsigned pricing) {
  american(numPaths, numSteps, spot, strike, vol, maturity, american_price, european_price, pricing);
}

void price_american_lsmc(int numPaths, int numSteps, float spot, float strike, float vol, float maturity,
                         float* american_price, float* european_price, unsigned pricing) {
  american(numPaths, numSteps, spot, strike, vol, maturity, american_price, european_price, pricing);
}
