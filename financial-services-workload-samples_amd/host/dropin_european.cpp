// Drop-in price_european (MonteCarloEuropeanOptions sample) on the engine's
// GBM kernels.  See include/fsmc_dropin.hpp for the contract.
#include <cstdio>
#include <limits>
#include <vector>

#include "dropin_common.hpp"
#include "fsmc_dropin.hpp"

namespace {

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
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx) return;
  std::vector<double> s(nOpt), x(nOpt), t(nOpt), res(nOpt), conf(nOpt);
  for (int i = 0; i < nOpt; ++i) {
    s[i] = (double)S[i];
    x[i] = (double)X[i];
    t[i] = (double)T[i];
  }
  const fsmc_gbm_params p = {nOpt, nPaths, (double)r, (double)sigma};
  const fsmc_mode m = {sizeof(Real) == 8 ? FSMC_FP64 : FSMC_FP32, FSMC_RNG_PHILOX, seed, 0, 0};
  const int st = fsmc_gbm_european(ctx, &p, &m, s.data(), x.data(), t.data(), nullptr, res.data(), conf.data(),
                                   nullptr);
  if (st != FSMC_OK) {
    fsmc_dropin::report("fsmc_gbm_european", st);
    return;
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
