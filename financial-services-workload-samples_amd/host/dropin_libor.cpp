// Drop-in liborSwaptionGreeks / generateRand (LiborMarketModel sample) on the
// engine's LIBOR kernels.  See include/fsmc_dropin.hpp for the contract.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "dropin_common.hpp"
#include "fsmc_dropin.hpp"

namespace {

constexpr int kNmat = 40;   // normals per path, params.h:28
constexpr int kLoops = 10;  // NLOOPS, liborSwaptionGreeks.cpp:27

// Fills num doubles with generateRand's stream: MRG32k3a seeded with seed,
// inverse-CDF normals, positions 0 .. num-1 (mklRng.cpp:44-139).
bool fill_normals(std::vector<double>& out, int num, double mean, double stddev, int seed) {
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx || num <= 0) return false;
  out.resize((size_t)num);
  const int st = fsmc_mrg32k3a_gaussian(ctx, (uint32_t)seed, 0, num, mean, stddev, out.data());
  if (st != FSMC_OK) {
    fsmc_dropin::report("fsmc_mrg32k3a_gaussian", st);
    return false;
  }
  return true;
}

template <typename Real>
void generate(Real* z, int num, Real mean, Real stddev, int seed) {
  std::vector<double> buf;
  if (!fill_normals(buf, num, (double)mean, (double)stddev, seed)) {
    for (int i = 0; i < num; ++i) z[i] = std::numeric_limits<Real>::quiet_NaN();
    return;
  }
  for (int i = 0; i < num; ++i) z[i] = (Real)buf[i];
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

template <typename Real>
void greeks(Real* z, int num, Real mean, Real std_dev, Real* d_v, Real* d_Lb, int numPaths, int precision) {
  const Real nan = std::numeric_limits<Real>::quiet_NaN();
  *d_v = *d_Lb = nan;
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx || numPaths < 2 || num < kNmat * numPaths) {
    std::fprintf(stderr, "fsmc: liborSwaptionGreeks needs a device, numPaths >= 2 and num >= 40*numPaths\n");
    return;
  }
  auto t = std::chrono::steady_clock::now();
  generate(z, num, mean, std_dev, 12345);  // the sample ignores its seed argument
  const double rngtime = ms_since(t);
  std::vector<double> zd((size_t)kNmat * numPaths);
  for (size_t i = 0; i < zd.size(); ++i) zd[i] = (double)z[i];
  t = std::chrono::steady_clock::now();
  // lambda = 0.2 and L0 = 0.051, liborSwaptionGreeks.cpp:327-329
  fsmc_libor_params p = {numPaths, 0.051, 0.2};
  fsmc_mode m = {precision, FSMC_RNG_BUFFER, 0u, 0u, 0u};
  double sumv = 0.0, sumlb = 0.0, value[2], deltas[160];
  for (int loop = 0; loop < kLoops; ++loop) {
    const int st = fsmc_libor_swaption_greeks(ctx, &p, &m, zd.data(), value, deltas);
    if (st != FSMC_OK) {
      fsmc_dropin::report("fsmc_libor_swaption_greeks", st);
      return;
    }
    sumv += value[0];
    sumlb += deltas[79];
  }
  *d_v = (Real)sumv;
  *d_Lb = (Real)sumlb;
  const double proctime = ms_since(t);
  std::printf("Vector setup time: 0\nRNG time         : %g\nInit time        : 0\nProc time        : %g\n"
              "Number of Loops    : %d\n",
              rngtime, proctime, kLoops);
}

}  // namespace

void generateRand(double* z, int num, double mean, double stddev, int seed) { generate(z, num, mean, stddev, seed); }
void generateRand(float* z, int num, float mean, float stddev, int seed) { generate(z, num, mean, stddev, seed); }

void liborSwaptionGreeks(double* z, int num, double mean, double std_dev, int, double* d_v, double* d_Lb,
                         int numPaths) {
  greeks(z, num, mean, std_dev, d_v, d_Lb, numPaths, FSMC_FP64);
}

void liborSwaptionGreeks(float* z, int num, float mean, float std_dev, int, float* d_v, float* d_Lb,
                         int numPaths) {
  greeks(z, num, mean, std_dev, d_v, d_Lb, numPaths, FSMC_FP32);
}
