// fsmc_dropin.hpp - C++ drop-in entry points with the reference samples' own
// names and signatures, implemented on the engine's C ABI (fsmc.h).  Each is
// overloaded for REAL = double and REAL = float, so a sample built with or
// without -DUSE_DOUBLE links against libfsmc.so instead of its CPU source.
// They run synchronously on device 0, write the caller's host buffers and
// keep the sample's observable semantics; they never exit the process (on a
// failure they print to stderr and write NaN results).
#ifndef FSMC_DROPIN_HPP_
#define FSMC_DROPIN_HPP_

// LiborMarketModel/mklRng.h:9 -- z[k] = mean + stddev * icdf(u_k), k < num,
// with u_k the MRG32k3a stream of vslNewStream(VSL_BRNG_MRG32K3A, seed)
// (mklRng.cpp:44-139), generated on the GPU: uniforms as MKL's bit for bit,
// normals within ~2.4e-9 relative of MKL's ICDF (fsmc_mrg32k3a_gaussian).
void generateRand(double* z, int num, double mean, double stddev, int seed = 12345);
void generateRand(float* z, int num, float mean, float stddev, int seed = 12345);

// LiborMarketModel/liborSwaptionGreeks.hpp:12-13 -- regenerates z (seed
// argument ignored, as liborSwaptionGreeks.cpp:319 does), prices the swaption
// portfolio for numPaths paths NLOOPS = 10 times and returns the sums divided
// by numPaths: *d_v = 10 x mean value, *d_Lb = 10 x mean dV/dL0_79
// (liborSwaptionGreeks.cpp:336-356).
void liborSwaptionGreeks(double* z, int num, double mean, double std_dev, int seed, double* d_v, double* d_Lb,
                         int numPaths);
void liborSwaptionGreeks(float* z, int num, float mean, float std_dev, int seed, float* d_v, float* d_Lb,
                         int numPaths);

namespace fsmc_dropin {

// AsianOptionsPricing/MCAsianOptions.cpp:23-43 -- the sample's compile-time
// constants, with its values; set them before a call to price other shapes.
struct AsianConfig {
  int opt_n = 100;         // OPT_N: length of every array argument
  int rand_n = 131072;     // RAND_N: paths per option, a multiple of 64 (VECLEN)
  int simsteps = 128;      // SIMSTEPS
  int num_loops = 2;       // NUM_LOOPS: pricings per call; the last one is returned
  double riskfree = 0.03;  // RISKFREE
  double rho = -0.65;      // RHO
  unsigned seed = 42;      // RANDSEED: MT19937 stream 1; stream 2 takes seed + 1
  bool refcompat = true;   // the sample's arithmetic: variance as the volatility, running val2
};
AsianConfig& asian_config();

}  // namespace fsmc_dropin

// AsianOptionsPricing/MCAsianOptions.cpp:76-86 -- prices asian_config().opt_n
// Heston arithmetic-average Asian calls as the sample does (:86-212): two
// MKL-compatible MT19937 streams (seeds 42, 43) created per call, drawn as
// Box-Muller normals option by option, 64-path chunk by chunk, for
// NUM_LOOPS pricings; writes the last pricing's prices and 95 % confidence
// half-widths and prints the sample's three report lines (and the rate on
// stderr).  The double overload runs the FP64 engine path, the float one FP32.
void MonteCarloTimeStep(double* h_CallResult, double* h_CallConfidence, double* S, double* X, double* T,
                        double* initialVol, double* kappaVals, double* thetaVals, double* xiVals);
void MonteCarloTimeStep(float* h_CallResult, float* h_CallConfidence, float* S, float* X, float* T,
                        float* initialVol, float* kappaVals, float* thetaVals, float* xiVals);

#endif
