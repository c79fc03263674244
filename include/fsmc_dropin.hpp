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
// normals within ~7.3e-9 relative of MKL's ICDF (fsmc_mrg32k3a_gaussian).
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

namespace fsmc_dropin {

// MonteCarloEuropeanOptions/MonteCarloInsideBlockingDP.cpp:141-148, 229 --
// the normals price_european draws.  sample_stream: the sample's own
// streams, SFMT19937 seeded seed + t for sample thread t of `threads`, which
// prices the t-th contiguous block of nOpt / threads options drawing each
// option's paths in order (MKL's seeding and inverse-CDF mapping, generated
// on the GPU); otherwise the engine's Philox stream of `seed`.
struct EuropeanConfig {
  bool sample_stream = true;
  int threads = 1;  // NTHREADS of the sample's command line; must divide nOpt
};
EuropeanConfig& european_config();

}  // namespace fsmc_dropin

// MonteCarloEuropeanOptions: the sample prices inside main
// (MonteCarloInsideBlockingDP.cpp:229-270), so this entry point is new: it
// prices nOpt GBM European calls (spot S, strike X, expiry T, rate r,
// volatility sigma) with nPaths paths each and writes the discounted mean
// payoff and the 95 % confidence half-width, the sample's CallResultList and
// CallConfidenceList (:251-268).  blockLen is validated as the sample
// validates RAND_BLOCK_LENGTH (:165-174: at least 16, a multiple of 16, at
// most nPaths, dividing nPaths); the GPU does not block (with inverse-CDF
// normals the stream order does not depend on it).  Normals as
// european_config() says (by default the sample's SFMT19937 streams of
// seed 777 + thread).  The double overload runs FP64, the float one FP32.
void price_european(const double* S, const double* X, const double* T, double r, double sigma, int nOpt,
                    int nPaths, int blockLen, double* CallResult, double* CallConfidence,
                    unsigned long long seed = 777);
void price_european(const float* S, const float* X, const float* T, float r, float sigma, int nOpt, int nPaths,
                    int blockLen, float* CallResult, float* CallConfidence, unsigned long long seed = 777);

namespace fsmc_dropin {

// MonteCarloAmericanOptions/utilities.h:23-24 and main.cpp:103 -- the shape
// perform_sims fills and its streams.  `pricing` counts perform_sims calls:
// call c reads run c of numPaths draws of every stream, as the sample's
// streams advance by numPaths per loop; each call increments it.
struct AmericanConfig {
  int num_paths = 131072 * 2 * 2;  // numPaths
  int num_steps = 256;             // numSteps
  unsigned seed = 12345;           // vslNewStream(VSL_BRNG_MRG32K3A, 12345)
  unsigned pricing = 0;            // perform_sims calls so far
};
AmericanConfig& american_config();

}  // namespace fsmc_dropin

// MonteCarloAmericanOptions/utilities.h:34 -- fills a[s][p], s = 1..numSteps,
// p < numPaths, with a[s][p] = a[s-1][p] * exp2(fwdFactor + volSqrtdt * z)
// (utilities.cpp:42-88) on the GPU.  a[0] is left as it is and must hold one
// spot price in every column (main.cpp:84-87 fills it so); otherwise nothing
// is written and an error is printed.  z is the sample's single-thread draw:
// MRG32k3a seed 12345, step s on stream s-1 skipped ahead by
// (numPaths / 1024) (s-1) (main.cpp:93-106), with ICDF normals generated on
// the GPU.  `streams` and `nthreads` are not read: MKL stream state is opaque,
// so the streams' position is american_config().pricing.
void perform_sims(double** a, void** streams, int nthreads, double fwdFactor, double volSqrtdt);
void perform_sims(float** a, void** streams, int nthreads, float fwdFactor, float volSqrtdt);

// MonteCarloAmericanOptions/utilities.h:32 -- a = x^-1 for a 3x3 matrix by
// cofactors over the determinant (utilities.cpp:90-123); prints "Trying to
// invert non invertible matrix" when |det| < 1e-8 and divides regardless.
// numRows and numCols are not read.
void invert3_arr(double a[][3], double x[][3], int numRows, int numCols);
void invert3_arr(float a[][3], float x[][3], int numRows, int numCols);

// MonteCarloAmericanOptions: the sweep is inlined in the sample's main
// (main.cpp:117-279), so this entry point is new: one pricing of the sample's
// algorithm (paths, lagged quadratic regression with ridge 0.01, cofactor
// inverse) for an American put, writing the sample's "American" and
// "European" prices (main.cpp:294-295).  Normals as perform_sims draws them
// with `pricing` loops already drawn.  The double overload runs FP64, the
// float one FP32.
void price_american_lsmc(int numPaths, int numSteps, double spot, double strike, double vol, double maturity,
                         double* american, double* european, unsigned pricing = 0);
void price_american_lsmc(int numPaths, int numSteps, float spot, float strike, float vol, float maturity,
                         float* american, float* european, unsigned pricing = 0);

#endif
