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

#endif
