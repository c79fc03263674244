// Closed-form Black-Scholes and the Cox-Ross-Rubinstein lattice, batched on
// the GPU: the analytic step beside the Monte Carlo path (the BlackScholes
// sample, BlackScholesDP.cpp:177-201) and the Binomial sample's lattice
// (BinomialOptions/binomial_cpu.cpp:62-101), extended to puts and American
// exercise so that it prices the LSMC batches' lattice references.
//
//  black_scholes_kernel  one option per lane: d1, d2, the normal CDF (erfc,
//                        or the sample's 5-term polynomial under
//                        FSMC_QUIRK_BS_POLY_CND), call, put by parity, and
//                        the call's delta, vega, rho.  HBM-bound: 32 B in,
//                        16 B (+24 B Greeks) out per option.
//  lattice_kernel        one workgroup per option; the lattice values of the
//                        current step live in LDS (one buffer, at most 64 KB:
//                        8000 steps in FP64), node j handled by thread
//                        j mod 256, which computes all its new values into
//                        registers between two barriers; for American exercise
//                        each thread keeps the spots of its nodes in registers
//                        and moves them one step back per step
//                        (S(i-1,j) = S(i,j) u).
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int kBsBlock = 256;
constexpr int kLatBlock = 256;
constexpr int kMaxLatticeSteps = 8000;                                              // (8001 doubles) <= 64 KB of LDS
constexpr int kNodesPerThread = (kMaxLatticeSteps + 1 + kLatBlock - 1) / kLatBlock;  // 32

// Abramowitz & Stegun 26.2.17, as BlackScholesDP.cpp:84-113 evaluates it.
template <typename Real>
__device__ inline Real cnd_poly(Real x) {
  const Real ax = x < Real(0) ? -x : x;
  const Real pdf = Real(0.39894228040143270286) * (Real)exp(-(double)(ax * ax) * 0.5);
  const Real k = Real(1) / (Real(1) + Real(0.2316419) * ax);
  const Real poly = k * (Real(0.319381530) +
                         k * (Real(-0.356563782) + k * (Real(1.781477937) + k * (Real(-1.821255978) + k * Real(1.330274429)))));
  const Real c = Real(1) - pdf * poly;
  return x < Real(0) ? Real(1) - c : c;
}

template <typename Real>
__device__ inline Real cnd_exact(Real x) {
  return Real(0.5) * (Real)erfc(-(double)x * 0.70710678118654752440);
}

struct BsArgs {
  int n;
  double r;
  const double *S, *X, *T, *sigma;
  double *call, *put, *greeks;  // greeks: [n][3] (call delta, vega, rho), may be NULL
};

template <typename Real, bool kPoly>
__global__ __launch_bounds__(kBsBlock) void black_scholes_kernel(BsArgs a) {
  const int i = blockIdx.x * kBsBlock + threadIdx.x;
  if (i >= a.n) return;
  const Real S = (Real)a.S[i], X = (Real)a.X[i], T = (Real)a.T[i], v = (Real)a.sigma[i], r = (Real)a.r;
  const Real sqT = sqrt(T);
  const Real vsq = v * sqT;
  const Real d1 = (log(S / X) + (r + Real(0.5) * v * v) * T) / vsq;
  const Real d2 = d1 - vsq;
  const Real n1 = kPoly ? cnd_poly(d1) : cnd_exact(d1);
  const Real n2 = kPoly ? cnd_poly(d2) : cnd_exact(d2);
  const Real xe = X * exp(-r * T);
  const Real call = S * n1 - xe * n2;
  a.call[i] = (double)call;
  if (a.put) a.put[i] = (double)(call + xe - S);
  if (a.greeks) {
    const Real pdf1 = Real(0.39894228040143270286) * exp(Real(-0.5) * d1 * d1);
    a.greeks[3 * (size_t)i + 0] = (double)n1;
    a.greeks[3 * (size_t)i + 1] = (double)(S * pdf1 * sqT);
    a.greeks[3 * (size_t)i + 2] = (double)(xe * T * n2);
  }
}

struct LatArgs {
  int n, steps;
  double r;
  const double *S, *X, *T, *sigma;
  double* price;
};

// Backward induction on the CRR lattice with u = e^{sigma sqrt(dt)}, d = 1/u,
// pu = (e^{r dt} - d) / (u - d), discount e^{-r dt} (binomial_cpu.cpp:66-99).
// Leaves S u^{2j - N}; node (i, j) holds max(continuation, exercise) when
// American.
template <typename Real, bool kAmerican, bool kPut>
__global__ __launch_bounds__(kLatBlock) void lattice_kernel(LatArgs a) {
  extern __shared__ unsigned char lat_lds[];
  Real* V = (Real*)lat_lds;
  const int o = blockIdx.x;
  const int N = a.steps;
  const double Sd = a.S[o], Xd = a.X[o], Td = a.T[o], vol = a.sigma[o];
  const double dt = Td / N;
  const double vDt = vol * sqrt(dt);
  const double rDt = a.r * dt;
  const double u = exp(vDt), d = exp(-vDt);
  const double pu = (exp(rDt) - d) / (u - d);
  const double Df = exp(-rDt);
  const Real puDf = (Real)(pu * Df), pdDf = (Real)((1.0 - pu) * Df), X = (Real)Xd, ur = (Real)u;
  Real Sj[kNodesPerThread], nv[kNodesPerThread];
#pragma unroll
  for (int k = 0; k < kNodesPerThread; ++k) {
    const int j = threadIdx.x + k * kLatBlock;
    if (j <= N) {
      const Real s = (Real)(Sd * exp(vDt * (double)(2 * j - N)));
      Sj[k] = s;
      const Real pay = kPut ? X - s : s - X;
      V[j] = pay > Real(0) ? pay : Real(0);
    }
  }
  __syncthreads();
  for (int i = N - 1; i >= 0; --i) {  // step i has nodes 0..i
#pragma unroll
    for (int k = 0; k < kNodesPerThread; ++k) {
      const int j = threadIdx.x + k * kLatBlock;
      if (j <= i) {
        Real v = puDf * V[j + 1] + pdDf * V[j];
        if (kAmerican) {
          const Real s = Sj[k] * ur;  // S(i, j) = S(i+1, j) u
          Sj[k] = s;
          const Real ex = kPut ? X - s : s - X;
          v = v > ex ? v : ex;
        }
        nv[k] = v;
      }
    }
    __syncthreads();  // every read of step i+1 done
#pragma unroll
    for (int k = 0; k < kNodesPerThread; ++k) {
      const int j = threadIdx.x + k * kLatBlock;
      if (j <= i) V[j] = nv[k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) a.price[o] = (double)V[0];
}

int bs_validate(int n, const fsmc_mode* m) {
  FSMC_TRY(validate_mode(m, false));
  FSMC_REQUIRE(n >= 1, "n %d < 1", n);
  FSMC_REQUIRE((m->quirks & ~FSMC_QUIRK_BS_POLY_CND) == 0, "quirks 0x%x not for Black-Scholes", m->quirks);
  return FSMC_OK;
}

int lattice_validate(int n, const fsmc_mode* m, int steps, unsigned flags) {
  FSMC_TRY(validate_mode(m, false));
  FSMC_REQUIRE(m->quirks == 0, "quirks are not supported by the lattice");
  FSMC_REQUIRE(n >= 1, "n %d < 1", n);
  FSMC_REQUIRE(steps >= 1 && steps <= kMaxLatticeSteps, "steps %d outside [1, %d]", steps, kMaxLatticeSteps);
  FSMC_REQUIRE((flags & ~(FSMC_LATTICE_AMERICAN | FSMC_LATTICE_PUT)) == 0, "unknown lattice flags 0x%x", flags);
  return FSMC_OK;
}

template <typename Real>
void bs_launch(const BsArgs& a, bool poly, hipStream_t s) {
  const dim3 grid((a.n + kBsBlock - 1) / kBsBlock);
  if (poly)
    hipLaunchKernelGGL((black_scholes_kernel<Real, true>), grid, dim3(kBsBlock), 0, s, a);
  else
    hipLaunchKernelGGL((black_scholes_kernel<Real, false>), grid, dim3(kBsBlock), 0, s, a);
}

template <typename Real>
void lattice_launch(const LatArgs& a, unsigned flags, hipStream_t s) {
  const size_t lds = (size_t)(a.steps + 1) * sizeof(Real);
  const dim3 grid(a.n);
  const bool am = flags & FSMC_LATTICE_AMERICAN, put = flags & FSMC_LATTICE_PUT;
  if (am && put)
    hipLaunchKernelGGL((lattice_kernel<Real, true, true>), grid, dim3(kLatBlock), lds, s, a);
  else if (am)
    hipLaunchKernelGGL((lattice_kernel<Real, true, false>), grid, dim3(kLatBlock), lds, s, a);
  else if (put)
    hipLaunchKernelGGL((lattice_kernel<Real, false, true>), grid, dim3(kLatBlock), lds, s, a);
  else
    hipLaunchKernelGGL((lattice_kernel<Real, false, false>), grid, dim3(kLatBlock), lds, s, a);
}

}  // namespace
}  // namespace fsmc

using namespace fsmc;

extern "C" {

int fsmc_black_scholes_async(int n, const fsmc_mode* m, const double* d_S, const double* d_X, const double* d_T,
                             const double* d_sigma, double r, double* d_call, double* d_put, double* d_greeks,
                             void* stream) {
  FSMC_TRY(bs_validate(n, m));
  FSMC_REQUIRE(d_S && d_X && d_T && d_sigma && d_call, "NULL device pointer");
  BsArgs a{n, r, d_S, d_X, d_T, d_sigma, d_call, d_put, d_greeks};
  const bool poly = (m->quirks & FSMC_QUIRK_BS_POLY_CND) != 0;
  if (m->precision == FSMC_FP64)
    bs_launch<double>(a, poly, (hipStream_t)stream);
  else
    bs_launch<float>(a, poly, (hipStream_t)stream);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int fsmc_black_scholes(fsmc_ctx* ctx, int n, const fsmc_mode* m, const double* S, const double* X, const double* T,
                       const double* sigma, double r, double* call, double* put, double* greeks) {
  FSMC_REQUIRE(ctx && S && X && T && sigma && call, "NULL argument");
  FSMC_TRY(bs_validate(n, m));
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t nb = (size_t)n * sizeof(double);
  FSMC_TRY(ctx_prepare(ctx, 4 * nb, 0, 5 * nb, 0));
  hipStream_t s = ctx->stream;
  double* in = (double*)ctx->inputs.ptr;
  double* out = (double*)ctx->outputs.ptr;
  const double* src[4] = {S, X, T, sigma};
  for (int k = 0; k < 4; ++k) FSMC_HIP(hipMemcpyAsync(in + k * (size_t)n, src[k], nb, hipMemcpyHostToDevice, s));
  FSMC_TRY(fsmc_black_scholes_async(n, m, in, in + n, in + 2 * (size_t)n, in + 3 * (size_t)n, r, out,
                                    out + n, greeks ? out + 2 * (size_t)n : nullptr, s));
  FSMC_HIP(hipMemcpyAsync(call, out, nb, hipMemcpyDeviceToHost, s));
  if (put) FSMC_HIP(hipMemcpyAsync(put, out + n, nb, hipMemcpyDeviceToHost, s));
  if (greeks) FSMC_HIP(hipMemcpyAsync(greeks, out + 2 * (size_t)n, 3 * nb, hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

int fsmc_lattice_async(int n, const fsmc_mode* m, const double* d_S, const double* d_X, const double* d_T,
                       const double* d_sigma, double r, int steps, unsigned flags, double* d_price, void* stream) {
  FSMC_TRY(lattice_validate(n, m, steps, flags));
  FSMC_REQUIRE(d_S && d_X && d_T && d_sigma && d_price, "NULL device pointer");
  LatArgs a{n, steps, r, d_S, d_X, d_T, d_sigma, d_price};
  if (m->precision == FSMC_FP64)
    lattice_launch<double>(a, flags, (hipStream_t)stream);
  else
    lattice_launch<float>(a, flags, (hipStream_t)stream);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int fsmc_lattice(fsmc_ctx* ctx, int n, const fsmc_mode* m, const double* S, const double* X, const double* T,
                 const double* sigma, double r, int steps, unsigned flags, double* price) {
  FSMC_REQUIRE(ctx && S && X && T && sigma && price, "NULL argument");
  FSMC_TRY(lattice_validate(n, m, steps, flags));
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t nb = (size_t)n * sizeof(double);
  FSMC_TRY(ctx_prepare(ctx, 4 * nb, 0, nb, 0));
  hipStream_t s = ctx->stream;
  double* in = (double*)ctx->inputs.ptr;
  const double* src[4] = {S, X, T, sigma};
  for (int k = 0; k < 4; ++k) FSMC_HIP(hipMemcpyAsync(in + k * (size_t)n, src[k], nb, hipMemcpyHostToDevice, s));
  FSMC_TRY(fsmc_lattice_async(n, m, in, in + n, in + 2 * (size_t)n, in + 3 * (size_t)n, r, steps, flags,
                              (double*)ctx->outputs.ptr, s));
  FSMC_HIP(hipMemcpyAsync(price, ctx->outputs.ptr, nb, hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"
