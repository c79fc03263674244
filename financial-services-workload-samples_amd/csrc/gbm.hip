// GBM European call: one fused kernel generates normals (Philox + Box-Muller,
// or reads the reference's float samples), evolves the single terminal step
// S_T = S * 2^(MuByT + VBySqrtT * z) in the log2 domain exactly as
// MonteCarloInsideBlockingDP.cpp:239-257 does, and reduces payoff sums (and
// optional pathwise Greeks) per workgroup in FP64.  A second tiny kernel
// combines the per-workgroup partials in a fixed order and applies the
// estimator of MonteCarloInsideBlockingDP.cpp:266-269.
//
// Optional control variate (not in the sample): u = S_T - S e^{rT} has known
// mean 0 under the pricing measure, so with y = (S_T - K)^+ the estimator
// e^{-rT} (mean(y) - beta mean(u)), beta = cov(y, u) / var(u) from the same
// paths, has variance var(y) (1 - corr(y, u)^2) / N.  The kernel adds the
// sums of u, u^2 and u y over every path.
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"

namespace fsmc {
namespace {

constexpr int kBlock = 256;
constexpr double kLog2e = 1.4426950408889634074;

struct GbmArgs {
  int n_paths;
  double r, sigma;
  const double *S, *X, *T;
  const float* samples;  // buffer mode only
  uint32_t k0, k1, w;    // Philox key and counter word w
  double* partials;      // [n_options][gridDim.x][NV]
};

template <typename Real, bool kGreeks, bool kCV>
struct GbmAcc {
  // Payoff sum and sum of squares; with Greeks also the pathwise delta and
  // vega kernels (and their squares) and the in-the-money count (rho); with
  // the control variate the sums of u, u^2 and u y, accumulated in FP64 (an
  // FP64 add issues at the FP32 rate here: one conversion and one add per
  // term cost less than a compensated FP32 sum).
  CompSum<Real> v0, v1, dl, dl2, vg, vg2;
  PathSum cu, cuu, cuy;
  uint32_t itm = 0;
  Real F = 0;  // S e^{rT}, the mean of S_T
  __device__ __forceinline__ void control(Real st, Real c) {
    if constexpr (kCV) {
      const Real u = st - F;
      cu.add((double)u);
      cuu.add((double)(u * u));
      cuy.add(c > Real(0) ? (double)(u * c) : 0.0);
    }
  }
  __device__ __forceinline__ void hit(Real c, Real st, Real dlog) {
    v0.add(c);
    v1.add(c * c);
    if constexpr (kGreeks) {
      const Real g = st * dlog;
      dl.add(st);
      dl2.add(st * st);
      vg.add(g);
      vg2.add(g * g);
      ++itm;
    }
  }
  __device__ __forceinline__ void path(Real Y, Real Z, Real mu, Real vol, Real sqrtT, Real sigT, Real z) {
    const Real st = Y * fexp2(mu + vol * z);
    const Real c = st - Z;
    if (c > Real(0)) hit(c, st, sqrtT * z - sigT);
    control(st, c);
  }
};

template <typename Real, bool kBuffer, bool kGreeks, bool kCV>
__global__ __launch_bounds__(kBlock) void gbm_paths_kernel(GbmArgs a) {
  constexpr int NG = kGreeks ? 7 : 2;  // the control-variate sums follow at NG
  constexpr int NV = NG + (kCV ? 3 : 0);
  __shared__ double lds[NV * (kBlock / 64)];
  const int o = blockIdx.y;
  const double Td = a.T[o];
  const double sqrtTd = sqrt(Td);
  const Real mu = (Real)(kLog2e * (a.r - 0.5 * a.sigma * a.sigma) * Td);
  const Real vol = (Real)(kLog2e * a.sigma * sqrtTd);
  const Real Y = (Real)a.S[o];
  const Real Z = (Real)a.X[o];
  const Real sqrtT = (Real)sqrtTd;
  const Real sigT = (Real)(a.sigma * Td);
  GbmAcc<Real, kGreeks, kCV> acc;
  acc.F = (Real)(a.S[o] * exp(a.r * Td));
  const int t = blockIdx.x * kBlock + threadIdx.x;
  const int stride = gridDim.x * kBlock;
  if (kBuffer) {
    const float* src = a.samples + (size_t)o * a.n_paths;
    for (int p = t; p < a.n_paths; p += stride) {
      // z is recovered only for the Greeks; the price uses s as given.
      const Real s = (Real)src[p];
      const Real st = Y * fexp2(s);
      const Real c = st - Z;
      if (c > Real(0)) acc.hit(c, st, kGreeks ? sqrtT * ((s - mu) / vol) - sigT : Real(0));
      acc.control(st, c);
    }
  } else {
    const int groups = (a.n_paths + 3) >> 2;
    for (int q = t; q < groups; q += stride) {
      Real z[4];
      normals4<Real>(U4{(uint32_t)q, 0u, (uint32_t)o, a.w}, a.k0, a.k1, z[0], z[1], z[2], z[3]);
      const int base = q << 2;
      if (base + 4 <= a.n_paths) {
#pragma unroll
        for (int j = 0; j < 4; ++j) acc.path(Y, Z, mu, vol, sqrtT, sigT, z[j]);
      } else {
        for (int j = 0; j < 4; ++j)
          if (base + j < a.n_paths) acc.path(Y, Z, mu, vol, sqrtT, sigT, z[j]);
      }
    }
  }
  double v[NV];
  v[0] = acc.v0.value();
  v[1] = acc.v1.value();
  if constexpr (kGreeks) {
    v[2] = acc.dl.value();
    v[3] = acc.dl2.value();
    v[4] = acc.vg.value();
    v[5] = acc.vg2.value();
    v[6] = (double)acc.itm;
  }
  if constexpr (kCV) {
    v[NG] = acc.cu.value();
    v[NG + 1] = acc.cuu.value();
    v[NG + 2] = acc.cuy.value();
  }
  block_sum<NV>(v, lds);
  if (threadIdx.x == 0) {
    double* out = a.partials + ((size_t)o * gridDim.x + blockIdx.x) * NV;
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = v[k];
  }
}

// Mean and standard error of a per-path quantity from its sum and sum of
// squares, both scaled by f.
__device__ __forceinline__ void mean_se(double s, double s2, double N, double f, double* mean, double* se) {
  *mean = f * s / N;
  const double var = (N * s2 - s * s) / (N * (N - 1.0));
  *se = fabs(f) * sqrt(var > 0.0 ? var : 0.0) / sqrt(N);
}

// One wavefront per option (workgroup of 64): the option's workgroup
// partials combined by wave_combine (fixed order).  cv: the last three
// partials are the control-variate sums.
__global__ __launch_bounds__(64) void gbm_finalize_kernel(int nbx, int nv, int cv, int n_paths, double r,
                                                         const double* S, const double* X, const double* T,
                                                         const double* partials, double* result,
                                                         double* confidence, double* greeks) {
  const int o = blockIdx.x;
  double s[10];
  wave_combine<10>(partials + (size_t)o * nbx * nv, nbx, nv, s);
  if (threadIdx.x != 0) return;
  const double N = (double)n_paths;
  // MonteCarloInsideBlockingDP.cpp:266-269.
  const double exprt = exp2(-r * kLog2e * T[o]);
  const double vy = (N * s[1] - s[0] * s[0]) * (1.0 / (N * (N - 1.0)));
  if (cv) {
    const double su = s[nv - 3], suu = s[nv - 2], suy = s[nv - 1];
    const double vu = (N * suu - su * su) * (1.0 / (N * (N - 1.0)));
    const double cov = (N * suy - su * s[0]) * (1.0 / (N * (N - 1.0)));
    const double beta = vu > 0.0 ? cov / vu : 0.0;
    const double v = vy - beta * cov;
    result[o] = exprt * (s[0] - beta * su) * (1.0 / N);
    confidence[o] = exprt * sqrt(v > 0.0 ? v : 0.0) * (1.96 / sqrt(N));
  } else {
    result[o] = exprt * s[0] * (1.0 / N);
    confidence[o] = exprt * sqrt(vy) * (1.96 / sqrt(N));
  }
  if (greeks) {
    double* g = greeks + 6 * (size_t)o;
    // delta_p = e^{-rT} 1{S_T>K} S_T / S0
    mean_se(s[2], s[3], N, exprt / S[o], g + 0, g + 3);
    // vega_p = e^{-rT} 1{S_T>K} S_T (sqrt(T) z - sigma T)
    mean_se(s[4], s[5], N, exprt, g + 1, g + 4);
    // rho_p = e^{-rT} 1{S_T>K} K T  (a Bernoulli count times K T)
    const double kt = exprt * X[o] * T[o];
    mean_se(s[6], s[6], N, kt, g + 2, g + 5);
  }
}

struct GbmPlan {
  int nbx, nv;
  size_t ws;
};

GbmPlan gbm_plan(const fsmc_gbm_params* p, const fsmc_mode* m, bool greeks) {
  GbmPlan g;
  g.nv = (greeks ? 7 : 2) + (p->control_variate ? 3 : 0);
  // About 16 paths per thread: enough to amortise the per-thread setup.
  const long long per_block = (long long)kBlock * 16;
  long long nbx = (p->n_paths + per_block - 1) / per_block;
  if (nbx < 1) nbx = 1;
  if (nbx > 65535) nbx = 65535;
  g.nbx = (int)nbx;
  g.ws = (size_t)p->n_options * g.nbx * g.nv * sizeof(double);
  (void)m;
  return g;
}

int gbm_validate(const fsmc_gbm_params* p, const fsmc_mode* m) {
  FSMC_REQUIRE(p != nullptr, "params is NULL");
  FSMC_TRY(validate_mode(m, true));
  FSMC_REQUIRE(p->n_options >= 1 && p->n_options <= 65535, "n_options %d outside [1, 65535]", p->n_options);
  FSMC_REQUIRE(p->n_paths >= 2 && p->n_paths <= (1 << 30), "n_paths %d outside [2, 2^30]", p->n_paths);
  FSMC_REQUIRE(p->volatility > 0.0, "volatility must be > 0");
  FSMC_REQUIRE(p->control_variate == 0 || p->control_variate == 1, "control_variate %d is neither 0 nor 1",
               p->control_variate);
  return FSMC_OK;
}

template <typename Real, bool B, bool G, bool C>
void launch(const GbmArgs& a, int n_options, int nbx, hipStream_t s) {
  hipLaunchKernelGGL((gbm_paths_kernel<Real, B, G, C>), dim3(nbx, n_options), dim3(kBlock), 0, s, a);
}

template <typename Real, bool B>
void launch_gc(const GbmArgs& a, bool greeks, bool cv, int n_options, int nbx, hipStream_t s) {
  if (greeks)
    cv ? launch<Real, B, true, true>(a, n_options, nbx, s) : launch<Real, B, true, false>(a, n_options, nbx, s);
  else
    cv ? launch<Real, B, false, true>(a, n_options, nbx, s) : launch<Real, B, false, false>(a, n_options, nbx, s);
}

}  // namespace
}  // namespace fsmc

using namespace fsmc;

extern "C" {

size_t fsmc_gbm_european_workspace(const fsmc_gbm_params* p, const fsmc_mode* m) {
  if (gbm_validate(p, m) != FSMC_OK) return 0;
  return gbm_plan(p, m, true).ws;
}

int fsmc_gbm_european_async(const fsmc_gbm_params* p, const fsmc_mode* m, const double* d_S, const double* d_X,
                            const double* d_T, const float* d_samples, double* d_result, double* d_conf,
                            double* d_greeks, void* d_ws, size_t ws_bytes, void* stream) {
  FSMC_TRY(gbm_validate(p, m));
  FSMC_REQUIRE(d_S && d_X && d_T && d_result && d_conf, "NULL device pointer");
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  FSMC_REQUIRE(!buffer || d_samples, "buffer mode needs samples");
  const bool greeks = d_greeks != nullptr;
  const GbmPlan g = gbm_plan(p, m, greeks);
  if (d_ws == nullptr || ws_bytes < g.ws) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, g.ws);
    return FSMC_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  GbmArgs a;
  a.n_paths = p->n_paths;
  a.r = p->riskfree;
  a.sigma = p->volatility;
  a.S = d_S;
  a.X = d_X;
  a.T = d_T;
  a.samples = d_samples;
  a.k0 = (uint32_t)m->seed;
  a.k1 = (uint32_t)(m->seed >> 32);
  a.w = counter_w(kTagGbm, m->stream);
  a.partials = (double*)d_ws;
  const bool f64 = m->precision == FSMC_FP64;
  const bool cv = p->control_variate != 0;
  if (f64) {
    if (buffer) launch_gc<double, true>(a, greeks, cv, p->n_options, g.nbx, s);
    else launch_gc<double, false>(a, greeks, cv, p->n_options, g.nbx, s);
  } else {
    if (buffer) launch_gc<float, true>(a, greeks, cv, p->n_options, g.nbx, s);
    else launch_gc<float, false>(a, greeks, cv, p->n_options, g.nbx, s);
  }
  FSMC_HIP(hipGetLastError());
  hipLaunchKernelGGL(gbm_finalize_kernel, dim3(p->n_options), dim3(64), 0, s, g.nbx, g.nv, cv ? 1 : 0,
                     p->n_paths, p->riskfree, d_S, d_X, d_T, a.partials, d_result, d_conf, d_greeks);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

int fsmc_gbm_european(fsmc_ctx* ctx, const fsmc_gbm_params* p, const fsmc_mode* m, const double* S,
                      const double* X, const double* T, const float* samples, double* call_result,
                      double* call_confidence, double* greeks) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(gbm_validate(p, m));
  FSMC_REQUIRE(S && X && T && call_result && call_confidence, "NULL host pointer");
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  FSMC_REQUIRE(!buffer || samples, "buffer mode needs samples");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t n = (size_t)p->n_options;
  const size_t in_bytes = 3 * n * sizeof(double);
  const size_t z_bytes = buffer ? n * (size_t)p->n_paths * sizeof(float) : 0;
  const size_t out_bytes = 8 * n * sizeof(double);
  const size_t ws = gbm_plan(p, m, greeks != nullptr).ws;
  FSMC_TRY(ctx_prepare(ctx, in_bytes, z_bytes, out_bytes, ws));
  double* dS = (double*)ctx->inputs.ptr;
  double* dX = dS + n;
  double* dT = dX + n;
  double* dres = (double*)ctx->outputs.ptr;
  double* dconf = dres + n;
  double* dgr = greeks ? dconf + n : nullptr;
  hipStream_t s = ctx->stream;
  FSMC_HIP(hipMemcpyAsync(dS, S, n * sizeof(double), hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(dX, X, n * sizeof(double), hipMemcpyHostToDevice, s));
  FSMC_HIP(hipMemcpyAsync(dT, T, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (buffer) FSMC_HIP(hipMemcpyAsync(ctx->normals.ptr, samples, z_bytes, hipMemcpyHostToDevice, s));
  int st = fsmc_gbm_european_async(p, m, dS, dX, dT, (const float*)ctx->normals.ptr, dres, dconf, dgr,
                                   ctx->workspace.ptr, ctx->workspace.bytes, s);
  if (st != FSMC_OK) {
    (void)hipStreamSynchronize(s);
    return st;
  }
  FSMC_HIP(hipMemcpyAsync(call_result, dres, n * sizeof(double), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipMemcpyAsync(call_confidence, dconf, n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (greeks) FSMC_HIP(hipMemcpyAsync(greeks, dgr, 6 * n * sizeof(double), hipMemcpyDeviceToHost, s));
  FSMC_HIP(hipStreamSynchronize(s));
  return FSMC_OK;
}

}  // extern "C"
