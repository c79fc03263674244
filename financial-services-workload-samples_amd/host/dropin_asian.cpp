// Drop-in MonteCarloTimeStep (AsianOptionsPricing sample) on the engine's
// Heston kernels.  See include/fsmc_dropin.hpp for the contract.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <limits>
#include <vector>

#include "dropin_common.hpp"
#include "fsmc_dropin.hpp"

namespace fsmc_dropin {

AsianConfig& asian_config() {
  static AsianConfig c;
  return c;
}

}  // namespace fsmc_dropin

namespace {

// Device buffers of one MonteCarloTimeStep call, released on every return.
struct AsianBuffers {
  double *prm = nullptr, *z1 = nullptr, *z2 = nullptr, *res = nullptr, *conf = nullptr;
  void* ws = nullptr;
  hipStream_t s = nullptr, s2 = nullptr;  // s2 generates the second stream's normals
  std::vector<hipEvent_t> ev;
  fsmc_mt19937 *g1 = nullptr, *g2 = nullptr;
  ~AsianBuffers() {
    if (s) (void)hipStreamSynchronize(s);
    if (s2) (void)hipStreamSynchronize(s2);
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
    if (s2) (void)hipStreamDestroy(s2);
    for (void* p : {(void*)prm, (void*)z1, (void*)z2, (void*)res, (void*)conf, ws})
      if (p) (void)hipFree(p);
    fsmc_mt19937_destroy(g1);
    fsmc_mt19937_destroy(g2);
  }
};

bool hip_ok(hipError_t e, const char* where) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "fsmc: %s failed: %s\n", where, hipGetErrorString(e));
  return false;
}

bool fsmc_ok(int st, const char* where) {
  if (st == FSMC_OK) return true;
  fsmc_dropin::report(where, st);
  return false;
}

// Prices with the sample's loop structure (MCAsianOptions.cpp:86-212):
// streams created per call, NUM_LOOPS pricings of all options, each option
// drawing its paths x steps normals from both streams (the two independent
// sequential generators run concurrently on two HIP streams).  Returns
// false on a failure (results then NaN).
bool price(const double* prm, int n, int precision, double* res, double* conf) {
  const fsmc_dropin::AsianConfig& c = fsmc_dropin::asian_config();
  fsmc_ctx* ctx = fsmc_dropin::context();
  if (!ctx) return false;
  AsianBuffers b;
  const fsmc_heston_params p1 = {1, c.rand_n, c.simsteps, c.riskfree, c.rho};
  const fsmc_mode m = {precision, FSMC_RNG_BUFFER, 0, 0,
                       c.refcompat ? FSMC_QUIRK_ASIAN_VAR_AS_VOL | FSMC_QUIRK_ASIAN_RUNNING_VAL2 : 0u};
  const size_t ws_bytes = fsmc_heston_asian_workspace(&p1, &m);
  if (ws_bytes == 0) return fsmc_ok(FSMC_ERR_INVALID_ARG, "fsmc_heston_asian_workspace");
  const long long per_option = (long long)c.rand_n * c.simsteps;
  if (!fsmc_ok(fsmc_mt19937_create(ctx, c.seed, &b.g1), "fsmc_mt19937_create") ||
      !fsmc_ok(fsmc_mt19937_create(ctx, c.seed + 1, &b.g2), "fsmc_mt19937_create") ||
      !hip_ok(hipMalloc((void**)&b.prm, 7 * (size_t)n * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.z1, per_option * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.z2, per_option * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.res, n * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc((void**)&b.conf, n * sizeof(double)), "hipMalloc") ||
      !hip_ok(hipMalloc(&b.ws, ws_bytes), "hipMalloc") || !hip_ok(hipStreamCreate(&b.s), "hipStreamCreate") ||
      !hip_ok(hipStreamCreate(&b.s2), "hipStreamCreate") ||
      !hip_ok(hipMemcpy(b.prm, prm, 7 * (size_t)n * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy"))
    return false;
  // per option: [0] generation start, [1] generation end (both streams), [2] s2's part done
  b.ev.resize(3 * (size_t)n);
  for (auto& e : b.ev)
    if (!hip_ok(hipEventCreateWithFlags(&e, hipEventDefault), "hipEventCreate")) return false;
  double total = 0.0, rng = 0.0;
  for (int l = 0; l < c.num_loops; ++l) {
    const auto t0 = std::chrono::steady_clock::now();
    for (int o = 0; o < n; ++o) {
      const double* q = b.prm + o;
      const char* gen = "fsmc_mt19937_gaussian_async";
      hipEvent_t* e = &b.ev[3 * (size_t)o];
      // the two streams are independent: the second one's normals are drawn
      // on s2, after the previous option's pricing has read z2
      if (!hip_ok(hipEventRecord(e[0], b.s), "hipEventRecord") ||
          !hip_ok(hipStreamWaitEvent(b.s2, e[0], 0), "hipStreamWaitEvent") ||
          !fsmc_ok(fsmc_mt19937_gaussian_async(b.g2, per_option, 0.0, 1.0, b.z2, b.s2), gen) ||
          !hip_ok(hipEventRecord(e[2], b.s2), "hipEventRecord") ||
          !fsmc_ok(fsmc_mt19937_gaussian_async(b.g1, per_option, 0.0, 1.0, b.z1, b.s), gen) ||
          !hip_ok(hipStreamWaitEvent(b.s, e[2], 0), "hipStreamWaitEvent") ||
          !hip_ok(hipEventRecord(e[1], b.s), "hipEventRecord") ||
          !fsmc_ok(fsmc_heston_asian_async(&p1, &m, q, q + n, q + 2 * n, q + 3 * n, q + 4 * n, q + 5 * n, q + 6 * n,
                                           b.z1, b.z2, b.res + o, b.conf + o, nullptr, b.ws, ws_bytes, b.s),
                   "fsmc_heston_asian_async"))
        return false;
    }
    if (!hip_ok(hipStreamSynchronize(b.s), "hipStreamSynchronize")) return false;
    total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int o = 0; o < n; ++o) {
      float ms = 0.0f;
      if (!hip_ok(hipEventElapsedTime(&ms, b.ev[3 * (size_t)o], b.ev[3 * (size_t)o + 1]), "hipEventElapsedTime"))
        return false;
      rng += 1e-3 * ms;
    }
  }
  if (!hip_ok(hipMemcpy(res, b.res, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy") ||
      !hip_ok(hipMemcpy(conf, b.conf, n * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy"))
    return false;
  // the sample's report, MCAsianOptions.cpp:208-212
  const double time = total / c.num_loops;
  std::printf("Completed RNG %8.4f seconds \n", rng / c.num_loops);
  std::printf("Completed MC by timestep in %8.4f seconds \n", time);
  std::printf("Computation rate MC by time Step - %8.4f.\n", n / time);
  std::fprintf(stderr, "%8.4f\n", n / time);
  return true;
}

template <typename Real>
void time_step(Real* h_CallResult, Real* h_CallConfidence, const Real* S, const Real* X, const Real* T,
               const Real* initialVol, const Real* kappaVals, const Real* thetaVals, const Real* xiVals) {
  const int n = fsmc_dropin::asian_config().opt_n;
  std::vector<double> prm(7 * (size_t)n), res(n), conf(n);
  const Real* in[7] = {S, X, T, initialVol, kappaVals, thetaVals, xiVals};
  for (int k = 0; k < 7; ++k)
    for (int i = 0; i < n; ++i) prm[(size_t)k * n + i] = (double)in[k][i];
  if (!price(prm.data(), n, sizeof(Real) == 8 ? FSMC_FP64 : FSMC_FP32, res.data(), conf.data())) {
    for (int i = 0; i < n; ++i) h_CallResult[i] = h_CallConfidence[i] = std::numeric_limits<Real>::quiet_NaN();
    return;
  }
  for (int i = 0; i < n; ++i) {
    h_CallResult[i] = (Real)res[i];
    h_CallConfidence[i] = (Real)conf[i];
  }
}

}  // namespace

void MonteCarloTimeStep(double* h_CallResult, double* h_CallConfidence, double* S, double* X, double* T,
                        double* initialVol, double* kappaVals, double* thetaVals, double* xiVals) {
  time_step(h_CallResult, h_CallConfidence, S, X, T, initialVol, kappaVals, thetaVals, xiVals);
}

void MonteCarloTimeStep(float* h_CallResult, float* h_CallConfidence, float* S, float* X, float* T,
                        float* initialVol, float* kappaVals, float* thetaVals, float* xiVals) {
  time_step(h_CallResult, h_CallConfidence, S, X, T, initialVol, kappaVals, thetaVals, xiVals);
}
