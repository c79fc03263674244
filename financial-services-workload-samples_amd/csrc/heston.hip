// Heston arithmetic-average Asian call, one path per lane, the whole time
// loop in registers.  Each step follows MCAsianOptions.cpp: correlate the
// two normals (:153-158), advance the spot with the variance at the start of
// the step (:51-62), then the full-truncation Euler variance (:65-74), and
// accumulate the arithmetic average over steps 1..N (:168-182).
//
// Correct Heston uses sqrt(v+) as the volatility; FSMC_QUIRK_ASIAN_VAR_AS_VOL
// reproduces the sample's v+ (:57-60).  The estimator is per-path sums
// (sum, sum of squares); FSMC_QUIRK_ASIAN_RUNNING_VAL2 instead reproduces the
// sample's chunk-of-64 running val2 (:184-201) from per-wavefront chunk sums.
#include "fsmc_device.hpp"
#include "fsmc_host.hpp"
#include "path_models.hpp"

namespace fsmc {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 64;  // VECLEN of MCAsianOptions.cpp:29 == one wavefront

struct HestonArgs {
  int n_paths, n_steps;
  double r, rho;
  const double *S, *X, *T, *v0, *kappa, *theta, *xi;
  const double *z1, *z2;  // buffer mode
  uint32_t k0, k1, w;
  double* partials;  // [n_options][gridDim.x][2]
  double* chunks;    // [n_options][n_paths / 64]  (running-val2 estimator)
};

template <typename Real, bool kBuffer, bool kVarAsVol, bool kChunks>
__global__ __launch_bounds__(kBlock) void heston_paths_kernel(HestonArgs a) {
  __shared__ double lds[2 * (kBlock / 64)];
  const int o = blockIdx.y;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  const HestonConsts<Real> c(a.T[o], a.n_steps, a.r, a.kappa[o], a.xi[o], a.theta[o], a.rho);
  double pay = 0.0;
  if (p < a.n_paths) {
    Real S = (Real)a.S[o];
    Real v = (Real)a.v0[o];
    PathSum avg;
    if (kBuffer) {
      const int nchunks = a.n_paths / kChunk;
      const size_t base = ((size_t)o * nchunks + (p / kChunk)) * (size_t)a.n_steps * kChunk + (p % kChunk);
      const double* z1 = a.z1 + base;
      const double* z2 = a.z2 + base;
      for (int k = 0; k < a.n_steps; ++k) {
        heston_step<Real, kVarAsVol>(c, S, v, (Real)z1[(size_t)k * kChunk], (Real)z2[(size_t)k * kChunk]);
        avg.add(S);
      }
    } else {
      const int nfull = a.n_steps & ~1;
      for (int k = 0; k < nfull; k += 2)
        heston_philox_pair<Real, kVarAsVol>(c, S, v, avg, U4{(uint32_t)p, (uint32_t)(k >> 1), (uint32_t)o, a.w},
                                            a.k0, a.k1, true);
      if (a.n_steps & 1)
        heston_philox_pair<Real, kVarAsVol>(c, S, v, avg, U4{(uint32_t)p, (uint32_t)(nfull >> 1), (uint32_t)o, a.w},
                                            a.k0, a.k1, false);
    }
    const double mean = avg.value() / (double)a.n_steps;
    const double x = mean - a.X[o];
    pay = x > 0.0 ? x : 0.0;
  }
  if (kChunks) {
    // One wavefront == one chunk of 64 paths (n_paths % 64 == 0 enforced).
    const double cs = wave_sum(pay);
    if ((threadIdx.x & 63) == 0 && p < a.n_paths) {
      const int nchunks = a.n_paths / kChunk;
      a.chunks[(size_t)o * nchunks + p / kChunk] = cs;
    }
  }
  double v2[2] = {pay, pay * pay};
  block_sum<2>(v2, lds);
  if (threadIdx.x == 0) {
    double* out = a.partials + ((size_t)o * gridDim.x + blockIdx.x) * 2;
    out[0] = v2[0];
    out[1] = v2[1];
  }
}

__device__ __forceinline__ void mean_se(double s, double s2, double N, double f, double* mean, double* se) {
  *mean = f * s / N;
  const double var = (N * s2 - s * s) / (N * (N - 1.0));
  *se = fabs(f) * sqrt(var > 0.0 ? var : 0.0) / sqrt(N);
}

// One wavefront per option (workgroup of 64).
__global__ __launch_bounds__(64) void heston_finalize_kernel(int nbx, int n_paths, double r, const double* T,
                                                            const double* partials, const double* chunks,
                                                            int running_val2, double* result, double* confidence) {
  const int o = blockIdx.x;
  const double N = (double)n_paths;
  double s[2];
  double val = 0.0, val2 = 0.0;
  if (running_val2) {  // uniform branch
    // MCAsianOptions.cpp:184-201 in meaning: val accumulates the chunk
    // means, val2 the square of the running val after each chunk.  One
    // wavefront: lane l owns the contiguous chunks [l m, (l + 1) m); the
    // running val before them is the exclusive prefix of the lanes' sums
    // (a wave scan), so each lane walks its chunks once more adding the
    // squares.  Differs from the sequential sum only in rounding order.
    const int lane = threadIdx.x & 63;
    const int nchunks = n_paths / kChunk;
    const int m = (nchunks + 63) / 64;
    const int c0 = min(nchunks, lane * m), c1 = min(nchunks, c0 + m);
    const double* cs = chunks + (size_t)o * nchunks;
    double loc = 0.0;
    for (int ci = c0; ci < c1; ++ci) loc += cs[ci] * (1.0 / kChunk);
    double incl = loc;  // inclusive scan over lanes
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const double up = __shfl_up(incl, off, 64);
      if (lane >= off) incl += up;
    }
    double run = __shfl_up(incl, 1, 64), sq = 0.0;  // exclusive prefix
    if (lane == 0) run = 0.0;
    for (int ci = c0; ci < c1; ++ci) {
      run += cs[ci] * (1.0 / kChunk);
      sq += run * run;
    }
    val2 = wave_sum(sq);
    val = __shfl(incl, 63, 64);
  } else {
    wave_combine<2>(partials + (size_t)o * nbx * 2, nbx, 2, s);
  }
  if (threadIdx.x != 0) return;
  const double exprt = exp(-r * T[o]);
  if (running_val2) {
    result[o] = exprt * val * (1.0 / N) * kChunk;
    const double stddev = sqrt((N * val2 - val * val) * (1.0 / (N * (N - 1.0))));
    confidence[o] = exprt * 1.96 * stddev * (1.0 / sqrt(N));
  } else {
    const double s0 = s[0], s1 = s[1];
    result[o] = exprt * s0 / N;
    const double stddev = sqrt((N * s1 - s0 * s0) * (1.0 / (N * (N - 1.0))));
    confidence[o] = exprt * 1.96 * stddev * (1.0 / sqrt(N));
  }
}

// Tangent kernel: one path per lane through heston_tangent_path (the same
// Greeks carried forward with the path; no checkpoints), partials as the
// adjoint kernel's.
template <typename Real>
__global__ __launch_bounds__(kBlock) void heston_tangent_kernel(HestonArgs a) {
  __shared__ double lds[8 * (kBlock / 64)];
  const int o = blockIdx.y;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (p < a.n_paths) {
    const HestonConsts<Real> c(a.T[o], a.n_steps, a.r, a.kappa[o], a.xi[o], a.theta[o], a.rho);
    double g[3];
    const double x = heston_tangent_path<Real>(c, a.S[o], a.X[o], a.T[o], a.v0[o], a.r, a.n_steps, (uint32_t)p,
                                               (uint32_t)o, a.w, a.k0, a.k1, g);
    out[0] = x;
    out[1] = x * x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out[2 + 2 * k] = g[k];
      out[3 + 2 * k] = g[k] * g[k];
    }
  }
  block_sum<8>(out, lds);
  if (threadIdx.x == 0) {
    double* dst = a.partials + ((size_t)o * gridDim.x + blockIdx.x) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = out[k];
  }
}

// Adjoint kernel: one path per lane through heston_adjoint_path (reverse
// sweep over kHestonSeg-step checkpoints kept in the workspace as
// [option][segment][S|v][path]); the forward sweep also yields the price.
template <typename Real>
__global__ __launch_bounds__(kBlock) void heston_adjoint_kernel(HestonArgs a, Real* ckpt) {
  __shared__ double lds[8 * (kBlock / 64)];
  const int o = blockIdx.y;
  const int p = blockIdx.x * kBlock + threadIdx.x;
  double out[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (p < a.n_paths) {
    const HestonConsts<Real> c(a.T[o], a.n_steps, a.r, a.kappa[o], a.xi[o], a.theta[o], a.rho);
    const int nseg = (a.n_steps + kHestonSeg - 1) / kHestonSeg;
    Real* ck = ckpt + (size_t)o * nseg * 2 * a.n_paths + p;
    double g[3];
    const double x = heston_adjoint_path<Real>(c, a.S[o], a.X[o], a.T[o], a.v0[o], a.r, a.n_steps, (uint32_t)p,
                                               (uint32_t)o, a.w, a.k0, a.k1, ck, (size_t)a.n_paths, g);
    out[0] = x;
    out[1] = x * x;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      out[2 + 2 * k] = g[k];
      out[3 + 2 * k] = g[k] * g[k];
    }
  }
  block_sum<8>(out, lds);
  if (threadIdx.x == 0) {
    double* dst = a.partials + ((size_t)o * gridDim.x + blockIdx.x) * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) dst[k] = out[k];
  }
}

// One wavefront per option (workgroup of 64); the price sums are combined
// in the same order as heston_finalize_kernel's, so both report the same
// price for the same paths.
__global__ __launch_bounds__(64) void heston_greeks_finalize_kernel(int nbx, int n_paths, double r, const double* T,
                                                                   const double* partials, double* result,
                                                                   double* confidence, double* greeks) {
  const int o = blockIdx.x;
  double s[8];
  wave_combine<8>(partials + (size_t)o * nbx * 8, nbx, 8, s);
  if (threadIdx.x != 0) return;
  const double N = (double)n_paths;
  const double exprt = exp(-r * T[o]);
  double se;
  mean_se(s[0], s[1], N, exprt, result + o, &se);
  confidence[o] = 1.96 * se;
  double* g = greeks + 6 * (size_t)o;
  for (int k = 0; k < 3; ++k) mean_se(s[2 + 2 * k], s[3 + 2 * k], N, 1.0, g + k, g + 3 + k);
}

struct HestonPlan {
  int nbx;
  size_t partial_bytes, chunk_bytes, ws;  // forward estimator
  size_t gpartial_bytes, ckpt_bytes, gws;  // adjoint
};

HestonPlan heston_plan(const fsmc_heston_params* p, const fsmc_mode* m) {
  HestonPlan h;
  h.nbx = (p->n_paths + kBlock - 1) / kBlock;
  h.partial_bytes = align_up((size_t)p->n_options * h.nbx * 2 * sizeof(double), 256);
  h.chunk_bytes = align_up((size_t)p->n_options * (p->n_paths / kChunk + 1) * sizeof(double), 256);
  h.ws = h.partial_bytes + h.chunk_bytes;
  const size_t real = m->precision == FSMC_FP64 ? sizeof(double) : sizeof(float);
  const size_t nseg = (size_t)(p->n_steps + kHestonSeg - 1) / kHestonSeg;
  h.gpartial_bytes = align_up((size_t)p->n_options * h.nbx * 8 * sizeof(double), 256);
  h.ckpt_bytes = align_up((size_t)p->n_options * nseg * 2 * (size_t)p->n_paths * real, 256);
  h.gws = h.gpartial_bytes + h.ckpt_bytes;
  return h;
}

int heston_validate(const fsmc_heston_params* p, const fsmc_mode* m) {
  FSMC_REQUIRE(p != nullptr, "params is NULL");
  FSMC_TRY(validate_mode(m, true));
  FSMC_REQUIRE(p->n_options >= 1 && p->n_options <= 65535, "n_options %d outside [1, 65535]", p->n_options);
  FSMC_REQUIRE(p->n_paths >= 2 && p->n_paths <= (1 << 30), "n_paths %d outside [2, 2^30]", p->n_paths);
  FSMC_REQUIRE(p->n_steps >= 1, "n_steps %d < 1", p->n_steps);
  FSMC_REQUIRE(p->rho >= -1.0 && p->rho <= 1.0, "rho %g outside [-1, 1]", p->rho);
  const bool needs64 = m->rng == FSMC_RNG_BUFFER || (m->quirks & FSMC_QUIRK_ASIAN_RUNNING_VAL2);
  FSMC_REQUIRE(!needs64 || p->n_paths % kChunk == 0,
               "buffer mode and the running-val2 estimator need n_paths %% 64 == 0 (got %d)", p->n_paths);
  return FSMC_OK;
}

template <typename Real, bool B, bool Q4, bool CH>
void launch(const HestonArgs& a, int n_options, int nbx, hipStream_t s) {
  hipLaunchKernelGGL((heston_paths_kernel<Real, B, Q4, CH>), dim3(nbx, n_options), dim3(kBlock), 0, s, a);
}

template <typename Real, bool B>
void launch_q(const HestonArgs& a, bool q4, bool ch, int n_options, int nbx, hipStream_t s) {
  if (q4) ch ? launch<Real, B, true, true>(a, n_options, nbx, s) : launch<Real, B, true, false>(a, n_options, nbx, s);
  else ch ? launch<Real, B, false, true>(a, n_options, nbx, s) : launch<Real, B, false, false>(a, n_options, nbx, s);
}

}  // namespace

}  // namespace fsmc

using namespace fsmc;

extern "C" {

size_t fsmc_heston_asian_workspace(const fsmc_heston_params* p, const fsmc_mode* m) {
  if (heston_validate(p, m) != FSMC_OK) return 0;
  const HestonPlan h = heston_plan(p, m);
  // Philox mode without quirks may be asked for Greeks: size for the adjoint.
  const bool greeks_possible = m->rng == FSMC_RNG_PHILOX && m->quirks == 0;
  return greeks_possible && h.gws > h.ws ? h.gws : h.ws;
}

}  // extern "C"

namespace fsmc {
namespace {

int heston_check_args(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S, const double* d_X,
                      const double* d_T, const double* d_v0, const double* d_kappa, const double* d_theta,
                      const double* d_xi, const double* d_z1, const double* d_z2, const double* d_result,
                      const double* d_conf, const double* d_greeks) {
  FSMC_TRY(heston_validate(p, m));
  FSMC_REQUIRE(d_S && d_X && d_T && d_v0 && d_kappa && d_theta && d_xi && d_result && d_conf,
               "NULL device pointer");
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  FSMC_REQUIRE(!buffer || (d_z1 && d_z2), "buffer mode needs z1 and z2");
  FSMC_REQUIRE(!d_greeks || (!buffer && m->quirks == 0), "Greeks need Philox mode without quirks");
  return FSMC_OK;
}

// Enqueues the path (or adjoint) kernel and its finalize on stream s with the
// workspace d_ws; the arguments have passed heston_check_args and d_ws holds
// at least the plan's forward (ws) or adjoint (gws) bytes.  Every workspace
// word the kernels read is written earlier in the same launch sequence, so
// the workspace needs no clearing.
int heston_enqueue(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S, const double* d_X,
                   const double* d_T, const double* d_v0, const double* d_kappa, const double* d_theta,
                   const double* d_xi, const double* d_z1, const double* d_z2, double* d_result, double* d_conf,
                   double* d_greeks, void* d_ws, hipStream_t s, bool tangent = false) {
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  const HestonPlan h = heston_plan(p, m);
  HestonArgs a;
  a.n_paths = p->n_paths;
  a.n_steps = p->n_steps;
  a.r = p->riskfree;
  a.rho = p->rho;
  a.S = d_S;
  a.X = d_X;
  a.T = d_T;
  a.v0 = d_v0;
  a.kappa = d_kappa;
  a.theta = d_theta;
  a.xi = d_xi;
  a.z1 = d_z1;
  a.z2 = d_z2;
  a.k0 = (uint32_t)m->seed;
  a.k1 = (uint32_t)(m->seed >> 32);
  a.w = counter_w(kTagHeston, m->stream);
  a.partials = (double*)d_ws;
  a.chunks = (double*)((char*)d_ws + h.partial_bytes);
  if (d_greeks && tangent) {
    a.partials = (double*)d_ws;
    if (m->precision == FSMC_FP64)
      hipLaunchKernelGGL(heston_tangent_kernel<double>, dim3(h.nbx, p->n_options), dim3(kBlock), 0, s, a);
    else
      hipLaunchKernelGGL(heston_tangent_kernel<float>, dim3(h.nbx, p->n_options), dim3(kBlock), 0, s, a);
    FSMC_HIP(hipGetLastError());
    hipLaunchKernelGGL(heston_greeks_finalize_kernel, dim3(p->n_options), dim3(64), 0, s, h.nbx, p->n_paths,
                       p->riskfree, d_T, a.partials, d_result, d_conf, d_greeks);
    FSMC_HIP(hipGetLastError());
    return FSMC_OK;
  }
  if (d_greeks) {
    // One launch: the adjoint kernel's forward sweep also yields the price.
    a.partials = (double*)d_ws;
    if (m->precision == FSMC_FP64)
      hipLaunchKernelGGL(heston_adjoint_kernel<double>, dim3(h.nbx, p->n_options), dim3(kBlock), 0, s, a,
                         (double*)((char*)d_ws + h.gpartial_bytes));
    else
      hipLaunchKernelGGL(heston_adjoint_kernel<float>, dim3(h.nbx, p->n_options), dim3(kBlock), 0, s, a,
                         (float*)((char*)d_ws + h.gpartial_bytes));
    FSMC_HIP(hipGetLastError());
    hipLaunchKernelGGL(heston_greeks_finalize_kernel, dim3(p->n_options), dim3(64), 0, s, h.nbx, p->n_paths,
                       p->riskfree, d_T, a.partials, d_result, d_conf, d_greeks);
    FSMC_HIP(hipGetLastError());
    return FSMC_OK;
  }
  const bool q4 = (m->quirks & FSMC_QUIRK_ASIAN_VAR_AS_VOL) != 0;
  const bool ch = (m->quirks & FSMC_QUIRK_ASIAN_RUNNING_VAL2) != 0;
  if (m->precision == FSMC_FP64) {
    buffer ? launch_q<double, true>(a, q4, ch, p->n_options, h.nbx, s)
           : launch_q<double, false>(a, q4, ch, p->n_options, h.nbx, s);
  } else {
    buffer ? launch_q<float, true>(a, q4, ch, p->n_options, h.nbx, s)
           : launch_q<float, false>(a, q4, ch, p->n_options, h.nbx, s);
  }
  FSMC_HIP(hipGetLastError());
  hipLaunchKernelGGL(heston_finalize_kernel, dim3(p->n_options), dim3(64), 0, s, h.nbx, p->n_paths, p->riskfree,
                     d_T, a.partials, a.chunks, ch ? 1 : 0, d_result, d_conf);
  FSMC_HIP(hipGetLastError());
  return FSMC_OK;
}

// Workspace of fsmc_heston_asian_noalloc: a fixed device array in this
// library's code object.  Each GPU holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device (the array's name is a host-side stand-in).
constexpr size_t kNoallocScratchBytes = size_t(16) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

}  // namespace
}  // namespace fsmc

static __device__ __align__(256) unsigned char scratch[fsmc::kNoallocScratchBytes];

namespace fsmc {
namespace {

// The scratch array's address on the current device, looked up once per
// device (with that device current) and kept.
int scratch_on_current_device(void** out) {
  int dev = -1;
  FSMC_HIP(hipGetDevice(&dev));
  FSMC_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d outside [0, %d)", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    FSMC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    FSMC_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned NULL for the scratch array on device %d", dev);
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return FSMC_OK;
}

}  // namespace
}  // namespace fsmc

extern "C" {

int fsmc_heston_asian_async(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                            const double* d_X, const double* d_T, const double* d_v0, const double* d_kappa,
                            const double* d_theta, const double* d_xi, const double* d_z1, const double* d_z2,
                            double* d_result, double* d_conf, double* d_greeks, void* d_ws, size_t ws_bytes,
                            void* stream) {
  FSMC_TRY(heston_check_args(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, d_z1, d_z2, d_result, d_conf,
                             d_greeks));
  const size_t need = fsmc_heston_asian_workspace(p, m);
  if (d_ws == nullptr || ws_bytes < need) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, need);
    return FSMC_ERR_WORKSPACE;
  }
  return heston_enqueue(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, d_z1, d_z2, d_result, d_conf, d_greeks,
                        d_ws, (hipStream_t)stream);
}

size_t fsmc_heston_asian_tangent_workspace(const fsmc_heston_params* p, const fsmc_mode* m) {
  if (heston_validate(p, m) != FSMC_OK) return 0;
  return heston_plan(p, m).gpartial_bytes;
}

int fsmc_heston_asian_tangent_async(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                                    const double* d_X, const double* d_T, const double* d_v0, const double* d_kappa,
                                    const double* d_theta, const double* d_xi, double* d_result, double* d_conf,
                                    double* d_greeks, void* d_ws, size_t ws_bytes, void* stream) {
  FSMC_REQUIRE(d_greeks != nullptr, "the tangent entry point computes Greeks: d_greeks is NULL");
  FSMC_TRY(heston_check_args(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, nullptr, nullptr, d_result, d_conf,
                             d_greeks));
  const size_t need = fsmc_heston_asian_tangent_workspace(p, m);
  if (d_ws == nullptr || ws_bytes < need) {
    set_error("workspace %zu bytes < required %zu", ws_bytes, need);
    return FSMC_ERR_WORKSPACE;
  }
  return heston_enqueue(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, nullptr, nullptr, d_result, d_conf,
                        d_greeks, d_ws, (hipStream_t)stream, true);
}

size_t fsmc_heston_asian_noalloc_limit(void) { return kNoallocScratchBytes; }

int fsmc_heston_asian_noalloc(const fsmc_heston_params* p, const fsmc_mode* m, const double* d_S,
                              const double* d_X, const double* d_T, const double* d_v0, const double* d_kappa,
                              const double* d_theta, const double* d_xi, const double* d_z1, const double* d_z2,
                              double* d_result, double* d_conf, double* d_greeks, void* stream) {
  FSMC_TRY(heston_check_args(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, d_z1, d_z2, d_result, d_conf,
                             d_greeks));
  const HestonPlan h = heston_plan(p, m);
  const size_t need = d_greeks ? h.gws : h.ws;
  if (need > kNoallocScratchBytes) {
    set_error("this pricing needs %zu bytes of workspace; the allocation-free entry point has %zu", need,
              kNoallocScratchBytes);
    return FSMC_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  FSMC_TRY(scratch_on_current_device(&ws));
  return heston_enqueue(p, m, d_S, d_X, d_T, d_v0, d_kappa, d_theta, d_xi, d_z1, d_z2, d_result, d_conf, d_greeks,
                        ws, (hipStream_t)stream);
}

int fsmc_heston_asian(fsmc_ctx* ctx, const fsmc_heston_params* p, const fsmc_mode* m, const double* S,
                      const double* X, const double* T, const double* v0, const double* kappa, const double* theta,
                      const double* xi, const double* z1, const double* z2, double* call_result,
                      double* call_confidence, double* greeks) {
  FSMC_REQUIRE(ctx != nullptr, "ctx is NULL");
  FSMC_TRY(heston_validate(p, m));
  FSMC_REQUIRE(S && X && T && v0 && kappa && theta && xi && call_result && call_confidence, "NULL host pointer");
  const bool buffer = m->rng == FSMC_RNG_BUFFER;
  FSMC_REQUIRE(!buffer || (z1 && z2), "buffer mode needs z1 and z2");
  std::lock_guard<std::mutex> lock(ctx->mu);
  const size_t n = (size_t)p->n_options;
  const size_t zn = buffer ? n * (size_t)p->n_paths * (size_t)p->n_steps : 0;
  FSMC_TRY(ctx_prepare(ctx, 7 * n * sizeof(double), 2 * zn * sizeof(double), 8 * n * sizeof(double),
                       fsmc_heston_asian_workspace(p, m)));
  double* d = (double*)ctx->inputs.ptr;
  const double* host[7] = {S, X, T, v0, kappa, theta, xi};
  hipStream_t s = ctx->stream;
  for (int k = 0; k < 7; ++k)
    FSMC_HIP(hipMemcpyAsync(d + k * n, host[k], n * sizeof(double), hipMemcpyHostToDevice, s));
  double* dz1 = (double*)ctx->normals.ptr;
  double* dz2 = dz1 ? dz1 + zn : nullptr;
  if (buffer) {
    FSMC_HIP(hipMemcpyAsync(dz1, z1, zn * sizeof(double), hipMemcpyHostToDevice, s));
    FSMC_HIP(hipMemcpyAsync(dz2, z2, zn * sizeof(double), hipMemcpyHostToDevice, s));
  }
  double* dres = (double*)ctx->outputs.ptr;
  double* dconf = dres + n;
  double* dgr = greeks ? dconf + n : nullptr;
  int st = fsmc_heston_asian_async(p, m, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n, dz1, dz2,
                                   dres, dconf, dgr, ctx->workspace.ptr, ctx->workspace.bytes, s);
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
