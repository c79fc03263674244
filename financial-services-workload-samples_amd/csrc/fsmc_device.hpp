// Device building blocks shared by every Monte Carlo kernel of the engine:
//   * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
//     "Parallel random numbers: as easy as 1, 2, 3", SC'11),
//   * the Box-Muller transform from two 32-bit words to two N(0,1) samples,
//   * branch-free two-sum compensated FP32 accumulation,
//   * wavefront (64-lane) and workgroup reductions in FP64.
//
// The counter layout and the uniform mapping below are a contract: the CPU
// oracle (oracle/fsmc_oracle.c) reproduces them bit for bit, so a Philox-mode
// run on the GPU and the oracle consume identical normals.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fsmc {

// ---------------------------------------------------------------------------
// Philox4x32-10
// ---------------------------------------------------------------------------
struct U4 {
  uint32_t x, y, z, w;
};

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// Stream tags occupy the top 4 bits of counter word w so that the workloads
// never share a counter even with the same seed.
constexpr uint32_t kTagGbm = 1u;
constexpr uint32_t kTagHeston = 2u;
constexpr uint32_t kTagLibor = 3u;
constexpr uint32_t kTagLsmc = 4u;

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // One v_mad_u64_u32 per product gives both halves.
    const uint64_t p0 = (uint64_t)kPhiloxM0 * c.x;
    const uint64_t p1 = (uint64_t)kPhiloxM1 * c.z;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    if (r < 3) {
      // Early rounds often see wave-uniform words (a step counter, the
      // option index, the keys): plain XORs let the uniform parts go to the
      // scalar unit instead of being copied into vector registers.
      c.x = hi1 ^ (c.y ^ k0);
      c.z = hi0 ^ (c.w ^ k1);
    } else {
      // three-input XOR in one v_bitop3_b32 (truth table 0x96)
      c.x = __builtin_amdgcn_bitop3_b32(hi1, c.y, k0, 0x96);
      c.z = __builtin_amdgcn_bitop3_b32(hi0, c.w, k1, 0x96);
    }
    c.y = lo1;
    c.w = lo0;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return c;
}

// max(v, 0).  In FP32 one v_med3_f32 (clamp to [0, FLT_MAX]): a plain max of
// a loop-carried value costs a second instruction that quiets signalling NaNs.
__device__ __forceinline__ float positive_part(float v) { return __builtin_amdgcn_fmed3f(v, 0.f, 3.40282347e38f); }
__device__ __forceinline__ double positive_part(double v) { return v > 0.0 ? v : 0.0; }

// Counter word w: tag in bits 28..31, caller stream id in bits 0..27.
__host__ __device__ __forceinline__ uint32_t counter_w(uint32_t tag, uint32_t stream) {
  return (tag << 28) ^ (stream & 0x0FFFFFFFu);
}

// ---------------------------------------------------------------------------
// Box-Muller.  u1 = fl(a | 1) 2^-32 in (0, 1], u2 = fl(b) 2^-32 in [0, 1],
// with fl() the round-to-nearest-even conversion of a 32-bit word to FP32:
// one conversion each (the low bit of a keeps u1 off zero), and both values
// are exact in FP32 and FP64, so the two precisions see the same uniforms.
// z0 = r cos(2 pi u2), z1 = r sin(2 pi u2), r = sqrt(-2 ln u1) <= 6.66.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& z0, float& z1) {
  // u1 = f 2^-32 with f = fl(a | 1); v_log_f32 is log2, so
  // -2 ln u1 = -2 ln2 (log2 f - 32) = fma(-2 ln2, log2 f, 64 ln2).
  const float f = (float)(a | 1u);
  const float u2 = (float)b * 0x1.0p-32f;
  const float r = __builtin_amdgcn_sqrtf(__builtin_fmaf(-1.3862943611198906f, __builtin_amdgcn_logf(f),
                                                       44.361419555836500f));
  // v_sin_f32 / v_cos_f32 take their argument in revolutions.
  z0 = r * __builtin_amdgcn_cosf(u2);
  z1 = r * __builtin_amdgcn_sinf(u2);
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double& z0, double& z1) {
  const double u1 = (double)(float)(a | 1u) * 0x1.0p-32;
  const double u2 = (double)(float)b * 0x1.0p-32;
  const double r = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  z0 = r * c;
  z1 = r * s;
}

// The same pair in polar form: z0 = r c, z1 = r s returned as (r^2, c, s),
// r^2 = -2 ln u1, for consumers that multiply both normals by a common
// factor and can fold r into its square root (the c, s and r^2 that
// box_muller's FP32 form uses).
__device__ __forceinline__ void box_muller_polar(uint32_t a, uint32_t b, float& r2, float& c, float& s) {
  const float f = (float)(a | 1u);
  const float u2 = (float)b * 0x1.0p-32f;
  r2 = __builtin_fmaf(-1.3862943611198906f, __builtin_amdgcn_logf(f), 44.361419555836500f);
  c = __builtin_amdgcn_cosf(u2);
  s = __builtin_amdgcn_sinf(u2);
}

// Four normals from one Philox call.
template <typename Real>
__device__ __forceinline__ void normals4(U4 ctr, uint32_t k0, uint32_t k1, Real& n0, Real& n1, Real& n2,
                                         Real& n3) {
  const U4 r = philox4x32_10(ctr, k0, k1);
  box_muller(r.x, r.y, n0, n1);
  box_muller(r.z, r.w, n2, n3);
}

// ---------------------------------------------------------------------------
// Compensated accumulation.  In FP32 mode a lane accumulates with Kahan's
// compensated summation (error bounded by 2 ulp of the sum, independent of
// the number of terms, four operations per term); the compensation is folded
// in before the FP64 cross-lane combine.  Contraction into FMA is switched
// off locally so that the error term survives -O3.
// ---------------------------------------------------------------------------
template <typename Real>
struct CompSum {
  Real s = 0, c = 0;  // c = (what was added) - (what should have been added)
  __device__ __forceinline__ void add(Real x) {
#pragma clang fp contract(off)
    const Real y = x - c;
    const Real t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ double value() const { return (double)s - (double)c; }
};

template <>
struct CompSum<double> {
  double s = 0;
  __device__ __forceinline__ void add(double x) { s += x; }
  __device__ __forceinline__ double value() const { return s; }
};

// Running sum of one path's monitored spots (the arithmetic-average leg): an
// FP64 accumulator fed FP32 or FP64 terms.  On this GPU an FP64 add issues at
// the FP32 rate, so one conversion and one add per term cost less than the
// four operations of an FP32 compensated sum and lose nothing.
struct PathSum {
  double s = 0.0;
  __device__ __forceinline__ void add(double x) { s += x; }
  __device__ __forceinline__ double value() const { return s; }
};

// ---------------------------------------------------------------------------
// Reductions (fixed order => bitwise-reproducible for a fixed launch shape).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums n records of nv <= NV doubles (record b at p[b * nv]) with one
// wavefront: lane l adds records l, l + 64, ... in order, then the xor tree
// (every lane ends with the same bits).  The order depends only on n, so the
// result is reproducible; called by whole waves.
template <int NV>
__device__ __forceinline__ void wave_combine(const double* p, int n, int nv, double (&s)[NV]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = 0.0;
  for (int b = lane; b < n; b += 64) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k < nv) s[k] += p[(size_t)b * nv + k];
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = wave_sum(s[k]);
}

// Sums NV values across a workgroup of blockDim.x threads (multiple of 64).
// Result valid in thread 0.  `lds` must hold NV * (blockDim.x / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nwaves + wave] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < nwaves; ++w) s += lds[k * nwaves + w];
      v[k] = s;
    }
  }
}

// FP32 wave sum: four adds with a DPP operand inside each row of 16 lanes
// (xor 1, xor 2 by quad permutes, then the half-row and row mirrors), then
// two cross-row shuffles.  Every lane ends with the total; the tree keeps the
// rounding of a 64-term sum within 6 ulps.
template <int kCtrl>
__device__ __forceinline__ float dpp_f32(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xF, 0xF, true));
}
__device__ __forceinline__ float wave_sum(float v) {
  v += dpp_f32<0xB1>(v);   // quad_perm [1,0,3,2]: lane ^ 1
  v += dpp_f32<0x4E>(v);   // quad_perm [2,3,0,1]: lane ^ 2
  v += dpp_f32<0x141>(v);  // row_half_mirror: lane i <-> 7 - i within 8
  v += dpp_f32<0x140>(v);  // row_mirror: lane i <-> 15 - i within 16
  v += __shfl_xor(v, 16, 64);
  v += __shfl_xor(v, 32, 64);
  return v;
}

// Workgroup sum of NV FP32 thread values: each wave sums in FP32 (above),
// the wave totals are combined in FP64.  Result (FP64) valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(const float (&v)[NV], double (&out)[NV], double* lds) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  float w[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k * nwaves + wave] = (double)w[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double t = 0.0;
      for (int i = 0; i < nwaves; ++i) t += lds[k * nwaves + i];
      out[k] = t;
    }
  }
}

// Precision-dependent exponential and division.  FP32 uses the hardware
// v_exp_f32 (base 2, so exponents are pre-scaled by kE = log2 e) and
// v_rcp_f32; FP64 uses the full-precision library exp and IEEE division, so
// its results follow the CPU oracle to rounding.
template <typename Real>
struct MathOps;
template <>
struct MathOps<float> {
  static constexpr float kE = 1.4426950408889634f;
  __device__ static float ex(float x) { return __builtin_amdgcn_exp2f(x); }
  __device__ static float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};
template <>
struct MathOps<double> {
  static constexpr double kE = 1.0;
  __device__ static double ex(double x) { return exp(x); }
  __device__ static double div(double a, double b) { return a / b; }
};

// exp2 in the working precision: hardware v_exp_f32 for FP32.
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double fexp2(double x) { return exp2(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double fsqrt(double x) { return sqrt(x); }
// 1/sqrt(x) in FP32: the hardware reciprocal square root (one instruction, ~1 ulp).
__device__ __forceinline__ float frsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
// Explicit fused multiply-add: a * b + c with one rounding.
__device__ __forceinline__ float ffma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double ffma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ double frcp(double x) { return 1.0 / x; }

// Wichura, "Algorithm AS 241: The percentage points of the normal
// distribution", Appl. Statist. 37 (1988) 477-484 (PPND16), in FP64.
__device__ inline double icdf(double p) {
  const double q = p - 0.5;
  if (fabs(q) <= 0.425) {
    const double r = 0.180625 - q * q;
    const double num = ((((((2.5090809287301226727e+3 * r + 3.3430575583588128105e+4) * r +
                            6.7265770927008700853e+4) * r + 4.5921953931549871457e+4) * r +
                          1.3731693765509461125e+4) * r + 1.9715909503065514427e+3) * r +
                        1.3314166789178437745e+2) * r + 3.3871328727963666080e+0;
    const double den = ((((((5.2264952788528545610e+3 * r + 2.8729085735721942674e+4) * r +
                            3.9307895800092710610e+4) * r + 2.1213794301586595867e+4) * r +
                          5.3941960214247511077e+3) * r + 6.8718700749205790830e+2) * r +
                        4.2313330701600911252e+1) * r + 1.0;
    return q * num / den;
  }
  double r = sqrt(-log(q < 0.0 ? p : 1.0 - p));
  double x;
  if (r <= 5.0) {
    r -= 1.6;
    const double num = ((((((7.74545014278341407640e-4 * r + 2.27238449892691845833e-2) * r +
                            2.41780725177450611770e-1) * r + 1.27045825245236838258e+0) * r +
                          3.64784832476320460504e+0) * r + 5.76949722146069140550e+0) * r +
                        4.63033784615654529590e+0) * r + 1.42343711074968357734e+0;
    const double den = ((((((1.05075007164441684324e-9 * r + 5.47593808499534494600e-4) * r +
                            1.51986665636164571966e-2) * r + 1.48103976427480074590e-1) * r +
                          6.89767334985100004550e-1) * r + 1.67638483018380384940e+0) * r +
                        2.05319162663775882187e+0) * r + 1.0;
    x = num / den;
  } else {
    r -= 5.0;
    const double num = ((((((2.01033439929228813265e-7 * r + 2.71155556874348757815e-5) * r +
                            1.24266094738807843860e-3) * r + 2.65321895265761230930e-2) * r +
                          2.96560571828504891230e-1) * r + 1.78482653991729133580e+0) * r +
                        5.46378491116411436990e+0) * r + 6.65790464350110377720e+0;
    const double den = ((((((2.04426310338993978564e-15 * r + 1.42151175831644588870e-7) * r +
                            1.84631831751005468180e-5) * r + 7.86869131145613259100e-4) * r +
                          1.48753612908506148525e-2) * r + 1.36929880922735805310e-1) * r +
                        5.99832206555887937690e-1) * r + 1.0;
    x = num / den;
  }
  return q < 0.0 ? -x : x;
}

}  // namespace fsmc
