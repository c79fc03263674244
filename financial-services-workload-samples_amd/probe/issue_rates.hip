// Issue-rate probe (measurement infrastructure for bench.py, not part of the
// pricing library): the sustained whole-GPU rate of one VALU instruction
// class, in wave64 instructions per second, from kernels made of 8
// independent dependency chains of that instruction per lane (no memory
// traffic inside the timed loop), launched on enough waves to fill every
// SIMD.  The instruction-mix ceiling of a path kernel is then its per-class
// instruction counts divided by these rates, summed.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

enum Kind : int {
  kFmaF32 = 0,   // v_fma_f32
  kExpF32 = 1,   // v_exp_f32
  kLogF32 = 2,   // v_log_f32
  kSqrtF32 = 3,  // v_sqrt_f32
  kRcpF32 = 4,   // v_rcp_f32
  kSinF32 = 5,   // v_sin_f32
  kMadU64 = 6,   // v_mad_u64_u32
  kFmaF64 = 7,   // v_fma_f64
  kMix7Fma1Exp = 8,  // 7 v_fma_f32 + 1 v_exp_f32 per chain step (overlap test)
  kFmaF32Vgpr = 9,   // v_fma_f32 with three VGPR operands (VOP3 encoding)
  kAddF32 = 10,      // v_add_f32 (VOP2)
  kCvtF32U32 = 11,   // v_cvt_f32_u32
  kKinds = 12
};

constexpr int kChains = 8;
constexpr int kUnroll = 4;

// Wave instructions of the measured class issued per loop trip per wave.
__host__ __device__ constexpr int insts_per_trip(int kind) {
  return kind == kMix7Fma1Exp ? kChains * kChains * kUnroll : kChains * kUnroll;
}

template <int K>
__global__ __launch_bounds__(256) void probe_kernel(int trips, float seed, float* out) {
  const float t = seed + (float)threadIdx.x * 1e-7f;
  if constexpr (K == kMadU64) {
    uint64_t r[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) r[c] = (uint64_t)(threadIdx.x + 17u * c);
    const uint32_t m = 0xD2511F53u ^ (uint32_t)blockIdx.x;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) r[c] = (uint64_t)(uint32_t)(r[c] >> 32) * m + r[c];
      }
    }
    uint64_t s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += r[c];
    if (s == 0x123456789ull) out[blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
  } else if constexpr (K == kFmaF64) {
    double v[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) v[c] = (double)t + c;
    const double a = 0.999999 + 1e-9 * seed, b = 1e-7;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) v[c] = __builtin_fma(v[c], a, b);
      }
    }
    double s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += v[c];
    if (s == 12345.0) out[blockIdx.x * blockDim.x + threadIdx.x] = (float)s;
  } else {
    float v[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) v[c] = t + 0.01f * c;
    const float a = 0.9999f + 1e-9f * seed, b = 1e-6f;
    // per-lane operands: the VOP3 form needs them in VGPRs
    const float av = 0.9999f + 1e-12f * (float)threadIdx.x, bv = 1e-6f * (float)(threadIdx.x & 7);
    uint32_t w[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) w[c] = threadIdx.x * 2654435761u + 977u * c;
    for (int i = 0; i < trips; ++i) {
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
          if constexpr (K == kFmaF32) v[c] = __builtin_fmaf(v[c], a, b);
          if constexpr (K == kExpF32) v[c] = __builtin_amdgcn_exp2f(-v[c]);
          if constexpr (K == kLogF32) v[c] = __builtin_amdgcn_logf(__builtin_fabsf(v[c]) + 0.0f);
          if constexpr (K == kSqrtF32) v[c] = __builtin_amdgcn_sqrtf(v[c]);
          if constexpr (K == kRcpF32) v[c] = __builtin_amdgcn_rcpf(v[c]);
          if constexpr (K == kSinF32) v[c] = __builtin_amdgcn_sinf(v[c]);
          if constexpr (K == kFmaF32Vgpr) v[c] = __builtin_fmaf(v[c], av, bv);
          if constexpr (K == kAddF32) v[c] = v[c] + av;
          if constexpr (K == kCvtF32U32) {
            // one conversion per step; the integer chain advances through the
            // float's bits (a move, not an ALU instruction on this target)
            v[c] = (float)w[c];
            w[c] = __builtin_bit_cast(uint32_t, v[c]);
          }
          if constexpr (K == kMix7Fma1Exp) {
            // one chain step: 7 dependent FMAs then one exp; the 8 chains
            // are independent, so 7 of every 8 instructions are FMAs
            if (c == 0) {
#pragma unroll
              for (int k = 0; k < kChains; ++k) v[k] = __builtin_amdgcn_exp2f(-v[k]);
            } else {
#pragma unroll
              for (int k = 0; k < kChains; ++k) v[k] = __builtin_fmaf(v[k], a, b);
            }
          }
        }
      }
    }
    float s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += v[c];
    if (s == 12345.0f) out[blockIdx.x * blockDim.x + threadIdx.x] = s;
  }
}

template <int K>
hipError_t launch(int blocks, int trips, float* out, hipStream_t s) {
  hipLaunchKernelGGL(probe_kernel<K>, dim3(blocks), dim3(256), 0, s, trips, 1.0f, out);
  return hipGetLastError();
}

hipError_t launch_kind(int kind, int blocks, int trips, float* out, hipStream_t s) {
  switch (kind) {
    case kFmaF32: return launch<kFmaF32>(blocks, trips, out, s);
    case kExpF32: return launch<kExpF32>(blocks, trips, out, s);
    case kLogF32: return launch<kLogF32>(blocks, trips, out, s);
    case kSqrtF32: return launch<kSqrtF32>(blocks, trips, out, s);
    case kRcpF32: return launch<kRcpF32>(blocks, trips, out, s);
    case kSinF32: return launch<kSinF32>(blocks, trips, out, s);
    case kMadU64: return launch<kMadU64>(blocks, trips, out, s);
    case kFmaF64: return launch<kFmaF64>(blocks, trips, out, s);
    case kMix7Fma1Exp: return launch<kMix7Fma1Exp>(blocks, trips, out, s);
    case kFmaF32Vgpr: return launch<kFmaF32Vgpr>(blocks, trips, out, s);
    case kAddF32: return launch<kAddF32>(blocks, trips, out, s);
    case kCvtF32U32: return launch<kCvtF32U32>(blocks, trips, out, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

extern "C" {

int fsmc_probe_kinds(void) { return kKinds; }

// rates[k] = wave64 instructions per second of class k over the whole GPU
// (for the mixed kind: all of its instructions).  The grid is 8 workgroups
// of 4 waves per CU (8 waves per SIMD); every launch is timed with HIP
// events after one untimed launch.  Returns 0, or a negative status.
int fsmc_probe_issue_rates(int trips, double* rates, int n) {
  if (trips < 1 || rates == nullptr || n < 1) return -1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return -2;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) return -2;
  const int blocks = cus * 8;  // 8 workgroups of 4 waves per CU: 8 waves per SIMD
  float* out = nullptr;
  if (hipMalloc(&out, sizeof(float) * (size_t)blocks * 256) != hipSuccess) return -3;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int st = 0;
  if (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    st = -2;
  for (int k = 0; k < n && k < kKinds && st == 0; ++k) {
    float ms = 0.f;
    if (launch_kind(k, blocks, trips, out, s) != hipSuccess || hipEventRecord(e0, s) != hipSuccess ||
        launch_kind(k, blocks, trips, out, s) != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0.f) {
      st = -2;
      break;
    }
    const double waves = (double)blocks * 4.0;
    rates[k] = waves * (double)trips * insts_per_trip(k) / (ms * 1e-3);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  (void)hipFree(out);
  return st;
}

}  // extern "C"
