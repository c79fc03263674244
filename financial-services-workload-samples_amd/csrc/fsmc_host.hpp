// Host-side plumbing shared by the C-ABI translation units: status handling,
// the per-device context with its grow-only device buffers, and argument
// validation helpers.  Nothing here is visible through include/fsmc.h except
// the opaque fsmc_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fsmc.h"

namespace fsmc {

// Thread-local detail message for the last failing call on this thread.
void set_error(const char* fmt, ...);

#define FSMC_HIP(expr)                                                                        \
  do {                                                                                        \
    hipError_t fsmc_e_ = (expr);                                                              \
    if (fsmc_e_ != hipSuccess) {                                                              \
      ::fsmc::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(fsmc_e_), __FILE__, \
                        __LINE__);                                                            \
      (void)hipGetLastError(); /* reported here: not again by a later check */                \
      return FSMC_ERR_DEVICE;                                                                 \
    }                                                                                         \
  } while (0)

#define FSMC_REQUIRE(cond, ...)      \
  do {                               \
    if (!(cond)) {                   \
      ::fsmc::set_error(__VA_ARGS__); \
      return FSMC_ERR_INVALID_ARG;   \
    }                                \
  } while (0)

#define FSMC_TRY(expr)                \
  do {                                \
    int fsmc_s_ = (expr);             \
    if (fsmc_s_ != FSMC_OK) return fsmc_s_; \
  } while (0)

// Grow-only device allocation owned by a context.
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  int reserve(size_t want);
  void release();
};

int validate_mode(const fsmc_mode* m, bool buffer_allowed, bool mrg_allowed = false);
// Enqueues positions offset .. offset+n-1 of the MKL-compatible MRG32k3a
// normal stream (mean + stddev * icdf(u)) into d_out on stream s.
int mrg_gaussian_enqueue(uint32_t seed, uint64_t offset, long long n, double mean, double stddev, double* d_out,
                         hipStream_t s);

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace fsmc

struct fsmc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  // Slots reused by the synchronous host-pointer entry points.
  fsmc::DeviceBuffer inputs, normals, outputs, workspace;
};

namespace fsmc {
// Makes ctx's device current and reserves the four slots.
int ctx_prepare(fsmc_ctx* ctx, size_t in_bytes, size_t normal_bytes, size_t out_bytes, size_t ws_bytes);
}  // namespace fsmc
