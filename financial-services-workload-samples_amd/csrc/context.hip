// Context, status strings and shared host helpers of the C ABI.
#include "fsmc_host.hpp"

namespace fsmc {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int DeviceBuffer::reserve(size_t want) {
  if (want <= bytes) return FSMC_OK;
  if (ptr) {
    // The previous contents are dead, but work enqueued earlier on any
    // stream may still use the block (the stream-ordered generator entry
    // points grow their slots without synchronising): wait for the device
    // before freeing it.
    (void)hipDeviceSynchronize();
    (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  const size_t n = align_up(want, 1u << 16);
  if (hipMalloc(&ptr, n) != hipSuccess) {
    (void)hipGetLastError();
    ptr = nullptr;
    set_error("hipMalloc(%zu) failed", n);
    return FSMC_ERR_ALLOC;
  }
  bytes = n;
  return FSMC_OK;
}

void DeviceBuffer::release() {
  if (ptr) (void)hipFree(ptr);
  ptr = nullptr;
  bytes = 0;
}

int validate_mode(const fsmc_mode* m, bool buffer_allowed, bool mrg_allowed) {
  FSMC_REQUIRE(m != nullptr, "mode is NULL");
  FSMC_REQUIRE(m->precision == FSMC_FP32 || m->precision == FSMC_FP64, "precision %d is neither FP32 nor FP64",
               m->precision);
  FSMC_REQUIRE(m->rng == FSMC_RNG_PHILOX || (buffer_allowed && m->rng == FSMC_RNG_BUFFER) ||
                   (mrg_allowed && m->rng == FSMC_RNG_MRG32K3A),
               "rng mode %d not supported here", m->rng);
  return FSMC_OK;
}

int ctx_prepare(fsmc_ctx* ctx, size_t in_bytes, size_t normal_bytes, size_t out_bytes, size_t ws_bytes) {
  FSMC_HIP(hipSetDevice(ctx->device));
  FSMC_TRY(ctx->inputs.reserve(in_bytes));
  FSMC_TRY(ctx->normals.reserve(normal_bytes));
  FSMC_TRY(ctx->outputs.reserve(out_bytes));
  FSMC_TRY(ctx->workspace.reserve(ws_bytes));
  return FSMC_OK;
}

}  // namespace fsmc

extern "C" {

int fsmc_ctx_create(int device, fsmc_ctx** out) {
  FSMC_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  int count = 0;
  FSMC_HIP(hipGetDeviceCount(&count));
  FSMC_REQUIRE(device >= 0 && device < count, "device %d out of range (%d devices)", device, count);
  FSMC_HIP(hipSetDevice(device));
  fsmc_ctx* c = new fsmc_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    fsmc::set_error("hipStreamCreate failed");
    delete c;
    return FSMC_ERR_DEVICE;
  }
  *out = c;
  return FSMC_OK;
}

int fsmc_ctx_destroy(fsmc_ctx* ctx) {
  if (!ctx) return FSMC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->inputs.release();
  ctx->normals.release();
  ctx->outputs.release();
  ctx->workspace.release();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FSMC_OK;
}

const char* fsmc_status_string(int status) {
  switch (status) {
    case FSMC_OK: return "ok";
    case FSMC_ERR_INVALID_ARG: return "invalid argument";
    case FSMC_ERR_DEVICE: return "device error";
    case FSMC_ERR_ALLOC: return "allocation failed";
    case FSMC_ERR_WORKSPACE: return "workspace too small";
    default: return "unknown status";
  }
}

const char* fsmc_last_error(void) { return fsmc::g_last_error; }

}  // extern "C"
