#include "dropin_common.hpp"

#include <mutex>

namespace fsmc_dropin {

fsmc_ctx* context() {
  static std::once_flag once;
  static fsmc_ctx* ctx = nullptr;
  std::call_once(once, [] {
    const int st = fsmc_ctx_create(0, &ctx);
    if (st != FSMC_OK) {
      report("fsmc_ctx_create", st);
      ctx = nullptr;
    }
  });
  return ctx;
}

void report(const char* where, int status) {
  std::fprintf(stderr, "fsmc: %s failed: %s (%d): %s\n", where, fsmc_status_string(status), status,
               fsmc_last_error());
}

}  // namespace fsmc_dropin
