// Process-wide context shared by the C++ drop-in entry points.
#pragma once
#include <cstdio>

#include "fsmc.h"

namespace fsmc_dropin {

// Lazily created context on device 0; nullptr (with a message) if no device.
fsmc_ctx* context();

// Prints the failing call and the library's detail message to stderr.
void report(const char* where, int status);

}  // namespace fsmc_dropin
