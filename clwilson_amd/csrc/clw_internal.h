// clw_internal.h -- shared declarations of the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clwilson.h"

#define CLW_SP_COUNT 169u  // odd primes 7..1021 used for trial division

struct clw_ctx {
  int device;
  hipStream_t stream;
  hipDeviceProp_t prop;
};

struct clw_event {
  hipEvent_t ev;
};

// Records an error message for clw_last_error() and returns `code`.
int clw_fail(int code, const char* fmt, ...);
int clw_fail_hip(hipError_t e, const char* what);
// Checks hipGetLastError() after a launch.
int clw_check_launch(const char* kernel);
static inline hipStream_t clw_stream(clw_ctx* ctx) { return ctx->stream; }
int clw_upload_sieve_tables();
