// common.h -- shared helpers of the HIP library (error handling, status).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace obit {

// Status codes of the ObitHIP* extension entry points (ObitHIPVis.h).
enum Status { kOK = 0, kInvalid = 1, kHipError = 2, kNoMemory = 3, kState = 4 };

void set_last_error(const std::string &msg);
int fail(int code, const char *fmt, ...);

// Count of host-to-device copies made through ObitCUDAUtilHost2GPU /
// ObitCUDAUtilHostAny2GPU: a cached host copy of a device table is stale
// once it changes.
unsigned long long upload_generation();

}  // namespace obit

// ObitCUDA* entry points abort on a HIP error, as the reference's
// checkCudaErrors does; the message names the call and the source line.
#define OBIT_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "obit_hip: %s failed at %s:%d: %s\n", #expr,       \
                   __FILE__, __LINE__, hipGetErrorString(e_));                 \
      std::abort();                                                            \
    }                                                                          \
  } while (0)

// ObitHIP* entry points return a status instead.
#define OBIT_HIP_TRY(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return obit::fail(obit::kHipError, "%s failed at %s:%d: %s", #expr,     \
                        __FILE__, __LINE__, hipGetErrorString(e_));            \
  } while (0)
