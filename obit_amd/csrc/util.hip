// util.hip -- device, stream, event and memory helpers with the C ABI of
// Obit's CUDA utility layer (reference src/ObitCUDAUtil.cu:51-402).
#include "common.h"
#include "ObitCUDAUtil.h"
#include "ObitHIPVis.h"

#include <cstdarg>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <atomic>

namespace obit {

static std::mutex g_err_mu;
static std::string g_last_error;

void set_last_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_last_error = msg;
}

int fail(int code, const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  set_last_error(buf);
  return code;
}

}  // namespace obit

static std::atomic<unsigned long long> g_uploads{0};
namespace obit {
unsigned long long upload_generation() { return g_uploads.load(); }
}  // namespace obit

extern "C" {

const char *ObitHIPLastError(void) {
  static thread_local std::string copy;
  {
    std::lock_guard<std::mutex> lk(obit::g_err_mu);
    copy = obit::g_last_error;
  }
  return copy.c_str();
}

float CUDAMagicF(void) {
  union { char s[4]; float f; } fb;
  fb.s[0] = 'I'; fb.s[1] = 'N'; fb.s[2] = 'D'; fb.s[3] = 'E';
  return fb.f;
}

long ObitCUDACheckGPU(int cuda_device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return -1;
  if (cuda_device < 0 || cuda_device >= count) return -1;
  return 1;
}

void ObitCUDASetGPU(int cuda_device) {
  int device = cuda_device < 0 ? 0 : cuda_device;
  OBIT_HIP_CHECK(hipSetDevice(device));
}

void ObitCUDAResetGPU(void) { OBIT_HIP_CHECK(hipDeviceSynchronize()); }

void ObitCUDADeviceSynchronize(int *event) {
  (void)event;
  OBIT_HIP_CHECK(hipDeviceSynchronize());
}

int *ObitCUDAStreamCreate(void) {
  hipStream_t s;
  OBIT_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return reinterpret_cast<int *>(s);
}

void ObitCUDAStreamDestroy(int *stream) {
  if (stream) OBIT_HIP_CHECK(hipStreamDestroy(reinterpret_cast<hipStream_t>(stream)));
}

int *ObitCUDAEventCreate(void) {
  hipEvent_t e;
  OBIT_HIP_CHECK(hipEventCreate(&e));
  return reinterpret_cast<int *>(e);
}

void ObitCUDAEventDestroy(int *event) {
  if (event) OBIT_HIP_CHECK(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
}

void ObitCUDAEventRecord(int *event, int *stream) {
  OBIT_HIP_CHECK(hipEventRecord(reinterpret_cast<hipEvent_t>(event),
                                reinterpret_cast<hipStream_t>(stream)));
}

void ObitCUDAEventSynchronize(int *event) {
  OBIT_HIP_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
}

float *ObitCUDAUtilAllocHost(int memsize) {
  void *p = nullptr;
  OBIT_HIP_CHECK(hipHostMalloc(&p, (size_t)(memsize > 0 ? memsize : 1), hipHostMallocDefault));
  return static_cast<float *>(p);
}

void ObitCUDAUtilFreeHost(float *host) {
  if (host) OBIT_HIP_CHECK(hipHostFree(host));
}

float *ObitCUDAUtilAllocGPU(int memsize) {
  void *p = nullptr;
  OBIT_HIP_CHECK(hipMalloc(&p, (size_t)(memsize > 0 ? memsize : 1)));
  return static_cast<float *>(p);
}

void ObitCUDAUtilFreeGPU(float *GPU) {
  if (GPU) OBIT_HIP_CHECK(hipFree(GPU));
}

static void copy_any(void *dst, const void *src, size_t n, hipMemcpyKind kind, int *stream) {
  if (n == 0) return;
  if (stream) {
    OBIT_HIP_CHECK(hipMemcpyAsync(dst, src, n, kind, reinterpret_cast<hipStream_t>(stream)));
  } else {
    OBIT_HIP_CHECK(hipMemcpy(dst, src, n, kind));
  }
}

void ObitCUDAUtilHost2GPU(float *GPU, float *host, int memsize, int *stream) {
  copy_any(GPU, host, (size_t)memsize, hipMemcpyHostToDevice, stream);
  ++g_uploads;
}
void ObitCUDAUtilHostAny2GPU(void *GPU, void *host, int memsize, int *stream) {
  copy_any(GPU, host, (size_t)memsize, hipMemcpyHostToDevice, stream);
  ++g_uploads;
}
void ObitCUDAUtilGPU2Host(float *host, float *GPU, int memsize, int *stream) {
  copy_any(host, GPU, (size_t)memsize, hipMemcpyDeviceToHost, stream);
}
void ObitCUDAUtilGPUAny2Host(void *host, void *GPU, int memsize, int *stream) {
  copy_any(host, GPU, (size_t)memsize, hipMemcpyDeviceToHost, stream);
}

size_t ObitCUDAUtilMemory(int cuda_device) {
  hipDeviceProp_t prop;
  OBIT_HIP_CHECK(hipGetDeviceProperties(&prop, cuda_device < 0 ? 0 : cuda_device));
  return prop.totalGlobalMem;
}

void ObitCUDAUtilPointerType(void *ptr, char *label) {
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
    (void)hipGetLastError();
    std::fprintf(stderr, "%s type: unregistered host\n", label ? label : "");
    return;
  }
  const char *t = attr.type == hipMemoryTypeDevice ? "device"
                : attr.type == hipMemoryTypeHost   ? "host"
                                                   : "other";
  std::fprintf(stderr, "%s type: %s\n", label ? label : "", t);
}

// ---- size_t helpers for whole-data-set residency (ObitHIPVis.h) ----

int ObitHIPMalloc(void **ptr, size_t bytes) {
  if (!ptr) return obit::fail(obit::kInvalid, "ObitHIPMalloc: null ptr");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess)
    return obit::fail(obit::kNoMemory, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return obit::kOK;
}

int ObitHIPFree(void *ptr) {
  if (ptr) OBIT_HIP_TRY(hipFree(ptr));
  return obit::kOK;
}

int ObitHIPMemcpyH2D(void *dst, const void *src, size_t bytes) {
  if (bytes) OBIT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return obit::kOK;
}

int ObitHIPMemcpyD2H(void *dst, const void *src, size_t bytes) {
  if (bytes) OBIT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return obit::kOK;
}

int ObitHIPMemcpyD2D(void *dst, const void *src, size_t bytes) {
  if (bytes) OBIT_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return obit::kOK;
}

int ObitHIPDeviceSynchronize(void) {
  OBIT_HIP_TRY(hipDeviceSynchronize());
  return obit::kOK;
}

float ObitHIPEventElapsedMs(int *start, int *stop) {
  float ms = -1.0f;
  if (hipEventElapsedTime(&ms, reinterpret_cast<hipEvent_t>(start),
                          reinterpret_cast<hipEvent_t>(stop)) != hipSuccess) {
    (void)hipGetLastError();
    return -1.0f;
  }
  return ms;
}

int ObitHIPDeviceInfo(int device, char *name, int nameLen, int *nCU, int *clockKHz,
                      long long *memBytes) {
  hipDeviceProp_t prop;
  OBIT_HIP_TRY(hipGetDeviceProperties(&prop, device));
  if (name && nameLen > 0) {
    std::snprintf(name, (size_t)nameLen, "%s", prop.gcnArchName);
  }
  if (nCU) *nCU = prop.multiProcessorCount;
  if (clockKHz) *clockKHz = prop.clockRate;
  if (memBytes) *memBytes = (long long)prop.totalGlobalMem;
  return obit::kOK;
}

// Plain device copy kernel (16-byte vectors): the known-size reference the
// HBM-traffic counters are calibrated against.
__global__ void k_copy16(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int ObitHIPCopyKernel(void *dst, const void *src, size_t bytes, void *stream) {
  if ((bytes & 15) || ((size_t)dst & 15) || ((size_t)src & 15))
    return obit::fail(obit::kInvalid, "ObitHIPCopyKernel: 16-byte aligned sizes only");
  const size_t n = bytes / 16;
  if (n == 0) return obit::kOK;
  const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
  hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, (hipStream_t)stream,
                     (const float4 *)src, (float4 *)dst, n);
  OBIT_HIP_TRY(hipGetLastError());
  return obit::kOK;
}

}  // extern "C"
