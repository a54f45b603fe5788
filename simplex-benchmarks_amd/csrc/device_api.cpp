// Device plumbing of the C ABI: allocation, copies, streams and events on the
// HIP runtime the operators use.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <stdexcept>
#include <string>

#include "blas.h"
#include "op.h"

using namespace smf;

#define MF_HIP_CALL(call)                                                   \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));         \
      return -1;                                                            \
    }                                                                       \
  } while (0)

extern "C" {

int mf_device_count(int32_t* n) {
  int c = 0;
  MF_HIP_CALL(hipGetDeviceCount(&c));
  *n = c;
  return 0;
}

int mf_device_name(char* buf, int32_t len) {
  hipDeviceProp_t prop;
  int dev = 0;
  MF_HIP_CALL(hipGetDevice(&dev));
  MF_HIP_CALL(hipGetDeviceProperties(&prop, dev));
  std::string s = std::string(prop.name) + " " + prop.gcnArchName;
  std::strncpy(buf, s.c_str(), len - 1);
  buf[len - 1] = 0;
  return 0;
}

int mf_malloc(void** ptr, uint64_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) return 0;
  MF_HIP_CALL(hipMalloc(ptr, bytes));
  return 0;
}
int mf_free(void* ptr) {
  if (ptr) MF_HIP_CALL(hipFree(ptr));
  return 0;
}
int mf_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
  if (bytes) MF_HIP_CALL(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return 0;
}
int mf_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
  if (bytes) MF_HIP_CALL(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return 0;
}
int mf_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* stream) {
  if (bytes) MF_HIP_CALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
  return 0;
}
int mf_copy_kernel(void* dst, const void* src, uint64_t n, void* stream) {
  blas_convert<double, double>((int64_t)n, (const double*)src, (double*)dst, (hipStream_t)stream);
  MF_HIP_CALL(hipGetLastError());
  return 0;
}
int mf_memset(void* ptr, int32_t value, uint64_t bytes, void* stream) {
  if (bytes) MF_HIP_CALL(hipMemsetAsync(ptr, value, bytes, (hipStream_t)stream));
  return 0;
}
int mf_synchronize(void) {
  MF_HIP_CALL(hipDeviceSynchronize());
  return 0;
}
int mf_stream_create(void** stream) {
  hipStream_t s;
  MF_HIP_CALL(hipStreamCreate(&s));
  *stream = s;
  return 0;
}
int mf_stream_destroy(void* stream) {
  MF_HIP_CALL(hipStreamDestroy((hipStream_t)stream));
  return 0;
}
int mf_stream_synchronize(void* stream) {
  MF_HIP_CALL(hipStreamSynchronize((hipStream_t)stream));
  return 0;
}
int mf_event_create(void** event) {
  hipEvent_t e;
  MF_HIP_CALL(hipEventCreate(&e));
  *event = e;
  return 0;
}
int mf_event_destroy(void* event) {
  MF_HIP_CALL(hipEventDestroy((hipEvent_t)event));
  return 0;
}
int mf_event_record(void* event, void* stream) {
  MF_HIP_CALL(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
  return 0;
}
int mf_event_elapsed_ms(void* start, void* stop, float* ms) {
  MF_HIP_CALL(hipEventSynchronize((hipEvent_t)stop));
  MF_HIP_CALL(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return 0;
}

}  // extern "C"
