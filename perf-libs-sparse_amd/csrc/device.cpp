// device.cpp — HIP runtime plumbing: device buffers, the library stream,
// host/device pointer detection and error recording.
#include <cstdio>
#include <cstring>

#include "plsparse.hpp"

namespace pls {

namespace {
hipStream_t g_stream = nullptr;
std::string g_last_error;
int g_device_state = 0;  // 0 unknown, 1 ok, -1 unavailable
}  // namespace

hipError_t DevBuf::alloc(size_t bytes) {
  reset();
  if (bytes == 0) bytes = 256;  // keep a valid pointer for empty arrays
  hipError_t e = hipMalloc(&p_, bytes);
  if (e != hipSuccess) {
    p_ = nullptr;
    return e;
  }
  bytes_ = bytes;
  return hipSuccess;
}

hipError_t DevBuf::upload(const void *host, size_t bytes) {
  hipError_t e = alloc(bytes);
  if (e != hipSuccess || bytes == 0 || host == nullptr) return e;
  return hipMemcpy(p_, host, bytes, hipMemcpyHostToDevice);
}

void DevBuf::reset() {
  if (p_) (void)hipFree(p_);
  p_ = nullptr;
  bytes_ = 0;
}

hipStream_t stream() { return g_stream; }
void set_stream(hipStream_t s) { g_stream = s; }

hipError_t ensure_device() {
  if (g_device_state == 1) return hipSuccess;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    g_device_state = -1;
    (void)hipGetLastError();
    return e != hipSuccess ? e : hipErrorNoDevice;
  }
  g_device_state = 1;
  return hipSuccess;
}

bool is_device_ptr(const void *p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t attr;
  std::memset(&attr, 0, sizeof(attr));
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

perflibs_status_t hip_fail(perflibs_spmat_top_t *A, hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  set_last_error(msg);
  if (A) {
    A->err.type = PERFLIBS_STATUS_EXECUTION_FAILURE;
    A->err.code = 0;
    A->err.msg = msg;
  }
  return PERFLIBS_STATUS_EXECUTION_FAILURE;
}

perflibs_status_t exec_fail(perflibs_spmat_top_t *A, const std::string &msg) {
  set_last_error(msg);
  if (A) {
    A->err.type = PERFLIBS_STATUS_EXECUTION_FAILURE;
    A->err.code = 0;
    A->err.msg = msg;
  }
  return PERFLIBS_STATUS_EXECUTION_FAILURE;
}

perflibs_status_t param_fail(perflibs_spmat_top_t *A, int64_t code, const std::string &msg) {
  if (!msg.empty()) set_last_error(msg);
  if (A) {
    A->err.type = PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
    A->err.code = code;
    A->err.msg = msg;
  }
  return PERFLIBS_STATUS_INPUT_PARAMETER_ERROR;
}

}  // namespace pls

extern "C" {

perflibs_status_t perflibs_amd_set_stream(void *stream) {
  pls::set_stream(static_cast<hipStream_t>(stream));
  return PERFLIBS_STATUS_SUCCESS;
}

void *perflibs_amd_get_stream(void) { return static_cast<void *>(pls::stream()); }

const char *perflibs_amd_last_error(void) { return pls::g_last_error.c_str(); }

}  // extern "C"
