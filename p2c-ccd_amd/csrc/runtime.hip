// Library-wide C-ABI plumbing: error text, version, device probe.
#include "runtime.h"

#include <cstring>

namespace p2c {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return set_error(P2C_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace p2c

extern "C" {

const char* p2c_last_error(void) { return p2c::g_last_error.c_str(); }

int p2c_version(void) { return 1; }

int p2c_device_available(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

int p2c_stream_synchronize(void* stream) {
  P2C_HIP_CHECK(hipStreamSynchronize(p2c::as_stream(stream)));
  return P2C_OK;
}

}  // extern "C"
