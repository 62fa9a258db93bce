// Library-wide C-ABI plumbing: error text, version, device probe.
#include "runtime.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

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

namespace {
struct Pending {
  hipEvent_t ev;
  size_t lo, hi;
};
struct StagingRing {
  char* buf = nullptr;
  size_t cap = 0, head = 0;
  std::vector<Pending> inflight;  // copies that may still read [lo, hi) of buf
  std::vector<hipEvent_t> spare;
};
thread_local std::unordered_map<int, StagingRing> g_rings;

int retire(StagingRing& r, size_t lo, size_t hi) {
  // drop finished copies; wait for the ones that read the stretch about to be written
  size_t keep = 0;
  for (size_t i = 0; i < r.inflight.size(); ++i) {
    Pending& p = r.inflight[i];
    const bool overlaps = p.lo < hi && lo < p.hi;
    hipError_t q = overlaps ? hipEventSynchronize(p.ev) : hipEventQuery(p.ev);
    if (q == hipErrorNotReady) {
      r.inflight[keep++] = p;
      continue;
    }
    if (q != hipSuccess) return hip_fail(q, "staging ring event");
    r.spare.push_back(p.ev);
  }
  r.inflight.resize(keep);
  return P2C_OK;
}
}  // namespace

int upload_table(const void* host, size_t bytes, void* dev, hipStream_t s) {
  if (bytes == 0) return P2C_OK;
  int d = 0;
  P2C_HIP_CHECK(hipGetDevice(&d));
  StagingRing& r = g_rings[d];
  const size_t need = (bytes + 255) & ~size_t(255);
  if (need > r.cap) {
    if (int rc = retire(r, 0, SIZE_MAX)) return rc;  // waits for every copy out of the old buffer
    if (r.buf != nullptr) P2C_HIP_CHECK(hipHostFree(r.buf));
    r.buf = nullptr;
    r.cap = std::max<size_t>(4 * need, size_t(1) << 20);
    P2C_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&r.buf), r.cap, hipHostMallocDefault));
    r.head = 0;
  }
  if (r.head + need > r.cap) r.head = 0;
  const size_t lo = r.head, hi = lo + need;
  if (int rc = retire(r, lo, hi)) return rc;
  std::memcpy(r.buf + lo, host, bytes);
  P2C_HIP_CHECK(hipMemcpyAsync(dev, r.buf + lo, bytes, hipMemcpyHostToDevice, s));
  hipEvent_t ev;
  if (r.spare.empty()) {
    P2C_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  } else {
    ev = r.spare.back();
    r.spare.pop_back();
  }
  P2C_HIP_CHECK(hipEventRecord(ev, s));
  r.inflight.push_back(Pending{ev, lo, hi});
  r.head = hi;
  return P2C_OK;
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
