// libnpbx runtime: errors, device selection, memory, streams, events, the
// polling sync, and the library-owned per-device workspace.
#include <hip/hip_runtime.h>

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <mutex>
#include <sched.h>

#include "common.h"

#ifndef NPBX_VERSION_STRING
#define NPBX_VERSION_STRING "0.1.0"
#endif

namespace npbx {
namespace {
thread_local char g_last_error[512] = "";

int64_t monotonic_ns() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return static_cast<int64_t>(ts.tv_sec) * 1000000000LL + ts.tv_nsec;
}

size_t copy_str(char* buf, size_t len, const char* s) {
  if (!buf || len == 0) return 0;
  size_t n = strlen(s);
  if (n >= len) n = len - 1;
  memcpy(buf, s, n);
  buf[n] = '\0';
  return n;
}

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
constexpr int kMaxDevices = 64;
Workspace g_workspace[kMaxDevices];
std::mutex g_workspace_mu;
}  // namespace

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t err, const char* what, const char* file, int line) {
  const char* base = strrchr(file, '/');
  return fail(err == hipErrorOutOfMemory ? NPBX_ENOMEM : NPBX_EHIP, "%s failed: %s (%s:%d)", what,
              hipGetErrorString(err), base ? base + 1 : file, line);
}

void clear_error() { g_last_error[0] = '\0'; }

int check_launch(const char* kernel, const char* file, int line) {
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return hip_fail(err, kernel, file, line);
  return NPBX_OK;
}

int workspace(size_t bytes, void** ptr) {
  int dev = 0;
  NPBX_HIP(hipGetDevice(&dev));
  NPBX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_workspace_mu);
  Workspace& ws = g_workspace[dev];
  if (ws.bytes < bytes) {
    // Work already queued may still read the old buffer: drain the device
    // before releasing it.
    if (ws.ptr) {
      NPBX_HIP(hipDeviceSynchronize());
      NPBX_HIP(hipFree(ws.ptr));
      ws.ptr = nullptr;
      ws.bytes = 0;
    }
    size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
    NPBX_HIP(hipMalloc(&ws.ptr, want));
    ws.bytes = want;
  }
  *ptr = ws.ptr;
  return NPBX_OK;
}

}  // namespace npbx

using npbx::fail;

extern "C" {

int npbx_version(char* buf, size_t len) {
  NPBX_REQUIRE(buf && len > 0, "npbx_version: null buffer");
  npbx::copy_str(buf, len, NPBX_VERSION_STRING " gfx950");
  return NPBX_OK;
}

int npbx_last_error(char* buf, size_t len) {
  if (!buf || len == 0) return NPBX_EINVAL;
  npbx::copy_str(buf, len, npbx::g_last_error);
  return NPBX_OK;
}

int npbx_device_count(int* count) {
  NPBX_REQUIRE(count, "npbx_device_count: null pointer");
  int n = 0;
  hipError_t err = hipGetDeviceCount(&n);
  if (err == hipErrorNoDevice) {
    *count = 0;
    return NPBX_OK;
  }
  NPBX_HIP(err);
  *count = n;
  return NPBX_OK;
}

int npbx_init(int device) {
  int n = 0;
  NPBX_HIP(hipGetDeviceCount(&n));
  NPBX_REQUIRE(device >= 0 && device < n, "device %d not present (%d visible)", device, n);
  NPBX_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  NPBX_HIP(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(NPBX_EUNSUPPORTED, "device %d is %s; libnpbx is built for gfx950 only", device,
                prop.gcnArchName);
  // Set-up of the allocation-free entry point: its fixed scratch array's
  // address on this device.
  return npbx::adi_noalloc_setup();
}

// Releases the library-owned workspace of the current device (after the
// device has drained); the next call that needs it allocates again.  Plans
// and queues are released by their own destroy calls.
int npbx_finalize(void) {
  int dev = 0;
  NPBX_HIP(hipGetDevice(&dev));
  NPBX_REQUIRE(dev >= 0 && dev < npbx::kMaxDevices, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lock(npbx::g_workspace_mu);
  npbx::Workspace& ws = npbx::g_workspace[dev];
  if (ws.ptr) {
    NPBX_HIP(hipDeviceSynchronize());
    NPBX_HIP(hipFree(ws.ptr));
    ws.ptr = nullptr;
    ws.bytes = 0;
  }
  return NPBX_OK;
}

int npbx_malloc(void** ptr, size_t bytes) {
  NPBX_REQUIRE(ptr, "npbx_malloc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return NPBX_OK;
  NPBX_HIP(hipMalloc(ptr, bytes));
  return NPBX_OK;
}

int npbx_free(void* ptr) {
  if (ptr) NPBX_HIP(hipFree(ptr));
  return NPBX_OK;
}

int npbx_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return NPBX_OK;
  NPBX_REQUIRE(dst && src, "npbx_memcpy_h2d: null pointer");
  NPBX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, npbx::as_stream(stream)));
  return NPBX_OK;
}

int npbx_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return NPBX_OK;
  NPBX_REQUIRE(dst && src, "npbx_memcpy_d2h: null pointer");
  NPBX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, npbx::as_stream(stream)));
  return NPBX_OK;
}

int npbx_stream_create(void** stream) {
  NPBX_REQUIRE(stream, "npbx_stream_create: null pointer");
  hipStream_t s = nullptr;
  // Default priority and non-blocking with respect to the null stream.
  NPBX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return NPBX_OK;
}

int npbx_stream_destroy(void* stream) {
  if (stream) NPBX_HIP(hipStreamDestroy(npbx::as_stream(stream)));
  return NPBX_OK;
}

int npbx_sync(void* stream, int64_t deadline_ns) {
  hipStream_t s = npbx::as_stream(stream);
  int polls = 0;
  for (;;) {
    hipError_t q = hipStreamQuery(s);
    if (q == hipSuccess) return NPBX_OK;
    if (q != hipErrorNotReady) return npbx::hip_fail(q, "hipStreamQuery", __FILE__, __LINE__);
    if (deadline_ns > 0 && npbx::monotonic_ns() >= deadline_ns)
      return fail(NPBX_ETIMEDOUT, "npbx_sync: deadline passed with work still queued");
    // Spin briefly for short kernels, then back off to keep a core free.
    if (++polls < 2000) {
      sched_yield();
    } else {
      timespec ts{0, 20000};
      nanosleep(&ts, nullptr);
    }
  }
}

int npbx_event_create(void** event) {
  NPBX_REQUIRE(event, "npbx_event_create: null pointer");
  hipEvent_t e = nullptr;
  NPBX_HIP(hipEventCreate(&e));
  *event = e;
  return NPBX_OK;
}

int npbx_event_destroy(void* event) {
  if (event) NPBX_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return NPBX_OK;
}

int npbx_event_record(void* event, void* stream) {
  NPBX_REQUIRE(event, "npbx_event_record: null event");
  NPBX_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), npbx::as_stream(stream)));
  return NPBX_OK;
}

int npbx_event_elapsed_ms(void* start, void* stop, float* ms) {
  NPBX_REQUIRE(start && stop && ms, "npbx_event_elapsed_ms: null pointer");
  NPBX_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  NPBX_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                               reinterpret_cast<hipEvent_t>(stop)));
  return NPBX_OK;
}

}  // extern "C"
