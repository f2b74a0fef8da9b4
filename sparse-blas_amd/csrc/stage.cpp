// stage.cpp -- copies between the caller's pageable host arrays and device
// memory through a pinned double buffer.
//
// A pageable hipMemcpy moves about 30 GB/s on the MI355X box; copying each
// chunk into pinned memory with several host threads while the DMA engine
// moves the previous chunk reaches about 37 GB/s, and the device-to-host
// direction gains more.  The drop-in entry points move the whole matrix on
// every call, so this is the bulk of their time.  The pinned buffers (two
// chunks) and the copy stream are created on first use and kept for the
// process; one transfer runs at a time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "spx_internal.h"

namespace spx {

namespace {

constexpr size_t kChunk = 32u << 20;   // bytes per pinned buffer
constexpr size_t kDirect = 4u << 20;   // smaller copies go straight through hipMemcpy
constexpr int kThreads = 4;            // host threads per chunk copy

std::mutex g_mu;
char *g_pin = nullptr;  // 2 * kChunk pinned bytes
hipStream_t g_stream = nullptr;
hipEvent_t g_ev[2] = {nullptr, nullptr};
int g_dev = -1;

hipError_t setup() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (g_pin && g_dev == dev) return hipSuccess;
  if (g_pin) {  // a different device is current: its stream and events are per device
    (void)hipStreamDestroy(g_stream);
    (void)hipEventDestroy(g_ev[0]);
    (void)hipEventDestroy(g_ev[1]);
    g_stream = nullptr;
  } else if ((e = hipHostMalloc((void **)&g_pin, 2 * kChunk, hipHostMallocDefault)) != hipSuccess) {
    g_pin = nullptr;
    return e;
  }
  if ((e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&g_ev[0], hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&g_ev[1], hipEventDisableTiming)) != hipSuccess) return e;
  g_dev = dev;
  return hipSuccess;
}

// memcpy of len bytes split over kThreads host threads
void par_copy(char *dst, const char *src, size_t len) {
  if (len < (1u << 20)) {
    std::memcpy(dst, src, len);
    return;
  }
  std::thread th[kThreads - 1];
  for (int k = 1; k < kThreads; k++) {
    const size_t a = len * k / kThreads, b = len * (k + 1) / kThreads;
    th[k - 1] = std::thread([=] { std::memcpy(dst + a, src + a, b - a); });
  }
  std::memcpy(dst, src, len / kThreads);
  for (auto &t : th) t.join();
}

}  // namespace

hipError_t stage_h2d(void *dst, const void *src, size_t bytes) {
  if (bytes < kDirect) return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  std::lock_guard<std::mutex> lk(g_mu);
  hipError_t e = setup();
  if (e != hipSuccess) return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  const char *s = static_cast<const char *>(src);
  char *d = static_cast<char *>(dst);
  size_t i = 0;
  for (size_t off = 0; off < bytes && e == hipSuccess; off += kChunk, i++) {
    const size_t len = std::min(kChunk, bytes - off);
    char *buf = g_pin + (i & 1) * kChunk;
    if (i >= 2 && (e = hipEventSynchronize(g_ev[i & 1])) != hipSuccess) break;  // buffer free again
    par_copy(buf, s + off, len);
    if ((e = hipMemcpyAsync(d + off, buf, len, hipMemcpyHostToDevice, g_stream)) != hipSuccess) break;
    e = hipEventRecord(g_ev[i & 1], g_stream);
  }
  const hipError_t e2 = hipStreamSynchronize(g_stream);
  return e != hipSuccess ? e : e2;
}

hipError_t stage_d2h(void *dst, const void *src, size_t bytes) {
  if (bytes < kDirect) return bytes ? hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) : hipSuccess;
  std::lock_guard<std::mutex> lk(g_mu);
  hipError_t e = setup();
  if (e != hipSuccess) return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  const char *s = static_cast<const char *>(src);
  char *d = static_cast<char *>(dst);
  const size_t n = (bytes + kChunk - 1) / kChunk;
  // chunk i lands in buffer i & 1; the DMA of chunk i + 1 runs while the
  // host copies chunk i out
  auto issue = [&](size_t i) {
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    hipError_t r = hipMemcpyAsync(g_pin + (i & 1) * kChunk, s + off, len, hipMemcpyDeviceToHost, g_stream);
    return r == hipSuccess ? hipEventRecord(g_ev[i & 1], g_stream) : r;
  };
  e = issue(0);
  for (size_t i = 0; i < n && e == hipSuccess; i++) {
    if (i + 1 < n && (e = issue(i + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(g_ev[i & 1])) != hipSuccess) break;
    const size_t off = i * kChunk, len = std::min(kChunk, bytes - off);
    par_copy(d + off, g_pin + (i & 1) * kChunk, len);
  }
  const hipError_t e2 = hipStreamSynchronize(g_stream);
  return e != hipSuccess ? e : e2;
}

}  // namespace spx
