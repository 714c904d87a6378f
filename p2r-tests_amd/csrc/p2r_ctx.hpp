// Context object behind the opaque p2r_ctx handle of include/p2r.h.
#pragma once
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/p2r.h"

// Makes `device` current for the scope and restores the caller's device.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int device) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != device) (void)hipSetDevice(device);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

// Optional roctx ranges around the entry points (recorded by `rocprofv3 --marker-trace`), enabled with
// P2R_ROCTX=1 in the environment at p2r_create.  The marker library is opened at run time, so the
// library does not link against it and costs nothing when the ranges are off.
struct Roctx {
  int (*push)(const char*) = nullptr;
  int (*pop)() = nullptr;
  void* handle = nullptr;
  void load() {
    const char* e = std::getenv("P2R_ROCTX");
    if (!e || std::strcmp(e, "1") != 0) return;
    handle = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!handle) return;
    push = reinterpret_cast<int (*)(const char*)>(dlsym(handle, "roctxRangePushA"));
    pop = reinterpret_cast<int (*)()>(dlsym(handle, "roctxRangePop"));
    if (!push || !pop) push = nullptr, pop = nullptr;
  }
};

struct p2r_ctx {
  int device = 0;
  Roctx roctx;
  int num_cus = 0;
  char msg[256] = "no error";
  // Launch choices (p2r_set_tuning / the autotuners); defaults below.
  p2r_tuning tune = {256, 0, P2R_RAGGED_PAIR_STREAM, 256, 4, 0};
  // Streams of p2r_forward_streamed, created on first use (default priority).
  hipStream_t streams[16] = {};
  hipStream_t stream(int i) {
    if (!streams[i] && hipStreamCreateWithFlags(&streams[i], hipStreamNonBlocking) != hipSuccess) streams[i] = nullptr;
    return streams[i];
  }
  void release_streams() {
    for (hipStream_t& s : streams)
      if (s) {
        (void)hipStreamDestroy(s);
        s = nullptr;
      }
  }
  // Workspace owned by the context (device memory), grown on demand.
  void* ws = nullptr;
  size_t ws_bytes = 0;
  // This device's copy of the library's static scratch (p2r_forward_adjoint_noalloc), resolved at creation.
  void* scratch = nullptr;
  int scratch_status = P2R_ERR_HIP;

  int fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    return status;
  }
  int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(P2R_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return P2R_OK;
  }
  // Ensures at least `bytes` of workspace; returns nullptr on failure.  Growing it waits for the device
  // first: kernels of earlier calls, on any stream, may still be reading the old buffer.
  void* workspace(size_t bytes) {
    if (bytes <= ws_bytes) return ws;
    if (ws) (void)hipDeviceSynchronize();
    release_workspace();
    if (hipMalloc(&ws, bytes) != hipSuccess) {
      ws = nullptr;
      return nullptr;
    }
    ws_bytes = bytes;
    return ws;
  }
  void release_workspace() {
    if (ws) (void)hipFree(ws);
    ws = nullptr;
    ws_bytes = 0;
  }
};

// True when the byte ranges [a, a + na) and [b, b + nb) share a byte.
inline bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const char *x = static_cast<const char*>(a), *y = static_cast<const char*>(b);
  return na > 0 && nb > 0 && x < y + nb && y < x + na;
}

// A roctx range for the scope of one entry point (no-op when the context's ranges are off).
struct RangeScope {
  const Roctx* r;
  RangeScope(const p2r_ctx* c, const char* name) : r(c && c->roctx.push ? &c->roctx : nullptr) {
    if (r) r->push(name);
  }
  ~RangeScope() {
    if (r) r->pop();
  }
};
