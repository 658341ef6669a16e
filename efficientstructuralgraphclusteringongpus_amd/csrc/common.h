// Shared host-side helpers: status codes, thread-local error text, timers.
#pragma once

#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include <chrono>
#include <string>

#include "sgc.h"

namespace sgc {

// Records `code` and a printf-style message as this thread's last error and
// returns `code`, so call sites can write `return fail(SGC_EINVAL, ...)`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline double now_ms() {
  using namespace std::chrono;
  return duration_cast<duration<double, std::milli>>(steady_clock::now().time_since_epoch()).count();
}

// Stage progress of long host/device runs on stderr when SGC_PROGRESS is set.
void progress(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

// Plain or pinned host allocation; pinned memory comes from hipHostMalloc so
// that partitions can be streamed to the device with asynchronous copies.
void* host_alloc(size_t bytes, bool pinned);
void host_free(void* p, bool pinned);

}  // namespace sgc
