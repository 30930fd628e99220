// sre_host.h -- host-side helpers shared by the C ABI translation units:
// the thread-local last-error message, HIP error propagation and typed
// device allocation.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sre.h"

namespace sre {

// Records printf-style text as the calling thread's last error; returns s.
sre_status fail(sre_status s, const char *fmt, ...);

// Bound positions of one replica of an engine (device pointer to its first
// atom, atom count, stride in floats) and its box.
sre_status engine_replica_positions(sre_engine *e, int replica, const float **pos,
                                    int *n, int *stride, float box[3]);

// The stream the engine's kernels run on (work another stream does on the
// engine's buffers waits for it through an event).
hipStream_t engine_stream(sre_engine *e);

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void **)p, count * sizeof(T));
}
template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

}  // namespace sre

#define HIP_TRY(call)                                                      \
  do {                                                                     \
    hipError_t _e = (call);                                                \
    if (_e != hipSuccess)                                                  \
      return ::sre::fail(SRE_ERR_HIP, "%s: %s (%s:%d)", #call,             \
                         hipGetErrorString(_e), __FILE__, __LINE__);       \
  } while (0)
