// Timing events for measuring launches on a stream.  They are created with
// hipEventDisableSystemFence: recording one does not write back and
// invalidate the caches, so it neither delays nor slows the work that
// follows (a plain event record does both).  Such an event only gives
// timestamps: it says nothing about the visibility of the work before it.
#include "common.h"

extern "C" int flux_timer_create(void **ev) {
  FLUX_CHECK(ev, "flux_timer_create: NULL out pointer");
  hipEvent_t e = nullptr;
  FLUX_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  *ev = (void *)e;
  return 0;
}

extern "C" int flux_timer_record(void *ev, void *stream) {
  FLUX_CHECK(ev, "flux_timer_record: NULL event");
  FLUX_TRY_HIP(hipEventRecord((hipEvent_t)ev, flux::as_stream(stream)));
  return 0;
}

extern "C" int flux_timer_elapsed(void *start, void *end, float *ms) {
  FLUX_CHECK(start && end && ms, "flux_timer_elapsed: NULL argument");
  FLUX_TRY_HIP(hipEventSynchronize((hipEvent_t)end));
  FLUX_TRY_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)end));
  return 0;
}

extern "C" int flux_timer_destroy(void *ev) {
  if (ev) FLUX_TRY_HIP(hipEventDestroy((hipEvent_t)ev));
  return 0;
}
