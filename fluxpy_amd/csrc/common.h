// Shared helpers for libfluxhip: error reporting, launch checks, dtype
// dispatch.  Every kernel here targets gfx950 (wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdarg.h>
#include <string.h>

#include "../../include/fluxhip.h"

namespace flux {

// Thread-local message behind flux_last_error().
void set_error(const char *fmt, ...);

#define FLUX_TRY_HIP(expr)                                                   \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::flux::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                \
      return -1;                                                            \
    }                                                                       \
  } while (0)

#define FLUX_CHECK(cond, ...)                                               \
  do {                                                                      \
    if (!(cond)) {                                                          \
      ::flux::set_error(__VA_ARGS__);                                       \
      return -2;                                                            \
    }                                                                       \
  } while (0)

#define FLUX_LAUNCH_CHECK() FLUX_TRY_HIP(hipGetLastError())

static inline hipStream_t as_stream(void *s) { return (hipStream_t)s; }

static inline unsigned int ceil_div_u(int64_t a, int64_t b) {
  return (unsigned int)((a + b - 1) / b);
}

// Workgroups of a grid-stride launch over `total` elements: the dispatch
// packet counts work-items in 32 bits, so a launch of one thread per element
// must stay far below 2^32 threads; kernels launched this way loop.
constexpr int64_t kMaxGrid1D = 1 << 20;
static inline unsigned int grid_1d(int64_t total, int64_t tb) {
  const int64_t g = (total + tb - 1) / tb;
  return (unsigned int)(g < 1 ? 1 : (g < kMaxGrid1D ? g : kMaxGrid1D));
}

constexpr int kWave = 64;

}  // namespace flux

// Scene layout shared by the ray kernels (scene.hip builds it).
// BVH node layouts per scene: one per octant of ray directions (scene.hip)
constexpr int kOctants = 8;

struct flux_scene {
  int device;
  int64_t nv, nf;
  int32_t nnodes;
  double *tri;    // nf * 9: v0, v1, v2 in FP64, original facet order
  double *cen;    // nf * 3: ((v0+v1)+v2)/3 in FP64
  float4 *nodes;  // kOctants layouts of 2 * nnodes: {lo.xyz, escape bits}, {hi.xyz, tri bits}
  float pad;      // absolute box padding used at build
};
