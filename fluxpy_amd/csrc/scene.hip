// Scene build: linear BVH over the triangle mesh, entirely on the GPU.
//
// Replaces the CGAL AABB tree build (cgal/aabb_wrapper.cpp:40-57) and the
// Embree scene commit (shape.py:296-344).  Steps:
//   1. FP64 triangle copy + FP64 centroids ((v0+v1)+v2)/3 (the ray origins of
//      CGAL mode, aabb_wrapper.cpp:69-77);
//   2. 30-bit Morton codes of the centroids, keyed with the facet index so
//      every key is unique (keeps Karras' construction well defined);
//   3. bitonic sort in LDS + global merge passes;
//   4. Karras (2012) "Maximizing parallelism in the construction of BVHs"
//      hierarchy, one thread per internal node;
//   5. bottom-up refit with per-node arrival counters; leaf boxes are the
//      FP64 triangle boxes rounded outward to FP32 and padded, so the FP32
//      slab test never culls a triangle the FP64 triangle test would hit;
//   6. preorder layouts with escape links (stackless traversal): a node's
//      escape is its preorder position plus its subtree size.  There are
//      eight layouts, one per octant of ray directions: in the layout of
//      octant o (bit c set: direction component c negative) the child whose
//      box centre lies earlier along s_o = (+-1, +-1, +-1) comes first, so a
//      ray walks its octant's layout roughly front to back and an any-hit
//      query meets a blocker near its origin sooner.  Every layout holds
//      the same boxes, so results do not depend on which one a ray walks.
#include "common.h"

namespace flux {

template <typename T>
__global__ void k_tri_setup(const T *__restrict__ V, const int64_t *__restrict__ F,
                            int64_t nf, int64_t nv, double *__restrict__ tri,
                            double *__restrict__ cen, int *__restrict__ bad) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  double v[9];
  for (int k = 0; k < 3; ++k) {
    int64_t vi = F[f * 3 + k];
    if (vi < 0 || vi >= nv) {
      atomicOr(bad, 1);
      vi = 0;
    }
    for (int c = 0; c < 3; ++c) v[k * 3 + c] = (double)V[vi * 3 + c];
  }
  for (int k = 0; k < 9; ++k) tri[f * 9 + k] = v[k];
  for (int c = 0; c < 3; ++c) cen[f * 3 + c] = ((v[c] + v[3 + c]) + v[6 + c]) / 3.0;
}

// Single-workgroup reduction: bounds of the centroids (Morton frame) and of
// the vertices (padding scale).  out: cmin[3], cmax[3], vmin[3], vmax[3].
__global__ void k_bounds(const double *__restrict__ tri, const double *__restrict__ cen,
                         int64_t nf, double *__restrict__ out) {
  __shared__ double s[12][256];
  double r[12];
  for (int c = 0; c < 3; ++c) {
    r[c] = INFINITY; r[3 + c] = -INFINITY; r[6 + c] = INFINITY; r[9 + c] = -INFINITY;
  }
  for (int64_t f = threadIdx.x; f < nf; f += blockDim.x) {
    for (int c = 0; c < 3; ++c) {
      double x = cen[f * 3 + c];
      r[c] = fmin(r[c], x);
      r[3 + c] = fmax(r[3 + c], x);
      for (int k = 0; k < 3; ++k) {
        double v = tri[f * 9 + k * 3 + c];
        r[6 + c] = fmin(r[6 + c], v);
        r[9 + c] = fmax(r[9 + c], v);
      }
    }
  }
  for (int q = 0; q < 12; ++q) s[q][threadIdx.x] = r[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      for (int q = 0; q < 12; ++q) {
        double a = s[q][threadIdx.x], b = s[q][threadIdx.x + w];
        bool is_min = (q < 3) || (q >= 6 && q < 9);
        s[q][threadIdx.x] = is_min ? fmin(a, b) : fmax(a, b);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 12) out[threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint32_t expand_bits10(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

__global__ void k_morton(const double *__restrict__ cen, int64_t nf, int64_t P,
                         double ox, double oy, double oz, double sx, double sy,
                         double sz, uint64_t *__restrict__ keys) {
  int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= P) return;
  if (f >= nf) {
    keys[f] = ~0ull;
    return;
  }
  double q[3] = {(cen[f * 3 + 0] - ox) * sx, (cen[f * 3 + 1] - oy) * sy,
                 (cen[f * 3 + 2] - oz) * sz};
  uint32_t code = 0;
  for (int c = 0; c < 3; ++c) {
    double x = fmin(fmax(q[c], 0.0), 1023.0);
    code |= expand_bits10((uint32_t)x) << (2 - c);
  }
  keys[f] = ((uint64_t)code << 32) | (uint64_t)f;
}

// Bitonic sort.  The direction of stage k at element i is (i & k) == 0
// (ascending) with the GLOBAL index, so tiles sorted in LDS merge correctly.
__global__ void k_bitonic_local(uint64_t *__restrict__ keys, int64_t tile,
                                int64_t kmin, int64_t kmax, int64_t jstart) {
  extern __shared__ uint64_t s[];
  int64_t base = (int64_t)blockIdx.x * tile;
  for (int64_t t = threadIdx.x; t < tile; t += blockDim.x) s[t] = keys[base + t];
  __syncthreads();
  for (int64_t k = kmin; k <= kmax; k <<= 1) {
    for (int64_t j = (k == kmin ? jstart : k >> 1); j >= 1; j >>= 1) {
      for (int64_t t = threadIdx.x; t < tile / 2; t += blockDim.x) {
        int64_t lo = ((t / j) * 2 * j) + (t % j);
        int64_t hi = lo + j;
        int64_t gi = base + lo;
        bool asc = (gi & k) == 0;
        uint64_t a = s[lo], b = s[hi];
        if ((a > b) == asc) {
          s[lo] = b;
          s[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  for (int64_t t = threadIdx.x; t < tile; t += blockDim.x) keys[base + t] = s[t];
}

__global__ void k_bitonic_global(uint64_t *__restrict__ keys, int64_t P, int64_t k,
                                 int64_t j) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P / 2) return;
  int64_t lo = ((t / j) * 2 * j) + (t % j);
  int64_t hi = lo + j;
  bool asc = (lo & k) == 0;
  uint64_t a = keys[lo], b = keys[hi];
  if ((a > b) == asc) {
    keys[lo] = b;
    keys[hi] = a;
  }
}

__device__ inline int kdelta(const uint64_t *keys, int n, int a, int b) {
  if (b < 0 || b >= n) return -1;
  return __clzll(keys[a] ^ keys[b]);
}

// Unified node ids: internal nodes 0..n-2, leaf s has id n-1+s.
__global__ void k_karras(const uint64_t *__restrict__ keys, int n, int *__restrict__ lch,
                         int *__restrict__ rch, int *__restrict__ rlo,
                         int *__restrict__ rhi, int *__restrict__ parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = kdelta(keys, n, i, i - d);
  int lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
  int l = 0;
  for (int t = lmax / 2; t >= 1; t /= 2)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = kdelta(keys, n, i, j);
  int s = 0, div = 2, t;
  do {
    t = (l + div - 1) / div;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    div *= 2;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int lo = min(i, j), hi = max(i, j);
  int left = (lo == gamma) ? (n - 1 + gamma) : gamma;
  int right = (hi == gamma + 1) ? (n - 1 + gamma + 1) : gamma + 1;
  lch[i] = left;
  rch[i] = right;
  rlo[i] = lo;
  rhi[i] = hi;
  parent[left] = i;
  parent[right] = i;
}

__global__ void k_refit(const double *__restrict__ tri, const uint64_t *__restrict__ keys,
                        int n, float pad, const int *__restrict__ parent,
                        const int *__restrict__ lch, const int *__restrict__ rch,
                        float *boxes, int *flags) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int f = (int)(keys[s] & 0xffffffffull);
  const double *v = tri + (int64_t)f * 9;
  float b[6];
  for (int c = 0; c < 3; ++c) {
    double lo = fmin(fmin(v[c], v[3 + c]), v[6 + c]);
    double hi = fmax(fmax(v[c], v[3 + c]), v[6 + c]);
    b[c] = __double2float_rd(lo) - pad;
    b[3 + c] = __double2float_ru(hi) + pad;
  }
  int uid = n - 1 + s;
  for (int q = 0; q < 6; ++q) boxes[(int64_t)uid * 6 + q] = b[q];
  __threadfence();
  int u = parent[uid];
  for (int steps = 0; u >= 0 && steps < 256; ++steps) {
    int old = atomicAdd(&flags[u], 1);
    if (old == 0) return;
    __threadfence();
    int a = lch[u], c2 = rch[u];
    const volatile float *ba = boxes + (int64_t)a * 6;
    const volatile float *bb = boxes + (int64_t)c2 * 6;
    for (int c = 0; c < 3; ++c) {
      b[c] = fminf(ba[c], bb[c]);
      b[3 + c] = fmaxf(ba[3 + c], bb[3 + c]);
    }
    for (int q = 0; q < 6; ++q) boxes[(int64_t)u * 6 + q] = b[q];
    __threadfence();
    u = parent[u];
  }
}

__device__ inline int subtree_size(int u, int n, const int *rlo, const int *rhi) {
  if (u >= n - 1) return 1;
  return 2 * (rhi[u] - rlo[u] + 1) - 1;
}

// true if child l precedes child r in the layout of octant oct: its box
// centre has the smaller coordinate along s_oct (ties: l first)
__device__ inline bool child_first(const float *__restrict__ boxes, int l, int r, int oct) {
  float kl = 0.0f, kr = 0.0f;
  for (int c = 0; c < 3; ++c) {
    const float sg = ((oct >> c) & 1) ? -1.0f : 1.0f;
    kl += sg * (boxes[(int64_t)l * 6 + c] + boxes[(int64_t)l * 6 + 3 + c]);
    kr += sg * (boxes[(int64_t)r * 6 + c] + boxes[(int64_t)r * 6 + 3 + c]);
  }
  return kl <= kr;
}

// pos[oct * n_all + u] = preorder position of node u in the layout of octant oct
__global__ void k_preorder(const int *__restrict__ parent, const int *__restrict__ lch,
                           const int *__restrict__ rch, const int *__restrict__ rlo,
                           const int *__restrict__ rhi, const float *__restrict__ boxes, int n,
                           int *__restrict__ pos) {
  const int uid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_all = 2 * n - 1;
  if (uid >= n_all) return;
  for (int oct = 0; oct < kOctants; ++oct) {
    int p = 0, u = uid;
    // A Karras tree over unique 64-bit keys is at most 64 levels deep; the
    // step bound only guarantees termination.
    for (int steps = 0; parent[u] >= 0 && steps < 256; ++steps) {
      const int par = parent[u];
      const int l = lch[par], r = rch[par];
      const int first = child_first(boxes, l, r, oct) ? l : r;
      p += (first == u) ? 1 : 1 + subtree_size(first, n, rlo, rhi);
      u = par;
    }
    pos[(int64_t)oct * n_all + uid] = p;
  }
}

__global__ void k_emit(const uint64_t *__restrict__ keys, const float *__restrict__ boxes,
                       const int *__restrict__ pos, const int *__restrict__ rlo,
                       const int *__restrict__ rhi, int n, float4 *__restrict__ nodes) {
  const int uid = blockIdx.x * blockDim.x + threadIdx.x;
  const int n_all = 2 * n - 1;
  if (uid >= n_all) return;
  const int sz = subtree_size(uid, n, rlo, rhi);
  const int t = (uid >= n - 1) ? (int)(keys[uid - (n - 1)] & 0xffffffffull) : -1;
  const float *b = boxes + (int64_t)uid * 6;
  for (int oct = 0; oct < kOctants; ++oct) {
    const int p = pos[(int64_t)oct * n_all + uid];
    float4 *nd = nodes + (int64_t)oct * 2 * n_all;
    nd[2 * p] = make_float4(b[0], b[1], b[2], __int_as_float(p + sz));
    nd[2 * p + 1] = make_float4(b[3], b[4], b[5], __int_as_float(t));
  }
}

static void scene_free(flux_scene *s) {
  if (!s) return;
  if (s->tri) (void)hipFree(s->tri);
  if (s->cen) (void)hipFree(s->cen);
  if (s->nodes) (void)hipFree(s->nodes);
  delete s;
}

struct TempBuf {
  void *p = nullptr;
  ~TempBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace flux

using namespace flux;

extern "C" int flux_scene_create(const void *V, int64_t nv, const int64_t *F,
                                 int64_t nf, int dtype, void *stream_,
                                 flux_scene **out) {
  FLUX_CHECK(out != nullptr, "flux_scene_create: out is NULL");
  *out = nullptr;
  FLUX_CHECK(nf >= 0 && nv >= 0, "flux_scene_create: negative sizes");
  FLUX_CHECK(nf < (int64_t)1 << 30, "flux_scene_create: too many facets (%lld)",
             (long long)nf);
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64,
             "flux_scene_create: V must be float32 or float64");
  FLUX_CHECK(nf == 0 || (V && F), "flux_scene_create: NULL V or F");
  hipStream_t st = as_stream(stream_);
  flux_scene *s = new flux_scene();
  FLUX_TRY_HIP(hipGetDevice(&s->device));
  s->nv = nv;
  s->nf = nf;
  s->nnodes = nf > 0 ? (int32_t)(2 * nf - 1) : 0;
  if (nf == 0) {
    *out = s;
    return 0;
  }
  int rc = 0;
  TempBuf tb_keys, tb_int, tb_box, tb_bounds;
  do {
    if (hipMalloc(&s->tri, nf * 9 * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->cen, nf * 3 * sizeof(double)) != hipSuccess ||
        hipMalloc(&s->nodes, (size_t)kOctants * s->nnodes * 2 * sizeof(float4)) != hipSuccess) {
      set_error("flux_scene_create: out of device memory (nf=%lld)", (long long)nf);
      rc = -1;
      break;
    }
    int64_t P = 1;
    while (P < nf) P <<= 1;
    int64_t n_int = nf - 1;
    int64_t n_all = 2 * nf - 1;
    if (hipMalloc(&tb_keys.p, P * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&tb_int.p, (size_t)((5 + kOctants) * n_all + nf + 16) * sizeof(int)) !=
            hipSuccess ||
        hipMalloc(&tb_box.p, (size_t)n_all * 6 * sizeof(float)) != hipSuccess ||
        hipMalloc(&tb_bounds.p, 16 * sizeof(double)) != hipSuccess) {
      set_error("flux_scene_create: out of device memory for temporaries");
      rc = -1;
      break;
    }
    int *ib = (int *)tb_int.p;
    int *lch = ib;
    int *rch = lch + n_all;
    int *rlo = rch + n_all;
    int *rhi = rlo + n_all;
    int *parent = rhi + n_all;
    int *flags = parent + n_all;
    int *pos = flags + nf;               // kOctants * n_all
    int *bad = pos + kOctants * n_all;
    (void)n_int;
    if (hipMemsetAsync(bad, 0, sizeof(int), st) != hipSuccess ||
        hipMemsetAsync(flags, 0, nf * sizeof(int), st) != hipSuccess) {
      set_error("flux_scene_create: memset failed");
      rc = -1;
      break;
    }
    const int TB = 256;
    if (dtype == FLUX_F64)
      hipLaunchKernelGGL(k_tri_setup<double>, dim3(ceil_div_u(nf, TB)), dim3(TB), 0, st,
                         (const double *)V, F, nf, nv, s->tri, s->cen, bad);
    else
      hipLaunchKernelGGL(k_tri_setup<float>, dim3(ceil_div_u(nf, TB)), dim3(TB), 0, st,
                         (const float *)V, F, nf, nv, s->tri, s->cen, bad);
    double *bounds = (double *)tb_bounds.p;
    hipLaunchKernelGGL(k_bounds, dim3(1), dim3(256), 0, st, s->tri, s->cen, nf, bounds);
    double hb[12];
    int hbad = 0;
    if (hipMemcpyAsync(hb, bounds, 12 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      set_error("flux_scene_create: bounds readback failed: %s",
                hipGetErrorString(hipGetLastError()));
      rc = -1;
      break;
    }
    if (hbad) {
      set_error("flux_scene_create: face index out of range [0, %lld)", (long long)nv);
      rc = -2;
      break;
    }
    double maxabs = 0.0, extent = 0.0;
    for (int c = 0; c < 3; ++c) {
      maxabs = fmax(maxabs, fmax(fabs(hb[6 + c]), fabs(hb[9 + c])));
      extent = fmax(extent, hb[9 + c] - hb[6 + c]);
    }
    if (!(maxabs < 1e30)) {
      set_error("flux_scene_create: non-finite vertex coordinates");
      rc = -2;
      break;
    }
    // 2^-20 of the scene scale: far above every FP32 rounding error of the
    // slab test (2^-24 relative) yet small next to any facet.
    s->pad = (float)(fmax(maxabs, extent) * 0x1p-20) + 1e-30f;
    double sc[3];
    for (int c = 0; c < 3; ++c) {
      double w = hb[3 + c] - hb[c];
      sc[c] = w > 0 ? 1023.0 / w : 0.0;
    }
    uint64_t *keys = (uint64_t *)tb_keys.p;
    hipLaunchKernelGGL(k_morton, dim3(ceil_div_u(P, TB)), dim3(TB), 0, st, s->cen, nf, P,
                       hb[0], hb[1], hb[2], sc[0], sc[1], sc[2], keys);
    // sort
    int64_t tile = P < 1024 ? P : 1024;
    if (P >= 2) {
      int thr = (int)(tile / 2 < 256 ? tile / 2 : 256);
      hipLaunchKernelGGL(k_bitonic_local, dim3((unsigned)(P / tile)), dim3(thr),
                         tile * sizeof(uint64_t), st, keys, tile, (int64_t)2, tile,
                         (int64_t)1);
      for (int64_t k = 2 * tile; k <= P; k <<= 1) {
        for (int64_t j = k >> 1; j >= tile; j >>= 1)
          hipLaunchKernelGGL(k_bitonic_global, dim3(ceil_div_u(P / 2, TB)), dim3(TB), 0, st,
                             keys, P, k, j);
        hipLaunchKernelGGL(k_bitonic_local, dim3((unsigned)(P / tile)), dim3(thr),
                           tile * sizeof(uint64_t), st, keys, tile, k, k, tile >> 1);
      }
    }
    int n = (int)nf;
    if (hipMemsetAsync(parent, 0xff, n_all * sizeof(int), st) != hipSuccess) {
      set_error("flux_scene_create: memset failed");
      rc = -1;
      break;
    }
    if (n >= 2)
      hipLaunchKernelGGL(k_karras, dim3(ceil_div_u(n - 1, TB)), dim3(TB), 0, st, keys, n,
                         lch, rch, rlo, rhi, parent);
    float *boxes = (float *)tb_box.p;
    hipLaunchKernelGGL(k_refit, dim3(ceil_div_u(n, TB)), dim3(TB), 0, st, s->tri, keys, n,
                       s->pad, parent, lch, rch, boxes, flags);
    hipLaunchKernelGGL(k_preorder, dim3(ceil_div_u(n_all, TB)), dim3(TB), 0, st, parent,
                       lch, rch, rlo, rhi, boxes, n, pos);
    hipLaunchKernelGGL(k_emit, dim3(ceil_div_u(n_all, TB)), dim3(TB), 0, st, keys, boxes,
                       pos, rlo, rhi, n, s->nodes);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_error("flux_scene_create: build kernels failed: %s", hipGetErrorString(e));
      rc = -1;
      break;
    }
  } while (0);
  if (rc != 0) {
    scene_free(s);
    return rc;
  }
  *out = s;
  return 0;
}

extern "C" int flux_scene_destroy(flux_scene *scene) {
  scene_free(scene);
  return 0;
}

extern "C" int flux_scene_num_faces(const flux_scene *scene, int64_t *nf) {
  FLUX_CHECK(scene && nf, "flux_scene_num_faces: NULL argument");
  *nf = scene->nf;
  return 0;
}

extern "C" int flux_scene_num_nodes(const flux_scene *scene, int64_t *nnodes) {
  FLUX_CHECK(scene && nnodes, "flux_scene_num_nodes: NULL argument");
  *nnodes = scene->nnodes;
  return 0;
}

extern "C" int flux_scene_centroids(const flux_scene *scene, const double **cen) {
  FLUX_CHECK(scene && cen, "flux_scene_centroids: NULL argument");
  *cen = scene->cen;
  return 0;
}

extern "C" int flux_scene_copy_nodes(const flux_scene *scene, float *nodes_out,
                                     void *stream) {
  FLUX_CHECK(scene && (nodes_out || scene->nnodes == 0),
             "flux_scene_copy_nodes: NULL argument");
  if (scene->nnodes == 0) return 0;
  FLUX_TRY_HIP(hipMemcpyAsync(nodes_out, scene->nodes,
                              (size_t)kOctants * scene->nnodes * 2 * sizeof(float4),
                              hipMemcpyDeviceToDevice, as_stream(stream)));
  return 0;
}
