// Visibility, sun occlusion and single-ray queries on a scene.
#include "ray.h"

namespace flux {

// One thread per (p, q) pair; rays of a workgroup share the source facet
// I[p], so their traversals stay coherent.
template <typename T>
__global__ void k_vis_mn(flux_scene S, const int64_t *__restrict__ I, int64_t m,
                         const int64_t *__restrict__ J, int64_t n, int mode,
                         const T *__restrict__ Pm, uint8_t *__restrict__ out) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t p = blockIdx.y + (int64_t)blockIdx.z * 65535;
  if (q >= n || p >= m) return;
  if (I[p] < 0 || I[p] >= S.nf || J[q] < 0 || J[q] >= S.nf) {
    out[p * n + q] = 0;
    return;
  }
  int i = (int)I[p], j = (int)J[q];
  T pi[3] = {0, 0, 0}, pj[3] = {0, 0, 0};
  if (mode == FLUX_MODE_EMBREE) {
    for (int c = 0; c < 3; ++c) {
      pi[c] = Pm[(int64_t)i * 3 + c];
      pj[c] = Pm[(int64_t)j * 3 + c];
    }
  }
  out[p * n + q] = face_to_face_vis(S, i, j, mode, pi, pj) ? 1 : 0;
}

// occlusion: CGAL mode ray from the FP64 centroid, facet skipped
// (aabb_wrapper.cpp:95-108); EMBREE mode from P + 1e-3 N, nothing skipped
// (shape.py:400-421).
template <typename T>
__global__ void k_occluded(flux_scene S, const int64_t *__restrict__ I, int64_t m,
                           const double *__restrict__ D, int64_t M, int per_face,
                           int mode, const T *__restrict__ Nrm,
                           const T *__restrict__ Pm, uint8_t *__restrict__ out) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t cols = per_face ? 1 : M;
  if (idx >= m * cols) return;
  int64_t p = idx / cols, q = idx % cols;
  if (I[p] < 0 || I[p] >= S.nf) {
    out[idx] = 0;
    return;
  }
  int i = (int)I[p];
  const double *dq = D + (per_face ? p : q) * 3;
  double o[3], d[3] = {dq[0], dq[1], dq[2]};
  int skip = -1;
  if (mode == FLUX_MODE_CGAL) {
    const double *ci = S.cen + (int64_t)i * 3;
    o[0] = ci[0];
    o[1] = ci[1];
    o[2] = ci[2];
    skip = i;
  } else {
    embree_sun_ray(Pm + (int64_t)i * 3, Nrm + (int64_t)i * 3, dq, o, d);
  }
  Ray64 r;
  ray_setup(r, o, d);
  RayBox rb;
  raybox_setup(rb, o, d);
  out[idx] = any_blocker(S, r, rb, INFINITY, 0x7fffffff, skip, -1) ? 1 : 0;
}

__global__ void k_intersect1(flux_scene S, const double *__restrict__ x,
                             const double *__restrict__ d, int64_t nr,
                             int64_t *__restrict__ hit, double *__restrict__ xt) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nr) return;
  double o[3] = {x[k * 3], x[k * 3 + 1], x[k * 3 + 2]};
  double dd[3] = {d[k * 3], d[k * 3 + 1], d[k * 3 + 2]};
  Ray64 r;
  ray_setup(r, o, dd);
  RayBox rb;
  raybox_setup(rb, o, dd);
  double t;
  int h = closest_hit(S, r, rb, -1, &t);
  hit[k] = h;
  for (int c = 0; c < 3; ++c) xt[k * 3 + c] = h >= 0 ? o[c] + t * dd[c] : 0.0;
}

// E[f, k] = F0 * max(0, N_f . d_k) if facet f sees sun k, else 0: the 1-D
// sun branch of get_direct_irradiance (shape.py:224-226) for M suns at once.
// The dot product is taken in FP64 (N . Dsun of a mesh-dtype N with an FP64
// direction promotes to FP64 in NumPy) and stored in the mesh dtype.
template <typename T>
__global__ void k_direct_irradiance(flux_scene S, const T *__restrict__ Nrm,
                                    const T *__restrict__ Pm, const double *__restrict__ D,
                                    int64_t M, double F0, int mode, T *__restrict__ E) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= S.nf * M) return;
  const int64_t f = idx / M, k = idx % M;
  const double d[3] = {D[k * 3], D[k * 3 + 1], D[k * 3 + 2]};
  double nf3[3];
  load3(Nrm, f, nf3);
  double o[3];
  int skip = -1;
  if (mode == FLUX_MODE_CGAL) {
    for (int c = 0; c < 3; ++c) o[c] = S.cen[f * 3 + c];
    skip = (int)f;
  }
  double dr[3] = {d[0], d[1], d[2]};  // the ray's direction (FP32 in EMBREE mode)
  if (mode != FLUX_MODE_CGAL) embree_sun_ray(Pm + f * 3, Nrm + f * 3, d, o, dr);
  Ray64 r;
  ray_setup(r, o, dr);
  RayBox rb;
  raybox_setup(rb, o, dr);
  const bool occ = any_blocker(S, r, rb, INFINITY, 0x7fffffff, skip, -1);
  const double cosi = (nf3[0] * d[0] + nf3[1] * d[1]) + nf3[2] * d[2];
  E[idx] = occ ? (T)0 : (T)(F0 * (cosi > 0.0 ? cosi : 0.0));
}

}  // namespace flux

using namespace flux;

extern "C" int flux_direct_irradiance(const flux_scene *scene, const void *P, const void *Nrm,
                                      int dtype, const double *D, int64_t M, double F0,
                                      int mode, void *E, void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_direct_irradiance: NULL scene");
  FLUX_CHECK(M >= 0, "flux_direct_irradiance: negative M");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_direct_irradiance: bad dtype");
  FLUX_CHECK(mode == FLUX_MODE_CGAL || mode == FLUX_MODE_EMBREE,
             "flux_direct_irradiance: bad mode");
  const int64_t total = scene->nf * M;
  if (total == 0) return 0;
  FLUX_CHECK(Nrm && D && E && (mode == FLUX_MODE_CGAL || P),
             "flux_direct_irradiance: NULL array");
  const int TB = 256;
  hipStream_t st = as_stream(stream);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_direct_irradiance<double>, dim3(ceil_div_u(total, TB)), dim3(TB), 0, st,
                       *scene, (const double *)Nrm, (const double *)P, D, M, F0, mode,
                       (double *)E);
  else
    hipLaunchKernelGGL(k_direct_irradiance<float>, dim3(ceil_div_u(total, TB)), dim3(TB), 0, st,
                       *scene, (const float *)Nrm, (const float *)P, D, M, F0, mode,
                       (float *)E);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_vis_mn(const flux_scene *scene, const int64_t *I, int64_t m,
                           const int64_t *J, int64_t n, int mode, const void *P,
                           int dtype, uint8_t *out, void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_vis_mn: NULL scene");
  FLUX_CHECK(m >= 0 && n >= 0, "flux_vis_mn: negative sizes");
  FLUX_CHECK(mode == FLUX_MODE_CGAL || mode == FLUX_MODE_EMBREE, "flux_vis_mn: bad mode %d",
             mode);
  FLUX_CHECK(mode == FLUX_MODE_CGAL || P != nullptr, "flux_vis_mn: EMBREE mode needs P");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_vis_mn: bad dtype");
  if (m == 0 || n == 0) return 0;
  FLUX_CHECK(I && J && out, "flux_vis_mn: NULL array");
  FLUX_CHECK(m <= (int64_t)65535 * 65535, "flux_vis_mn: m too large");
  hipStream_t st = as_stream(stream);
  const int TB = 256;
  dim3 grid(ceil_div_u(n, TB), (unsigned)(m < 65535 ? m : 65535),
            (unsigned)((m + 65534) / 65535));
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_vis_mn<double>, grid, dim3(TB), 0, st, *scene, I, m, J, n, mode,
                       (const double *)P, out);
  else
    hipLaunchKernelGGL(k_vis_mn<float>, grid, dim3(TB), 0, st, *scene, I, m, J, n, mode,
                       (const float *)P, out);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_occluded(const flux_scene *scene, const int64_t *I, int64_t m,
                             const double *D, int64_t M, int per_face, int mode,
                             const void *P, const void *Nrm, int dtype, uint8_t *out,
                             void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_occluded: NULL scene");
  FLUX_CHECK(m >= 0 && M >= 0, "flux_occluded: negative sizes");
  FLUX_CHECK(mode == FLUX_MODE_CGAL || mode == FLUX_MODE_EMBREE,
             "flux_occluded: bad mode %d", mode);
  FLUX_CHECK(!per_face || M == m, "flux_occluded: per_face needs M == m");
  FLUX_CHECK(dtype == FLUX_F32 || dtype == FLUX_F64, "flux_occluded: bad dtype");
  if (m == 0 || M == 0) return 0;
  FLUX_CHECK(I && D && out, "flux_occluded: NULL array");
  FLUX_CHECK(mode == FLUX_MODE_CGAL || (Nrm != nullptr && P != nullptr),
             "flux_occluded: EMBREE mode needs P and N");
  hipStream_t st = as_stream(stream);
  const int TB = 256;
  int64_t total = m * (per_face ? 1 : M);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_occluded<double>, dim3(ceil_div_u(total, TB)), dim3(TB), 0, st,
                       *scene, I, m, D, M, per_face, mode, (const double *)Nrm,
                       (const double *)P, out);
  else
    hipLaunchKernelGGL(k_occluded<float>, dim3(ceil_div_u(total, TB)), dim3(TB), 0, st,
                       *scene, I, m, D, M, per_face, mode, (const float *)Nrm,
                       (const float *)P, out);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_intersect1(const flux_scene *scene, const double *x, const double *d,
                               int64_t n_rays, int64_t *hit_index, double *xt,
                               void *stream) {
  FLUX_CHECK(scene != nullptr, "flux_intersect1: NULL scene");
  FLUX_CHECK(n_rays >= 0, "flux_intersect1: negative size");
  if (n_rays == 0) return 0;
  FLUX_CHECK(x && d && hit_index && xt, "flux_intersect1: NULL array");
  const int TB = 256;
  hipLaunchKernelGGL(k_intersect1, dim3(ceil_div_u(n_rays, TB)), dim3(TB), 0,
                     as_stream(stream), *scene, x, d, n_rays, hit_index, xt);
  FLUX_LAUNCH_CHECK();
  return 0;
}
