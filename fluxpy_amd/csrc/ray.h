// Device-side ray machinery shared by the visibility, occlusion and
// form-factor kernels.
//
// Triangle test: the watertight ray/triangle test of Woop, Benthin & Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013, evaluated in FP64
// with every operation rounded separately (this file must be compiled with
// -ffp-contract=off) so its decisions are reproducible bit for bit on the
// host.  Hits count for t >= 0 (the ray source included, as a CGAL Ray_3
// does); edges are shared (a ray through a common edge hits both facets).
//
// Traversal: stackless walk over a preorder node array of scene.hip with
// escape links (the layout of the ray's direction octant).  Node boxes are FP32, rounded outward and padded at build,
// so the FP32 slab test is conservative: the set of triangles tested always
// contains every triangle the FP64 test hits on the queried segment.  The
// result of every query is therefore independent of the tree shape.
#pragma once
#include "common.h"

namespace flux {

struct Ray64 {
  double o[3];
  double Sx, Sy, Sz;
  int kx, ky, kz;
};

__device__ __forceinline__ void ray_setup(Ray64 &r, const double o[3], const double d[3]) {
  r.o[0] = o[0];
  r.o[1] = o[1];
  r.o[2] = o[2];
  double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]);
  int kz = 0;
  double mx = ax;
  if (ay > mx) {
    kz = 1;
    mx = ay;
  }
  if (az > mx) kz = 2;
  int kx = kz + 1;
  if (kx == 3) kx = 0;
  int ky = kx + 1;
  if (ky == 3) ky = 0;
  if (d[kz] < 0.0) {
    int t = kx;
    kx = ky;
    ky = t;
  }
  r.kx = kx;
  r.ky = ky;
  r.kz = kz;
  r.Sx = d[kx] / d[kz];
  r.Sy = d[ky] / d[kz];
  r.Sz = 1.0 / d[kz];
}

__device__ __forceinline__ double sel3(const double *v, int k) {
  return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]);
}

// Returns t >= 0 of the hit, or -1 for a miss.  v: 9 doubles (A, B, C).
__device__ __forceinline__ double tri_hit(const Ray64 &r, const double *__restrict__ v) {
  const double A_x = sel3(v, r.kx) - sel3(r.o, r.kx);
  const double A_y = sel3(v, r.ky) - sel3(r.o, r.ky);
  const double A_z = sel3(v, r.kz) - sel3(r.o, r.kz);
  const double B_x = sel3(v + 3, r.kx) - sel3(r.o, r.kx);
  const double B_y = sel3(v + 3, r.ky) - sel3(r.o, r.ky);
  const double B_z = sel3(v + 3, r.kz) - sel3(r.o, r.kz);
  const double C_x = sel3(v + 6, r.kx) - sel3(r.o, r.kx);
  const double C_y = sel3(v + 6, r.ky) - sel3(r.o, r.ky);
  const double C_z = sel3(v + 6, r.kz) - sel3(r.o, r.kz);
  const double Ax = A_x - r.Sx * A_z;
  const double Ay = A_y - r.Sy * A_z;
  const double Bx = B_x - r.Sx * B_z;
  const double By = B_y - r.Sy * B_z;
  const double Cx = C_x - r.Sx * C_z;
  const double Cy = C_y - r.Sy * C_z;
  const double U = Cx * By - Cy * Bx;
  const double V = Ax * Cy - Ay * Cx;
  const double W = Bx * Ay - By * Ax;
  if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return -1.0;
  const double det = U + V + W;
  if (det == 0.0) return -1.0;
  const double Az = r.Sz * A_z;
  const double Bz = r.Sz * B_z;
  const double Cz = r.Sz * C_z;
  const double T = U * Az + V * Bz + W * Cz;
  const double t = T / det;
  if (!(t >= 0.0)) return -1.0;
  return t;
}

// FP32 ray for the slab test.  Zero or tiny direction components are
// replaced by +-1e-30 so that no 0*inf NaN can appear in the slab test.
struct RayBox {
  float o[3];
  float inv[3];
};

__device__ __forceinline__ void raybox_setup(RayBox &rb, const double o[3], const double d[3]) {
  for (int c = 0; c < 3; ++c) {
    rb.o[c] = (float)o[c];
    float dc = (float)d[c];
    float iv = 1.0f / dc;
    if (!(fabsf(iv) <= 1e30f)) iv = (dc < 0.0f || (dc == 0.0f && signbit(dc))) ? -1e30f : 1e30f;
    rb.inv[c] = iv;
  }
}

__device__ __forceinline__ bool slab(const RayBox &rb, const float4 &lo, const float4 &hi,
                                     float tmin, float tmax) {
  float t0x = (lo.x - rb.o[0]) * rb.inv[0], t1x = (hi.x - rb.o[0]) * rb.inv[0];
  float t0y = (lo.y - rb.o[1]) * rb.inv[1], t1y = (hi.y - rb.o[1]) * rb.inv[1];
  float t0z = (lo.z - rb.o[2]) * rb.inv[2], t1z = (hi.z - rb.o[2]) * rb.inv[2];
  float tn = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), tmin));
  float tf = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fminf(fmaxf(t0z, t1z), tmax));
  return tn <= tf;
}

// The node layout of the ray's direction octant (scene.hip): children in
// roughly front-to-back order for this ray.
__device__ __forceinline__ const float4 *octant_nodes(const flux_scene &S, const RayBox &rb) {
  const int oct = (rb.inv[0] < 0.0f ? 1 : 0) | (rb.inv[1] < 0.0f ? 2 : 0) |
                  (rb.inv[2] < 0.0f ? 4 : 0);
  return S.nodes + (int64_t)oct * 2 * S.nnodes;
}

// Any-hit query: is there a facet k (k != skip1, k != skip2) hit at t with
// (t, k) lexicographically before (tlim, klim)?  tlim = +inf / klim = INT_MAX
// gives "anything at t >= 0".
__device__ inline bool any_blocker(const flux_scene &S, const Ray64 &r, const RayBox &rb,
                                   double tlim, int klim, int skip1, int skip2) {
  const float tmax = isinf(tlim) ? 3.0e38f : (float)(tlim * (1.0 + 1e-5)) + 1e-30f;
  const float4 *__restrict__ nodes = octant_nodes(S, rb);
  const int nn = S.nnodes;
  int n = 0;
  while (n < nn) {
    const float4 lo = nodes[2 * n];
    const float4 hi = nodes[2 * n + 1];
    // escape links always point forward; max() only guarantees progress
    const int esc = max(__float_as_int(lo.w), n + 1);
    if (slab(rb, lo, hi, 0.0f, tmax)) {
      const int t = __float_as_int(hi.w);
      if (t >= 0) {
        if (t != skip1 && t != skip2 && t < S.nf) {
          double th = tri_hit(r, S.tri + (int64_t)t * 9);
          if (th >= 0.0 && (th < tlim || (th == tlim && t < klim))) return true;
        }
        n = esc;
      } else {
        n = n + 1;
      }
    } else {
      n = esc;
    }
  }
  return false;
}

// Closest hit (t, k) in lexicographic order, k != skip.  Returns k or -1.
__device__ inline int closest_hit(const flux_scene &S, const Ray64 &r, const RayBox &rb,
                                  int skip, double *t_out) {
  double best_t = INFINITY;
  int best_k = -1;
  const float4 *__restrict__ nodes = octant_nodes(S, rb);
  const int nn = S.nnodes;
  int n = 0;
  while (n < nn) {
    const float4 lo = nodes[2 * n];
    const float4 hi = nodes[2 * n + 1];
    // escape links always point forward; max() only guarantees progress
    const int esc = max(__float_as_int(lo.w), n + 1);
    const float tmax = best_k < 0 ? 3.0e38f : (float)(best_t * (1.0 + 1e-5)) + 1e-30f;
    if (slab(rb, lo, hi, 0.0f, tmax)) {
      const int t = __float_as_int(hi.w);
      if (t >= 0) {
        if (t != skip && t < S.nf) {
          double th = tri_hit(r, S.tri + (int64_t)t * 9);
          if (th >= 0.0 && (th < best_t || (th == best_t && t < best_k))) {
            best_t = th;
            best_k = t;
          }
        }
        n = esc;
      } else {
        n = n + 1;
      }
    } else {
      n = esc;
    }
  }
  *t_out = best_t;
  return best_k;
}

// eps of the Embree backend (shape.py:351, 405): 1e3 times
// np.finfo(np.float32).resolution, the float32 1e-6 times a Python float,
// an FP64 number (NumPy 1.x).
constexpr double kEmbreeEps = 1e3 * (double)1e-6f;

// Embree's ray for the pair (i, j) as fluxpy forms it (shape.py:355-385):
// d = P_j - P_i, |d| = sqrt((d0^2 + d1^2) + d2^2), the unit direction and
// the origin P_i + eps d/|d| in the model dtype T (NumPy applies the FP64
// eps to a T array in T), then stored into Embree's FP32 ray arrays.
// Returns false when |d| <= eps (the pair is visible by default).
template <typename T>
__device__ inline bool embree_pair_ray(const T *Pi, const T *Pj, double o[3], double d[3]) {
  const T e = (T)kEmbreeEps;
  const T dd[3] = {Pj[0] - Pi[0], Pj[1] - Pi[1], Pj[2] - Pi[2]};
  const T nrm = sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
  if (!(nrm > e)) return false;
  for (int c = 0; c < 3; ++c) {
    const T dh = dd[c] / nrm;
    const T oc = Pi[c] + e * dh;
    o[c] = (double)(float)oc;
    d[c] = (double)(float)dh;
  }
  return true;
}

// Embree's occlusion ray (shape.py:410-411): origin P + eps N in the model
// dtype, direction D (FP64), both stored as FP32.
template <typename T>
__device__ inline void embree_sun_ray(const T *Pi, const T *Ni, const double *D, double o[3],
                                      double d[3]) {
  const T e = (T)kEmbreeEps;
  for (int c = 0; c < 3; ++c) {
    const T oc = Pi[c] + e * Ni[c];
    o[c] = (double)(float)oc;
    d[c] = (double)(float)D[c];
  }
}

// Facet-to-facet visibility of facet i -> facet j.
//   CGAL mode (aabb_wrapper.cpp:64-93): origin = FP64 centroid of i,
//   direction = centroid_j - centroid_i, i skipped, visible iff j is the
//   closest hit; i == j -> false.
//   EMBREE mode (shape.py:349-398): the ray of embree_pair_ray from the
//   model centroids Pi, Pj (read in this mode only), nothing skipped,
//   visible iff j is the closest hit.
template <typename T>
__device__ inline bool face_to_face_vis(const flux_scene &S, int i, int j, int mode,
                                        const T *Pi, const T *Pj) {
  double o[3], d[3];
  int skip = -1;
  if (mode == FLUX_MODE_CGAL) {
    if (i == j) return false;
    const double *ci = S.cen + (int64_t)i * 3;
    const double *cj = S.cen + (int64_t)j * 3;
    for (int c = 0; c < 3; ++c) {
      o[c] = ci[c];
      d[c] = cj[c] - ci[c];
    }
    skip = i;
  } else {
    if (!embree_pair_ray(Pi, Pj, o, d)) return true;
  }
  Ray64 r;
  ray_setup(r, o, d);
  double tj = tri_hit(r, S.tri + (int64_t)j * 9);
  if (tj < 0.0) return false;
  RayBox rb;
  raybox_setup(rb, o, d);
  return !any_blocker(S, r, rb, tj, j, skip, j);
}

template <typename T>
__device__ __forceinline__ void load3(const T *p, int64_t k, double out[3]) {
  out[0] = (double)p[k * 3 + 0];
  out[1] = (double)p[k * 3 + 1];
  out[2] = (double)p[k * 3 + 2];
}

}  // namespace flux
