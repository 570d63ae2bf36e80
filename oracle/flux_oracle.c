/*
 * flux_oracle.c -- CPU oracle for fluxpy's visibility / occlusion /
 * form-factor path.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and
 * the CPU-baseline leg of bench.py use it as the checker; the product never
 * calls it.
 *
 * It restates, in plain C (FP64, OpenMP over rows, -ffp-contract=off):
 *   - CGAL visibility   cgal/aabb_wrapper.cpp:64-93 (+ cgal/aabb.pyx:58-70:
 *                        i == j skipped -> false)
 *   - CGAL occlusion    cgal/aabb_wrapper.cpp:95-108, cgal/aabb.pyx:72-87
 *   - CGAL intersect1   cgal/aabb_wrapper.cpp:110-129
 *   - Embree visibility shape.py:349-398 (1e-3 offset, |d| <= 1e-3 visible)
 *   - Embree occlusion  shape.py:400-421 (origin P + 1e-3 N)
 *   - form factors      form_factors.py:11-72 (eps prefilter, CSR rows)
 * The ray/triangle test is the watertight test of Woop, Benthin & Wald
 * (JCGT 2013) in FP64; decisions use closest hit with ties broken by the
 * smaller facet index.  Two search structures are provided: brute force over
 * every facet, and a CPU BVH (median split, FP64 boxes padded outward) that
 * must agree with brute force exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double o[3];
  double Sx, Sy, Sz;
  int kx, ky, kz;
} oray;

static void oray_init(oray *r, const double o[3], const double d[3]) {
  double ax = fabs(d[0]), ay = fabs(d[1]), az = fabs(d[2]), mx;
  int kz = 0, kx, ky, t;
  mx = ax;
  if (ay > mx) {
    kz = 1;
    mx = ay;
  }
  if (az > mx) kz = 2;
  kx = (kz + 1) % 3;
  ky = (kx + 1) % 3;
  if (d[kz] < 0.0) {
    t = kx;
    kx = ky;
    ky = t;
  }
  memcpy(r->o, o, sizeof(r->o));
  r->kx = kx;
  r->ky = ky;
  r->kz = kz;
  r->Sx = d[kx] / d[kz];
  r->Sy = d[ky] / d[kz];
  r->Sz = 1.0 / d[kz];
}

/* t >= 0 of the hit, or -1. */
static double otri(const oray *r, const double *v) {
  const int kx = r->kx, ky = r->ky, kz = r->kz;
  double A_x = v[kx] - r->o[kx], A_y = v[ky] - r->o[ky], A_z = v[kz] - r->o[kz];
  double B_x = v[3 + kx] - r->o[kx], B_y = v[3 + ky] - r->o[ky], B_z = v[3 + kz] - r->o[kz];
  double C_x = v[6 + kx] - r->o[kx], C_y = v[6 + ky] - r->o[ky], C_z = v[6 + kz] - r->o[kz];
  double Ax = A_x - r->Sx * A_z, Ay = A_y - r->Sy * A_z;
  double Bx = B_x - r->Sx * B_z, By = B_y - r->Sy * B_z;
  double Cx = C_x - r->Sx * C_z, Cy = C_y - r->Sy * C_z;
  double U = Cx * By - Cy * Bx;
  double V = Ax * Cy - Ay * Cx;
  double W = Bx * Ay - By * Ax;
  double det, Az, Bz, Cz, T, t;
  if ((U < 0.0 || V < 0.0 || W < 0.0) && (U > 0.0 || V > 0.0 || W > 0.0)) return -1.0;
  det = U + V + W;
  if (det == 0.0) return -1.0;
  Az = r->Sz * A_z;
  Bz = r->Sz * B_z;
  Cz = r->Sz * C_z;
  T = U * Az + V * Bz + W * Cz;
  t = T / det;
  if (!(t >= 0.0)) return -1.0;
  return t;
}

/* ------------------------------------------------------------------ */
/* Scene                                                               */
/* ------------------------------------------------------------------ */

typedef struct {
  double lo[3], hi[3];
  int left, right; /* children, or -1 */
  int first, count; /* leaf range into perm */
} onode;

typedef struct {
  int64_t nf;
  double *tri; /* nf*9 */
  double *cen; /* nf*3 */
  int *perm;
  onode *nodes;
  int nnodes;
  int use_bvh;
} oscene;

static double *g_keyc; /* sort helper (single-threaded build) */
static int g_axis;
static int cmp_axis(const void *a, const void *b) {
  int ia = *(const int *)a, ib = *(const int *)b;
  double xa = g_keyc[ia * 3 + g_axis], xb = g_keyc[ib * 3 + g_axis];
  if (xa < xb) return -1;
  if (xa > xb) return 1;
  return ia < ib ? -1 : (ia > ib);
}

static int build_rec(oscene *s, int first, int count, double pad) {
  int id = s->nnodes++, c, k, f;
  onode *n = &s->nodes[id];
  for (c = 0; c < 3; ++c) {
    n->lo[c] = INFINITY;
    n->hi[c] = -INFINITY;
  }
  for (k = first; k < first + count; ++k) {
    f = s->perm[k];
    for (c = 0; c < 3; ++c) {
      int q;
      for (q = 0; q < 3; ++q) {
        double x = s->tri[(int64_t)f * 9 + q * 3 + c];
        if (x < n->lo[c]) n->lo[c] = x;
        if (x > n->hi[c]) n->hi[c] = x;
      }
    }
  }
  for (c = 0; c < 3; ++c) {
    n->lo[c] -= pad;
    n->hi[c] += pad;
  }
  n->first = first;
  n->count = count;
  n->left = n->right = -1;
  if (count > 4) {
    double ext[3], cmin[3], cmax[3];
    int axis = 0, half = count / 2, l, r;
    for (c = 0; c < 3; ++c) {
      cmin[c] = INFINITY;
      cmax[c] = -INFINITY;
    }
    for (k = first; k < first + count; ++k)
      for (c = 0; c < 3; ++c) {
        double x = s->cen[(int64_t)s->perm[k] * 3 + c];
        if (x < cmin[c]) cmin[c] = x;
        if (x > cmax[c]) cmax[c] = x;
      }
    for (c = 0; c < 3; ++c) ext[c] = cmax[c] - cmin[c];
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    g_keyc = s->cen;
    g_axis = axis;
    qsort(s->perm + first, (size_t)count, sizeof(int), cmp_axis);
    l = build_rec(s, first, half, pad);
    r = build_rec(s, first + half, count - half, pad);
    s->nodes[id].left = l;
    s->nodes[id].right = r;
  }
  return id;
}

void *oracle_scene_create(const double *V, int64_t nv, const int64_t *F, int64_t nf,
                          int use_bvh) {
  oscene *s = (oscene *)calloc(1, sizeof(oscene));
  int64_t f;
  int c, k;
  double maxabs = 0.0;
  s->nf = nf;
  s->use_bvh = use_bvh;
  s->tri = (double *)malloc(sizeof(double) * 9 * (nf > 0 ? nf : 1));
  s->cen = (double *)malloc(sizeof(double) * 3 * (nf > 0 ? nf : 1));
  for (f = 0; f < nf; ++f) {
    for (k = 0; k < 3; ++k) {
      int64_t vi = F[f * 3 + k];
      if (vi < 0 || vi >= nv) vi = 0;
      for (c = 0; c < 3; ++c) {
        s->tri[f * 9 + k * 3 + c] = V[vi * 3 + c];
        if (fabs(V[vi * 3 + c]) > maxabs) maxabs = fabs(V[vi * 3 + c]);
      }
    }
    for (c = 0; c < 3; ++c)
      s->cen[f * 3 + c] =
          ((s->tri[f * 9 + c] + s->tri[f * 9 + 3 + c]) + s->tri[f * 9 + 6 + c]) / 3.0;
  }
  if (use_bvh && nf > 0) {
    s->perm = (int *)malloc(sizeof(int) * nf);
    for (f = 0; f < nf; ++f) s->perm[f] = (int)f;
    s->nodes = (onode *)malloc(sizeof(onode) * 2 * nf);
    s->nnodes = 0;
    build_rec(s, 0, (int)nf, maxabs * 1e-12 + 1e-300);
  }
  return s;
}

void oracle_scene_destroy(void *p) {
  oscene *s = (oscene *)p;
  if (!s) return;
  free(s->tri);
  free(s->cen);
  free(s->perm);
  free(s->nodes);
  free(s);
}

const double *oracle_scene_centroids(void *p) { return ((oscene *)p)->cen; }

/* Segment [0, tmax] of the ray against a FP64 box (robust enough: the
 * boxes are padded outward). */
static int obox(const double o[3], const double inv[3], const onode *n, double tmax) {
  double tn = 0.0, tf = tmax;
  int c;
  for (c = 0; c < 3; ++c) {
    double t0 = (n->lo[c] - o[c]) * inv[c], t1 = (n->hi[c] - o[c]) * inv[c];
    double a = t0 < t1 ? t0 : t1, b = t0 < t1 ? t1 : t0;
    if (t0 != t0 || t1 != t1) { /* 0 * inf: ray lies in the slab plane */
      a = -INFINITY;
      b = INFINITY;
    }
    if (a > tn) tn = a;
    if (b < tf) tf = b;
  }
  return tn <= tf * (1.0 + 1e-9) + 1e-300;
}

/* Visit every facet whose box the segment [0, tmax] may touch; cb returns
 * nonzero to stop.  Brute force visits every facet. */
typedef int (*visit_fn)(void *ctx, int k);

static void ovisit(const oscene *s, const double o[3], const double d[3], double tmax,
                   visit_fn cb, void *ctx) {
  int64_t k;
  if (!s->use_bvh) {
    for (k = 0; k < s->nf; ++k)
      if (cb(ctx, (int)k)) return;
    return;
  }
  {
    double inv[3];
    int stack[256], sp = 0, c;
    for (c = 0; c < 3; ++c) inv[c] = 1.0 / d[c];
    if (s->nnodes == 0) return;
    stack[sp++] = 0;
    while (sp > 0) {
      const onode *n = &s->nodes[stack[--sp]];
      if (!obox(o, inv, n, tmax)) continue;
      if (n->left < 0) {
        int q;
        for (q = n->first; q < n->first + n->count; ++q)
          if (cb(ctx, s->perm[q])) return;
      } else {
        stack[sp++] = n->right;
        stack[sp++] = n->left;
      }
    }
  }
}

typedef struct {
  const oscene *s;
  oray r;
  double tlim;
  int klim, skip1, skip2, found;
} blk_ctx;

static int blk_cb(void *p, int k) {
  blk_ctx *c = (blk_ctx *)p;
  double t;
  if (k == c->skip1 || k == c->skip2) return 0;
  t = otri(&c->r, c->s->tri + (int64_t)k * 9);
  if (t >= 0.0 && (t < c->tlim || (t == c->tlim && k < c->klim))) {
    c->found = 1;
    return 1;
  }
  return 0;
}

static int oany(const oscene *s, const double o[3], const double d[3], double tlim, int klim,
                int skip1, int skip2) {
  blk_ctx c;
  c.s = s;
  oray_init(&c.r, o, d);
  c.tlim = tlim;
  c.klim = klim;
  c.skip1 = skip1;
  c.skip2 = skip2;
  c.found = 0;
  ovisit(s, o, d, isinf(tlim) ? INFINITY : tlim * (1.0 + 1e-6), blk_cb, &c);
  return c.found;
}

typedef struct {
  const oscene *s;
  oray r;
  double best_t;
  int best_k, skip;
} cls_ctx;

static int cls_cb(void *p, int k) {
  cls_ctx *c = (cls_ctx *)p;
  double t;
  if (k == c->skip) return 0;
  t = otri(&c->r, c->s->tri + (int64_t)k * 9);
  if (t >= 0.0 && (t < c->best_t || (t == c->best_t && k < c->best_k))) {
    c->best_t = t;
    c->best_k = k;
  }
  return 0;
}

/* eps of shape.py:351, 405: 1e3 * np.finfo(np.float32).resolution, the
 * float32 1e-6 times a Python float, an FP64 number (NumPy 1.x). */
static const double kEmb = 1e3 * (double)1e-6f;

/* Embree's ray for the pair (i, j) as fluxpy forms it (shape.py:355-385):
 * d = P_j - P_i, |d| = sqrt((d0^2 + d1^2) + d2^2), the unit direction and
 * the origin P_i + eps d/|d| in the model dtype (f32 != 0: FP32; NumPy
 * applies the FP64 eps to an FP32 array in FP32), then stored into
 * Embree's FP32 ray arrays.  Returns 0 when |d| <= eps (visible by
 * default).  pi, pj: model centroids widened to FP64 (exact). */
static int oemb_pair(const double *pi, const double *pj, int f32, double o[3], double d[3]) {
  int c;
  if (f32) {
    float e = (float)kEmb, dd[3], nrm;
    for (c = 0; c < 3; ++c) dd[c] = (float)pj[c] - (float)pi[c];
    nrm = sqrtf((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
    if (!(nrm > e)) return 0;
    for (c = 0; c < 3; ++c) {
      float dh = dd[c] / nrm;
      float oc = (float)pi[c] + e * dh;
      o[c] = (double)oc;
      d[c] = (double)dh;
    }
  } else {
    double dd[3], nrm;
    for (c = 0; c < 3; ++c) dd[c] = pj[c] - pi[c];
    nrm = sqrt((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
    if (!(nrm > kEmb)) return 0;
    for (c = 0; c < 3; ++c) {
      double dh = dd[c] / nrm;
      o[c] = (double)(float)(pi[c] + kEmb * dh);
      d[c] = (double)(float)dh;
    }
  }
  return 1;
}

/* Embree's occlusion ray (shape.py:410-411): origin P + eps N in the model
 * dtype, direction D, both stored as FP32. */
static void oemb_sun(const double *pi, const double *ni, const double *D, int f32, double o[3],
                     double d[3]) {
  int c;
  for (c = 0; c < 3; ++c) {
    if (f32) {
      float oc = (float)pi[c] + (float)kEmb * (float)ni[c];
      o[c] = (double)oc;
    } else {
      o[c] = (double)(float)(pi[c] + kEmb * ni[c]);
    }
    d[c] = (double)(float)D[c];
  }
}

static int ovis(const oscene *s, int i, int j, int mode, const double *Pm, int f32) {
  double o[3], d[3], tj;
  int skip = -1, c;
  oray r;
  if (mode == 0) {
    if (i == j) return 0;
    for (c = 0; c < 3; ++c) {
      o[c] = s->cen[(int64_t)i * 3 + c];
      d[c] = s->cen[(int64_t)j * 3 + c] - o[c];
    }
    skip = i;
  } else {
    if (!oemb_pair(Pm + (int64_t)i * 3, Pm + (int64_t)j * 3, f32, o, d)) return 1;
  }
  oray_init(&r, o, d);
  tj = otri(&r, s->tri + (int64_t)j * 9);
  if (tj < 0.0) return 0;
  return !oany(s, o, d, tj, j, skip, j);
}

/* The Embree rays of m pairs (pi[p], pj[p]) (oemb_pair): o, d (m x 3) and
 * ok[p] = 0 where |d| <= eps (visible without a ray). */
void oracle_embree_pair_rays(const double *pi, const double *pj, int64_t m, int f32, double *o,
                             double *d, uint8_t *ok) {
  int64_t p;
  for (p = 0; p < m; ++p) ok[p] = (uint8_t)oemb_pair(pi + p * 3, pj + p * 3, f32, o + p * 3, d + p * 3);
}

/* The Embree occlusion rays of m facets (oemb_sun) toward one direction D. */
void oracle_embree_sun_rays(const double *pi, const double *ni, int64_t m, const double *D,
                            int f32, double *o, double *d) {
  int64_t p;
  for (p = 0; p < m; ++p) oemb_sun(pi + p * 3, ni + p * 3, D, f32, o + p * 3, d + p * 3);
}

/* out[p*n+q] = vis(I[p], J[q]).  Pm: model centroids widened to FP64,
 * EMBREE only; f32: the model dtype is FP32. */
void oracle_vis_mn(void *sp, const int64_t *I, int64_t m, const int64_t *J, int64_t n,
                   int mode, const double *Pm, int f32, uint8_t *out) {
  const oscene *s = (const oscene *)sp;
  int64_t p;
#pragma omp parallel for schedule(dynamic, 1)
  for (p = 0; p < m; ++p) {
    int64_t q;
    for (q = 0; q < n; ++q) out[p * n + q] = (uint8_t)ovis(s, (int)I[p], (int)J[q], mode, Pm, f32);
  }
}

void oracle_occluded(void *sp, const int64_t *I, int64_t m, const double *D, int64_t M,
                     int per_face, int mode, const double *Pm, const double *Nm, int f32,
                     uint8_t *out) {
  const oscene *s = (const oscene *)sp;
  int64_t cols = per_face ? 1 : M, p;
#pragma omp parallel for schedule(dynamic, 16)
  for (p = 0; p < m; ++p) {
    int64_t q;
    int i = (int)I[p], c;
    for (q = 0; q < cols; ++q) {
      const double *dq = D + (per_face ? p : q) * 3;
      double o[3], d[3] = {dq[0], dq[1], dq[2]};
      int skip = -1;
      if (mode == 0) {
        for (c = 0; c < 3; ++c) o[c] = s->cen[(int64_t)i * 3 + c];
        skip = i;
      } else {
        oemb_sun(Pm + (int64_t)i * 3, Nm + (int64_t)i * 3, dq, f32, o, d);
      }
      out[p * cols + q] = (uint8_t)oany(s, o, d, INFINITY, 0x7fffffff, skip, -1);
    }
  }
}

void oracle_intersect1(void *sp, const double *x, const double *d, int64_t nr, int64_t *hit,
                       double *xt) {
  const oscene *s = (const oscene *)sp;
  int64_t k;
  for (k = 0; k < nr; ++k) {
    cls_ctx c;
    int q;
    c.s = s;
    oray_init(&c.r, x + k * 3, d + k * 3);
    c.best_t = INFINITY;
    c.best_k = -1;
    c.skip = -1;
    ovisit(s, x + k * 3, d + k * 3, INFINITY, cls_cb, &c);
    hit[k] = c.best_k;
    for (q = 0; q < 3; ++q) xt[k * 3 + q] = c.best_k >= 0 ? x[k * 3 + q] + c.best_t * d[k * 3 + q] : 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* Form factors, form_factors.py:11-72, in the model dtype.            */
/* ------------------------------------------------------------------ */

#define FF_IMPL(T, SUF)                                                                   \
  static T dot3_##SUF(const T *a, const T *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; } \
  static T pos_##SUF(T x) { return x > (T)0 ? x : (T)0; }                                \
  /* Row p: writes up to n entries (ascending q) into cols/vals; returns count. */        \
  static int64_t ff_row_##SUF(const oscene *s, const T *P, const T *N, const T *A,        \
                              int i, const int64_t *J, int64_t n, T eps, int mode,        \
                              int32_t *cols, T *vals) {                                   \
    int64_t q, cnt = 0;                                                                   \
    const T *Pi = P + (int64_t)i * 3, *Ni = N + (int64_t)i * 3;                           \
    double pm_i[3], pm_j[3];                                                              \
    for (q = 0; q < n; ++q) {                                                             \
      int j = (int)J[q];                                                                  \
      const T *Pj = P + (int64_t)j * 3, *Nj = N + (int64_t)j * 3;                         \
      T d[3] = {Pj[0] - Pi[0], Pj[1] - Pi[1], Pj[2] - Pi[2]};                             \
      T npj = dot3_##SUF(Nj, Pj);                                                         \
      T c = pos_##SUF(dot3_##SUF(Ni, d)) * pos_##SUF(dot3_##SUF(Pi, Nj) - npj);           \
      int vis;                                                                            \
      if (i == j || !(c > eps)) continue;                                                 \
      if (mode == 0) {                                                                    \
        vis = ovis(s, i, j, 0, NULL, 0);                                                  \
      } else {                                                                            \
        int k;                                                                            \
        for (k = 0; k < 3; ++k) {                                                         \
          pm_i[k] = (double)Pi[k];                                                        \
          pm_j[k] = (double)Pj[k];                                                        \
        }                                                                                 \
        vis = ovis_embree_ids(s, j, pm_i, pm_j, sizeof(T) == 4);                          \
      }                                                                                   \
      if (!vis) continue;                                                                 \
      {                                                                                   \
        T e0 = Pi[0] - Pj[0], e1 = Pi[1] - Pj[1], e2 = Pi[2] - Pj[2];                     \
        T s2 = (e0 * e0 + e1 * e1) + e2 * e2;                                             \
        T sv = (T)3.141592653589793 * (s2 * s2);                                          \
        cols[cnt] = (int32_t)q;                                                           \
        vals[cnt] = sv == (T)0 ? (T)0 : (c * A[j]) / sv;                                  \
        ++cnt;                                                                            \
      }                                                                                   \
    }                                                                                     \
    return cnt;                                                                           \
  }

/* Embree-mode visibility of facet j from model centroid pi (shape.py:349-398). */
static int ovis_embree_ids(const oscene *s, int j, const double *pi, const double *pj, int f32) {
  double o[3], d[3], tj;
  oray r;
  if (!oemb_pair(pi, pj, f32, o, d)) return 1;
  oray_init(&r, o, d);
  tj = otri(&r, s->tri + (int64_t)j * 9);
  if (tj < 0.0) return 0;
  return !oany(s, o, d, tj, j, -1, j);
}

FF_IMPL(double, f64)
FF_IMPL(float, f32)

/* Form-factor CSR.  Pass 1 (cols == NULL): row_nnz[p] only.  Pass 2: fills
 * cols/vals at indptr[p]. dtype 1 = f64, 0 = f32. */
void oracle_ff_csr(void *sp, const void *P, const void *N, const void *A, int dtype,
                   const int64_t *I, int64_t m, const int64_t *J, int64_t n, double eps,
                   int mode, int64_t *row_nnz, const int64_t *indptr, int32_t *cols,
                   void *vals) {
  const oscene *s = (const oscene *)sp;
  int64_t p;
#pragma omp parallel
  {
    int32_t *tc = (int32_t *)malloc(sizeof(int32_t) * (n > 0 ? n : 1));
    void *tv = malloc(8 * (n > 0 ? n : 1));
#pragma omp for schedule(dynamic, 1)
    for (p = 0; p < m; ++p) {
      int64_t cnt;
      if (dtype == 1)
        cnt = ff_row_f64(s, (const double *)P, (const double *)N, (const double *)A, (int)I[p],
                         J, n, (double)eps, mode, tc, (double *)tv);
      else
        cnt = ff_row_f32(s, (const float *)P, (const float *)N, (const float *)A, (int)I[p], J,
                         n, (float)eps, mode, tc, (float *)tv);
      if (row_nnz) row_nnz[p] = cnt;
      if (cols) {
        memcpy(cols + indptr[p], tc, sizeof(int32_t) * cnt);
        if (dtype == 1)
          memcpy((double *)vals + indptr[p], tv, sizeof(double) * cnt);
        else
          memcpy((float *)vals + indptr[p], tv, sizeof(float) * cnt);
      }
    }
    free(tc);
    free(tv);
  }
}
