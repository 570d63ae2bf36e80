// Column-blocked vector kernels for the radiosity solvers (Jacobi, CG,
// GMRES, iterative refinement).  Vectors are n x nrhs row-major blocks;
// every reduction is per column, in FP64, two-stage with a fixed order, so
// results are deterministic.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <mutex>

#include "common.h"

namespace flux {

template <typename T>
__global__ void __launch_bounds__(256) k_dot_partial(int64_t n, int64_t nrhs,
                                                     const T *__restrict__ x,
                                                     const T *__restrict__ y,
                                                     double *__restrict__ part, int mode) {
  // mode 0: sum x*y; mode 1: max |x - y|; mode 2: max |x|
  __shared__ double s[256];
  const int64_t col = blockIdx.y;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = (double)x[i * nrhs + col];
    if (mode == 0) {
      acc += a * (double)y[i * nrhs + col];
    } else if (mode == 1) {
      acc = fmax(acc, fabs(a - (double)y[i * nrhs + col]));
    } else {
      acc = fmax(acc, fabs(a));
    }
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      s[threadIdx.x] = mode == 0 ? s[threadIdx.x] + s[threadIdx.x + w]
                                 : fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[col * gridDim.x + blockIdx.x] = s[0];
}

__global__ void k_dot_final(int64_t nrhs, int nb, const double *__restrict__ part,
                            double *__restrict__ out, int mode) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= nrhs) return;
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double v = part[col * nb + b];
    acc = mode == 0 ? acc + v : fmax(acc, v);
  }
  out[col] = acc;
}

// y[:, c] = alpha[c] * x[:, c] + beta[c] * y[:, c]  (alpha/beta device
// arrays of nrhs doubles; NULL means 1 for beta and 1 for alpha).
template <typename T>
__global__ void k_axpby(int64_t n, int64_t nrhs, const double *__restrict__ alpha,
                        const T *__restrict__ x, const double *__restrict__ beta,
                        T *__restrict__ y) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nrhs) return;
  const int64_t c = k % nrhs;
  const double a = alpha ? alpha[c] : 1.0;
  const double b = beta ? beta[c] : 1.0;
  y[k] = (T)(a * (double)x[k] + b * (double)y[k]);
}

// out[i, c] = s_i * x[i, c] with s_i = rho[i] or rho_s; `one_minus` uses
// (1 - s_i); `add` adds `add_s` afterwards.
template <typename T>
__global__ void k_rowscale(int64_t n, int64_t nrhs, const T *__restrict__ rho, double rho_s,
                           int one_minus, double add_s, const T *__restrict__ x,
                           T *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nrhs) return;
  const int64_t i = k / nrhs;
  double s = rho ? (double)rho[i] : rho_s;
  if (one_minus) s = 1.0 - s;
  out[k] = (T)(s * (double)x[k] + add_s);
}

// out = max(0, x) in place or copy.
template <typename T>
__global__ void k_clamp0(int64_t total, const T *__restrict__ x, T *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= total) return;
  const T v = x[k];
  out[k] = v > (T)0 ? v : (T)0;
}

// Equilibrium temperature (model.py:19-24 / :264-270):
//   tot = (1 - rho_i) B + emiss_i Q + fsurf;  clamp -> max(0, tot)
//   T   = (tot / (emiss_i sigma)) ** 0.25, in FP64, stored in T's dtype.
template <typename T>
__global__ void k_temperature(int64_t n, int64_t nrhs, const T *__restrict__ B,
                              const T *__restrict__ Q, const T *__restrict__ rho, double rho_s,
                              const T *__restrict__ emiss, double emiss_s, double fsurf,
                              int clamp, double sigma, T *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n * nrhs) return;
  const int64_t i = k / nrhs;
  const double r = rho ? (double)rho[i] : rho_s;
  const double e = emiss ? (double)emiss[i] : emiss_s;
  double tot = (1.0 - r) * (double)B[k] + e * (double)Q[k] + fsurf;
  if (clamp) tot = tot > 0.0 ? tot : 0.0;
  out[k] = (T)pow(tot / (e * sigma), 0.25);
}

// Dtype conversion between FP64 / FP32 blocks and FP32 <-> BF16 / FP16
// (round to nearest even).
__device__ __forceinline__ void cvt(float x, __hip_bfloat16 &y) { y = __float2bfloat16(x); }
__device__ __forceinline__ void cvt(float x, __half &y) { y = __float2half(x); }
__device__ __forceinline__ void cvt(__hip_bfloat16 x, float &y) { y = __bfloat162float(x); }
__device__ __forceinline__ void cvt(__half x, float &y) { y = __half2float(x); }
__device__ __forceinline__ void cvt(float x, float &y) { y = x; }
__device__ __forceinline__ void cvt(float x, double &y) { y = (double)x; }
__device__ __forceinline__ void cvt(double x, float &y) { y = (float)x; }
__device__ __forceinline__ void cvt(double x, double &y) { y = x; }

template <typename TI, typename TO>
__global__ void k_convert(int64_t total, const TI *__restrict__ x, TO *__restrict__ y) {
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += step)
    cvt(x[k], y[k]);
}

}  // namespace flux

using namespace flux;

extern "C" int flux_vec_reduce(int64_t n, int64_t nrhs, const void *x, const void *y,
                               int dtype, int mode, double *out, double *work, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0 && mode >= 0 && mode <= 2, "flux_vec_reduce: bad args");
  if (nrhs == 0) return 0;
  FLUX_CHECK(out && work && (n == 0 || x), "flux_vec_reduce: NULL array");
  FLUX_CHECK(mode == 2 || n == 0 || y, "flux_vec_reduce: y required");
  hipStream_t st = as_stream(stream);
  int nb = (int)((n + 255) / 256);
  if (nb > 64) nb = 64;
  if (nb < 1) nb = 1;
  FLUX_CHECK(nrhs <= 65535, "flux_vec_reduce: nrhs too large");
  dim3 grid(nb, (unsigned)nrhs);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_dot_partial<double>, grid, dim3(256), 0, st, n, nrhs,
                       (const double *)x, (const double *)y, work, mode);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_dot_partial<float>, grid, dim3(256), 0, st, n, nrhs, (const float *)x,
                       (const float *)y, work, mode);
  else
    FLUX_CHECK(false, "flux_vec_reduce: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dot_final, dim3(ceil_div_u(nrhs, 256)), dim3(256), 0, st, nrhs, nb, work,
                     out, mode);
  FLUX_LAUNCH_CHECK();
  return 0;
}

// ---- the same reduction without a caller workspace --------------------
// The 64 partial results per column live in `scratch`, a file-scope device
// array: one copy per GPU, held until the process ends.  In host code the
// name `scratch` is a host-side stand-in, not a device address; the device
// address is looked up once per device with hipGetSymbolAddress and only
// that address is passed to kernels.  Every call on one GPU shares the
// array: calls that overlap on one GPU (other streams, other threads) race.
// Nothing in it has to start at zero: k_dot_partial writes every partial
// that k_dot_final reads.
namespace flux {

constexpr int64_t kReduceScratchCols = 4096;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[64 * kReduceScratchCols];

static std::mutex g_scratch_mutex;
static double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
static int scratch_address(double **out) {
  int dev = 0;
  FLUX_TRY_HIP(hipGetDevice(&dev));
  FLUX_CHECK(dev >= 0 && dev < kMaxDevices, "flux_vec_reduce_noalloc: device %d out of range",
             dev);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void *p = nullptr;
    FLUX_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    FLUX_CHECK(p != nullptr, "flux_vec_reduce_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace flux

extern "C" int flux_vec_reduce_noalloc(int64_t n, int64_t nrhs, const void *x, const void *y,
                                       int dtype, int mode, double *out, void *stream) {
  using namespace flux;
  FLUX_CHECK(nrhs >= 0 && nrhs <= kReduceScratchCols,
             "flux_vec_reduce_noalloc: %lld columns exceed the %lld of the built-in scratch "
             "(use flux_vec_reduce with a workspace)",
             (long long)nrhs, (long long)kReduceScratchCols);
  if (nrhs == 0) return 0;
  double *work = nullptr;
  const int rc = scratch_address(&work);
  if (rc != 0) return rc;
  return flux_vec_reduce(n, nrhs, x, y, dtype, mode, out, work, stream);
}

extern "C" int flux_vec_axpby(int64_t n, int64_t nrhs, const double *alpha, const void *x,
                              const double *beta, void *y, int dtype, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_vec_axpby: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(x && y, "flux_vec_axpby: NULL array");
  hipStream_t st = as_stream(stream);
  const unsigned g = ceil_div_u(n * nrhs, 256);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(g), dim3(256), 0, st, n, nrhs, alpha,
                       (const double *)x, beta, (double *)y);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_axpby<float>, dim3(g), dim3(256), 0, st, n, nrhs, alpha,
                       (const float *)x, beta, (float *)y);
  else
    FLUX_CHECK(false, "flux_vec_axpby: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_vec_rowscale(int64_t n, int64_t nrhs, const void *rho, double rho_s,
                                 int one_minus, double add_s, const void *x, void *out,
                                 int dtype, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_vec_rowscale: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(x && out, "flux_vec_rowscale: NULL array");
  hipStream_t st = as_stream(stream);
  const unsigned g = ceil_div_u(n * nrhs, 256);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_rowscale<double>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const double *)rho, rho_s, one_minus, add_s, (const double *)x,
                       (double *)out);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_rowscale<float>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const float *)rho, rho_s, one_minus, add_s, (const float *)x,
                       (float *)out);
  else
    FLUX_CHECK(false, "flux_vec_rowscale: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_vec_clamp0(int64_t total, const void *x, void *out, int dtype,
                               void *stream) {
  FLUX_CHECK(total >= 0, "flux_vec_clamp0: bad size");
  if (total == 0) return 0;
  FLUX_CHECK(x && out, "flux_vec_clamp0: NULL array");
  hipStream_t st = as_stream(stream);
  const unsigned g = ceil_div_u(total, 256);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_clamp0<double>, dim3(g), dim3(256), 0, st, total, (const double *)x,
                       (double *)out);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_clamp0<float>, dim3(g), dim3(256), 0, st, total, (const float *)x,
                       (float *)out);
  else
    FLUX_CHECK(false, "flux_vec_clamp0: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_vec_convert(int64_t total, const void *x, int xdtype, void *y, int ydtype,
                                void *stream) {
  FLUX_CHECK(total >= 0, "flux_vec_convert: bad size");
  if (total == 0) return 0;
  FLUX_CHECK(x && y, "flux_vec_convert: NULL array");
  hipStream_t st = as_stream(stream);
  const unsigned g = grid_1d(total, 256);
  if (xdtype == FLUX_F64 && ydtype == FLUX_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, st, total,
                       (const double *)x, (float *)y);
  else if (xdtype == FLUX_F32 && ydtype == FLUX_F64)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, st, total,
                       (const float *)x, (double *)y);
  else if (xdtype == FLUX_F64 && ydtype == FLUX_F64)
    hipLaunchKernelGGL((k_convert<double, double>), dim3(g), dim3(256), 0, st, total,
                       (const double *)x, (double *)y);
  else if (xdtype == FLUX_F32 && ydtype == FLUX_F32)
    hipLaunchKernelGGL((k_convert<float, float>), dim3(g), dim3(256), 0, st, total,
                       (const float *)x, (float *)y);
  else if (xdtype == FLUX_F32 && ydtype == FLUX_BF16)
    hipLaunchKernelGGL((k_convert<float, __hip_bfloat16>), dim3(g), dim3(256), 0, st, total,
                       (const float *)x, (__hip_bfloat16 *)y);
  else if (xdtype == FLUX_F32 && ydtype == FLUX_F16)
    hipLaunchKernelGGL((k_convert<float, __half>), dim3(g), dim3(256), 0, st, total,
                       (const float *)x, (__half *)y);
  else if (xdtype == FLUX_BF16 && ydtype == FLUX_F32)
    hipLaunchKernelGGL((k_convert<__hip_bfloat16, float>), dim3(g), dim3(256), 0, st, total,
                       (const __hip_bfloat16 *)x, (float *)y);
  else if (xdtype == FLUX_F16 && ydtype == FLUX_F32)
    hipLaunchKernelGGL((k_convert<__half, float>), dim3(g), dim3(256), 0, st, total,
                       (const __half *)x, (float *)y);
  else
    FLUX_CHECK(false, "flux_vec_convert: unsupported dtypes");
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_temperature(int64_t n, int64_t nrhs, const void *B, const void *Q,
                                const void *rho, double rho_s, const void *emiss,
                                double emiss_s, double fsurf, int clamp, double sigma,
                                void *T, int dtype, void *stream) {
  FLUX_CHECK(n >= 0 && nrhs >= 0, "flux_temperature: bad sizes");
  if (n * nrhs == 0) return 0;
  FLUX_CHECK(B && Q && T, "flux_temperature: NULL array");
  hipStream_t st = as_stream(stream);
  const unsigned g = ceil_div_u(n * nrhs, 256);
  if (dtype == FLUX_F64)
    hipLaunchKernelGGL(k_temperature<double>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const double *)B, (const double *)Q, (const double *)rho, rho_s,
                       (const double *)emiss, emiss_s, fsurf, clamp, sigma, (double *)T);
  else if (dtype == FLUX_F32)
    hipLaunchKernelGGL(k_temperature<float>, dim3(g), dim3(256), 0, st, n, nrhs,
                       (const float *)B, (const float *)Q, (const float *)rho, rho_s,
                       (const float *)emiss, emiss_s, fsurf, clamp, sigma, (float *)T);
  else
    FLUX_CHECK(false, "flux_temperature: dtype must be F32/F64");
  FLUX_LAUNCH_CHECK();
  return 0;
}
