// Reverse mode of the steady-state temperature map, around one adjoint
// radiosity solve.  Forward (per sun s, per facet i):
//   (I - diag(rho) F) B = E                      radiosity, model.py:8-42 /
//                                                solve.py:25-45 with 'left'
//   Phi = E + F B                                incident flux
//   T   = ((1 - rho) Phi / (emiss sigma))^(1/4)  equilibrium temperature
// Given gT = dL/dT, with A = T^4 the absorbed flux over emiss sigma:
//   gA   = gT / (4 T^3)           (0 where T = 0: the clamp of the forward)
//   gPhi = gA (1 - rho) / (emiss sigma)
//   d/drho   through A:  -gT T / (4 (1 - rho))
//   d/demiss through A:  -gT T / (4 emiss)
// then gB = F^T gPhi, the adjoint solve (I - F^T diag(rho)) lam = gB, and
//   dL/dE   = gPhi + lam
//   dL/drho = (sum over suns of the terms above) + sum_s lam (F B).
// Vectors are n x k row-major (facet-major, suns contiguous), FP64.
#include "common.h"

namespace flux {

// One thread per facet: gPhi (n x k), and the per-facet sums through A.
__global__ void k_tadj_seed(int64_t n, int64_t k, const double *__restrict__ T,
                            const double *__restrict__ gT, const double *__restrict__ rho,
                            const double *__restrict__ emiss, double sigma,
                            double *__restrict__ gPhi, double *__restrict__ grho,
                            double *__restrict__ gemiss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i], e = emiss[i];
  double sr = 0.0, se = 0.0;
  for (int64_t s = 0; s < k; ++s) {
    const double t = T[i * k + s], g = gT[i * k + s];
    double gphi = 0.0;
    if (t > 0.0) {
      const double ga = g / (4.0 * t * t * t);
      gphi = ga * (1.0 - r) / (e * sigma);
      const double gtt = g * t * 0.25;
      sr -= gtt / (1.0 - r);
      se -= gtt / e;
    }
    gPhi[i * k + s] = gphi;
  }
  grho[i] = sr;
  gemiss[i] = se;
}

// gE = gPhi + lam;  grho[i] += sum_s lam (F B)
__global__ void k_tadj_finish(int64_t n, int64_t k, const double *__restrict__ gPhi,
                              const double *__restrict__ lam, const double *__restrict__ FB,
                              double *__restrict__ gE, double *__restrict__ grho) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t s = 0; s < k; ++s) {
    const double l = lam[i * k + s];
    gE[i * k + s] = gPhi[i * k + s] + l;
    acc += l * FB[i * k + s];
  }
  grho[i] += acc;
}

}  // namespace flux

using namespace flux;

extern "C" int flux_temperature_adjoint_seed(int64_t n, int64_t k, const double *T,
                                             const double *gT, const double *rho,
                                             const double *emiss, double sigma, double *gPhi,
                                             double *grho, double *gemiss, void *stream) {
  FLUX_CHECK(n >= 0 && k >= 0, "flux_temperature_adjoint_seed: bad sizes");
  if (n == 0) return 0;
  FLUX_CHECK(T && gT && rho && emiss && gPhi && grho && gemiss,
             "flux_temperature_adjoint_seed: NULL array");
  FLUX_CHECK(sigma > 0.0, "flux_temperature_adjoint_seed: sigma must be positive");
  hipLaunchKernelGGL(k_tadj_seed, dim3(ceil_div_u(n, 256)), dim3(256), 0, as_stream(stream),
                     n, k, T, gT, rho, emiss, sigma, gPhi, grho, gemiss);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_temperature_adjoint_finish(int64_t n, int64_t k, const double *gPhi,
                                               const double *lam, const double *FB, double *gE,
                                               double *grho, void *stream) {
  FLUX_CHECK(n >= 0 && k >= 0, "flux_temperature_adjoint_finish: bad sizes");
  if (n == 0) return 0;
  FLUX_CHECK(gPhi && lam && FB && gE && grho, "flux_temperature_adjoint_finish: NULL array");
  hipLaunchKernelGGL(k_tadj_finish, dim3(ceil_div_u(n, 256)), dim3(256), 0, as_stream(stream),
                     n, k, gPhi, lam, FB, gE, grho);
  FLUX_LAUNCH_CHECK();
  return 0;
}
