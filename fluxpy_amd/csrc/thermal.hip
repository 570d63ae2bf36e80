// Time-dependent thermal model: 1-D subsurface conduction for every facet
// (pcc/conductionQ.c:10-186, pcc/conductionT.c:4-80 with the Thomas solve of
// pcc/tridag.c:5-34, which thermal.pyx:218-246 calls one facet at a time) and
// the explicit radiance update of model.py:44-69 around it.
//
// One thread per facet, one wave per workgroup (a facet's solve is a serial
// recurrence, so the work is spread over as many CUs as there are waves).
// Profiles are stored layer-major, T[l * nf + f] for layer l = 0 (surface)
// .. nz of facet f, so a wave's loads of one layer are contiguous.  The step reads T_in and writes T_out (the caller swaps them),
// which keeps the previous profile for the predictor-corrector repeats of
// conductionQ and leaves the state untouched when the caller rejects a step.
//
// The coefficients depend on z, ti, rhoc and dt only; k_conduction_coef
// evaluates them once per step with the reference's operation order, and the
// step kernels read them with uniform (scalar) loads.  This file is compiled
// without FMA contraction: each facet's result equals the reference's
// sequence of separately rounded FP64 operations.
//
// Status words (device int[2], cleared by the coefficient kernel):
//   status[0] bits: 1 nonfinite input, 2 negative input, 4 zero first pivot
//                   ("tridag: rewrite equations"), 8 zero pivot ("tridag
//                   failed"), 16 nonfinite T, 32 negative T;
//   status[1]: facets whose predictor-corrector ran 9 or more repeats (the
//              reference prints "consider taking shorter time steps").
#include "common.h"

namespace flux {
namespace {

constexpr double kSigmaPcc = 5.6704e-8;  // conductionQ.c:60

enum : int { kBadInput = 1, kNegInput = 2, kPivot1 = 4, kPivot = 8, kNanT = 16, kNegT = 32 };

// coef layout (n1 = nz + 1): alpha[0..nz], gamma[n1 + 0..nz], then k[1] and
// the surface conductance k[1] / (2 z[1]) (conductionQ's k1).
__global__ void k_conduction_coef(int nz, const double *__restrict__ z, double dt,
                                  const double *__restrict__ ti,
                                  const double *__restrict__ rhoc, int bc_q,
                                  double *__restrict__ coef, int *__restrict__ status) {
  const int n1 = nz + 1;
  double *alpha = coef, *gamma = coef + n1;
  if (threadIdx.x == 0) {
    status[0] = 0;
    status[1] = 0;
  }
  for (int i = threadIdx.x + 1; i <= nz; i += blockDim.x) {
    const double ki = ti[i] * ti[i] / rhoc[i];
    if (i == 1) {
      const double k2 = ti[2] * ti[2] / rhoc[2];
      if (bc_q) {
        const double dz = 2. * z[1];
        const double beta = dt / rhoc[1] / (2. * dz * dz);
        alpha[1] = beta * k2;
        gamma[1] = beta * ki;
        coef[2 * n1 + 1] = ki / dz;
      } else {
        alpha[1] = k2 * dt / rhoc[1] / (z[2] - z[1]) / z[2];
        gamma[1] = ki * dt / rhoc[1] / z[1] / z[2];
        coef[2 * n1 + 1] = 0.0;
      }
      coef[2 * n1] = ki;
    } else if (i < nz) {
      const double kp = ti[i + 1] * ti[i + 1] / rhoc[i + 1];
      const double buf = dt / (z[i + 1] - z[i - 1]);
      alpha[i] = 2 * kp * buf / (rhoc[i] + rhoc[i + 1]) / (z[i + 1] - z[i]);
      gamma[i] = 2 * ki * buf / (rhoc[i] + rhoc[i + 1]) / (z[i] - z[i - 1]);
    } else {
      const double buf = dt / (z[nz] - z[nz - 1]) / (z[nz] - z[nz - 1]);
      alpha[nz] = 0.0;
      gamma[nz] = ki * buf / (2 * rhoc[nz]);
    }
  }
}

// Forward elimination and back substitution of one facet's system, with the
// first row (b1, r1) given and rows 2..nz built from the old profile Tin:
//   r_j = gamma_j T_{j-1} + (1 - alpha_j - gamma_j) T_j + alpha_j T_{j+1}
//   r_nz = gamma_nz T_{nz-1} + (1 - gamma_nz) T_nz + dt Fgeo / (rhoc_nz dz_nz)
// Writes the solution to Tout[1..nz] and the multipliers to gamw; returns
// u_1 and flags nonfinite / negative values and zero pivots in `bad`.
__device__ __forceinline__ double conduction_solve(int64_t nf, int64_t f, int nz,
                                                   const double *__restrict__ z, double dt,
                                                   const double *__restrict__ rhoc,
                                                   const double *__restrict__ alpha,
                                                   const double *__restrict__ gamma, double fgeo,
                                                   double b1, double r1,
                                                   const double *__restrict__ Tin,
                                                   double *__restrict__ Tout,
                                                   double *__restrict__ gamw, int &bad) {
#pragma clang fp contract(off)
  double bet = b1;
  if (bet == 0.0) bad |= kPivot1;
  double u = r1 / bet;
  Tout[nf + f] = u;
  double tm = Tin[nf + f];                       // T_{j-1}
  double tc = Tin[2 * nf + f];                   // T_j
  // rows 2 .. nz-1; the loads of Tin do not depend on the recurrence, so the
  // unrolled body issues them ahead of the divisions
#pragma unroll 8
  for (int j = 2; j < nz; ++j) {
    const double tp = Tin[(int64_t)(j + 1) * nf + f];
    const double aj = -gamma[j];
    const double bj = 1. + alpha[j] + gamma[j];
    const double rj = gamma[j] * tm + (1 - alpha[j] - gamma[j]) * tc + alpha[j] * tp;
    tm = tc;
    tc = tp;
    const double gam = -alpha[j - 1] / bet;
    bet = bj - aj * gam;
    if (bet == 0.0) bad |= kPivot;
    u = (rj - aj * u) / bet;
    gamw[(int64_t)j * nf + f] = gam;
    Tout[(int64_t)j * nf + f] = u;
  }
  {  // bottom row: geothermal flux
    const double aj = -gamma[nz];
    const double bj = 1. + gamma[nz];
    const double rj =
        gamma[nz] * tm + (1. - gamma[nz]) * tc + dt / rhoc[nz] * fgeo / (z[nz] - z[nz - 1]);
    const double gam = -alpha[nz - 1] / bet;
    bet = bj - aj * gam;
    if (bet == 0.0) bad |= kPivot;
    u = (rj - aj * u) / bet;
    gamw[(int64_t)nz * nf + f] = gam;
    Tout[(int64_t)nz * nf + f] = u;
  }
  if (!isfinite(u)) bad |= kNanT;
  if (u < 0.0) bad |= kNegT;
#pragma unroll 8
  for (int j = nz - 1; j >= 1; --j) {
    u = Tout[(int64_t)j * nf + f] - gamw[(int64_t)(j + 1) * nf + f] * u;
    Tout[(int64_t)j * nf + f] = u;
    if (!isfinite(u)) bad |= kNanT;
    if (u < 0.0) bad |= kNegT;
  }
  return u;
}

__device__ __forceinline__ int input_flags(double x) {
  return (isfinite(x) ? 0 : kBadInput) | (x < 0.0 ? kNegInput : 0);
}

__global__ __launch_bounds__(64) void k_conduction_q(
    int64_t nf, int nz, const double *__restrict__ z, double dt, const double *__restrict__ Qn,
    const double *__restrict__ Qnp1, const double *__restrict__ Tin, double *__restrict__ Tout,
    const double *__restrict__ rhoc, const double *__restrict__ emiss,
    const double *__restrict__ fgeo, const double *__restrict__ coef,
    double *__restrict__ Fsurf, double *__restrict__ gamw, int *__restrict__ status) {
#pragma clang fp contract(off)
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int n1 = nz + 1;
  const double *alpha = coef, *gamma = coef + n1;
  const double kz1 = coef[2 * n1], k1 = coef[2 * n1 + 1];
  const double em = emiss[f], qn = Qn[f], qn1 = Qnp1[f], fg = fgeo[f];
  int flags = input_flags(qn1);
  double Tr = Tin[f];
  const double T1 = Tin[nf + f], T2 = Tin[2 * nf + f];
  int iter = 0;
  double T0, u1;
  for (;;) {
    const double arad = -3 * em * kSigmaPcc * Tr * Tr * Tr * Tr;
    const double brad = 2 * em * kSigmaPcc * Tr * Tr * Tr;
    const double ann = (qn - arad) / (k1 + brad);
    const double annp1 = (qn1 - arad) / (k1 + brad);
    const double bn = (k1 - brad) / (k1 + brad);
    const double b1 = 1. + alpha[1] + gamma[1] - gamma[1] * bn;
    const double r1 = gamma[1] * (annp1 + ann) +
                      (1. - alpha[1] - gamma[1] + gamma[1] * bn) * T1 + alpha[1] * T2;
    int bad = 0;
    u1 = conduction_solve(nf, f, nz, z, dt, rhoc, alpha, gamma, fg, b1, r1, Tin, Tout, gamw,
                          bad);
    T0 = 0.5 * (annp1 + bn * u1 + u1);
    if ((T0 > 1.2 * Tr || T0 < 0.8 * Tr) && iter < 10) {
      ++iter;
      Tr = sqrt(Tr * T0);
      continue;
    }
    flags |= bad;
    break;
  }
  if (!isfinite(T0)) flags |= kNanT;
  if (T0 < 0.0) flags |= kNegT;
  Tout[f] = T0;
  Fsurf[f] = -kz1 * (u1 - T0) / z[1];
  if (flags) atomicOr(status, flags);
  if (iter >= 9) atomicAdd(status + 1, 1);
}

__global__ __launch_bounds__(64) void k_conduction_t(
    int64_t nf, int nz, const double *__restrict__ z, double dt, const double *__restrict__ Tin,
    double *__restrict__ Tout, const double *__restrict__ Tsurf,
    const double *__restrict__ Tsurfp1, const double *__restrict__ rhoc,
    const double *__restrict__ fgeo, const double *__restrict__ coef,
    double *__restrict__ Fsurf, double *__restrict__ gamw, int *__restrict__ status) {
#pragma clang fp contract(off)
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nf) return;
  const int n1 = nz + 1;
  const double *alpha = coef, *gamma = coef + n1;
  const double kz1 = coef[2 * n1];
  const double ts = Tsurf[f], ts1 = Tsurfp1[f];
  int flags = input_flags(ts1);
  const double T0 = Tin[f], T1 = Tin[nf + f], T2 = Tin[2 * nf + f];
  const double b1 = 1. + alpha[1] + gamma[1];
  const double r1 = alpha[1] * T2 + (1. - alpha[1] - gamma[1]) * T1 + gamma[1] * (ts + ts1);
  int bad = 0;
  const double u1 =
      conduction_solve(nf, f, nz, z, dt, rhoc, alpha, gamma, fgeo[f], b1, r1, Tin, Tout, gamw,
                       bad);
  flags |= bad;
  // conductionT leaves T[0] as it was
  if (!isfinite(T0)) flags |= kNanT;
  if (T0 < 0.0) flags |= kNegT;
  Tout[f] = T0;
  Fsurf[f] = -kz1 * (u1 - ts1) / z[1];
  if (flags) atomicOr(status, flags);
}

// X[i, 0] = E_i + Qrefl_i and X[i, 1] = emiss_i sigma Ts_i^4 + (1 - emiss_i)
// QIR_i (model.py:60,64), Qrefl = QIR = 0 when `first`; X is n x 2.  sigma
// comes from the caller (model.py uses scipy.constants.Stefan_Boltzmann).
__global__ void k_thermal_sources(int64_t n, const double *__restrict__ E,
                                  const double *__restrict__ Qrefl,
                                  const double *__restrict__ QIR,
                                  const double *__restrict__ Ts,
                                  const double *__restrict__ emiss, double sigma, int first,
                                  double *__restrict__ X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double em = emiss[i], t = Ts[i];
  const double qr = first ? 0.0 : Qrefl[i], qi = first ? 0.0 : QIR[i];
  X[2 * i] = E[i] + qr;
  X[2 * i + 1] = em * sigma * ((t * t) * (t * t)) + (1 - em) * qi;
}

// From Y = F X (n x 2, FP32 or FP64): Qrefl = rho Y[:, 0], QIR = Y[:, 1]
// (each clamped at 0 if `clamp`, model.py:60-67) and the net absorbed flux
// Q = (1 - rho)(E + Qrefl) + emiss QIR (model.py:224).
template <typename TY>
__global__ void k_thermal_update(int64_t n, const TY *__restrict__ Y,
                                 const double *__restrict__ E, const double *__restrict__ rho,
                                 const double *__restrict__ emiss, int clamp,
                                 double *__restrict__ Qrefl, double *__restrict__ QIR,
                                 double *__restrict__ Q) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  double qr = r * (double)Y[2 * i];
  double qi = (double)Y[2 * i + 1];
  if (clamp) {
    qr = fmax(0.0, qr);
    qi = fmax(0.0, qi);
  }
  Qrefl[i] = qr;
  QIR[i] = qi;
  Q[i] = (1 - r) * (E[i] + qr) + emiss[i] * qi;
}

int conduction_check(int64_t nf, int nz, const void *z, const void *Tin, const void *Tout,
                     const void *coef, const void *gamw, const void *status, const char *what) {
  FLUX_CHECK(nf >= 0, "%s: bad facet count", what);
  FLUX_CHECK(nz >= 2, "conduction: need at least two layers below the surface (nz >= 2)");
  if (nf == 0) return 0;
  FLUX_CHECK(z && Tin && Tout && coef && gamw && status, "conduction: NULL array");
  FLUX_CHECK(Tin != Tout, "conduction: T_in and T_out must be different buffers");
  return 0;
}

}  // namespace
}  // namespace flux

using namespace flux;

extern "C" int64_t flux_conduction_work_bytes(int64_t nf, int nz) {
  if (nf < 0 || nz < 0) return -1;
  // coefficients (2 (nz + 1) + 2 doubles, rounded to 256 B), then the
  // elimination multipliers (nz + 1) x nf
  const int64_t coef = ((int64_t)(2 * (nz + 1) + 2) * 8 + 255) / 256 * 256;
  return coef + (int64_t)(nz + 1) * nf * 8;
}

static double *conduction_coef(int64_t nf, int nz, void *work, double **gamw) {
  const int64_t coef = ((int64_t)(2 * (nz + 1) + 2) * 8 + 255) / 256 * 256;
  *gamw = reinterpret_cast<double *>(static_cast<char *>(work) + coef);
  (void)nf;
  return static_cast<double *>(work);
}

extern "C" int flux_conduction_q(int64_t nf, int nz, const double *z, double dt,
                                 const double *Qn, const double *Qnp1, const double *T_in,
                                 double *T_out, const double *ti, const double *rhoc,
                                 const double *emiss, const double *Fgeotherm, double *Fsurf,
                                 void *work, int *status, void *stream) {
  int rc = conduction_check(nf, nz, z, T_in, T_out, work, work, status, "flux_conduction_q");
  if (rc || nf == 0) return rc;
  FLUX_CHECK(Qn && Qnp1 && ti && rhoc && emiss && Fgeotherm && Fsurf,
             "flux_conduction_q: NULL array");
  double *gamw;
  double *coef = conduction_coef(nf, nz, work, &gamw);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_conduction_coef, dim3(1), dim3(256), 0, st, nz, z, dt, ti, rhoc, 1, coef,
                     status);
  hipLaunchKernelGGL(k_conduction_q, dim3(ceil_div_u(nf, 64)), dim3(64), 0, st, nf, nz, z, dt,
                     Qn, Qnp1, T_in, T_out, rhoc, emiss, Fgeotherm, coef, Fsurf, gamw, status);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_conduction_t(int64_t nf, int nz, const double *z, double dt,
                                 const double *T_in, double *T_out, const double *Tsurf,
                                 const double *Tsurfp1, const double *ti, const double *rhoc,
                                 const double *Fgeotherm, double *Fsurf, void *work, int *status,
                                 void *stream) {
  int rc = conduction_check(nf, nz, z, T_in, T_out, work, work, status, "flux_conduction_t");
  if (rc || nf == 0) return rc;
  FLUX_CHECK(Tsurf && Tsurfp1 && ti && rhoc && Fgeotherm && Fsurf,
             "flux_conduction_t: NULL array");
  double *gamw;
  double *coef = conduction_coef(nf, nz, work, &gamw);
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_conduction_coef, dim3(1), dim3(256), 0, st, nz, z, dt, ti, rhoc, 0, coef,
                     status);
  hipLaunchKernelGGL(k_conduction_t, dim3(ceil_div_u(nf, 64)), dim3(64), 0, st, nf, nz, z, dt,
                     T_in, T_out, Tsurf, Tsurfp1, rhoc, Fgeotherm, coef, Fsurf, gamw, status);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_thermal_sources(int64_t n, const double *E, const double *Qrefl,
                                    const double *QIR, const double *Tsurf, const double *emiss,
                                    double sigma, int first, double *X, void *stream) {
  FLUX_CHECK(n >= 0, "flux_thermal_sources: bad size");
  if (n == 0) return 0;
  FLUX_CHECK(E && Tsurf && emiss && X && (first || (Qrefl && QIR)),
             "flux_thermal_sources: NULL array");
  hipLaunchKernelGGL(k_thermal_sources, dim3(ceil_div_u(n, 256)), dim3(256), 0,
                     as_stream(stream), n, E, Qrefl, QIR, Tsurf, emiss, sigma, first, X);
  FLUX_LAUNCH_CHECK();
  return 0;
}

extern "C" int flux_thermal_update(int64_t n, const void *Y, int y_dtype, const double *E,
                                   const double *rho, const double *emiss, int clamp,
                                   double *Qrefl, double *QIR, double *Q, void *stream) {
  FLUX_CHECK(n >= 0, "flux_thermal_update: bad size");
  if (n == 0) return 0;
  FLUX_CHECK(Y && E && rho && emiss && Qrefl && QIR && Q, "flux_thermal_update: NULL array");
  FLUX_CHECK(y_dtype == FLUX_F32 || y_dtype == FLUX_F64, "flux_thermal_update: Y must be F32/F64");
  hipStream_t st = as_stream(stream);
  if (y_dtype == FLUX_F32)
    hipLaunchKernelGGL(k_thermal_update<float>, dim3(ceil_div_u(n, 256)), dim3(256), 0, st, n,
                       static_cast<const float *>(Y), E, rho, emiss, clamp, Qrefl, QIR, Q);
  else
    hipLaunchKernelGGL(k_thermal_update<double>, dim3(ceil_div_u(n, 256)), dim3(256), 0, st, n,
                       static_cast<const double *>(Y), E, rho, emiss, clamp, Qrefl, QIR, Q);
  FLUX_LAUNCH_CHECK();
  return 0;
}
