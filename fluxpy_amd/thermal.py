"""1-D subsurface heat conduction for every facet (drop-in for flux.thermal).

``PccThermalModel1D`` keeps the interface of thermal.pyx:13-246: the same
constructor arguments, validation and messages, properties and ``step``.
The per-facet loop that calls conductionQ / conductionT
(thermal.pyx:218-246, pcc/conductionQ.c:10-186, pcc/conductionT.c:4-80) is
one GPU launch over all facets (``flux_conduction_q`` /
``flux_conduction_t``), bit for bit the reference's FP64 arithmetic.

State lives on the GPU: the profiles layer-major, ``T_device[l, f]``, so a
layer is contiguous.  The properties return host copies in the reference's
layout (``T`` is num_faces x num_layers); ``T_device`` and ``Tsurf_device``
give the device arrays without a copy.

Differences from the reference, each deliberate:
* a zero pivot in the tridiagonal solve raises RuntimeError with the
  reference's message ("tridag: rewrite equations" / "tridag failed") and
  leaves the state unchanged, where the reference exits the process;
* the reference prints "consider taking shorter time steps" for each facet
  whose linearisation was repeated 9 or more times; here one warning per
  step (``logging``) gives the number of such facets;
* inputs are converted to contiguous FP64 arrays (the reference's typed
  memoryviews reject other dtypes).

``setgrid``, ``analytT_`` and ``flux_noatm`` are the host functions of
thermal.pyx:248-354 with the same formulas.
"""
import logging
import math

import numpy as np

from . import _lib

_log = logging.getLogger(__name__)

_BAD_INPUT, _NEG_INPUT, _PIVOT1, _PIVOT, _NAN_T, _NEG_T = 1, 2, 4, 8, 16, 32


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class PccThermalModel1D:
    """Bundle of 1-D conduction models, one per facet (thermal.pyx:13-246)."""

    def __init__(self, z, T0, ti, rhoc, emiss, Fgeotherm, Q0, Tsurfprev, bcond='Q'):
        if bcond not in {'Q', 'T'}:
            raise ValueError('bcond should be "Q" or "T"')
        self._bcond = bcond
        z, T0, ti, rhoc = _f64(z), _f64(T0), _f64(ti), _f64(rhoc)
        emiss, Fgeotherm, Q0, Tsurfprev = _f64(emiss), _f64(Fgeotherm), _f64(Q0), _f64(Tsurfprev)

        self._num_faces = T0.shape[0]
        self._num_layers = z.size
        self._num_layers_interior = self._num_layers - 1
        self._current_time = 0

        if not all(0 <= _ <= 1 for _ in emiss):
            raise ValueError('emissivity values ("emiss") should be in range [0, 1]')
        if emiss.ndim != 1 or emiss.size != self._num_faces:
            raise ValueError('"emiss" should be a 1D array with length == T0.shape[0]')
        if z[0] != 0:
            raise RuntimeError('z[0] is required to be zero')
        if (ti <= 0.0).any():
            raise ValueError('thermal inertia ("ti") values should be positive')
        if ti.ndim != 1 or ti.size != self._num_layers:
            raise ValueError('"ti" should be a 1D array with length == z.size')
        if (rhoc <= 0).any():
            raise ValueError('volumetric heat capacity ("rhoc") values should be positive')
        if rhoc.ndim != 1 or rhoc.size != self._num_layers:
            raise ValueError('"rhoc" should be a 1D array with length == z.size')
        if (Fgeotherm < 0).any():
            raise ValueError('geothermal fluxes ("Fgeotherm") should be positive')
        if Fgeotherm.ndim != 1 or Fgeotherm.size != self._num_faces:
            raise ValueError('"Fgeotherm" should be a 1D array with length == T0.shape[0]')
        if (Tsurfprev < 0).any():
            raise ValueError('surface temperatures ("Tsurfprev") should be nonnegative')
        if Tsurfprev.ndim != 1 or Tsurfprev.size != self._num_faces:
            raise ValueError('"Tsurfprev" should be a 1D array with length == T0.shape[0]')
        if (Q0 < 0).any():
            raise ValueError('initial irradiance ("Q0") should be nonnegative')
        if Q0.ndim != 1 or Q0.size != self._num_faces:
            raise ValueError('"Q0" should be a 1D array with length == T0.shape[0]')
        if (T0 < 0).any():
            raise ValueError('initial temperature ("T0") should be nonnegative')
        if T0.ndim != 2 or T0.shape[1] != self._num_layers:
            raise ValueError('"T0" should be a 2D array with T0.shape[1] == z.size')
        if self._num_layers_interior < 2:
            raise ValueError('need at least two layers below the surface (z.size >= 3)')

        t = _lib.require_gpu()
        dev = _lib.device()
        nf = self._num_faces
        to = lambda a: t.from_numpy(a).to(dev)
        self._z, self._ti, self._rhoc = to(z), to(ti), to(rhoc)
        self._emiss, self._Fgeotherm = to(emiss), to(Fgeotherm)
        self._Qprev = to(Q0.copy())
        self._Tsurfprev = to(Tsurfprev.copy())
        self._Fsurf = t.full((nf,), float('nan'), dtype=t.float64, device=dev)
        self._Fsurf_next = t.empty_like(self._Fsurf)
        self._T = to(np.ascontiguousarray(T0.T))           # layer-major
        self._T_next = t.empty_like(self._T)
        nb = int(_lib.load().flux_conduction_work_bytes(nf, self._num_layers_interior))
        self._work = t.empty(max(nb, 8), dtype=t.uint8, device=dev)
        self._status = t.zeros(2, dtype=t.int32, device=dev)

    # -- properties (host copies, reference layout) ----------------------------

    @property
    def num_layers(self):
        return self._num_layers

    @property
    def num_faces(self):
        return self._num_faces

    @property
    def num_layers_interior(self):
        return self._num_layers_interior

    @property
    def current_time(self):
        return self._current_time

    @property
    def emiss(self):
        return self._emiss.cpu().numpy()

    @property
    def z(self):
        return self._z.cpu().numpy()

    @property
    def ti(self):
        return self._ti.cpu().numpy()

    @property
    def rhoc(self):
        return self._rhoc.cpu().numpy()

    @property
    def Fgeotherm(self):
        return self._Fgeotherm.cpu().numpy()

    @property
    def Qprev(self):
        return self._Qprev.cpu().numpy()

    @property
    def Fsurf(self):
        return self._Fsurf.cpu().numpy()

    @property
    def T(self):
        return self._T.t().contiguous().cpu().numpy()

    @property
    def Tsurfprev(self):
        return self._Tsurfprev.cpu().numpy()

    @property
    def bcond(self):
        return self._bcond

    # -- device views ----------------------------------------------------------

    @property
    def T_device(self):
        """Profiles on the device, layer-major: T_device[l, f] (no copy)."""
        return self._T

    @property
    def Tsurf_device(self):
        """Surface temperatures T[:, 0] on the device (a contiguous view)."""
        return self._T[0]

    @property
    def emiss_device(self):
        return self._emiss

    # -- stepping -----------------------------------------------------------------

    def _input(self, X):
        t = _lib.torch()
        if isinstance(X, t.Tensor):
            return X.to(device=self._T.device, dtype=t.float64).reshape(X.shape).contiguous()
        return t.from_numpy(_f64(X)).to(self._T.device)

    def step(self, dt, X):
        """Advance every facet by dt (thermal.pyx:207-246).  X: the absorbed
        flux at the new time (bcond 'Q') or the new surface temperature
        ('T'), one value per facet, host array or device tensor."""
        what = 'fluxes' if self._bcond == 'Q' else 'temps'
        t = _lib.torch()
        if isinstance(X, t.Tensor):
            ndim, size = X.dim(), X.numel()
        else:
            X = np.asarray(X)
            ndim, size = X.ndim, X.size
        if ndim != 1 or size != self._num_faces:
            raise RuntimeError(f'input flux of wrong size: X.size == {size}')
        if self._bcond not in ('Q', 'T'):
            raise RuntimeError(f'got unexpected BC mode: "{self._bcond}"')
        Xd = self._input(X)
        nf, nz = self._num_faces, self._num_layers_interior
        st = _lib.stream()
        if self._bcond == 'Q':
            _lib.call("flux_conduction_q", nf, nz, self._z.data_ptr(), float(dt),
                      self._Qprev.data_ptr(), Xd.data_ptr(), self._T.data_ptr(),
                      self._T_next.data_ptr(), self._ti.data_ptr(), self._rhoc.data_ptr(),
                      self._emiss.data_ptr(), self._Fgeotherm.data_ptr(),
                      self._Fsurf_next.data_ptr(), self._work.data_ptr(),
                      self._status.data_ptr(), st)
        else:
            _lib.call("flux_conduction_t", nf, nz, self._z.data_ptr(), float(dt),
                      self._T.data_ptr(), self._T_next.data_ptr(), self._Tsurfprev.data_ptr(),
                      Xd.data_ptr(), self._ti.data_ptr(), self._rhoc.data_ptr(),
                      self._Fgeotherm.data_ptr(), self._Fsurf_next.data_ptr(),
                      self._work.data_ptr(), self._status.data_ptr(), st)
        flags, slow = (int(v) for v in self._status.cpu().numpy())
        if flags & _BAD_INPUT:
            raise RuntimeError(f'trying to step with nonfinite input {what}')
        if flags & _NEG_INPUT:
            raise RuntimeError('trying to step with negative fluxes' if self._bcond == 'Q'
                               else 'trying to step with negative input temps')
        if flags & _PIVOT1:
            raise RuntimeError('tridag: rewrite equations')
        if flags & _PIVOT:
            raise RuntimeError('tridag failed')
        if slow:
            _log.warning('consider taking shorter time steps (%d facets repeated the '
                         'linearisation 9 or more times)', slow)
        self._T, self._T_next = self._T_next, self._T
        self._Fsurf, self._Fsurf_next = self._Fsurf_next, self._Fsurf
        # the model keeps its own copy (a caller may reuse its buffer)
        (self._Qprev if self._bcond == 'Q' else self._Tsurfprev).copy_(Xd)
        if flags & _NAN_T:
            raise RuntimeError('computed nonfinite temperatures while stepping')
        if flags & _NEG_T:
            raise RuntimeError('computed negative temperatures while stepping')
        self._current_time += dt


def setgrid(nz, zfac, zmax):
    """Regular (zfac <= 1) or geometric subsurface grid (thermal.pyx:248-273):
    z(n) - z(1) = 2 z(1) (zfac^(n-1) - 1) / (zfac - 1), with z(2) = 3 z(1)
    as conductionQ expects.  Returns the nz depths below the surface (the
    surface point z = 0 is not included)."""
    z = [0]
    if zfac > 1.:
        dz = zmax / (3. + 2. * zfac * (zfac ** (nz - 2) - 1.) / (zfac - 1.))
        z.append(dz)
        z.append(3 * z[1])
        for i in range(3, nz + 1):
            z.append((1 + zfac) * z[i - 1] - zfac * z[i - 2])
    else:
        dz = zmax / nz
        z.extend((i - 0.5) * dz for i in range(1, nz + 1))
    return np.array(z)[1:]


def analytT_(z, T0, rhoc, ti, P, dt):
    """Analytic temperatures of a semi-infinite half space under a sinusoidal
    surface temperature of amplitude 30 K and period P (thermal.pyx:276-295):
    T = T0 + 30 exp(-z/delta) sin(z/delta - 2 pi t / P), delta = ti/rhoc
    sqrt(P/pi), one row per t in arange(0, P, dt)."""
    Ta = 30.
    delta = ti / rhoc * np.sqrt(P / np.pi)
    w = 2 * np.pi / P
    return np.vstack([T0 + Ta * np.exp(-z / delta) * np.sin(z / delta - w * t)
                      for t in np.arange(0, P, dt)])


def flux_noatm(R, decl, latitude, HA, SlopeAngle, azFac):
    """Solar flux on a sloped surface with no atmosphere (thermal.pyx:297-354).

    R: distance from the sun (AU); decl: solar declination; latitude; HA:
    hour angle (radians from noon, clockwise); SlopeAngle >= 0; azFac:
    azimuth of the topographic gradient (radians east of north, 0 = south
    facing).  Solar constant 1365 W/m^2."""
    So = 1365.
    c1 = math.cos(latitude) * math.cos(decl)
    s1 = math.sin(latitude) * math.sin(decl)
    sinbeta = c1 * math.cos(HA) + s1            # sun elevation above the horizon
    cosbeta = math.sqrt(1 - sinbeta ** 2)
    buf = (math.sin(decl) - math.sin(latitude) * sinbeta) / (math.cos(latitude) * cosbeta)
    if buf > +1.:
        buf = +1.0                              # roundoff
    if buf < -1.:
        buf = -1.0
    azSun = math.acos(buf)
    if math.sin(HA) >= 0:
        azSun = 2 * math.pi - azSun
    sintheta = (math.cos(SlopeAngle) * sinbeta
                - math.sin(SlopeAngle) * cosbeta * math.cos(azSun - azFac))
    if cosbeta == 0.:
        sintheta = math.cos(SlopeAngle) * sinbeta
    if sintheta < 0.:
        sintheta = 0.
    if sinbeta < 0.:
        sintheta = 0.
    return sintheta * So / (R ** 2)
