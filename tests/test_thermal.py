"""Time-dependent thermal model on the CPU: the conduction oracle
(oracle/pcc_oracle.c) and the host functions of fluxpy_amd.thermal.

What pins them:
* the reference's own PCC fixture (tests/data/thermal/pcc_thermal_model_1d,
  kept as tests/golden/pcc_conductionQ.npz): 50001 radiative-boundary steps
  of one facet, run as the reference's tests/test_thermal.py:36-80 runs
  them, must reproduce the surface and bottom temperatures and the last
  profile to the fixture's 3 decimals;
* the closed-form periodic solution of the heat equation in a half space
  (thermal.pyx:276-295, ``analytT_``) for the prescribed-temperature step;
* closed forms of setgrid and flux_noatm;
* planted changes to the boundary terms, which the fixture must reject;
* for the ThermalModel iteration (model.py:108-234): under a constant sun its
  fixed point is the steady-state temperature of model.py:8-42 (solved here
  with a dense LU), which a planted change to the absorbed flux misses.
"""
import os

import numpy as np
import pytest

from fluxpy_amd.thermal import analytT_, flux_noatm, setgrid

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                      "pcc_conductionQ.npz")


def run_pcc_case(O, nsteps=50000, emissivity=1.0, Fgeotherm=0.2, conduction=None):
    """tests/test_thermal.py:36-80 of the reference with z[0] = 0 prepended
    (the layout conductionQ expects, pcc/conductionQ.c:45-49)."""
    conduction = conduction or O.conduction_q
    nz = 60
    z = np.concatenate([[0.0], setgrid(nz=nz, zfac=1.05, zmax=2.5)])
    period = 88775.244 * 670
    albedo, Rau, Decl, lat = 0.2, 1.52, 0., np.deg2rad(5.)
    Qn = (1 - albedo) * flux_noatm(Rau, Decl, lat, 0., 0., 0.)
    T = np.full((1, nz + 1), 210.0)
    ti, rhoc = np.full(nz + 1, 120.), np.full(nz + 1, 960000.)
    dt = 495661.77900000004
    surf = []
    for i in range(nsteps + 1):
        time = (i + 1) * dt
        HA = 2 * np.pi * (time / period % 1.)
        Qnp1 = (1 - albedo) * flux_noatm(Rau, Decl, lat, HA, 0., 0.)
        conduction(z, dt, Qn, Qnp1, T, ti, rhoc, emissivity, Fgeotherm)
        Qn = Qnp1
        if i % 3 == 0:
            surf.append([time / period, T[0, 0], T[0, nz]])
    return np.array(surf), T[0].copy()


def compare_to_fixture(surf, prof):
    g = np.load(GOLDEN)
    S = surf[g["Tsurface_step"] // 3]
    d_time = np.abs(np.round(S[:, 0], 6) - g["Tsurface"][:, 0]).max()
    d_surf = np.abs(np.round(S[:, 1:], 3) - g["Tsurface"][:, 1:]).max()
    d_prof = np.abs(np.round(prof, 3) - g["Tprofile_last"]).max()
    return d_time, d_surf, d_prof


@pytest.fixture(scope="module")
def pcc_run(oracle_mod):
    return run_pcc_case(oracle_mod)


def test_conduction_q_reproduces_pcc_fixture(pcc_run):
    d_time, d_surf, d_prof = compare_to_fixture(*pcc_run)
    # the fixture holds 3 decimals (6 for the time): one unit in the last
    # place allows for a value sitting on a rounding boundary
    assert d_time <= 1.01e-6
    assert d_surf <= 1.01e-3
    assert d_prof <= 1.01e-3


@pytest.mark.parametrize("change", ["emissivity", "geotherm"])
def test_fixture_rejects_planted_boundary_mistakes(oracle_mod, change):
    kw = {"emissivity": 0.98} if change == "emissivity" else {"Fgeotherm": 0.25}
    surf, prof = run_pcc_case(oracle_mod, **kw)
    _, d_surf, d_prof = compare_to_fixture(surf, prof)
    assert max(d_surf, d_prof) > 0.05


def test_conduction_t_matches_periodic_half_space(oracle_mod):
    """Prescribed surface temperature T0 + 30 sin(-w t): after three periods
    the profile follows T0 + 30 exp(-z/d) sin(z/d - w t), d = ti/rhoc
    sqrt(P/pi) (thermal.pyx:276-295)."""
    ti, rhoc, P, T0 = 120., 960000., 88775.244, 200.
    delta = ti / rhoc * np.sqrt(P / np.pi)
    nz = 80
    z = np.concatenate([[0.0], setgrid(nz, 1.04, 16 * delta)])
    nstep = 2000
    dt = P / nstep
    tiv, rv = np.full(nz + 1, ti), np.full(nz + 1, rhoc)
    T = analytT_(z, T0, rhoc, ti, P, dt)[0][None, :].copy()
    surf = lambda t: T0 + 30. * np.sin(-2 * np.pi / P * t)
    for i in range(3 * nstep):
        oracle_mod.conduction_t(z, dt, T, surf(i * dt), surf((i + 1) * dt), tiv, rv, 0.0)
    exact = analytT_(z, T0, rhoc, ti, P, dt)[0]
    err = np.abs(T[0, 1:] - exact[1:])
    # 30 K amplitude; the discretisation leaves a few mK
    assert err.max() < 0.01, err.max()


def test_conduction_q_energy_balance_at_steady_state(oracle_mod):
    """Constant absorbed flux Q and no geothermal flux: the steady state is
    isothermal with emiss sigma T^4 = Q (sigma of pcc/conductionQ.c)."""
    nz = 40
    z = np.concatenate([[0.0], setgrid(nz, 1.1, 0.5)])
    T = np.full((3, nz + 1), 150.0)
    Q = np.array([100., 300., 600.])
    em = np.array([1.0, 0.9, 0.95])
    ti, rhoc = np.full(nz + 1, 200.), np.full(nz + 1, 1.2e6)
    for _ in range(32000):
        oracle_mod.conduction_q(z, 3600., Q, Q, T, ti, rhoc, em, 0.0)
    Teq = (Q / (em * 5.6704e-8)) ** 0.25
    assert np.allclose(T, Teq[:, None], rtol=1e-10, atol=0)


def test_setgrid_closed_forms():
    z = setgrid(60, 1.05, 2.5)
    assert z.size == 60 and np.isclose(z[-1], 2.5, rtol=1e-12)
    assert np.isclose(z[1], 3 * z[0], rtol=1e-15)
    # z(n) - z(1) = 2 z(1) (zfac^(n-1) - 1) / (zfac - 1)
    n = np.arange(1, 61)
    assert np.allclose(z - z[0], 2 * z[0] * (1.05 ** (n - 1) - 1) / 0.05, rtol=1e-10)
    zr = setgrid(10, 1.0, 1.0)
    assert np.allclose(zr, (np.arange(1, 11) - 0.5) * 0.1)


def test_flux_noatm_closed_forms():
    lat = np.deg2rad(5.)
    # noon, flat: sin(elevation) = cos(latitude) for zero declination
    assert np.isclose(flux_noatm(1.0, 0., lat, 0., 0., 0.), 1365. * np.cos(lat))
    assert np.isclose(flux_noatm(2.0, 0., lat, 0., 0., 0.), 1365. * np.cos(lat) / 4)
    # night
    assert flux_noatm(1.0, 0., lat, np.pi, 0., 0.) == 0.0
    # a south-facing slope (azFac = 0) tilted by the latitude sees the noon
    # sun overhead
    assert np.isclose(flux_noatm(1.0, 0., lat, 1e-9, lat, 0.), 1365., rtol=1e-9)
    # morning and afternoon are mirror images on flat ground
    assert np.isclose(flux_noatm(1.0, 0.1, lat, 0.7, 0., 0.),
                      flux_noatm(1.0, 0.1, lat, -0.7, 0., 0.))


@pytest.fixture(scope="module")
def crater_case(oracle_mod):
    from fluxpy_amd import meshes
    V, F = meshes.spherical_crater(npts=24, seed=1001)
    V = V.astype(np.float64)
    P, N, A = oracle_mod.geometry(V, F)[:3]
    scene = oracle_mod.OracleScene(V, F)
    FF = scene.form_factor_matrix(P, N, A)
    d = np.array([np.cos(np.deg2rad(30)), 0.0, np.sin(np.deg2rad(30))])
    occ = np.asarray(scene.occluded(np.arange(len(F)), d)).reshape(-1)
    E = oracle_mod.direct_irradiance(N, occ, 1365., d)
    return FF, E


@pytest.mark.parametrize("planted", [None, "no_reflection", "no_emissivity"])
def test_thermal_model_fixed_point_is_steady_state(oracle_mod, crater_case, planted):
    """Constant sun, no geothermal flux, profiles started at the steady
    state: 400 steps of 10 h keep the surface within 2e-5 of it (the
    conduction's sigma, 5.6704e-8, differs from scipy's by 5e-6)."""
    FF, E = crater_case
    nf = FF.shape[0]
    rho, em = 0.12, 0.95
    Tss = oracle_mod.steady_state_temp(
        FF, E, rho, em, solver=lambda F_, E_, r_: oracle_mod.solve_radiosity_dense(F_, E_, r_))
    nz, nsteps, dt = 20, 400, 36000.
    z = np.concatenate([[0.0], setgrid(nz, 1.1, 0.3)])
    kw = {}
    if planted == "no_reflection":
        kw["absorbed"] = lambda E_, Qr, Qi, r, e: (1 - r) * E_ + e * Qi
    elif planted == "no_emissivity":
        kw["absorbed"] = lambda E_, Qr, Qi, r, e: (1 - r) * (E_ + Qr) + Qi
    run = oracle_mod.thermal_model_run(FF, [E] * (nsteps + 1), dt * np.arange(nsteps + 1), rho,
                                       np.full(nf, em), z, np.repeat(Tss[:, None], nz + 1, 1),
                                       np.full(nz + 1, 120.), np.full(nz + 1, 960000.), 0.0, E,
                                       **kw)
    for T in run:
        pass
    dev = np.abs(T[:, 0] / Tss - 1).max()
    if planted is None:
        assert dev < 2e-5, dev
    else:
        assert dev > 1e-3, dev
