"""GPU tests of the time-dependent thermal model: the conduction kernels
against the CPU oracle (same FP64 operation order, no FMA contraction on
either side), the reference's PCC fixture through PccThermalModel1D, the
error paths of thermal.pyx:207-246, and the ThermalModel iteration
(model.py:108-234) against the oracle's restatement."""
import os

import numpy as np
import pytest

from fluxpy_amd import meshes
from fluxpy_amd.thermal import PccThermalModel1D, setgrid

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    torch.cuda.set_device(0)
    return torch


def _state(seed, nf=1000, nz=60):
    rng = np.random.default_rng(seed)
    z = np.concatenate([[0.0], setgrid(nz, 1.05, 2.5)])
    T = rng.uniform(150, 350, (nf, 1)) + rng.uniform(-5, 5, (nf, nz + 1))
    ti = np.full(nz + 1, 120.)
    ti[5:] = 300.                               # a layered column
    rhoc = np.full(nz + 1, 960000.)
    rhoc[5:] = 1.3e6
    em = rng.uniform(0.8, 1.0, nf)
    fg = rng.uniform(0, 0.5, nf)
    Q0 = rng.uniform(0, 600, nf)
    return rng, z, np.ascontiguousarray(T), ti, rhoc, em, fg, Q0


def _max_rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


def test_conduction_q_matches_oracle(torch_dev, oracle_mod):
    rng, z, T, ti, rhoc, em, fg, Qn = _state(5)
    print("case: conduction_q, 1000 facets, 60 layers, 6 steps", flush=True)
    model = PccThermalModel1D(z, T, ti, rhoc, em, fg, Qn, T[:, 0].copy())
    To = T.copy()
    repeats = 0
    for s in range(6):
        Q = (rng.uniform(0, 600, T.shape[0]) if s % 2 else
             np.where(rng.random(T.shape[0]) < 0.3, rng.uniform(1500, 3000, T.shape[0]), 0.0))
        Fo, it = oracle_mod.conduction_q(z, 20000., Qn, Q, To, ti, rhoc, em, fg)
        repeats += int((it > 0).sum())
        model.step(20000., Q)
        Qn = Q
        Tg = model.T
        # same operations in the same order: equal up to the last bit of the
        # square root and divisions (both correctly rounded)
        assert _max_rel(Tg, To) <= 1e-14, s
        # Fsurf = -k (T1 - T0) / z1 cancels: compare on the scale of |Fsurf|
        assert np.abs(model.Fsurf - Fo).max() <= 1e-10 * np.abs(Fo).max(), s
    print(f"exact T entries: {np.mean(Tg == To):.4f}; repeated linearisations: {repeats}")
    assert repeats > 0                          # the corrector path ran
    assert model.current_time == 6 * 20000.


def test_conduction_t_matches_oracle(torch_dev, oracle_mod):
    rng, z, T, ti, rhoc, em, fg, _ = _state(6, nf=777, nz=33)
    print("case: conduction_t, 777 facets, 33 layers, 5 steps", flush=True)
    Ts = T[:, 0].copy()
    model = PccThermalModel1D(z, T, ti, rhoc, em, fg, np.zeros(777), Ts, bcond='T')
    To = T.copy()
    for s in range(5):
        Ts1 = Ts + rng.uniform(-20, 20, Ts.size)
        Fo = oracle_mod.conduction_t(z, 5000., To, Ts, Ts1, ti, rhoc, fg)
        model.step(5000., Ts1)
        Ts = Ts1
        assert _max_rel(model.T, To) <= 1e-14, s
        # Fsurf = -k (T1 - T0) / z1 cancels: compare on the scale of |Fsurf|
        assert np.abs(model.Fsurf - Fo).max() <= 1e-10 * np.abs(Fo).max(), s
    assert np.array_equal(model.T[:, 0], T[:, 0])   # conductionT leaves T[0]
    assert np.array_equal(model.Tsurfprev, Ts)


def test_pcc_fixture_through_model(torch_dev):
    """The reference's tests/test_thermal.py:36-80 through PccThermalModel1D
    on the GPU (z with the surface point z[0] = 0, T0 of z.size values)."""
    from fluxpy_amd.thermal import flux_noatm
    from test_thermal import compare_to_fixture
    print("case: PCC fixture, 1 facet, 50001 steps", flush=True)
    nz = 60
    z = np.concatenate([[0.0], setgrid(nz=nz, zfac=1.05, zmax=2.5)])
    period = 88775.244 * 670
    albedo, Rau, Decl, lat = 0.2, 1.52, 0., np.deg2rad(5.)
    Qn = (1 - albedo) * flux_noatm(Rau, Decl, lat, 0., 0., 0.)
    model = PccThermalModel1D(z, np.full((1, nz + 1), 210.0), np.full(nz + 1, 120.),
                              np.full(nz + 1, 960000.), np.array([1.0]), np.array([0.2]),
                              np.array([Qn]), np.array([210.0]))
    dt = 495661.77900000004
    surf = []
    for i in range(50001):
        time = (i + 1) * dt
        HA = 2 * np.pi * (time / period % 1.)
        Qnp1 = (1 - albedo) * flux_noatm(Rau, Decl, lat, HA, 0., 0.)
        model.step(dt, np.array([Qnp1]))
        if i % 3 == 0:
            T = model.T
            surf.append([time / period, T[0][0], T[0][nz]])
    d_time, d_surf, d_prof = compare_to_fixture(np.array(surf), model.T[0])
    assert d_time <= 1.01e-6 and d_surf <= 1.01e-3 and d_prof <= 1.01e-3


def test_step_errors_keep_state(torch_dev):
    _, z, T, ti, rhoc, em, fg, Q0 = _state(7, nf=50, nz=10)
    model = PccThermalModel1D(z, T, ti, rhoc, em, fg, Q0, T[:, 0].copy())
    before = model.T
    bad = Q0.copy()
    bad[3] = -1.0
    with pytest.raises(RuntimeError, match='trying to step with negative fluxes'):
        model.step(100., bad)
    bad[3] = np.nan
    with pytest.raises(RuntimeError, match='trying to step with nonfinite input fluxes'):
        model.step(100., bad)
    with pytest.raises(RuntimeError, match='input flux of wrong size'):
        model.step(100., Q0[:-1])
    assert np.array_equal(model.T, before) and model.current_time == 0
    assert np.array_equal(model.Qprev, Q0)
    # a device tensor input is copied: the caller may reuse its buffer
    import torch
    Xd = torch.tensor(Q0, device="cuda")
    model.step(100., Xd)
    Xd.fill_(7.0)
    assert np.array_equal(model.Qprev, Q0)


def test_thermal_model_matches_oracle(torch_dev, oracle_mod):
    """model.py:108-234 on an FP64 crater: the GPU iteration against the
    oracle's restatement with the same F (SciPy CSR) and insolation."""
    from fluxpy_amd import shape
    from fluxpy_amd.form_factors import get_form_factor_matrix
    from fluxpy_amd.model import ThermalModel
    print("case: ThermalModel, spherical crater (FP64), 12 steps", flush=True)
    V, F = meshes.spherical_crater(npts=24, seed=1001)
    sm = shape.CgalTrimeshShapeModel(V.astype(np.float64), F)
    FF = get_form_factor_matrix(sm)
    nf = sm.num_faces
    nt = 13
    t = np.linspace(0, 88775.244, nt)
    az = np.linspace(0.3, 2.8, nt)
    el = np.deg2rad(np.linspace(5, 40, nt))
    D = np.stack([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)], 1)
    F0 = np.full(nt, 1365.)
    rho = 0.12
    nz = 30
    z = np.concatenate([[0.0], setgrid(nz, 1.1, 0.3)])
    kw = dict(z=z, T0=200., ti=120., rhoc=960000., emiss=0.95, Fgeotherm=0.0,
              shape_model=sm)
    model = ThermalModel(FF, t, D, F0, rho, return_flux=True, **kw)
    E_steps = [sm.get_direct_irradiance(F0[i], D[i]) for i in range(nt)]
    ref = oracle_mod.thermal_model_run(FF, E_steps, t, rho, np.full(nf, 0.95), z,
                                       200., np.full(nz + 1, 120.), np.full(nz + 1, 960000.),
                                       0.0, E_steps[0], return_flux=True)
    nstep = 0
    for (Tg, Eg, Qrg, QIRg), (To, Eo, Qro, QIRo) in zip(model, ref):
        nstep += 1
        dq = np.abs(Qrg - Qro).max() / np.abs(Qro).max()
        di = np.abs(QIRg - QIRo).max() / np.abs(QIRo).max()
        print(f"step {nstep}: Qrefl {dq:.2e} QIR {di:.2e} T {_max_rel(Tg, To):.2e}", flush=True)
        assert np.array_equal(Eg, Eo)
        assert dq <= 1e-12 and di <= 1e-12, nstep
        assert _max_rel(Tg, To) <= 1e-12, nstep
    assert nstep == nt - 1
    assert np.isfinite(Tg).all() and Tg.min() > 0


def test_ff_npz_round_trip_drives_solves_and_thermal_model(torch_dev, tmp_path):
    """The reference's examples keep F as a SciPy .npz (scipy.sparse.save_npz
    of get_form_factor_matrix, examples/spherical_crater/collect_data.py:141)
    and hand the loaded matrix to solve_radiosity and ThermalModel
    (examples/gerlache/collect_time_dep_data_spinup.py:30-37).  The same
    file drives the GPU solve and the thermal model unchanged."""
    import scipy.sparse
    from fluxpy_amd import shape
    from fluxpy_amd.form_factors import get_form_factor_matrix
    from fluxpy_amd.model import ThermalModel
    from fluxpy_amd.solve import solve_radiosity
    print("case: FF.npz round trip, spherical crater", flush=True)
    V, F = meshes.spherical_crater(npts=20, seed=1003)
    sm = shape.CgalTrimeshShapeModel(V.astype(np.float64), F)
    FF = get_form_factor_matrix(sm)
    path = tmp_path / "FF.npz"
    scipy.sparse.save_npz(path, FF)
    FF2 = scipy.sparse.load_npz(path)
    E = sm.get_direct_irradiance(1365.0, np.array([0.6, 0.0, 0.8]))
    B1, n1 = solve_radiosity(FF, E, 0.12)
    B2, n2 = solve_radiosity(FF2, E, 0.12)
    assert n1 == n2 and np.array_equal(B1, B2)
    nz = 10
    z = np.concatenate([[0.0], setgrid(nz, 1.1, 0.2)])
    t = np.linspace(0, 3600.0, 3)
    D = np.array([[0.6, 0.0, 0.8]] * 3)
    runs = []
    for M in (FF, FF2):
        tm = ThermalModel(M, t, D, np.full(3, 1365.0), 0.12, z=z, T0=200.0, ti=120.0,
                          rhoc=960000.0, emiss=0.95, Fgeotherm=0.0, shape_model=sm)
        runs.append([T for T in tm])
    assert len(runs[0]) == 2 and all(np.array_equal(a, b) for a, b in zip(*runs))


def test_thermal_model_extended_source_matches_oracle(torch_dev, oracle_mod):
    """ThermalModel with several sun directions per time point (the
    extended-source branch of model.py:186-199, insolation averaged over the
    directions as shape.py:234-240) against the oracle's iteration."""
    from fluxpy_amd import shape
    from fluxpy_amd.form_factors import get_form_factor_matrix
    from fluxpy_amd.model import ThermalModel
    print("case: ThermalModel, extended source (4 directions), crater", flush=True)
    V, F = meshes.spherical_crater(npts=20, seed=1004)
    sm = shape.CgalTrimeshShapeModel(V.astype(np.float64), F)
    FF = get_form_factor_matrix(sm)
    nf = sm.num_faces
    nt, M = 5, 4
    t = np.linspace(0, 7200.0, nt)
    rng = np.random.default_rng(9)
    az = rng.uniform(0, 2 * np.pi, (nt, M))
    el = np.deg2rad(rng.uniform(20, 40, (nt, M)))
    D = np.stack([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)], -1)
    D = D.reshape(nt * M, 3)
    F0 = np.full((nt * M, 1), 1365.0)
    nz = 12
    z = np.concatenate([[0.0], setgrid(nz, 1.1, 0.2)])
    model = ThermalModel(FF, t, D, F0, 0.12, z=z, T0=210.0, ti=150.0, rhoc=1.0e6,
                         emiss=0.9, Fgeotherm=0.0, shape_model=sm)
    E_steps = [sm.get_direct_irradiance(F0[i * M:(i + 1) * M, :], D[i * M:(i + 1) * M])
               for i in range(nt)]
    assert E_steps[1].shape == (nf,)
    ref = oracle_mod.thermal_model_run(FF, E_steps, t, 0.12, np.full(nf, 0.9), z, 210.0,
                                       np.full(nz + 1, 150.0), np.full(nz + 1, 1.0e6), 0.0,
                                       E_steps[0])
    n = 0
    for Tg, To in zip(model, ref):
        n += 1
        assert _max_rel(Tg, To) <= 1e-12, n
    assert n == nt - 1
