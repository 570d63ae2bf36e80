"""End-to-end physics check on the GPU: the shadowed floor of a spherical
bowl has Ingersoll's closed-form equilibrium temperature (ingersoll.py:
24-34 of fluxpy).  The mesh (ingersoll.HemisphericalCrater), its
visibility, form factors, insolation, both radiosity solves and the
temperature (model.py:8-42) all run through this package; the discrete
temperatures of the shadowed facets must match T_gt to the mesh's
discretisation error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("backend", ["cgal", "embree"])
def test_shadowed_floor_matches_ingersoll(backend):
    import torch
    torch.cuda.set_device(0)
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.ingersoll import HemisphericalCrater
    from fluxpy_amd.model import compute_steady_state_temp
    from fluxpy_amd.shape import CgalTrimeshShapeModel, EmbreeTrimeshShapeModel
    e0 = np.deg2rad(10.0)
    c = HemisphericalCrater(np.deg2rad(40.0), 0.5, e0, 1365.0, 0.3, 0.95)
    V, F, parts = c.make_trimesh(0.04, contour_rim=True, return_parts=True)
    cls = CgalTrimeshShapeModel if backend == "cgal" else EmbreeTrimeshShapeModel
    sm = cls(V, F)
    sun = np.array([np.cos(e0), 0.0, np.sin(e0)])
    E = sm.get_direct_irradiance(c.F0, sun)
    FF = FormFactorMatrix.from_shape_model(sm)
    T = compute_steady_state_temp(FF, E, c.rho, c.emiss)
    bowl = parts[1]
    shadow = bowl[E[bowl] == 0]
    lit = bowl[E[bowl] > 0]
    assert len(shadow) > 100 and len(lit) > 100
    err = np.abs(T[shadow] - c.T_gt) / c.T_gt
    print(f"{backend}: {len(shadow)} shadowed facets, T_gt {c.T_gt:.2f} K, "
          f"median err {np.median(err):.2e}, max {err.max():.2e}")
    # measured: median 2.4e-3, max 1.1e-2 (h = 0.04)
    assert np.median(err) < 5e-3 and err.max() < 2e-2
