"""Mesh front-ends: the hemispherical crater of Ingersoll's test
(ingersoll.py:10-455 of fluxpy) and DEM triangulation
(examples/ceres/generate_mesh_from_topo.py:36-93)."""
import numpy as np
import pytest
from scipy.constants import sigma

from fluxpy_amd.dem import mesh_from_grid, mesh_from_xyz
from fluxpy_amd.ingersoll import HemisphericalCrater


def _areas(V, F):
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    return np.linalg.norm(C, axis=1) / 2, C


def test_crater_constants():
    """r, H, Sc, f and T_gt against their definitions, worked independently:
    beta = 90 degrees is a hemisphere (H = 0, f = 1/2, Sc = 2 pi rc^2)."""
    c = HemisphericalCrater(np.pi / 2, 0.5, np.deg2rad(10), 1365.0, 0.3, 0.9)
    assert abs(c.H) < 1e-15 and abs(c.r - 0.5) < 1e-15
    assert abs(c.Sc - 2 * np.pi * 0.25) < 1e-12 and abs(c.f - 0.5) < 1e-15
    T4 = 1365.0 * np.sin(np.deg2rad(10)) * 0.5 * 0.7 / 0.85 * (1 + 0.3 * 0.5 / 0.9) / sigma
    assert abs(c.T_gt - T4 ** 0.25) < 1e-10
    c = HemisphericalCrater(np.deg2rad(40), 0.4, np.deg2rad(15), 1365.0, 0.12, 0.95)
    assert abs(c.r ** 2 - c.H ** 2 - c.rc ** 2) < 1e-12      # the rim lies on the sphere
    assert abs(c._get_z(0.0, 0.0) - (c.H - c.r)) < 1e-15 and c._get_z(0.9, 0.0) == 0


@pytest.mark.parametrize("h", [0.1, 0.05])
def test_contoured_crater_mesh(h):
    """The bowl and the plane share the rim loop; the areas converge to the
    cap area Sc and to 4 - pi rc^2; bowl vertices lie on the sphere;
    normals face up (into the bowl)."""
    c = HemisphericalCrater(np.deg2rad(40), 0.4, np.deg2rad(15), 1365.0, 0.12, 0.95)
    V, F, (i0, i1, i2) = c.make_trimesh(h, contour_rim=True, return_part_indices=True)
    A, C = _areas(V, F)
    assert (A > 0).all() and (C[:, 2] > 0).all()
    plane, bowl = F[i0:i1], F[i1:i2]
    assert abs(A[i1:i2].sum() - c.Sc) < 0.02 * c.Sc
    assert abs(A[i0:i1].sum() - (4 - np.pi * c.rc ** 2)) < 0.02
    vb = np.unique(bowl)
    d = np.linalg.norm(V[vb] - [0, 0, c.H], axis=1)
    assert np.abs(d - c.r).max() < 1e-12
    rim = np.intersect1d(np.unique(plane), vb)
    assert len(rim) >= 2 * np.pi * c.rc / h and np.abs(V[rim, 2]).max() == 0
    # every edge of the rim loop is shared by one plane and one bowl facet
    edges = {}
    for f in F:
        for a, b in ((f[0], f[1]), (f[1], f[2]), (f[2], f[0])):
            edges[(min(a, b), max(a, b))] = edges.get((min(a, b), max(a, b)), 0) + 1
    inner = [e for e, k in edges.items() if e[0] in set(rim) and e[1] in set(rim)]
    assert all(edges[e] == 2 for e in inner)
    parts = c.make_trimesh(h, contour_rim=True, return_parts=True)[2]
    assert parts[0].size == i1 and parts[1].size == i2 - i1
    with pytest.raises(NotImplementedError):
        c.make_trimesh(h, contour_rim=True, contour_shadow=True)


def test_plain_crater_mesh_and_silhouette():
    c = HemisphericalCrater(np.deg2rad(40), 0.4, np.deg2rad(15), 1365.0, 0.12, 0.95)
    V, F = c.make_trimesh(0.1)
    A, C = _areas(V, F)
    assert (C[:, 2] > 0).all() and abs(A.sum() - 4) < 0.2
    # the shadow line at y = 0: the ray from the far rim (x0 = rc) toward the
    # sun meets the bowl where z(x) = -t0 (x0 - x)
    x = c._x_silhouette(0.0)
    assert -c.rc < x < c.rc
    assert abs(c._get_z(x, 0.0) + c.t0 * (c.rc - x)) < 1e-12
    assert c._x_silhouette(0.5) is None


def test_dem_grid_and_xyz(tmp_path):
    """DEM triangulations keep every sample as a vertex with its height,
    cover the grid's rectangle and face up."""
    x = np.linspace(0, 3, 7)
    y = np.linspace(-1, 1, 5)
    Z = np.sin(x)[None, :] + y[:, None] ** 2
    V, F = mesh_from_grid(x, y, Z, z_offset=100.0)
    X, Y = np.meshgrid(x, y)
    assert np.array_equal(V[:, :2], np.column_stack([X.ravel(), Y.ravel()]))
    assert np.allclose(V[:, 2], Z.ravel() + 100)
    A, C = _areas(np.column_stack([V[:, :2], 0 * V[:, 2]]), F)
    assert abs(A.sum() - 3 * 2) < 1e-12 and (C[:, 2] > 0).all()
    with pytest.raises(ValueError):
        mesh_from_grid(x, y, Z.T)
    rng = np.random.default_rng(1)
    a = np.column_stack([rng.uniform(0, 1, (50, 2)), rng.uniform(0, 0.1, 50)])
    p = tmp_path / "dem.xyz"
    np.savetxt(p, a)
    V2, F2 = mesh_from_xyz(str(p))
    assert np.array_equal(V2, np.loadtxt(p))
    assert F2.max() == 49 and len(F2) >= 50
