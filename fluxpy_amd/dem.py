"""Triangle meshes from digital elevation models (the DEM front-end of
fluxpy's examples/ceres/generate_mesh_from_topo.py:36-93).

Both forms triangulate the DEM's own (x, y) positions by Delaunay, so every
vertex keeps its measured height (no interpolation):

* ``mesh_from_grid(x, y, Z)``: a gridded DEM (a GMT .grd / netCDF file's x,
  y and z variables: Z[i, j] is the height at (x[j], y[i]));
* ``mesh_from_xyz(a)``: scattered samples, an (n, 3) array of x, y, z
  columns or the path of a text file holding them.

Triangles are oriented so their normals point up (+z): the form-factor and
insolation code treats the normal as the facet's outward side.
"""
import numpy as np
from scipy.spatial import Delaunay


def _orient_up(V, F):
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    F = F.copy()
    flip = C[:, 2] < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return F


def _triangulate(xy, z):
    tri = Delaunay(xy)
    V = np.column_stack([tri.points, z]).astype(np.float64)
    F = tri.simplices.astype(np.int64)
    return V, _orient_up(V, F)


def mesh_from_grid(x, y, Z, z_offset=0.0):
    """(V, F) of a gridded DEM; V[:, 2] = Z + z_offset (the reference adds a
    constant offset to make a bowl-shaped crater face up)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    y = np.asarray(y, dtype=np.float64).ravel()
    Z = np.asarray(Z, dtype=np.float64)
    if Z.shape != (y.size, x.size):
        raise ValueError(f"Z must be (len(y), len(x)) = ({y.size}, {x.size}), got {Z.shape}")
    X, Y = np.meshgrid(x, y)
    xy = np.column_stack([X.ravel(), Y.ravel()])
    return _triangulate(xy, Z.ravel() + z_offset)


def mesh_from_xyz(a):
    """(V, F) of scattered DEM samples: an (n, 3) array or a text file of x,
    y, z columns.  The vertices are the samples themselves, in order."""
    if isinstance(a, (str, bytes)) or hasattr(a, "__fspath__"):
        a = np.loadtxt(a)
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] < 3:
        raise ValueError("expected x, y, z columns")
    V, F = _triangulate(a[:, :2], a[:, 2])
    if not np.array_equal(V, a[:, :3]):
        raise RuntimeError("Delaunay reordered or moved the samples")
    return V, F
