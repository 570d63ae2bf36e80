"""Hemispherical crater meshes with Ingersoll's closed-form shadow
temperature (drop-in for flux.ingersoll, ingersoll.py:10-455 of fluxpy).

The crater is a spherical cap cut into the plane z = 0: rim radius rc,
sphere of radius r = rc / sin(beta) centred at height H = rc / tan(beta).
For a sun at elevation e0 the permanently shadowed part of the bowl has, by
Ingersoll et al. (Icarus 100, 1992; the form used at ingersoll.py:24-34),
the uniform equilibrium temperature

    T_gt^4 = F0 sin(e0) f (1 - rho) / (1 - rho f) (1 + rho (1 - f) / emiss) / sigma,
    f = (1 - cos beta) / 2,

which makes a mesh of this crater an end-to-end check of visibility, form
factors, the radiosity solve and the temperature.

Meshing.  fluxpy builds its meshes with Shewchuk's Triangle (meshpy), which
is not available here; this module triangulates with SciPy's Delaunay
instead, and keeps the interface:

* ``make_trimesh(h)``: a jittered grid of spacing h over [-1, 1]^2, lifted
  onto the crater surface;
* ``make_trimesh(h, contour_rim=True)``: the bowl as concentric rings of
  spacing h (the outer ring on the rim) and the plane as a grid of spacing
  h outside the rim, sharing the rim vertices, so the rim is an edge loop
  of the mesh and each part can be returned (``return_part_indices``,
  ``return_parts``);
* ``contour_shadow=True`` needs the shadow line as a constrained edge, which
  an unconstrained Delaunay triangulation cannot guarantee, and raises
  NotImplementedError.
"""
import numpy as np
from scipy.constants import sigma
from scipy.optimize import brentq
from scipy.spatial import Delaunay


def _orient_up(V, F):
    """Flip triangles so that their normal has a non-negative z component."""
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    F = F.copy()
    flip = C[:, 2] < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return F


class HemisphericalCrater:
    """Spherical-cap crater in a plane (ingersoll.py:10-34): beta is the
    angle between the crater bottom and the rim seen from the sphere
    centre, rc the rim radius, e0 the solar elevation (radians), F0 the
    solar constant, rho the albedo and emiss the emissivity."""

    def __init__(self, beta, rc, e0, F0, rho, emiss):
        self.beta, self.rc, self.e0 = beta, rc, e0
        self.F0, self.rho, self.emiss = F0, rho, emiss
        self.H = rc / np.tan(beta)                   # sphere centre above the plane
        self.r = rc / np.sin(beta)                   # sphere radius
        self.Sc = 2 * np.pi * self.r * (self.r - self.H)   # cap area
        self.f = (1 - np.cos(beta)) / 2              # view factor of the cap to itself
        self.b = self.f * (emiss + rho * (1 - self.f)) / (1 - rho * self.f)
        T4 = F0 * np.sin(e0) * self.f * (1 - rho) / (1 - rho * self.f) \
            * (1 + rho * (1 - self.f) / emiss) / sigma
        self.T_gt = T4 ** 0.25

    @property
    def t0(self):
        return np.tan(self.e0)

    def _get_z(self, x, y):
        """Height of the surface at (x, y): the cap inside the rim, 0 outside."""
        R = np.hypot(x, y)
        inside = R <= self.rc + np.finfo(np.float32).eps
        z = self.H - np.sqrt(np.maximum(self.r ** 2 - np.minimum(R, self.rc) ** 2, 0.0))
        return np.where(inside, z, 0.0)

    def _x_silhouette(self, y):
        """x of the shadow line at y: the point of the bowl where the ray
        toward the sun (along -x, elevation e0) grazes the rim at
        x0 = sqrt(rc^2 - y^2); None outside the crater or where no shadow
        line crosses."""
        if abs(y) > self.rc:
            return None
        x0 = np.sqrt(self.rc ** 2 - y ** 2)

        def g(x):
            return self.H - np.sqrt(self.r ** 2 - x ** 2 - y ** 2) + self.t0 * (x0 - x)

        X = np.linspace(-x0, x0, 201)
        G = g(X)
        s = np.flatnonzero(np.sign(G[:-1]) * np.sign(G[1:]) < 0)
        # the crossing away from the far rim point x0 itself
        s = [i for i in s if X[i + 1] < x0 - 1e-12]
        if not s:
            return None
        i = s[-1]
        return brentq(g, X[i], X[i + 1])

    # -- meshes --------------------------------------------------------------

    def make_trimesh(self, h, contour_rim=False, contour_shadow=False, **kwargs):
        """(V, F) of the crater in [-1, 1]^2 with edge length about h (plus
        the part indices or parts with ``return_part_indices`` /
        ``return_parts`` when the rim is contoured)."""
        if contour_rim and contour_shadow:
            raise NotImplementedError(
                "contouring the shadow line needs a constrained triangulation")
        if contour_shadow:
            raise ValueError("can only contour the shadow line if the crater rim is also "
                             "contoured")
        if contour_rim:
            return self._make_trimesh__contour_rim(h, **kwargs)
        return self._make_trimesh__no_contouring(h, **kwargs)

    def _make_trimesh__no_contouring(self, h, seed=0):
        n = int(np.ceil(2 / h)) + 1
        g = np.linspace(-1, 1, n)
        X, Y = np.meshgrid(g, g, indexing="ij")
        rng = np.random.default_rng(seed)
        inner = (np.abs(X) < 1) & (np.abs(Y) < 1)
        step = g[1] - g[0]
        X = X + np.where(inner, rng.uniform(-0.25, 0.25, X.shape) * step, 0)
        Y = Y + np.where(inner, rng.uniform(-0.25, 0.25, Y.shape) * step, 0)
        xy = np.column_stack([X.ravel(), Y.ravel()])
        F = Delaunay(xy).simplices.astype(np.int64)
        V = np.column_stack([xy, self._get_z(xy[:, 0], xy[:, 1])])
        return V, _orient_up(V, F)

    def _rings(self, h):
        """Disk points: the centre and rings of radius rc k / K, k = 1..K,
        with about 2 pi r_k / h points each; the last ring is the rim."""
        K = max(1, int(np.ceil(self.rc / h)))
        pts = [np.zeros((1, 2))]
        for k in range(1, K + 1):
            rk = self.rc * k / K
            nk = max(6, int(np.ceil(2 * np.pi * rk / h)))
            th = 2 * np.pi * (np.arange(nk) + 0.5 * (k % 2)) / nk
            pts.append(np.column_stack([rk * np.cos(th), rk * np.sin(th)]))
        return np.concatenate(pts), len(pts[-1])

    def _make_trimesh__contour_rim(self, h, return_part_indices=False, return_parts=False):
        if return_part_indices and return_parts:
            raise Exception("only one of return_part_indices or return_parts should be true")
        disk, nrim = self._rings(h)
        rim = disk[-nrim:]
        # the bowl: Delaunay of the disk points (their hull is the rim polygon)
        Fc = Delaunay(disk).simplices.astype(np.int64)
        # the plane: grid points away from the rim plus the rim itself;
        # triangles whose centroid falls inside the rim polygon are the hole
        n = int(np.ceil(2 / h)) + 1
        g = np.linspace(-1, 1, n)
        X, Y = np.meshgrid(g, g, indexing="ij")
        grid = np.column_stack([X.ravel(), Y.ravel()])
        grid = grid[np.hypot(grid[:, 0], grid[:, 1]) > self.rc + 0.5 * h]
        plane = np.concatenate([rim, grid])
        Fp = Delaunay(plane).simplices.astype(np.int64)
        cen = plane[Fp].mean(1)
        apothem = self.rc * np.cos(np.pi / nrim)
        Fp = Fp[np.hypot(cen[:, 0], cen[:, 1]) > apothem]
        # one vertex array: the disk, then the grid points (the plane's rim
        # points are the disk's last ring)
        nd = disk.shape[0]
        remap = np.concatenate([np.arange(nd - nrim, nd), nd + np.arange(grid.shape[0])])
        Fp = remap[Fp]
        xy = np.concatenate([disk, grid])
        V = np.column_stack([xy, self._get_z(xy[:, 0], xy[:, 1])])
        V[nd - nrim:nd, 2] = 0.0                     # the rim lies in the plane
        Fp, Fc = _orient_up(V, Fp), _orient_up(V, Fc)
        F = np.concatenate([Fp, Fc])
        inds = (0, Fp.shape[0], F.shape[0])
        if return_part_indices:
            return V, F, inds
        if return_parts:
            return V, F, (np.arange(inds[0], inds[1]), np.arange(inds[1], inds[2]))
        return V, F
