"""Seeded synthetic meshes for the workloads and tests.

* ``icosphere(n)``         -- geodesic sphere with 10 n^2 + 2 vertices and
                              20 n^2 faces (the shape of the reference's
                              tests/data/icosa_sphere*.npz fixtures, n = 4, 5)
* ``spherical_crater``     -- unit square [-1, 1]^2 with a spherical-cap bowl
                              (rim radius 0.5, depth 0.15), Delaunay of a
                              jittered 32 x 32 grid (~1.9k facets)
* ``fractal_terrain``      -- diamond-square height field (Hurst 0.8) with a
                              seeded crater field, relief 0.1 of the width,
                              sampled onto a regular vertex grid
* ``grid_mesh``            -- two triangles per cell of a height grid

Every generator orients facets so normals point up (+z) or outward.
"""
import numpy as np


def icosphere(n=4, radius=1.0):
    """Icosahedron with every face split into n^2 triangles, projected onto
    the sphere.  Returns (V float64 (10n^2+2, 3), F int64 (20n^2, 3))."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                     [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                     [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    faces = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
             [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
             [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
             [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = {}
    V = []

    def vid(key, pos):
        if key not in verts:
            verts[key] = len(V)
            V.append(pos)
        return verts[key]

    F = []
    for (a, b, c) in faces:
        A, B, C = base[a], base[b], base[c]

        def point(i, j):
            # barycentric lattice point i*(B-A)/n + j*(C-A)/n; shared points
            # get one id through an exact rational key
            if i == 0 and j == 0:
                return vid(("v", a), A)
            if i == n and j == 0:
                return vid(("v", b), B)
            if i == 0 and j == n:
                return vid(("v", c), C)
            w = {a: n - i - j, b: i, c: j}
            key = tuple(sorted((k, v) for k, v in w.items() if v > 0))
            pos = (A * (n - i - j) + B * i + C * j) / n
            return vid(("p",) + key, pos)

        for i in range(n):
            for j in range(n - i):
                F.append([point(i, j), point(i + 1, j), point(i, j + 1)])
                if i + j < n - 1:
                    F.append([point(i + 1, j), point(i + 1, j + 1), point(i, j + 1)])
    V = np.array(V)
    V = radius * V / np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(F, dtype=np.int64)
    # outward orientation
    P = V[F].mean(axis=1)
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = (C * P).sum(1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def _orient_up(V, F):
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = C[:, 2] < 0
    F = F.copy()
    F[flip] = F[flip][:, [0, 2, 1]]
    return F


def grid_mesh(x, y, Z):
    """Regular grid (len(x) x len(y) vertices, Z[i, j] at (x[i], y[j])) with
    each cell split along the same diagonal.  Facets face +z."""
    nx, ny = len(x), len(y)
    X, Y = np.meshgrid(x, y, indexing="ij")
    V = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()]).astype(np.float64)
    idx = np.arange(nx * ny).reshape(nx, ny)
    a = idx[:-1, :-1].ravel()
    b = idx[1:, :-1].ravel()
    c = idx[1:, 1:].ravel()
    d = idx[:-1, 1:].ravel()
    F = np.concatenate([np.column_stack([a, b, c]), np.column_stack([a, c, d])]).astype(np.int64)
    return V, _orient_up(V, F)


def spherical_crater(npts=32, seed=1001, rim=0.5, depth=0.15, jitter=0.3):
    """Config-1 crater: jittered npts x npts grid on [-1, 1]^2, Delaunay,
    spherical-cap bowl of rim radius `rim` and depth `depth` cut in."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    g = np.linspace(-1.0, 1.0, npts)
    h = g[1] - g[0]
    X, Y = np.meshgrid(g, g, indexing="ij")
    X, Y = X.copy(), Y.copy()
    interior = (np.abs(X) < 1) & (np.abs(Y) < 1)
    X[interior] += rng.uniform(-jitter * h, jitter * h, interior.sum())
    Y[interior] += rng.uniform(-jitter * h, jitter * h, interior.sum())
    XY = np.column_stack([X.ravel(), Y.ravel()])
    F = Delaunay(XY).simplices.astype(np.int64)
    # sphere through the rim circle with the given depth
    R = (rim ** 2 + depth ** 2) / (2 * depth)
    H = R - depth
    r = np.hypot(XY[:, 0], XY[:, 1])
    Z = np.where(r <= rim, H - np.sqrt(np.maximum(R ** 2 - r ** 2, 0.0)), 0.0)
    V = np.column_stack([XY, Z]).astype(np.float64)
    return V, _orient_up(V, F)


def batch_instances(count=32, nmin=1000, nmax=20000, seed=1003):
    """The batching configuration's meshes: `count` instances with facet
    counts log-uniform in [nmin, nmax], alternately spherical craters and
    fractal terrains (seeded shapes, depth / relief 0.05-0.3).  Returns a
    list of (V, F) pairs; the facet counts land within a few percent of the
    drawn target (the generators make 2 (g - 1)^2 facets on a g x g grid)."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))
    depths = rng.uniform(0.05, 0.3, count)
    out = []
    for i, (nt, dp) in enumerate(zip(targets, depths)):
        g = max(4, int(round(np.sqrt(nt / 2.0))) + 1)
        if i % 2 == 0:
            out.append(spherical_crater(npts=g, seed=seed + 17 * i, depth=dp))
        else:
            out.append(fractal_terrain(grid=g, levels=max(4, int(np.ceil(np.log2(g - 1)))),
                                       relief=dp / 2.0, n_craters=10, seed=seed + 17 * i))
    return out


def diamond_square(levels, hurst=0.8, seed=0):
    """(2^levels + 1)^2 height field, amplitude 2^-hurst per octave."""
    rng = np.random.default_rng(seed)
    n = 2 ** levels + 1
    Z = np.zeros((n, n))
    Z[0, 0], Z[0, -1], Z[-1, 0], Z[-1, -1] = rng.standard_normal(4)
    step = n - 1
    amp = 1.0
    while step > 1:
        half = step // 2
        # diamond step: centres of squares
        ii = np.arange(half, n, step)
        I, J = np.meshgrid(ii, ii, indexing="ij")
        Z[I, J] = 0.25 * (Z[I - half, J - half] + Z[I - half, J + half] +
                          Z[I + half, J - half] + Z[I + half, J + half]) \
            + amp * rng.standard_normal(I.shape)
        # square step: edge midpoints (rows on the step lattice take the odd
        # columns, rows between them the even ones)
        pts = []
        for i in range(0, n, half):
            start = half if (i // half) % 2 == 0 else 0
            for j in range(start, n, step):
                pts.append((i, j))
        P = np.array(pts)
        I, J = P[:, 0], P[:, 1]
        acc = np.zeros(len(P))
        cnt = np.zeros(len(P))
        for di, dj in ((-half, 0), (half, 0), (0, -half), (0, half)):
            a, b = I + di, J + dj
            ok = (a >= 0) & (a < n) & (b >= 0) & (b < n)
            acc[ok] += Z[a[ok], b[ok]]
            cnt[ok] += 1
        Z[I, J] = acc / cnt + amp * rng.standard_normal(len(P))
        amp *= 2.0 ** (-hurst)
        step = half
    return Z


def fractal_terrain(grid=145, levels=8, hurst=0.8, n_craters=30, relief=0.1, seed=1002,
                    width=2.0):
    """Config-2 terrain: diamond-square on (2^levels+1)^2 (257^2 by default)
    plus `n_craters` parabolic bowls (radius log-uniform in [0.03, 0.3] of the
    width / 2, depth/diameter 0.2), rescaled to relief * width, sampled
    bilinearly onto a grid x grid vertex mesh over [-w/2, w/2]^2.
    grid = 145 gives 2 * 144^2 = 41,472 facets."""
    rng = np.random.default_rng(seed)
    Z = diamond_square(levels, hurst, seed=seed)
    n = Z.shape[0]
    u = np.linspace(-1.0, 1.0, n)
    U, W = np.meshgrid(u, u, indexing="ij")
    # normalise the fractal to unit relief over the [-1, 1]^2 frame, then add
    # bowls whose depth is 0.2 x diameter in that frame
    Z = (Z - Z.min()) / max(np.ptp(Z), 1e-300)
    for _ in range(n_craters):
        cx, cy = rng.uniform(-1, 1, 2)
        rad = np.exp(rng.uniform(np.log(0.03), np.log(0.3)))
        d = np.hypot(U - cx, W - cy)
        Z = Z + np.where(d < rad, (d ** 2 / rad ** 2 - 1.0) * 0.4 * rad, 0.0)
    Z = Z - Z.min()
    Z = Z / max(Z.max(), 1e-300) * relief * width
    # bilinear resample onto grid x grid
    s = np.linspace(0.0, n - 1.0, grid)
    i0 = np.minimum(np.floor(s).astype(int), n - 2)
    f = s - i0
    Zr = (Z[i0][:, i0] * (1 - f)[:, None] * (1 - f)[None, :]
          + Z[i0 + 1][:, i0] * f[:, None] * (1 - f)[None, :]
          + Z[i0][:, i0 + 1] * (1 - f)[:, None] * f[None, :]
          + Z[i0 + 1][:, i0 + 1] * f[:, None] * f[None, :])
    x = np.linspace(-width / 2, width / 2, grid)
    return grid_mesh(x, x, Zr)


def sun_directions(m, elev_deg=(1.0, 15.0), seed=1002):
    """m unit sun vectors, elevation uniform in elev_deg, azimuth in [0, 360)."""
    rng = np.random.default_rng(seed)
    e = np.deg2rad(rng.uniform(elev_deg[0], elev_deg[1], m))
    a = np.deg2rad(rng.uniform(0.0, 360.0, m))
    return np.column_stack([np.cos(e) * np.cos(a), np.cos(e) * np.sin(a), np.sin(e)])
