"""Triangulations for the solve: generators, BEST's file format and a closedness check.

BEST reads ``NAR, TRUEAREA, TRUEVOL`` (list-directed) and then one line of 9 coordinates per
triangle, written by rcoal as ``format(9e17.8)`` (rcoal/scoalesce.f:895-899).  Vertex order is
the reference convention: counter-clockwise seen from outside, so that BEST's normal
``(v3 - v1) x (v2 - v1)`` (best/tgetgr.f:56-58) points inward and the signed volume computed
with it is negative, as on the reference example meshes.

Shapes (all star-shaped about the origin, meshed by mapping unit directions through a radial
function, so the topology of the direction mesh carries over):

* ``icosphere(level)``           -- the icosahedron refined ``level`` times (20 * 4**level faces);
* ``sphere_mesh``                -- radius ``r``;
* ``bumpy_ellipsoid_mesh``       -- semi-axes (a, b, c) times (1 + sum_{1<=l<=lmax} c_lm Y_lm),
                                    with orthonormal real spherical harmonics and c_lm ~ N(0, sd^2)
                                    from ``numpy.random.default_rng(seed)``;
* ``random_delaunay_directions`` -- N triangles from the convex hull (the spherical Delaunay
                                    triangulation) of (N + 4) / 2 seeded random directions.

``exact_area_volume`` integrates the smooth surface with tensor Gauss quadrature in (cos theta, phi)
so the header carries the area and volume of the surface itself, not of the polyhedron.
"""
from __future__ import annotations

import math

import numpy as np


# ------------------------------------------------------------------ topology

def icosphere(level: int):
    """Unit-sphere vertices (V x 3) and faces (F x 3 int) of a refined icosahedron, outward CCW."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in verts]
    f = list(faces)
    for _ in range(level):
        mid = {}

        def midpoint(i, j):
            key = (i, j) if i < j else (j, i)
            if key not in mid:
                p = v[i] + v[j]
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]

        nf = []
        for (a, b, c) in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(v)
    F = np.array(f, dtype=np.int64)
    return V, orient_outward(V, F)


def orient_outward(V, F):
    """Reorders each face so that (v2 - v1) x (v3 - v1) points away from the origin."""
    F = F.copy()
    p1, p2, p3 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nrm = np.cross(p2 - p1, p3 - p1)
    flip = np.einsum("ij,ij->i", nrm, p1 + p2 + p3) < 0
    F[flip, 1], F[flip, 2] = F[flip, 2].copy(), F[flip, 1].copy()
    return F


def random_delaunay_directions(ntri: int, seed: int):
    """Unit directions and faces of the spherical Delaunay triangulation of (ntri + 4) / 2 seeded
    random points (a closed genus-0 mesh has F = 2V - 4, so ntri must be even)."""
    from scipy.spatial import ConvexHull
    if ntri % 2:
        raise ValueError("a closed triangulation of a sphere has an even number of triangles")
    nv = (ntri + 4) // 2
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(nv, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    hull = ConvexHull(p)
    F = orient_outward(p, np.asarray(hull.simplices, dtype=np.int64))
    if len(F) != ntri:
        raise RuntimeError(f"hull has {len(F)} faces, expected {ntri}")
    return p, F


# ------------------------------------------------------------------ radial shapes

def real_sph_harm(lmax: int, u):
    """Orthonormal real spherical harmonics Y_lm(u) for 1 <= l <= lmax at unit vectors u (K x 3);
    returns (K x nharm) with columns ordered l = 1..lmax, m = -l..l."""
    from scipy.special import sph_harm_y
    theta = np.arccos(np.clip(u[:, 2], -1.0, 1.0))
    phi = np.arctan2(u[:, 1], u[:, 0])
    cols = []
    for l in range(1, lmax + 1):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), theta, phi)
            if m == 0:
                cols.append(y.real)
            elif m > 0:
                cols.append(math.sqrt(2.0) * (-1) ** m * y.real)
            else:
                cols.append(math.sqrt(2.0) * (-1) ** m * y.imag)
    return np.stack(cols, axis=1)


class RadialShape:
    """x(u) = diag(axes) u (1 + sum c_lm Y_lm(u)) for unit directions u."""

    def __init__(self, axes=(1.0, 1.0, 1.0), lmax=0, sd=0.0, seed=0):
        self.axes = np.asarray(axes, dtype=np.float64)
        self.lmax = int(lmax)
        if self.lmax > 0 and sd > 0:
            rng = np.random.default_rng(seed)
            self.coef = rng.normal(0.0, sd, size=(self.lmax + 1) ** 2 - 1)
        else:
            self.coef = np.zeros(0)

    def bump(self, u):
        if self.coef.size == 0:
            return np.ones(len(u))
        return 1.0 + real_sph_harm(self.lmax, u) @ self.coef

    def __call__(self, u):
        u = np.atleast_2d(u)
        return (u * self.axes[None, :]) * self.bump(u)[:, None]


def exact_area_volume(shape: RadialShape, nth=160, nph=320, h=1e-3):
    """Area and enclosed volume of the smooth surface x(theta, phi), by Gauss-Legendre in cos(theta)
    and the trapezoid rule in phi (spectral for periodic integrands), with 4th-order central
    differences for the tangents."""
    z, wz = np.polynomial.legendre.leggauss(nth)
    phis = np.arange(nph) * (2 * math.pi / nph)
    th = np.arccos(z)
    T, P = np.meshgrid(th, phis, indexing="ij")

    def X(t, p):
        u = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], axis=-1).reshape(-1, 3)
        return shape(u).reshape(t.shape + (3,))

    def d(fun, t, p, which):
        if which == 0:
            f = lambda s: fun(t + s, p)
        else:
            f = lambda s: fun(t, p + s)
        return (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)

    x = X(T, P)
    xt, xp = d(X, T, P, 0), d(X, T, P, 1)
    cr = np.cross(xt, xp)
    # dA = |x_theta x x_phi| dtheta dphi and dtheta = dz / sin(theta)
    jac = 1.0 / np.sin(T)
    wts = (wz[:, None] * jac) * (2 * math.pi / nph)
    area = float(np.sum(np.linalg.norm(cr, axis=-1) * wts))
    vol = float(np.sum(np.einsum("...k,...k->...", x, cr) * wts) / 3.0)
    return area, abs(vol)


# ------------------------------------------------------------------ BEST triangulations

class Triangulation:
    """BEST's input: v1, v2, v3 (NAR x 3) and the header's true area and volume."""

    def __init__(self, v1, v2, v3, truearea, truevol, name="mesh"):
        self.v1 = np.asarray(v1, dtype=np.float64)
        self.v2 = np.asarray(v2, dtype=np.float64)
        self.v3 = np.asarray(v3, dtype=np.float64)
        self.truearea = float(truearea)
        self.truevol = float(truevol)
        self.name = name

    @property
    def nar(self):
        return len(self.v1)

    @classmethod
    def from_vf(cls, V, F, truearea, truevol, name="mesh"):
        return cls(V[F[:, 0]], V[F[:, 1]], V[F[:, 2]], truearea, truevol, name)

    def signed_volume_best_normal(self):
        """Sum over triangles of centroid . (BEST's normal) / 6: negative for the reference
        orientation."""
        n = np.cross(self.v3 - self.v1, self.v2 - self.v1)
        c = (self.v1 + self.v2 + self.v3) / 3.0
        return float(np.einsum("ij,ij->", c, n) / 6.0)

    def polyhedral_area(self):
        return float(0.5 * np.linalg.norm(np.cross(self.v2 - self.v1, self.v3 - self.v1), axis=1).sum())

    def concat(self, other, name=None):
        return Triangulation(np.vstack([self.v1, other.v1]), np.vstack([self.v2, other.v2]),
                             np.vstack([self.v3, other.v3]), self.truearea + other.truearea,
                             self.truevol + other.truevol, name or self.name + "+" + other.name)

    def as_written(self):
        """The triangulation as BEST would read it back from its own file format (coordinates
        rounded to the 8 significant digits of format 9e17.8)."""
        return parse_best(format_best(self), self.name)


def vertex_index(t: Triangulation):
    """Shared vertices by exact coordinates (as COMPON matches them): (V unique vertices x 3,
    idx NAR x 3 with t.v1[i] == V[idx[i, 0]] and so on)."""
    allv = np.concatenate([t.v1, t.v2, t.v3])
    V, inv = np.unique(allv, axis=0, return_inverse=True)
    return V, inv.reshape(3, t.nar).T.copy()


def with_vertices(t: Triangulation, V, idx, name=None) -> Triangulation:
    """The triangulation with its shared vertices moved to V (the header kept)."""
    V = np.asarray(V, dtype=np.float64)
    return Triangulation(V[idx[:, 0]], V[idx[:, 1]], V[idx[:, 2]], t.truearea, t.truevol, name or t.name)


def vertex_gradient(grad, idx, nv):
    """Per-triangle vertex-copy gradients (..., NAR, 3 vertices, 3) summed onto the shared
    vertices: (..., nv, 3)."""
    g = np.asarray(grad)
    lead = g.shape[:-3]
    out = np.zeros(lead + (nv, 3))
    flat = g.reshape(lead + (-1, 3))
    for k in np.ndindex(*lead) if lead else [()]:
        np.add.at(out[k], idx.reshape(-1), flat[k])
    return out


def _fortran_e(v: float, w=17, d=8) -> str:
    """Fortran Ew.d: 0.dddddddd E+xx with d significant digits."""
    if v == 0.0 or not math.isfinite(v):
        s = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(v):.{d - 1}E}".split("E")
        digits = m.replace(".", "")
        ex = int(e) + 1
        s = ("-" if v < 0 else "") + "0." + digits + ("E+" if ex >= 0 else "E-") + f"{abs(ex):02d}"
    return s.rjust(w)


def format_best(t: Triangulation) -> str:
    """BEST's input file text: list-directed header, then format(9e17.8) per triangle."""
    lines = [f" {t.nar:12d}  {t.truearea!r}  {t.truevol!r}"]
    for i in range(t.nar):
        row = list(t.v1[i]) + list(t.v2[i]) + list(t.v3[i])
        lines.append("".join(_fortran_e(float(c)) for c in row))
    return "\n".join(lines) + "\n"


def parse_best(text: str, name="mesh") -> Triangulation:
    head, _, body = text.partition("\n")
    h = head.split()
    nar = int(h[0])
    ta = float(h[1]) if len(h) > 1 else 0.0
    tv = float(h[2]) if len(h) > 2 else 0.0
    vals = np.array(body.replace("D", "E").replace("d", "e").split(), dtype=np.float64)[: 9 * nar].reshape(nar, 9)
    return Triangulation(vals[:, 0:3], vals[:, 3:6], vals[:, 6:9], ta, tv, name)


def read_best(path: str) -> Triangulation:
    with open(path) as f:
        import os
        return parse_best(f.read(), os.path.basename(path))


def write_best(path: str, t: Triangulation) -> None:
    with open(path, "w") as f:
        f.write(format_best(t))


def closed_check(t: Triangulation):
    """(open edges, edges on more than two triangles, duplicate triangles) with vertices matched
    by exact coordinates, as test/meshcheck.pl checks rcoal's output."""
    keys = {}
    ids = np.zeros((t.nar, 3), dtype=np.int64)
    for k, arr in enumerate((t.v1, t.v2, t.v3)):
        for i, p in enumerate(map(tuple, arr)):
            ids[i, k] = keys.setdefault(p, len(keys))
    edges = {}
    tris = set()
    dup = 0
    for a, b, c in ids:
        key = tuple(sorted((a, b, c)))
        dup += key in tris
        tris.add(key)
        for p, q in ((a, b), (b, c), (c, a)):
            e = (p, q) if p < q else (q, p)
            edges[e] = edges.get(e, 0) + 1
    return sum(1 for v in edges.values() if v == 1), sum(1 for v in edges.values() if v > 2), dup


# ------------------------------------------------------------------ the workloads

def sphere_mesh(level=3, radius=15.0, name=None) -> Triangulation:
    """Icosphere of the given level, vertices on the sphere, exact area 4 pi r^2 and volume."""
    V, F = icosphere(level)
    return Triangulation.from_vf(V * radius, F, 4.0 * math.pi * radius ** 2, 4.0 / 3.0 * math.pi * radius ** 3,
                                 name or f"sphere_r{radius:g}_L{level}").as_written()


def bumpy_shape(axes=(30.0, 20.0, 15.0), lmax=6, sd=0.03, seed=20261203) -> RadialShape:
    return RadialShape(axes, lmax, sd, seed)


def shape_mesh(shape: RadialShape, V, F, name) -> Triangulation:
    area, vol = exact_area_volume(shape)
    return Triangulation.from_vf(shape(V), F, area, vol, name).as_written()


def bumpy_ellipsoid_mesh(level=4, seed=20261203, name=None) -> Triangulation:
    """Semi-axes 30/20/15 A with an l <= 6 radial bump field of sd 0.03, on an icosphere."""
    V, F = icosphere(level)
    return shape_mesh(bumpy_shape(seed=seed), V, F, name or f"bumpy_ellipsoid_L{level}")


SERIES_N = (2500, 3143, 3786, 4429, 5071, 5714, 6357, 7000)


def series_meshes(seed=20261203, sizes=SERIES_N):
    """The 1/N series: the bumpy ellipsoid remeshed by random-point spherical Delaunay at each size
    (odd sizes, impossible for a closed genus-0 mesh, are rounded up to the next even number)."""
    shape = bumpy_shape(seed=seed)
    area, vol = exact_area_volume(shape)
    out = []
    for k, n in enumerate(sizes):
        ne = n + (n % 2)
        V, F = random_delaunay_directions(ne, seed + 1000 + k)
        out.append(Triangulation.from_vf(shape(V), F, area, vol, f"bumpy_delaunay_{ne:05d}").as_written())
    return out


def two_surface_mesh(level=4, seed=20261203, gap=12.0, small_r=2.0) -> Triangulation:
    """The bumpy ellipsoid plus a 2 A sphere of 80 triangles whose surface is ``gap`` A beyond the
    ellipsoid's largest x along the +x axis."""
    big = bumpy_ellipsoid_mesh(level, seed)
    xmax = float(np.max(np.concatenate([big.v1[:, 0], big.v2[:, 0], big.v3[:, 0]])))
    V, F = icosphere(1)
    centre = np.array([xmax + gap + small_r, 0.0, 0.0])
    small = Triangulation.from_vf(V * small_r + centre, F, 4 * math.pi * small_r ** 2,
                                  4.0 / 3.0 * math.pi * small_r ** 3, "sphere2").as_written()
    return big.concat(small, f"two_surface_L{level}")


def hmatrix_sphere_mesh(level=6, seed=20261204) -> Triangulation:
    """40 A sphere times (1 + l <= 8 bump field of sd 0.05)."""
    V, F = icosphere(level)
    return shape_mesh(RadialShape((40.0, 40.0, 40.0), 8, 0.05, seed), V, F, f"bumpy_sphere40_L{level}")
