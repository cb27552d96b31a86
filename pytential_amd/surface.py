"""Triangle-panel discretizations of closed surfaces in 3D: the inputs of the
QBX path for surfaces (what meshmode/modepy and pytential's geometry
primitives produce in the reference).

* Mesh: an icosahedron subdivided ``nsub`` times (20 * 4**nsub triangles),
  each flat triangle mapped exactly onto the surface by radial projection
  (sphere) followed by an axis scaling (ellipsoid) -- the role of
  meshmode.mesh.generation.generate_sphere / generate_icosphere.
* Stage-1 nodes: the (p+1)(p+2)/2 equispaced barycentric nodes of order p on
  each triangle (15 for p = 4); density interpolation uses an orthonormal
  Dubiner basis, so any unisolvent node set would do.
* Quad stage 2: the collapsed (Stroud conical) product rule with n x n points
  per triangle, exact for total degree 2n - 1.
* Expansion radii (pytential/symbolic/primitives.py:1150-1192, 1280-1290):
  ``r = 0.5 * 0.5 * stretch`` for surfaces, where ``stretch`` is the largest
  singular value of the Jacobian of the equilateral(side 2)-reference ->
  surface map, i.e. the side length for an equilateral flat triangle.
  Centers ``c+- = x +- r n`` interleaved ``[c-(n0), c+(n0), ...]``; tunnel
  radius ``0.5 r`` (:1359-1366).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.special import eval_jacobi, roots_jacobi


# ---- meshes

def icosahedron():
    """Vertices on the unit sphere and 20 faces oriented counter-clockwise
    seen from outside."""
    g = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, g, 0], [1, g, 0], [-1, -g, 0], [1, -g, 0],
                  [0, -1, g], [0, 1, g], [0, -1, -g], [0, 1, -g],
                  [g, 0, -1], [g, 0, 1], [-g, 0, -1], [-g, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1)[:, None]
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
                  [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
                  [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def subdivide(vertices, faces, nsub: int, project=True):
    """Split every triangle into four ``nsub`` times (edge midpoints shared)."""
    v = [tuple(x) for x in vertices]
    f = faces
    for _ in range(nsub):
        mid = {}
        verts = list(v)

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                m = (np.array(verts[a]) + np.array(verts[b])) * 0.5
                if project:
                    m /= np.linalg.norm(m)
                mid[key] = len(verts)
                verts.append(tuple(m))
            return mid[key]
        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)]
        v, f = verts, np.array(nf, dtype=np.int64)
    return np.array(v, dtype=np.float64), np.asarray(f, dtype=np.int64)


def sphere_map(radius: float = 1.0):
    return ellipsoid_map(radius, radius, radius)


def ellipsoid_map(a: float, b: float, c: float):
    """P (on flat triangles, 3 x N) -> x = A P/|P| and its derivative applied
    to a direction dP: dx = A (I - u u^T) dP / |P|, u = P/|P|."""
    A = np.array([a, b, c], dtype=np.float64)

    def mapping(P, dPs):
        nP = np.linalg.norm(P, axis=0)
        u = P / nP
        x = A[:, None] * u
        outs = []
        for dP in dPs:
            t = dP - u * np.sum(u * dP, axis=0)
            outs.append(A[:, None] * t / nP)
        return x, outs
    return mapping


# ---- reference triangle {(s, t): s, t >= 0, s + t <= 1}

def triangle_nodes(p: int) -> np.ndarray:
    """Equispaced nodes of order p, shape (2, (p+1)(p+2)/2)."""
    if p == 0:
        return np.array([[1 / 3], [1 / 3]])
    pts = [(i / p, j / p) for j in range(p + 1) for i in range(p + 1 - j)]
    return np.array(pts).T


def triangle_quadrature(n: int):
    """Collapsed Gauss rule: s = a, t = (1 - a) b with Gauss-Jacobi(1, 0) in a
    and Gauss-Legendre in b; n*n points, weights summing to 1/2."""
    xa, wa = roots_jacobi(n, 1.0, 0.0)
    xb, wb = np.polynomial.legendre.leggauss(n)
    a = (1 + xa) / 2
    b = (1 + xb) / 2
    A, B = np.meshgrid(a, b, indexing="ij")
    W = np.outer(wa, wb) / 8.0
    s = A.reshape(-1)
    t = ((1 - A) * B).reshape(-1)
    return np.stack([s, t]), W.reshape(-1)


def dubiner_vandermonde(p: int, st: np.ndarray) -> np.ndarray:
    """Orthogonal (Dubiner) basis of total degree <= p at points st (2, N)."""
    s, t = st
    r = 2 * s - 1
    q = 2 * t - 1
    # collapsed coordinates on the bi-unit triangle
    with np.errstate(divide="ignore", invalid="ignore"):
        a = np.where(np.abs(1 - q) > 1e-14, 2 * (1 + r) / (1 - q) - 1, -1.0)
    cols = []
    for i in range(p + 1):
        for j in range(p + 1 - i):
            cols.append(eval_jacobi(i, 0, 0, a) * ((1 - q) / 2) ** i *
                        eval_jacobi(j, 2 * i + 1, 0, q))
    return np.stack(cols, axis=1)


# equilateral (side 2) reference -> unit reference triangle, linear part
_EQUI_TO_UNIT = np.array([[1.0, -1.0 / np.sqrt(3.0)], [0.0, 2.0 / np.sqrt(3.0)]]) * 0.5


@dataclass
class SurfaceDiscretization:
    """Stage-1 and quad-stage-2 discretization of a closed surface made of
    curved triangles (same fields as the curve discretization, 3D arrays)."""
    nelements: int
    order: int                                   # stage-1 polynomial order p
    fine_order: int                              # quad-stage-2 rule: n x n points
    nodes: np.ndarray = field(repr=False)        # (3, n1)
    normals: np.ndarray = field(repr=False)
    weights: np.ndarray = field(repr=False)
    expansion_radii: np.ndarray = field(repr=False)
    quad_nodes: np.ndarray = field(repr=False)   # (3, n2)
    quad_normals: np.ndarray = field(repr=False)
    quad_weights: np.ndarray = field(repr=False)
    upsample: np.ndarray = field(repr=False)     # (nf, n1) per element
    vertices: np.ndarray = field(default=None, repr=False)   # flat triangle corners
    faces: np.ndarray = field(default=None, repr=False)
    mapping: object = field(default=None, repr=False)
    ambient_dim: int = 3

    @property
    def nodes_per_element(self) -> int:
        return self.upsample.shape[1]

    @property
    def fine_nodes_per_element(self) -> int:
        return self.upsample.shape[0]

    @property
    def ndofs(self) -> int:
        return self.nodes.shape[1]

    @property
    def nsources(self) -> int:
        return self.quad_nodes.shape[1]

    @property
    def ncenters(self) -> int:
        return 2 * self.ndofs

    def centers(self) -> np.ndarray:
        cm = self.nodes - self.expansion_radii * self.normals
        cp = self.nodes + self.expansion_radii * self.normals
        out = np.empty((3, 2 * self.ndofs))
        out[:, 0::2] = cm
        out[:, 1::2] = cp
        return out

    def center_radii(self) -> np.ndarray:
        return np.repeat(self.expansion_radii, 2)

    def tunnel_radii(self) -> np.ndarray:
        return 0.5 * self.expansion_radii

    def element_to_source_starts(self) -> np.ndarray:
        return (np.arange(self.nelements + 1) * self.nodes_per_element).astype(np.int32)

    def upsample_density(self, sigma: np.ndarray) -> np.ndarray:
        s = np.asarray(sigma).reshape(self.nelements, self.nodes_per_element)
        return (s @ self.upsample.T).reshape(-1)

    def strengths(self, sigma: np.ndarray) -> np.ndarray:
        return self.quad_weights * self.upsample_density(sigma)


def _eval_on_triangles(corners, mapping, st):
    """Points, unit normals, area elements and stretch at reference points st
    (2, m) on every triangle; corners (nel, 3, 3) = [element][vertex][axis]."""
    nel = corners.shape[0]
    s, t = st
    V0, V1, V2 = corners[:, 0, :], corners[:, 1, :], corners[:, 2, :]
    P = (V0[:, None, :] + s[None, :, None] * (V1 - V0)[:, None, :] +
         t[None, :, None] * (V2 - V0)[:, None, :]).reshape(-1, 3).T
    m = st.shape[1]
    dPs = np.repeat((V1 - V0), m, axis=0).T
    dPt = np.repeat((V2 - V0), m, axis=0).T
    x, (xs, xt) = mapping(P, [dPs, dPt])
    cr = np.cross(xs.T, xt.T).T
    area = np.linalg.norm(cr, axis=0)
    nrm = cr / area
    nrm *= np.sign(np.sum(nrm * x, axis=0))        # outward (star-shaped surfaces)
    # stretch: largest singular value of [xs xt] @ equilateral map, times 2
    J = np.stack([xs, xt], axis=2)                  # (3, N, 2)
    Je = 2.0 * np.einsum("dnk,kl->dnl", J, _EQUI_TO_UNIT)
    JtJ = np.einsum("dnk,dnl->nkl", Je, Je)
    stretch = np.sqrt(np.linalg.eigvalsh(JtJ)[:, -1])
    del nel
    return x, nrm, area, stretch


def discretize_surface(vertices, faces, mapping, order: int, fine_order: int,
                       corners=None) -> SurfaceDiscretization:
    if corners is None:
        corners = vertices[faces]
    nel = corners.shape[0]
    st1 = triangle_nodes(order)
    stf, wf = triangle_quadrature(fine_order)
    V1 = dubiner_vandermonde(order, st1)
    Vf = dubiner_vandermonde(order, stf)
    up = Vf @ np.linalg.inv(V1)
    x1, n1, a1, str1 = _eval_on_triangles(corners, mapping, st1)
    x2, n2, a2, _ = _eval_on_triangles(corners, mapping, stf)
    w2 = np.tile(wf, nel) * a2
    # interpolatory stage-1 weights from the fine rule
    w1 = ((np.tile(wf, nel) * a2).reshape(nel, -1) @ up).reshape(-1)
    radii = 0.25 * str1
    return SurfaceDiscretization(
        nelements=nel, order=order, fine_order=fine_order, nodes=x1, normals=n1,
        weights=w1, expansion_radii=radii, quad_nodes=x2, quad_normals=n2,
        quad_weights=w2, upsample=up, vertices=vertices, faces=faces, mapping=mapping)


def icosphere(nsub: int, order: int = 4, fine_order: int = 6, radius: float = 1.0):
    v, f = icosahedron()
    v, f = subdivide(v, f, nsub)
    return discretize_surface(v, f, sphere_map(radius), order, fine_order)


def ellipsoid(nsub: int, axes=(1.0, 0.75, 0.5), order: int = 4, fine_order: int = 6):
    v, f = icosahedron()
    v, f = subdivide(v, f, nsub)
    return discretize_surface(v, f, ellipsoid_map(*axes), order, fine_order)
