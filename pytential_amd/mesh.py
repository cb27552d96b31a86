"""Discretizations whose geometry is computed on the GPU from a mesh.

The input is what a meshmode mesh holds: the nodal coordinates of every
element (curve panels at Gauss-Legendre nodes on [-1, 1], curved triangles at
the equispaced nodes of order p in unit coordinates).  The stage-1 and
quad-stage-2 nodes, normals, weights (area element times reference weight),
expansion radii and centers then come from ``qbx_element_geometry``: the
derivatives of the element's polynomial interpolant, as pytential computes
them through meshmode (pytential/symbolic/primitives.py:1150-1447), not the
analytic parametrization.  ``MeshDiscretization`` carries the same fields as
the host-built discretizations (as device tensors), so it can be handed to
``QBXLayerPotentialSource`` and ``UnregularizedLayerPotentialSource``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import torch

from ._lib import call
from .discretization import gauss_legendre, lagrange_matrix
from .geometry import Context, _p, _stream, as_dev
from .surface import _EQUI_TO_UNIT, triangle_nodes, triangle_quadrature


def lagrange_derivative_matrix(xs):
    """D[i][j] = l_j'(xs[i]) for the Lagrange basis on the nodes xs."""
    n = len(xs)
    wb = np.array([1.0 / np.prod([xs[j] - xs[k] for k in range(n) if k != j])
                   for j in range(n)])
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = (wb[j] / wb[i]) / (xs[i] - xs[j])
        D[i, i] = -D[i].sum()
    return D


def _monomials(p, st, ds=0, dt=0):
    """Monomials s^a t^b (a + b <= p) or their derivatives at st (2, N)."""
    s, t = st
    cols = []
    for a in range(p + 1):
        for b in range(p + 1 - a):
            ca = (np.prod(range(a - ds + 1, a + 1)) if a >= ds else 0.0) if ds else 1.0
            cb = (np.prod(range(b - dt + 1, b + 1)) if b >= dt else 0.0) if dt else 1.0
            pa = s ** max(a - ds, 0)
            pb = t ** max(b - dt, 0)
            cols.append(ca * cb * pa * pb)
    return np.stack(cols, axis=1)


def triangle_operators(p, fine_n):
    """Interpolation and unit-coordinate derivative matrices of the order-p
    triangle (nodes ``triangle_nodes(p)``) at its own nodes and at the
    collapsed Gauss rule of ``fine_n`` x ``fine_n`` points, with the
    reference weights of both node sets."""
    st1 = triangle_nodes(p)
    stf, wf = triangle_quadrature(fine_n)
    Vi = np.linalg.inv(_monomials(p, st1))
    U = _monomials(p, stf) @ Vi
    D1 = np.stack([_monomials(p, st1, 1, 0) @ Vi, _monomials(p, st1, 0, 1) @ Vi])
    Df = np.stack([_monomials(p, stf, 1, 0) @ Vi, _monomials(p, stf, 0, 1) @ Vi])
    w1 = wf @ U
    return st1, U, D1, Df, w1, wf


def curve_operators(order, fine_order):
    xr, wr = gauss_legendre(order)
    xf, wf = gauss_legendre(fine_order)
    U = lagrange_matrix(xr, xf)
    D1 = lagrange_derivative_matrix(xr)[None]
    Df = (U @ D1[0])[None]
    return xr, U, D1, Df, wr, wf


@dataclass
class MeshDiscretization:
    """Stage-1 and quad-stage-2 geometry on the device (same field names as
    the host-built curve and surface discretizations)."""
    ambient_dim: int
    nelements: int
    order: int
    fine_order: int
    nodes: torch.Tensor = field(repr=False)
    normals: torch.Tensor = field(repr=False)
    weights: torch.Tensor = field(repr=False)
    expansion_radii: torch.Tensor = field(repr=False)
    interleaved_centers: torch.Tensor = field(repr=False)
    quad_nodes: torch.Tensor = field(repr=False)
    quad_normals: torch.Tensor = field(repr=False)
    quad_weights: torch.Tensor = field(repr=False)
    upsample: np.ndarray = field(repr=False)

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

    def centers(self):
        return self.interleaved_centers

    def center_radii(self):
        return torch.repeat_interleave(self.expansion_radii, 2)

    def tunnel_radii(self):
        return 0.5 * self.expansion_radii

    def element_to_source_starts(self):
        return torch.arange(self.nelements + 1, dtype=torch.int32,
                            device="cuda") * self.nodes_per_element


def _geometry(ctx, dim, el_nodes, interp, deriv, refw, E, factor, want_centers):
    _, nel, n1 = el_nodes.shape
    m = deriv.shape[1]
    xe = as_dev(np.ascontiguousarray(el_nodes).reshape(dim, -1))
    ip = None if interp is None else as_dev(interp)
    dv, rw, em = as_dev(deriv), as_dev(refw), as_dev(np.asarray(E, np.float64).reshape(-1))
    n = nel * m
    x = torch.empty((dim, n), dtype=torch.float64, device="cuda")
    nv = torch.empty_like(x)
    w = torch.empty(n, dtype=torch.float64, device="cuda")
    r = torch.empty(n, dtype=torch.float64, device="cuda") if want_centers else None
    c = torch.empty((dim, 2 * n), dtype=torch.float64, device="cuda") if want_centers else None
    call("qbx_element_geometry", ctx.h, _stream(), dim, nel, n1, _p(xe), m, _p(ip), _p(dv),
         _p(rw), _p(em), float(factor), _p(x), _p(nv), _p(w), _p(r), _p(c))
    return x, nv, w, r, c


def discretize_mesh(el_nodes, order, fine_order, ctx=None):
    """``el_nodes`` (dim, nelements, n1): the mesh's nodal coordinates --
    curves (dim 2) at the ``order`` Gauss-Legendre nodes of each panel,
    surfaces (dim 3) at ``triangle_nodes(order)`` of each triangle, oriented
    so that J_s x J_t is the normal wanted.  ``fine_order``: quad-stage-2
    Gauss-Legendre points per panel, or the n of the n x n collapsed rule."""
    ctx = ctx or Context.get()
    el = np.asarray(el_nodes, np.float64)
    dim, nel, n1 = el.shape
    if dim == 2:
        _, U, D1, Df, w1, wf = curve_operators(order, fine_order)
        E, factor = np.array([[2.0]]), 0.5
    elif dim == 3:
        _, U, D1, Df, w1, wf = triangle_operators(order, fine_order)
        E, factor = 2.0 * _EQUI_TO_UNIT, 0.25
    else:
        raise ValueError("meshes of curves (dim 2) or surfaces (dim 3) only")
    if U.shape[1] != n1:
        raise ValueError(f"{n1} nodes per element do not match order {order}")
    x1, n1v, w1v, r1, c1 = _geometry(ctx, dim, el, None, D1, w1, E, factor, True)
    x2, n2v, w2v, _, _ = _geometry(ctx, dim, el, U, Df, wf, E, factor, False)
    return MeshDiscretization(dim, nel, order, fine_order, x1, n1v, w1v, r1, c1, x2, n2v, w2v,
                              U)


def curve_mesh_nodes(curve, nelements, order, t0=0.0):
    """Nodal coordinates of a curve mesh: the curve at the Gauss-Legendre
    nodes of uniform parameter panels (what meshmode's make_curve_mesh holds)."""
    xr, _ = gauss_legendre(order)
    edges = t0 + 2 * np.pi * np.arange(nelements + 1) / nelements
    t = (edges[:-1, None] + (xr[None, :] + 1) * 0.5 * np.diff(edges)[:, None]).reshape(-1)
    x, y, _, _ = curve(t)
    return np.stack([x, y]).reshape(2, nelements, order)


def surface_mesh_nodes(vertices, faces, mapping, order):
    """Nodal coordinates of a curved-triangle mesh: ``mapping`` of the flat
    triangles at ``triangle_nodes(order)``, each face oriented so that its
    flat normal points away from the origin (star-shaped surfaces)."""
    f = np.array(faces, copy=True)
    V = np.asarray(vertices, np.float64)
    a, b, c = V[f[:, 0]], V[f[:, 1]], V[f[:, 2]]
    flip = np.sum(np.cross(b - a, c - a) * (a + b + c), axis=1) < 0
    f[flip, 1], f[flip, 2] = f[flip, 2], f[flip, 1].copy()
    st = triangle_nodes(order)
    V0, V1, V2 = V[f[:, 0]], V[f[:, 1]], V[f[:, 2]]
    P = (V0[:, None, :] + st[0][None, :, None] * (V1 - V0)[:, None, :] +
         st[1][None, :, None] * (V2 - V0)[:, None, :])          # (nel, n1, 3)
    nel, n1 = P.shape[:2]
    zero = np.zeros((3, nel * n1))
    x, _ = mapping(P.reshape(-1, 3).T, [zero, zero])
    return x.reshape(3, nel, n1)

