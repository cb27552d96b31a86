"""Panel discretizations of closed curves: the inputs of the QBX path.

This module produces, on the host, what meshmode/modepy and pytential's
geometry primitives hand to the QBX-FMM in the reference: stage-1 (density)
nodes, the oversampled quad-stage-2 source nodes with their weights, normals,
expansion radii and the interleaved QBX centers.

Conventions (reference citations are to pytential):

* expansion radius ``r = 0.5 * stretch`` with ``stretch = |dx/dr| * 2`` on the
  bi-unit reference element, i.e. ``r = 0.5 * |x'(t)| * dt`` for a panel of
  parameter length ``dt`` (``symbolic/primitives.py:1280-1365``: the factor is
  0.5 for curves, 0.25 for surfaces).
* centers ``c± = x ± r n`` interleaved as ``[c-(n0), c+(n0), c-(n1), ...]``
  (``symbolic/primitives.py:1389-1407``, ``symbolic/dof_connection.py:84-105``).
* close-target tunnel radius ``0.5 r`` (``symbolic/primitives.py:1359-1366``).
* strengths ``= (area element * quadrature weight) * density`` on the quad
  stage 2 nodes (``qbx/__init__.py:1075-1096``).
* an element of "order p" carries p Gauss-Legendre nodes, which matches the
  node counts in the benchmark configurations (64 x 8 = 512, 4096 x 8 =
  32768); meshmode itself would use p+1.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def gauss_legendre(n: int):
    x, w = np.polynomial.legendre.leggauss(n)
    return x, w


def lagrange_matrix(xs: np.ndarray, xt: np.ndarray) -> np.ndarray:
    """Interpolation matrix from values at nodes ``xs`` to points ``xt``
    (barycentric form, stable for Gauss-Legendre nodes)."""
    n = len(xs)
    wb = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                wb[j] /= (xs[j] - xs[k])
    mat = np.zeros((len(xt), n))
    for i, x in enumerate(xt):
        d = x - xs
        hit = np.nonzero(d == 0)[0]
        if len(hit):
            mat[i, hit[0]] = 1.0
            continue
        tmp = wb / d
        mat[i] = tmp / tmp.sum()
    return mat


def starfish(arms: int = 5, amplitude: float = 0.25, radius: float = 1.0,
             center=(0.0, 0.0), phase: float = 0.0):
    """r(t) = radius * (1 + amplitude cos(arms (t - phase))), counter-clockwise."""
    cx, cy = center

    def curve(t):
        r = radius * (1 + amplitude * np.cos(arms * (t - phase)))
        dr = -radius * amplitude * arms * np.sin(arms * (t - phase))
        x = cx + r * np.cos(t)
        y = cy + r * np.sin(t)
        dx = dr * np.cos(t) - r * np.sin(t)
        dy = dr * np.sin(t) + r * np.cos(t)
        return x, y, dx, dy
    return curve


def ellipse(a: float = 1.0, b: float = 0.5):
    def curve(t):
        return a * np.cos(t), b * np.sin(t), -a * np.sin(t), b * np.cos(t)
    return curve


@dataclass
class CurveDiscretization:
    """Stage-1 and quad-stage-2 discretization of a closed 2D curve."""
    nelements: int
    order: int
    fine_order: int
    # stage 1
    t_nodes: np.ndarray = field(repr=False)
    nodes: np.ndarray = field(repr=False)          # (2, n1)
    normals: np.ndarray = field(repr=False)        # (2, n1)
    weights: np.ndarray = field(repr=False)        # (n1,)  area element * weight
    expansion_radii: np.ndarray = field(repr=False)  # (n1,)
    # quad stage 2
    quad_nodes: np.ndarray = field(repr=False)     # (2, n2)
    quad_normals: np.ndarray = field(repr=False)
    quad_weights: np.ndarray = field(repr=False)
    upsample: np.ndarray = field(repr=False)       # (fine_order, order) per panel
    t_edges: np.ndarray = field(default=None, repr=False)  # (nelements+1,)
    curve: object = field(default=None, repr=False)
    ambient_dim: int = 2

    @property
    def nodes_per_element(self) -> int:
        return self.order

    @property
    def fine_nodes_per_element(self) -> int:
        return self.fine_order

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
        """Interleaved centers ``[c-(n0), c+(n0), ...]``, shape (2, 2*n1)."""
        cm = self.nodes - self.expansion_radii * self.normals
        cp = self.nodes + self.expansion_radii * self.normals
        out = np.empty((2, 2 * self.ndofs))
        out[:, 0::2] = cm
        out[:, 1::2] = cp
        return out

    def center_radii(self) -> np.ndarray:
        return np.repeat(self.expansion_radii, 2)

    def tunnel_radii(self) -> np.ndarray:
        return 0.5 * self.expansion_radii

    def element_to_source_starts(self) -> np.ndarray:
        return (np.arange(self.nelements + 1) * self.order).astype(np.int32)

    def upsample_density(self, sigma: np.ndarray) -> np.ndarray:
        """Stage-1 nodal values -> quad-stage-2 nodal values (per panel)."""
        s = np.asarray(sigma).reshape(self.nelements, self.order)
        return (s @ self.upsample.T).reshape(-1)

    def strengths(self, sigma: np.ndarray) -> np.ndarray:
        return self.quad_weights * self.upsample_density(sigma)

    def refined(self, refine_flags) -> tuple["CurveDiscretization", "PanelConnection"]:
        """Split every flagged panel in two (parameter midpoint).  Returns the
        new discretization and the connection that carries stage-1 nodal data
        across (per-panel Lagrange interpolation, the role of meshmode's
        refinement connection)."""
        flags = np.asarray(refine_flags).astype(bool)
        te = self.t_edges
        new_edges = [te[0]]
        parent, which = [], []
        for j in range(self.nelements):
            if flags[j]:
                new_edges += [0.5 * (te[j] + te[j + 1]), te[j + 1]]
                parent += [j, j]
                which += [1, 2]
            else:
                new_edges.append(te[j + 1])
                parent.append(j)
                which.append(0)
        new = discretize_curve_edges(self.curve, np.array(new_edges), self.order,
                                     self.fine_order)
        xr, _ = gauss_legendre(self.order)
        mats = [np.eye(self.order), lagrange_matrix(xr, (xr - 1) * 0.5),
                lagrange_matrix(xr, (xr + 1) * 0.5)]
        return new, PanelConnection(np.array(parent), np.array(which), mats, self.order)


@dataclass
class PanelConnection:
    """Stage-1 data on parent panels -> refined panels (block interpolation)."""
    parent: np.ndarray
    which: np.ndarray      # 0 copy, 1 left half, 2 right half
    mats: list
    order: int

    def __call__(self, x):
        x = np.asarray(x)
        blocks = x.reshape(-1, self.order)[self.parent]
        out = np.empty((len(self.parent), self.order), dtype=x.dtype)
        for k in range(3):
            sel = self.which == k
            out[sel] = blocks[sel] @ self.mats[k].T
        return out.reshape(-1)


def discretize_curve_edges(curve, t_edges, order: int, fine_order: int) -> CurveDiscretization:
    t_edges = np.asarray(t_edges, dtype=np.float64)
    nelements = len(t_edges) - 1
    xr, wr = gauss_legendre(order)
    xf, wf = gauss_legendre(fine_order)
    tl = t_edges[:-1]
    dt = np.diff(t_edges)

    def panel_eval(xref):
        t = (tl[:, None] + (xref[None, :] + 1) * 0.5 * dt[:, None]).reshape(-1)
        x, y, dx, dy = curve(t)
        speed = np.hypot(dx, dy)
        nrm = np.stack([dy / speed, -dx / speed])
        return t, np.stack([x, y]), nrm, speed

    dt1 = np.repeat(dt, order)
    dt2 = np.repeat(dt, fine_order)
    t1, x1, n1, sp1 = panel_eval(xr)
    w1 = np.tile(wr, nelements) * sp1 * 0.5 * dt1
    radii = 0.5 * sp1 * dt1
    _, x2, n2, sp2 = panel_eval(xf)
    w2 = np.tile(wf, nelements) * sp2 * 0.5 * dt2
    return CurveDiscretization(
        nelements=nelements, order=order, fine_order=fine_order,
        t_nodes=t1, nodes=x1, normals=n1, weights=w1, expansion_radii=radii,
        quad_nodes=x2, quad_normals=n2, quad_weights=w2,
        upsample=lagrange_matrix(xr, xf), t_edges=t_edges, curve=curve)


def discretize_curve(curve, nelements: int, order: int, fine_order: int,
                     t0: float = 0.0) -> CurveDiscretization:
    """Uniform panels in the curve parameter t over [t0, t0 + 2 pi]."""
    edges = t0 + 2 * np.pi * np.arange(nelements + 1) / nelements
    return discretize_curve_edges(curve, edges, order, fine_order)


def grid_targets(n: int, lo: float = -2.0, hi: float = 2.0) -> np.ndarray:
    g = np.linspace(lo, hi, n)
    xx, yy = np.meshgrid(g, g, indexing="ij")
    return np.stack([xx.reshape(-1), yy.reshape(-1)])
