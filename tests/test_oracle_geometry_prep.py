"""Pins of the oracle's geometry preparation (oracle/geometry_prep.py,
restating pytential/symbolic/primitives.py:1150-1447) by closed forms:
straight panels and flat triangles (exact area element, normal, stretch and
expansion radius), and the circle and sphere (perimeter 2 pi, area 4 pi,
unit outward normals)."""
import numpy as np
import pytest


def _curve(el, order, fine):
    from oracle.geometry_prep import element_geometry

    from pytential_amd.mesh import curve_operators
    _, U, D1, Df, w1, wf = curve_operators(order, fine)
    s1 = element_geometry(2, el, None, D1, w1, [[2.0]], 0.5)
    s2 = element_geometry(2, el, U, Df, wf, [[2.0]], 0.5)
    return s1, s2


def test_straight_panels_exact():
    from pytential_amd.discretization import gauss_legendre
    P, Q, nel, order = np.array([0.3, -1.0]), np.array([2.3, 0.5]), 6, 7
    xr, _ = gauss_legendre(order)
    s = (np.arange(nel)[:, None] + (xr[None, :] + 1) / 2) / nel      # (nel, order)
    el = P[:, None, None] + (Q - P)[:, None, None] * s[None]
    (x, n, w, r, c), (x2, n2, w2, _, _) = _curve(el, order, 12)
    L = np.linalg.norm(Q - P)
    t = (Q - P) / L
    assert np.allclose(n, np.array([t[1], -t[0]])[:, None], atol=1e-14)
    assert np.allclose(r, 0.5 * L / nel, rtol=1e-13)
    assert abs(w.sum() - L) < 1e-13 and abs(w2.sum() - L) < 1e-13
    assert np.allclose(c[:, 1::2], x + r * n, atol=1e-14)
    assert np.allclose(c[:, 0::2], x - r * n, atol=1e-14)


def test_circle_perimeter_and_normals():
    from pytential_amd.discretization import ellipse
    from pytential_amd.mesh import curve_mesh_nodes
    el = curve_mesh_nodes(ellipse(1.0, 1.0), 32, 8)
    (x, n, w, r, _), (x2, n2, w2, _, _) = _curve(el, 8, 16)
    assert abs(w2.sum() - 2 * np.pi) < 1e-10
    assert abs(w.sum() - 2 * np.pi) < 1e-10
    assert np.allclose(np.sum(n2 * x2, axis=0) / np.linalg.norm(x2, axis=0), 1.0, atol=1e-9)
    assert np.allclose(r, 0.5 * 2 * np.pi / 32, rtol=1e-6)


def _tri(el, order, fine):
    from oracle.geometry_prep import element_geometry

    from pytential_amd.mesh import triangle_operators
    from pytential_amd.surface import _EQUI_TO_UNIT
    _, U, D1, Df, w1, wf = triangle_operators(order, fine)
    E = 2 * _EQUI_TO_UNIT
    return (element_geometry(3, el, None, D1, w1, E, 0.25),
            element_geometry(3, el, U, Df, wf, E, 0.25))


@pytest.mark.parametrize("corners,stretch", [
    (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float), np.sqrt(2.0)),
    (np.array([[0, 0, 0], [2, 0, 0], [1, np.sqrt(3.0), 0]], float), 2.0)])
def test_flat_triangle_exact(corners, stretch):
    from pytential_amd.surface import triangle_nodes
    order = 4
    st = triangle_nodes(order)
    rot = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))[0]
    V = corners @ rot.T + np.array([0.2, -0.4, 1.0])
    el = (V[0][:, None] + st[0][None] * (V[1] - V[0])[:, None]
          + st[1][None] * (V[2] - V[0])[:, None])[:, None, :]
    (x, n, w, r, _), (x2, n2, w2, _, _) = _tri(el, order, 6)
    cr = np.cross(V[1] - V[0], V[2] - V[0])
    area = 0.5 * np.linalg.norm(cr)
    assert abs(w.sum() - area) < 1e-12 and abs(w2.sum() - area) < 1e-12
    assert np.allclose(n, (cr / np.linalg.norm(cr))[:, None], atol=1e-13)
    assert np.allclose(r, 0.25 * stretch, rtol=1e-12)


def test_sphere_area_and_normals():
    from pytential_amd.mesh import surface_mesh_nodes
    from pytential_amd.surface import icosahedron, sphere_map, subdivide
    v, f = icosahedron()
    v, f = subdivide(v, f, 2)
    el = surface_mesh_nodes(v, f, sphere_map(1.0), 4)
    (x, n, w, r, _), (x2, n2, w2, _, _) = _tri(el, 4, 6)
    assert abs(w2.sum() - 4 * np.pi) < 1e-4
    assert np.allclose(np.sum(n2 * x2, axis=0), 1.0, atol=1e-3)
    assert (np.sum(n * x, axis=0) > 0.99).all()
