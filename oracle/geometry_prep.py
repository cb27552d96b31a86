"""Oracle (test infrastructure only): the geometry of discretization nodes
from a mesh's nodal coordinates, restating pytential's geometry primitives
(pytential/symbolic/primitives.py:1150-1447 -- the parametrization
derivative of the element interpolant, normal, area element,
_simplex_mapping_max_stretch_factor :1172-1213, _expansion_radii_factor
:1301-1306, expansion_radii / expansion_centers :1360-1395,
weights_and_area_elements :1435-1447) in numpy.  Pinned in
tests/test_oracle_geometry_prep.py by closed forms: straight segments and
flat triangles (exact area, normal, stretch), and the circle / sphere
(perimeter 2 pi, area 4 pi, unit normals orthogonal to the tangents).
"""
import numpy as np


def element_geometry(dim, el_nodes, interp, deriv, ref_weights, stretch_map, radius_factor):
    """el_nodes (dim, nel, n1); interp (m, n1) or None; deriv (dim-1, m, n1);
    ref_weights (m,); stretch_map (dim-1, dim-1).  Returns nodes, normals
    (dim, nel*m), weights, radii (nel*m,), centers (dim, 2*nel*m) interleaved
    (c-, c+)."""
    xe = np.asarray(el_nodes, np.float64)
    _, nel, n1 = xe.shape
    D = np.asarray(deriv, np.float64)
    m = D.shape[1]
    I = np.eye(n1) if interp is None else np.asarray(interp, np.float64)
    x = np.einsum("ij,aej->aei", I, xe).reshape(dim, -1)
    J = np.einsum("kij,aej->akei", D, xe).reshape(dim, dim - 1, -1)   # (dim, rd, N)
    if dim == 2:
        t = J[:, 0, :]
        area = np.sqrt(t[0] ** 2 + t[1] ** 2)
        n = np.stack([t[1], -t[0]]) / area
        stretch = abs(float(np.asarray(stretch_map).reshape(-1)[0])) * area
    else:
        cr = np.cross(J[:, 0, :].T, J[:, 1, :].T).T
        area = np.linalg.norm(cr, axis=0)
        n = cr / area
        Je = np.einsum("akn,kl->aln", J, np.asarray(stretch_map, np.float64))
        G = np.einsum("akn,aln->nkl", Je, Je)
        # largest eigenvalue of each symmetric 2 x 2
        h = 0.5 * (G[:, 0, 0] - G[:, 1, 1])
        lam = 0.5 * (G[:, 0, 0] + G[:, 1, 1]) + np.sqrt(h * h + G[:, 0, 1] ** 2)
        stretch = np.sqrt(lam)
    w = area * np.tile(np.asarray(ref_weights, np.float64), nel)
    r = radius_factor * stretch
    c = np.empty((dim, 2 * x.shape[1]))
    c[:, 0::2] = x - r * n
    c[:, 1::2] = x + r * n
    del m
    return x, n, w, r, c
