"""Pins of the oracle's refinement conditions (oracle/refinement.py,
pytential/qbx/refinement.py:155-263) against scipy's KD-tree ball query and
closed-form cases."""
import numpy as np
import pytest


def _kdtree_flags_c1(nodes, starts, centers, radii, tol):
    from scipy.spatial import cKDTree
    t = cKDTree(nodes.T)
    flags = np.zeros(len(starts) - 1, np.int32)
    for i in range(centers.shape[1]):
        if t.query_ball_point(centers[:, i], (1 - tol) * radii[i]):
            flags[np.searchsorted(starts, i // 2, side="right") - 1] = 1
    return flags


@pytest.mark.parametrize("nel", [8, 24])
def test_condition1_matches_kdtree(nel):
    from oracle.refinement import disturbed_expansion_flags

    from pytential_amd.discretization import discretize_curve, starfish
    d = discretize_curve(starfish(5, 0.3), nel, 8, 16)
    starts = d.element_to_source_starts()
    c, r = d.centers(), d.center_radii()
    got = disturbed_expansion_flags(d.nodes, starts, c, r, 0.025)
    assert np.array_equal(got, _kdtree_flags_c1(d.nodes, starts, c, r, 0.025))
    if nel == 8:
        assert got.sum() > 0        # a coarse starfish has disturbed disks


def test_condition1_straight_line_has_none():
    from oracle.refinement import disturbed_expansion_flags

    from pytential_amd.discretization import gauss_legendre
    xr, _ = gauss_legendre(8)
    nel = 10
    s = ((np.arange(nel)[:, None] + (xr[None, :] + 1) / 2) / nel).reshape(-1)
    nodes = np.stack([s, 0 * s])
    h = 1.0 / nel
    n = np.array([0.0, -1.0])
    r = np.full(nodes.shape[1], 0.5 * h)
    c = np.empty((2, 2 * nodes.shape[1]))
    c[:, 0::2] = nodes - r * n[:, None]
    c[:, 1::2] = nodes + r * n[:, None]
    starts = np.arange(nel + 1) * 8
    assert disturbed_expansion_flags(nodes, starts, c, np.repeat(r, 2), 0.025).sum() == 0
    # a center moved onto a node of element 3 flags element 3 only
    c2 = c.copy()
    c2[:, 2 * 30 + 1] = nodes[:, 5]
    f = disturbed_expansion_flags(nodes, starts, c2, np.repeat(r, 2), 0.025)
    assert f.tolist() == [1 if e == 3 else 0 for e in range(nel)]


def test_condition2_matches_kdtree():
    from scipy.spatial import cKDTree

    from oracle.refinement import source_quadrature_flags
    rng = np.random.default_rng(4)
    src = rng.uniform(-1, 1, (3, 600))
    starts = np.arange(0, 601, 15)
    rad = rng.uniform(0.02, 0.12, len(starts) - 1)
    ctr = rng.uniform(-1, 1, (3, 300))
    got = source_quadrature_flags(src, starts, rad, ctr)
    t = cKDTree(ctr.T)
    exp = np.zeros(len(starts) - 1, np.int32)
    for i in range(src.shape[1]):
        e = np.searchsorted(starts, i, side="right") - 1
        if t.query_ball_point(src[:, i], rad[e]):
            exp[e] = 1
    assert np.array_equal(got, exp) and 0 < got.sum() < len(got)
