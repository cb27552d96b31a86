"""Oracle (test infrastructure only): the QBX refinement conditions of
pytential/qbx/refinement.py restated in numpy by brute force over all
(ball, point) pairs:

* condition 1, expansion disks undisturbed by sources (:155-211, driven at
  :363-430): the element of a stage-1 center is flagged when some stage-1
  node lies at distance <= (1 - tolerance) * the center's danger radius;
* condition 2, sufficient source quadrature resolution (:214-263, driven at
  :433-501): the element of a stage-2 source is flagged when some center lies
  at distance <= that element's danger radius.

Pinned in tests/test_oracle_refinement.py against scipy's KD-tree ball
query (an independent search) and closed-form cases (a straight line of
panels has no disturbed disk; a center placed on a node is flagged).
"""
import numpy as np


def _element_of(starts, idx):
    return np.searchsorted(np.asarray(starts), idx, side="right") - 1


def _any_within(balls, radii, pts, chunk=2048):
    balls = np.asarray(balls, np.float64)
    pts = np.asarray(pts, np.float64)
    hit = np.zeros(balls.shape[1], bool)
    for i in range(0, balls.shape[1], chunk):
        b = balls[:, i:i + chunk]
        d = np.sqrt(np.sum((b[:, :, None] - pts[:, None, :]) ** 2, axis=0))
        hit[i:i + chunk] = np.any(d <= radii[i:i + chunk, None], axis=1)
    return hit


def disturbed_expansion_flags(nodes, element_to_node_starts, centers, danger_radii,
                              tolerance):
    nel = len(element_to_node_starts) - 1
    r = (1.0 - tolerance) * np.asarray(danger_radii, np.float64)
    hit = _any_within(centers, r, nodes)
    flags = np.zeros(nel, np.int32)
    flags[_element_of(element_to_node_starts, np.nonzero(hit)[0] // 2)] = 1
    return flags


def source_quadrature_flags(sources, element_to_source_starts, danger_radii_by_element,
                            centers):
    nel = len(element_to_source_starts) - 1
    el = _element_of(element_to_source_starts, np.arange(np.asarray(sources).shape[1]))
    r = np.asarray(danger_radii_by_element, np.float64)[el]
    hit = _any_within(sources, r, centers)
    flags = np.zeros(nel, np.int32)
    flags[el[hit]] = 1
    return flags
