"""QBX refinement on the GPU: the checkers and the stage-1 refinement loop of
pytential/qbx/refinement.py.

* ``check_expansion_disks_undisturbed_by_sources`` -- condition 1
  (refinement.py:155-211, 363-430) through
  ``qbx_check_expansion_disks_undisturbed``;
* ``check_sufficient_source_quadrature_resolution`` -- condition 2
  (:214-263, 433-501) through ``qbx_check_source_quadrature_resolution``;
* ``check_element_prop_threshold`` -- the kernel-length-scale / curvature
  threshold checker (:336-360, 503-529);
* ``refine_stage1`` -- ``_refine_qbx_stage1`` (:667-792) for curves: per
  iteration the element-size check against ``kernel_length_scale`` and, when
  nothing else is violated, condition 1; flagged panels are split in two,
  up to ``maxiter`` iterations (``RefinerNotConvergedWarning`` beyond).

The expansion disturbance tolerance defaults to 0.025 and ``maxiter`` to 10,
as in refine_geometry_collection (:926-936).
"""
from __future__ import annotations

import warnings

import torch

from ._lib import call
from .geometry import Context, _p, _stream, as_dev


class RefinerNotConvergedWarning(UserWarning):
    """Mirror of pytential.qbx.refinement.RefinerNotConvergedWarning."""


def _flags_and_found(nel):
    return (torch.zeros(max(nel, 1), dtype=torch.int32, device="cuda")[:nel],
            torch.zeros(1, dtype=torch.int32, device="cuda"))


def check_expansion_disks_undisturbed_by_sources(discr, expansion_disturbance_tolerance=0.025,
                                                 refine_flags=None, ctx=None):
    """Condition 1 on a stage-1 discretization: flags (int32 [nelements],
    or ``refine_flags`` updated in place) and whether any element was
    flagged."""
    ctx = ctx or Context.get()
    nodes = as_dev(discr.nodes)
    centers = as_dev(discr.centers())
    radii = as_dev(discr.center_radii())
    starts = as_dev(discr.element_to_source_starts(), torch.int32)
    nel = discr.nelements
    flags, found = _flags_and_found(nel)
    if refine_flags is not None:
        flags = refine_flags
    call("qbx_check_expansion_disks_undisturbed", ctx.h, _stream(), nodes.shape[0], nel,
         _p(starts), nodes.shape[1], _p(nodes), _p(centers), _p(radii),
         float(expansion_disturbance_tolerance), _p(flags), _p(found))
    return flags, bool(found.item())


def check_sufficient_source_quadrature_resolution(stage2_discr, centers, refine_flags=None,
                                                  ctx=None):
    """Condition 2: the stage-2 discretization's nodes against the stage-1
    ``centers`` (dim, n); the danger radius of an element is 0.75 times the
    largest stage-2 expansion radius on it (primitives.py:1322-1340)."""
    ctx = ctx or Context.get()
    src = as_dev(stage2_discr.nodes)
    nel = stage2_discr.nelements
    starts = as_dev(stage2_discr.element_to_source_starts(), torch.int32)
    rad = as_dev(stage2_discr.expansion_radii)
    n1 = src.shape[1] // max(nel, 1)
    danger = 0.75 * rad.reshape(nel, n1).max(dim=1).values.contiguous()
    ctr = as_dev(centers)
    flags, found = _flags_and_found(nel)
    if refine_flags is not None:
        flags = refine_flags
    call("qbx_check_source_quadrature_resolution", ctx.h, _stream(), src.shape[0], nel,
         _p(starts), src.shape[1], _p(src), _p(danger), ctr.shape[1], _p(ctr), _p(flags),
         _p(found))
    return flags, bool(found.item())


def check_element_prop_threshold(element_property, threshold, refine_flags):
    """Flag the elements whose property exceeds ``threshold``; True when
    any did."""
    over = as_dev(element_property) > threshold
    refine_flags[over] = 1
    return bool(over.any())


def quad_resolution_by_element(discr):
    """Element-wise max of the quadrature resolution (the mapping stretch
    factor): the expansion radius over its factor (0.5 on curves, 0.25 on
    surfaces, primitives.py:1301-1306)."""
    factor = 0.5 if discr.ambient_dim == 2 else 0.25
    r = as_dev(discr.expansion_radii) / factor
    return r.reshape(discr.nelements, -1).max(dim=1).values


def refine_stage1(discr, kernel_length_scale=None, expansion_disturbance_tolerance=0.025,
                  maxiter=10, ctx=None):
    """Split flagged panels of a curve discretization until the element-size
    and undisturbed-expansion conditions hold.  Returns ``(discr,
    connections, violated)``: the refined discretization, the per-iteration
    ``PanelConnection``s (stage-1 data from the input to the output) and the
    criteria violated in each iteration that refined."""
    if not hasattr(discr, "refined"):
        raise NotImplementedError("stage-1 refinement splits curve panels only")
    connections, violated = [], []
    for _ in range(maxiter):
        it = []
        flags = torch.zeros(discr.nelements, dtype=torch.int32, device="cuda")
        if kernel_length_scale is not None:
            if check_element_prop_threshold(quad_resolution_by_element(discr),
                                            kernel_length_scale, flags):
                it.append("kernel length scale")
        if not it:
            _, found = check_expansion_disks_undisturbed_by_sources(
                discr, expansion_disturbance_tolerance, refine_flags=flags, ctx=ctx)
            if found:
                it.append("disturbed expansions")
        if not it:
            return discr, connections, violated
        violated.append(" and ".join(it))
        discr, conn = discr.refined(flags.cpu().numpy())
        connections.append(conn)
    warnings.warn("QBX layer potential source refiner did not terminate after "
                  f"{maxiter} iterations (violated: {', '.join(violated)})",
                  RefinerNotConvergedWarning, stacklevel=2)
    return discr, connections, violated
