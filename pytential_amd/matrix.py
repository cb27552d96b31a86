"""Dense layer-potential matrices on the GPU.

Mirror of pytential's matrix builders for one layer potential of a density on
a ``QBXLayerPotentialSource`` (pytential/symbolic/matrix.py):

* ``build_matrix(source, kernel, k, source_kind, qbx_forced_limit=+1)`` --
  ``MatrixBuilder.map_int_g`` (:451-522): targets are the stage-1 nodes, each
  expanded about its own center on the side ``qbx_forced_limit`` (the
  ``expansion_centers`` / ``expansion_radii`` of the target, :470-477), the
  sources are the quad-stage-2 nodes with weights and area elements
  (:505-510), and the density enters through the stage-1 -> quad-stage-2
  resampling (``map_interpolation``, :401-449).  The result maps stage-1
  nodal densities to potentials at the stage-1 nodes.
* ``qbx_forced_limit=None`` with point ``targets`` -- the same resampled,
  weighted quadrature with the plain kernel (no QBX): the operator the
  QBX-FMM applies at targets far from the surface.
* ``p2p_matrix(source, kernel, k, source_kind, targets, weighted=False)`` --
  ``P2PMatrixBuilder.map_int_g`` (:553-620): the plain kernel between the
  stage-1 nodes and ``targets`` (default: the stage-1 nodes), coincident
  points excluded, weights only on request.

All three run one kernel (``qbx_lpot_matrix``) that fuses the per-entry
expansion (or kernel) with the resampling contraction, so the quad-stage-2
intermediate the reference forms on the host never exists.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import _lib
from ._lib import call
from .geometry import _p, _stream, as_dev


def _kernel_desc(dim, kernel, k, source_kind, qbx_order):
    code = {(2, "laplace"): _lib.LAPLACE2D, (2, "helmholtz"): _lib.HELMHOLTZ2D,
            (3, "laplace"): _lib.LAPLACE3D, (3, "helmholtz"): _lib.HELMHOLTZ3D}
    if (dim, kernel) not in code:
        raise ValueError(f"unknown kernel {kernel!r} in {dim}D")
    if source_kind not in ("S", "D"):
        raise ValueError(f"unknown source kind {source_kind!r}")
    lo = np.ones(1, dtype=np.int32)
    kd = _lib.KernelDesc(code[(dim, kernel)], float(np.real(k)),
                         _lib.SOURCE_S if source_kind == "S" else _lib.SOURCE_D, qbx_order,
                         1, lo.ctypes.data_as(ct.POINTER(ct.c_int)), 1)
    kd._lo = lo
    return kd


def _run(source, kd, nel, n1, nf, upsample, nodes, normals, weights, targets, centers, radii):
    nt = targets.shape[1]
    ncols = nel * n1
    mat = torch.empty((nt, ncols), dtype=torch.complex128, device="cuda")
    call("qbx_lpot_matrix", source.ctx.h, _stream(), ct.byref(kd), nel, n1, nf, _p(upsample),
         _p(nodes), _p(normals), _p(weights), nt, _p(targets), _p(centers), _p(radii),
         ncols, _p(mat))
    return mat


def build_matrix(source, kernel, k=0.0, source_kind="S", qbx_forced_limit=+1,
                 targets="self"):
    """Dense matrix of ``S`` or ``D`` on ``source`` (a QBXLayerPotentialSource):
    columns are the stage-1 nodal values of the density.  With
    ``qbx_forced_limit`` +1 / -1 the rows are the stage-1 nodes evaluated
    through QBX about their own centers on that side; with ``None`` the rows
    are ``targets`` (a (dim, n) array) evaluated with the plain quadrature.
    Laplace matrices are real-valued (returned complex)."""
    d = source.density_discr
    dev = source.dev
    kd = _kernel_desc(source.ambient_dim, kernel, k, source_kind, source.qbx_order)
    normals = dev["quad_normals"] if source_kind == "D" else None
    if qbx_forced_limit is None:
        if isinstance(targets, str):
            raise ValueError("plain quadrature at the stage-1 nodes is singular: give "
                             "qbx_forced_limit or point targets")
        tg = as_dev(targets)
        centers = radii = None
    else:
        if qbx_forced_limit not in (-1, +1):
            raise ValueError("qbx_forced_limit must be +1, -1 or None")
        if not (isinstance(targets, str) and targets == "self"):
            raise ValueError("the QBX matrix has the stage-1 nodes as targets")
        tg = dev["nodes"]
        off = 1 if qbx_forced_limit > 0 else 0
        centers = dev["centers"][:, off::2].contiguous()
        radii = dev["center_radii"][off::2].contiguous()
    return _run(source, kd, d.nelements, d.nodes_per_element, d.fine_nodes_per_element,
                dev["upsample"], dev["quad_nodes"], normals, dev["quad_weights"], tg,
                centers, radii)


def associated_qbx_matrix(source, geo, kernel, k=0.0, source_kind="S"):
    """Dense QBX matrix whose rows are the user targets of ``geo`` (a
    QBXFMMGeometryData of ``source``) that the target association gave a
    center, each through the QBX expansion about that center -- the operator
    the QBX-FMM approximates at those targets (pytential/qbx/fmm.py:549-557
    evaluates each target's expansion about its associated center).  Returns
    (row indices into the user targets, matrix)."""
    d = source.density_discr
    dev = source.dev
    nc = geo.ncenters
    ttc = geo.user_target_to_center()[nc:].long()
    rows = torch.nonzero(ttc >= 0).reshape(-1)
    c = ttc[rows]
    tg = geo.target_info()["targets"][:, nc:][:, rows].contiguous()
    centers = geo.flat_centers()[:, c].contiguous()
    radii = geo.flat_expansion_radii()[c].contiguous()
    kd = _kernel_desc(source.ambient_dim, kernel, k, source_kind, source.qbx_order)
    normals = dev["quad_normals"] if source_kind == "D" else None
    return rows, _run(source, kd, d.nelements, d.nodes_per_element, d.fine_nodes_per_element,
                      dev["upsample"], dev["quad_nodes"], normals, dev["quad_weights"], tg,
                      centers, radii)


def p2p_matrix(source, kernel, k=0.0, source_kind="S", targets="self", weighted=False):
    """Plain kernel matrix between the stage-1 nodes (columns) and ``targets``
    (rows; default the stage-1 nodes, whose coincident pairs are zero), times
    the stage-1 weights and area elements when ``weighted``."""
    d = source.density_discr
    dev = source.dev
    kd = _kernel_desc(source.ambient_dim, kernel, k, source_kind, 0)
    tg = dev["nodes"] if isinstance(targets, str) and targets == "self" else as_dev(targets)
    w = as_dev(d.weights) if weighted else torch.ones(d.ndofs, dtype=torch.float64,
                                                      device="cuda")
    normals = dev["normals"] if source_kind == "D" else None
    n1 = d.nodes_per_element
    return _run(source, kd, d.nelements, n1, n1, None, dev["nodes"], normals, w, tg, None,
                None)
