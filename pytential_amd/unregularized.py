"""Unregularized layer potentials: element-based Nystrom quadrature with the
plain (unmodified) kernel, evaluated directly or by a point FMM on the GPU.

Mirror of pytential's ``UnregularizedLayerPotentialSource``
(pytential/unregularized.py:66-322): the sources are the stage-1 nodes of the
density discretization with strengths ``weights_and_area_elements * density``
(:170-172, :236-240), and the targets are plain points (no QBX, so
``qbx_forced_limit`` must be ``None``, :226).

* ``fmm_order=False`` (the default, :91-105): every target sums every source,
  sumpy's P2P (:154-198) -> ``qbx_p2p_all``.
* otherwise the point FMM (:214-289): a tree over [stage-1 nodes, targets]
  with source refine weight 1, target weight 0, at most 32 per leaf
  (:407-429), a traversal without target extents (:387-394), and boxtree's
  ``drive_fmm`` with the sumpy wrangler -> the same HIP stages as the QBX-FMM
  with no QBX centers (``qbx_drive_fmm``: P2M, M2M, M2L, L2L, P2P, M2P, P2L,
  L2P; the QBX stages find no centers and do nothing).

A target that coincides with a source gets no contribution from it.
"""
from __future__ import annotations

import ctypes as ct
from functools import cached_property

import numpy as np
import torch

from . import _lib
from . import fmm as _fmm
from ._lib import call
from .geometry import Context, Traversal, Tree, _GeoHandle, _p, _stream, as_dev

MAX_LEAF_REFINE_WEIGHT = 32   # pytential/unregularized.py:407


class UnregularizedLayerPotentialSource:
    """pytential/unregularized.py:66-130 (constructor arguments and their
    validation, ``copy``)."""

    def __init__(self, density_discr, fmm_order=False, fmm_level_to_order=None,
                 expansion_factory=None, debug=True, device=None):
        if fmm_order is not False and fmm_level_to_order is not None:
            raise TypeError("may not specify both fmm_order and fmm_level_to_order")
        if fmm_level_to_order is None:
            if fmm_order is not False:
                order = int(fmm_order)

                def fmm_level_to_order(kernel, kernel_args, tree, level):
                    return order
            else:
                fmm_level_to_order = False
        if expansion_factory is not None:
            raise NotImplementedError("only the default expansions exist here")
        self.density_discr = density_discr
        self.fmm_level_to_order = fmm_level_to_order
        self.debug = debug
        self.ctx = Context.get(device)
        d = density_discr
        self.dev = {"nodes": as_dev(d.nodes), "normals": as_dev(d.normals),
                    "weights": as_dev(d.weights)}
        self._geo_cache = {}

    def lpot_weights(self):
        """weights_and_area_elements at the stage-1 nodes (the sources)."""
        return self.dev["weights"]

    @property
    def ambient_dim(self):
        return self.density_discr.ambient_dim

    def copy(self, density_discr=None, fmm_level_to_order=None, debug=None):
        return type(self)(density_discr or self.density_discr,
                          fmm_level_to_order=fmm_level_to_order or self.fmm_level_to_order,
                          debug=self.debug if debug is None else debug)

    # ---- geometry (pytential/unregularized.py:376-460)
    def fmm_geometry_data(self, targets):
        key = tuple(t if isinstance(t, str) else id(t) for t in targets)
        if key not in self._geo_cache:
            self._geo_cache[key] = (_FMMGeometryData(self, targets), targets)
        return self._geo_cache[key][0]

    def _target_points(self, t):
        return self.dev["nodes"] if isinstance(t, str) and t == "self" else as_dev(t)

    def _density(self, density):
        """Stage-1 nodal density on the device, complex, [ndofs] or [ndofs, nrhs]."""
        d = torch.as_tensor(density) if not isinstance(density, torch.Tensor) else density
        d = d.to(device="cuda", dtype=torch.complex128)
        if d.shape[0] != self.density_discr.ndofs:
            raise ValueError(f"density has {d.shape[0]} nodal values, expected "
                             f"{self.density_discr.ndofs}")
        return d.contiguous()

    def _kernel_desc(self, kernel, k, source_kind, nrhs, level_orders):
        code = {(2, "laplace"): _lib.LAPLACE2D, (2, "helmholtz"): _lib.HELMHOLTZ2D,
                (3, "laplace"): _lib.LAPLACE3D, (3, "helmholtz"): _lib.HELMHOLTZ3D}
        if (self.ambient_dim, kernel) not in code:
            raise ValueError(f"unknown kernel {kernel!r}")
        if source_kind not in ("S", "D"):
            raise ValueError(f"unknown source kind {source_kind!r}")
        lo = np.ascontiguousarray(level_orders, dtype=np.int32)
        kd = _lib.KernelDesc(code[(self.ambient_dim, kernel)], float(np.real(k)),
                             _lib.SOURCE_S if source_kind == "S" else _lib.SOURCE_D, 0,
                             len(lo), lo.ctypes.data_as(ct.POINTER(ct.c_int)), nrhs)
        kd._lo = lo
        return kd

    def exec_compute_potential(self, kernel, density, targets, k=0.0, source_kind="S"):
        """``S`` or ``D`` of ``density`` (stage-1 nodal values, ``[ndofs]`` or
        ``[ndofs, nrhs]``) at each entry of ``targets`` (``"self"`` or a (dim, n)
        array); one complex tensor per entry (real for Laplace)."""
        dens = self._density(density)
        if kernel == "laplace" and self.ambient_dim == 2 and bool((dens.imag != 0).any()):
            # the 2D Laplace expansions take real strengths: evaluate the real
            # and imaginary parts as two right-hand sides (linearity)
            two = torch.stack([dens.real, dens.imag], dim=-1).reshape(dens.shape[0], -1)
            outs = self.exec_compute_potential(kernel, two, targets, k, source_kind)
            res = []
            for o in outs:
                o = o.reshape(o.shape[0], -1, 2)
                c = torch.complex(o[..., 0], o[..., 1])
                res.append(c if dens.dim() == 2 else c[:, 0])
            return res
        nrhs = 1 if dens.dim() == 1 else dens.shape[1]
        normals = self.dev["normals"] if source_kind == "D" else None
        if self.fmm_level_to_order is False:
            w = self.dev["weights"]
            strengths = (dens * (w if dens.dim() == 1 else w[:, None])).contiguous()
            outs = [self._direct(kernel, k, source_kind, strengths, normals,
                                 self._target_points(t)) for t in targets]
        else:
            outs = self._fmm(kernel, k, source_kind, dens, normals, targets, nrhs)
        if kernel == "laplace":
            outs = [o.real for o in outs]
        return outs

    def _direct(self, kernel, k, source_kind, strengths, normals, tgt):
        """pytential/unregularized.py:154-198."""
        nrhs = 1 if strengths.dim() == 1 else strengths.shape[1]
        if self.ambient_dim == 3 and nrhs != 1:
            cols = [self._direct(kernel, k, source_kind, strengths[:, r].contiguous(), normals,
                                 tgt) for r in range(nrhs)]
            return torch.stack(cols, dim=1)
        kd = self._kernel_desc(kernel, k, source_kind, nrhs, [1])
        nt = tgt.shape[1]
        pot = torch.empty((nt,) if nrhs == 1 else (nt, nrhs), dtype=torch.complex128,
                          device="cuda")
        d = self.density_discr
        call("qbx_p2p_all", self.ctx.h, _stream(), ct.byref(kd), d.ndofs, _p(self.dev["nodes"]),
             _p(normals), _p(strengths), nt, _p(tgt), _p(pot))
        return pot

    def _fmm(self, kernel, k, source_kind, dens, normals, targets, nrhs):
        """pytential/unregularized.py:214-289."""
        if self.ambient_dim == 3 and kernel == "helmholtz":
            raise NotImplementedError("no 3D Helmholtz FMM expansions: use fmm_order=False")
        geo = self.fmm_geometry_data(tuple(targets))
        tree = geo.tree()
        args = {"k": k} if kernel == "helmholtz" else {}
        lo = [int(self.fmm_level_to_order(kernel, args, tree, lev))
              for lev in range(tree.nlevels)]
        w = _fmm.QBXExpansionWrangler(geo, kernel, k, source_kind, 0, lo, nrhs=nrhs)
        st, nrm = w.reorder_sources(dens, self.lpot_weights(), normals)
        pot = _fmm.drive_fmm_fused(w, st, nrm)
        starts = geo.target_info()["target_discr_starts"]
        return [pot[a:b] for a, b in zip(starts[:-1], starts[1:])]


class _FMMGeometryData:
    """Tree and traversal of the point FMM (pytential/unregularized.py:376-460),
    with the interface the wrangler reads (an empty set of QBX centers)."""

    ncenters = 0

    def __init__(self, lpot_source, targets):
        self.lpot_source = lpot_source
        self.target_discrs = targets
        self.ctx = lpot_source.ctx

    @cached_property
    def _target_info(self):
        pts, starts, n = [], [], 0
        for t in self.target_discrs:
            p = self.lpot_source._target_points(t)
            starts.append(n)
            n += p.shape[1]
            pts.append(p)
        starts.append(n)
        tg = torch.cat(pts, dim=1).contiguous() if pts else torch.zeros(
            (self.lpot_source.ambient_dim, 0), dtype=torch.float64, device="cuda")
        return {"targets": tg, "target_discr_starts": starts, "ntargets": n}

    def target_info(self):
        return self._target_info

    @cached_property
    def _tree(self):
        ti = self._target_info
        tw = torch.zeros(ti["ntargets"], dtype=torch.int32, device="cuda")
        return Tree(self.lpot_source.dev["nodes"], ti["targets"], None, None, tw,
                    MAX_LEAF_REFINE_WEIGHT, 0.0, "l2", ctx=self.ctx)

    def tree(self):
        return self._tree

    @cached_property
    def _traversal(self):
        return Traversal(self._tree, 2, False)

    def traversal(self):
        return self._traversal

    @cached_property
    def _geo(self):
        nt = self._target_info["ntargets"]
        uttc = torch.full((nt,), -1, dtype=torch.int32, device="cuda")
        h = ct.c_void_p()
        call("qbx_geo_build", self.ctx.h, _stream(), self._tree.h, self._traversal.h, 0,
             _p(None), _p(None), _p(uttc), ct.byref(h))
        g = _GeoHandle(h)
        g._keep = uttc
        return g

    def non_qbx_box_target_lists(self):
        g = self._geo
        return {"box_target_starts": g.field("nqbtl_box_target_starts"),
                "box_target_counts_nonchild": g.field("nqbtl_box_target_counts_nonchild"),
                "unfiltered_from_filtered_target_indices":
                    g.field("nqbtl_unfiltered_from_filtered_target_indices")}
