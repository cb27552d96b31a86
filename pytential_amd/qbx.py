"""QBXLayerPotentialSource for the MI355X path (mirror of
pytential/qbx/__init__.py:125-1072, FMM-backed evaluation only).

Constructor arguments, defaults and error behaviour follow the reference
(``__init__`` at :164-370); evaluation follows
``exec_compute_potential_insn_fmm`` (:694-793): geometry data is memoized per
target set, strengths are ``weights * upsampled density``, the FMM runs in
the reference's stage order, and results are sliced per target set.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import fmm as _fmm
from .discretization import CurveDiscretization
from .geometry import Context, QBXFMMGeometryData, as_dev


def fmm_level_to_order_for_tolerance(kernel, k, tol, root_extent, nlevels,
                                     max_order=100):
    """Per-level FMM orders reaching relative error ~``tol`` against direct QBX
    (the role of sumpy's SimpleExpansionOrderFinder).  Laplace: geometric
    convergence with ratio 0.3 for 2-away separation plus target stick-out;
    Helmholtz adds the excess-bandwidth term k*h*sqrt(2)*1.5 of the
    translated region."""
    base = int(math.ceil(math.log(tol / 0.15) / math.log(0.3)))
    orders = []
    for lev in range(nlevels):
        h = root_extent * 2.0 ** (-lev)
        p = base
        if kernel == "helmholtz":
            kd = k * h * math.sqrt(2) * 1.5
            p = int(math.ceil(kd + 1.8 * (-math.log10(tol)) ** (2 / 3) * max(kd, 1) ** (1 / 3)))
            p = max(p, base)
        orders.append(int(min(max(p, 1), max_order)))
    return orders


class _NotProvided:
    def __repr__(self):
        return "NOT_PROVIDED"


NOT_PROVIDED = _NotProvided()


class QBXLayerPotentialSource:
    """QBX layer potential source for a panel discretization of a curve (2D)
    or a triangulated surface (3D)."""

    def __init__(self, density_discr: CurveDiscretization, fine_order=None, qbx_order=None,
                 fmm_order=None, fmm_level_to_order=None,
                 target_association_tolerance=None, _max_leaf_refine_weight=None,
                 _expansion_stick_out_factor=0.5, _well_sep_is_n_away=2,
                 _from_sep_smaller_min_nsources_cumul=None, _box_extent_norm=None,
                 _expansions_in_tree_have_extent=True, _use_target_specific_qbx=False,
                 device=None):
        if fine_order is None:
            raise ValueError("'fine_order' must be provided.")
        if qbx_order is None:
            raise ValueError("'qbx_order' must be provided.")
        if (fmm_order is not None) == (fmm_level_to_order is not None):
            raise TypeError(
                "must specify exactly one of 'fmm_order' or 'fmm_level_to_order'.")
        # fmm_order=False: direct QBX, no FMM (qbx/__init__.py:866-1068); the
        # tree is still built for target association, its orders are unused
        self._direct = fmm_order is False or fmm_level_to_order is False
        if self._direct:
            fmm_order, fmm_level_to_order = None, (lambda kernel, kernel_args, tree, level: 1)
        if not _expansions_in_tree_have_extent:
            raise NotImplementedError("only trees with center extent are supported")
        if fmm_level_to_order is None:
            def fmm_level_to_order(kernel, kernel_args, tree, level):
                return fmm_order
        self.density_discr = density_discr
        self.fine_order = fine_order
        self.qbx_order = qbx_order
        self.fmm_level_to_order = fmm_level_to_order
        self.target_association_tolerance = (
            1.0e3 * float(np.finfo(np.float64).eps)
            if target_association_tolerance is None else target_association_tolerance)
        self.ambient_dim = getattr(density_discr, "ambient_dim", 2)
        # qbx/__init__.py:290-304: 64 in 2D, 512 in 3D
        self._max_leaf_refine_weight = (
            (64 if self.ambient_dim == 2 else 512) if _max_leaf_refine_weight is None
            else _max_leaf_refine_weight)
        self._expansion_stick_out_factor = _expansion_stick_out_factor
        self._well_sep_is_n_away = _well_sep_is_n_away
        # qbx/__init__.py:303-310: 15 for curves, 30 for surfaces
        self._from_sep_smaller_min_nsources_cumul = (
            (15 if self.ambient_dim == 2 else 30)
            if _from_sep_smaller_min_nsources_cumul is None
            else _from_sep_smaller_min_nsources_cumul)
        self._box_extent_norm = "l2" if _box_extent_norm is None else _box_extent_norm
        self._use_target_specific_qbx = bool(_use_target_specific_qbx)
        self._device = device
        self._ctx = None
        self._dev = None
        self._geo_cache = {}

    @property
    def ctx(self):
        """The library context of this source's device (created on first use,
        so constructing a source needs no GPU)."""
        if self._ctx is None:
            self._ctx = Context.get(self._device)
        return self._ctx

    @property
    def dev(self):
        """Device copies of the discretization arrays (made on first use)."""
        if self._dev is None:
            d = self.density_discr
            self._dev = {
                "nodes": as_dev(d.nodes), "normals": as_dev(d.normals),
                "tunnel": as_dev(d.tunnel_radii()), "centers": as_dev(d.centers()),
                "center_radii": as_dev(d.center_radii()),
                "element_to_source_starts": as_dev(d.element_to_source_starts(),
                                                   torch.int32),
                "quad_nodes": as_dev(d.quad_nodes), "quad_normals": as_dev(d.quad_normals),
                "quad_weights": as_dev(d.quad_weights), "upsample": as_dev(d.upsample),
            }
        return self._dev

    def copy(self, density_discr=None, fine_order=None, qbx_order=None,
             fmm_order=NOT_PROVIDED, fmm_level_to_order=NOT_PROVIDED,
             target_association_tolerance=NOT_PROVIDED,
             _expansions_in_tree_have_extent=NOT_PROVIDED,
             _expansion_stick_out_factor=NOT_PROVIDED, _max_leaf_refine_weight=None,
             _box_extent_norm=None, _use_target_specific_qbx=NOT_PROVIDED):
        """A new source with some parameters replaced
        (pytential/qbx/__init__.py:377-478): arguments left out keep this
        source's values; ``fmm_order`` and ``fmm_level_to_order`` are mutually
        exclusive (TypeError), and without either the level-to-order map is
        kept.  ``None`` keeps the value for the arguments whose ``None`` the
        reference reads as "not given"; ``False``/``0`` are valid values of
        the ``NOT_PROVIDED`` ones."""
        if fmm_order is not NOT_PROVIDED and fmm_level_to_order is not NOT_PROVIDED:
            raise TypeError("may not specify both 'fmm_order' and 'fmm_level_to_order'")
        kwargs = {}
        if fmm_order is not NOT_PROVIDED:
            kwargs["fmm_order"] = fmm_order
        elif fmm_level_to_order is not NOT_PROVIDED:
            kwargs["fmm_level_to_order"] = fmm_level_to_order
        elif self._direct:
            kwargs["fmm_order"] = False
        else:
            kwargs["fmm_level_to_order"] = self.fmm_level_to_order
        return type(self)(
            density_discr or self.density_discr,
            fine_order=fine_order if fine_order is not None else self.fine_order,
            qbx_order=qbx_order if qbx_order is not None else self.qbx_order,
            target_association_tolerance=(
                self.target_association_tolerance
                if target_association_tolerance is NOT_PROVIDED
                else target_association_tolerance),
            _expansions_in_tree_have_extent=(
                True if _expansions_in_tree_have_extent is NOT_PROVIDED
                else _expansions_in_tree_have_extent),
            _expansion_stick_out_factor=(
                self._expansion_stick_out_factor
                if _expansion_stick_out_factor is NOT_PROVIDED
                else _expansion_stick_out_factor),
            _well_sep_is_n_away=self._well_sep_is_n_away,
            _max_leaf_refine_weight=_max_leaf_refine_weight or self._max_leaf_refine_weight,
            _box_extent_norm=_box_extent_norm or self._box_extent_norm,
            _from_sep_smaller_min_nsources_cumul=self._from_sep_smaller_min_nsources_cumul,
            _use_target_specific_qbx=(
                self._use_target_specific_qbx
                if _use_target_specific_qbx is NOT_PROVIDED
                else _use_target_specific_qbx),
            device=self._device,
            **kwargs)

    def qbx_fmm_geometry_data(self, target_discrs_and_qbx_sides):
        key = tuple((id(t) if not isinstance(t, str) else t,
                     s if np.isscalar(s) else id(s)) for t, s in target_discrs_and_qbx_sides)
        if key not in self._geo_cache:
            self._geo_cache[key] = (QBXFMMGeometryData(self, target_discrs_and_qbx_sides),
                                    target_discrs_and_qbx_sides)
        return self._geo_cache[key][0]

    def level_orders(self, kernel, k, tree):
        args = {"k": k} if kernel == "helmholtz" else {}
        return [int(self.fmm_level_to_order(kernel, args, tree, lev))
                for lev in range(tree.nlevels)]

    def quad_density(self, density):
        """Stage-1 nodal density ([ndofs] or [ndofs, nrhs]) -> quad-stage-2
        nodal values (GPU)."""
        from ._lib import call
        from .geometry import _p, _stream
        d = self.density_discr
        dens = as_dev(density, torch.complex128) if torch.is_complex(
            torch.as_tensor(density)) else as_dev(density).to(torch.complex128)
        if dens.shape[0] != d.ndofs:
            raise ValueError(f"density has {dens.shape[0]} nodal values, expected {d.ndofs}")
        nrhs = 1 if dens.dim() == 1 else dens.shape[1]
        out = torch.empty((d.nsources,) + tuple(dens.shape[1:]), dtype=torch.complex128,
                          device="cuda")
        call("qbx_upsample", self.ctx.h, _stream(), d.nelements, d.nodes_per_element,
             d.fine_nodes_per_element, nrhs,
             _p(self.dev["upsample"]), _p(dens.contiguous()), _p(out))
        return out

    def make_wrangler(self, geo_data, kernel, k, source_kind, nrhs=1, private_context=False,
                      target_dirs=None):
        """The wrangler of one evaluation.  ``private_context`` gives it a
        library context of its own (its own workspaces and translation
        tables), as an operator captured into a HIP graph needs;
        ``target_dirs`` makes it evaluate target derivatives."""
        tree = geo_data.tree()
        ctx = Context(self.ctx.device) if private_context else None
        return _fmm.QBXExpansionWrangler(
            geo_data, kernel, k, source_kind, self.qbx_order, self.level_orders(kernel, k, tree),
            use_target_specific_qbx=self._use_target_specific_qbx, nrhs=nrhs, ctx=ctx,
            target_dirs=target_dirs)

    def exec_compute_potential(self, kernel, density, targets_and_sides, k=0.0,
                               source_kind="S", fused=True, target_derivative=False,
                               target_dirs=None):
        """Evaluate ``S`` or ``D`` of ``density`` (stage-1 nodal values,
        ``[ndofs]`` or ``[ndofs, nrhs]`` for several densities in one
        evaluation) at each ``(targets, qbx_forced_limit)`` pair; ``targets``
        is ``"self"`` (the stage-1 nodes) or a (dim, n) array.  Returns one
        tensor per pair ([n] or [n, nrhs]).  Target-association failure raises
        QBXTargetAssociationFailedError like the reference.

        ``target_derivative``: the derivative along a direction at each
        target instead (Sp from S, Dp from D): the node normals for "self",
        ``target_dirs[i]`` ((dim, n) per pair) otherwise (2D)."""
        if kernel not in ("laplace", "helmholtz"):
            raise ValueError(f"unknown kernel {kernel!r}")
        if source_kind not in ("S", "D"):
            raise ValueError(f"unknown source kind {source_kind!r}")
        sides = []
        for t, s in targets_and_sides:
            if s is None:
                s = 0
            if isinstance(s, str) and s == "avg":
                raise ValueError("'avg' must be split into +1 and -1 before evaluation")
            sides.append((t, s))
        geo = self.qbx_fmm_geometry_data(tuple(sides))
        dens = torch.as_tensor(density)
        if kernel == "laplace" and self.ambient_dim == 2 and torch.is_complex(dens) and \
                bool((dens.imag != 0).any()):
            # the 2D Laplace expansions take real strengths: evaluate the real
            # and imaginary parts as two right-hand sides (linearity)
            two = torch.stack([dens.real, dens.imag], dim=-1)
            outs = self.exec_compute_potential(kernel, two.reshape(dens.shape[0], -1),
                                               targets_and_sides, k, source_kind, fused,
                                               target_derivative, target_dirs)
            nr = dens.shape[1] if dens.dim() == 2 else 1
            res = []
            for o in outs:
                o = o.reshape(o.shape[0], nr, 2)
                c = torch.complex(o[..., 0], o[..., 1])
                res.append(c if dens.dim() == 2 else c[:, 0])
            return res
        nrhs = 1 if dens.dim() == 1 else dens.shape[1]
        tdirs = None
        if target_derivative:
            parts = []
            for i, (t, _) in enumerate(sides):
                if isinstance(t, str) and t == "self":
                    parts.append(self.dev["normals"])
                elif target_dirs is None or target_dirs[i] is None:
                    raise ValueError("target derivative at points targets needs target_dirs")
                else:
                    parts.append(as_dev(target_dirs[i]))
            tdirs = torch.cat(parts, dim=1)
        w = self.make_wrangler(geo, kernel, k, source_kind, nrhs, target_dirs=tdirs)
        qd = self.quad_density(density)
        normals = self.dev["quad_normals"] if source_kind == "D" else None
        strengths, nrm = w.reorder_sources(qd, self.dev["quad_weights"], normals)
        if self._direct:
            pot = _fmm.direct_qbx(w, strengths, nrm)
        elif fused and (not w.using_tsqbx or w.kd.target_specific_qbx):
            pot = _fmm.drive_fmm_fused(w, strengths, nrm)
        else:
            pot = _fmm.drive_fmm(w, strengths, nrm)
        starts = geo.target_info()["target_discr_starts"]
        nc = geo.ncenters
        outs = [pot[a - nc:b - nc] for a, b in zip(starts[:-1], starts[1:])]
        if kernel == "laplace":
            outs = [o.real for o in outs]
        return outs

    def exec_compute_potential_adjoint(self, kernel, pot_adj, targets_and_sides, k=0.0,
                                       source_kind="S"):
        """The conjugate transpose of ``exec_compute_potential`` for one
        density (QBX-FMM; the 2D kernels, and in 3D the Helmholtz single
        layer): from the adjoints of the potentials (one
        tensor per ``(targets, qbx_forced_limit)`` pair, as the forward
        returns them) the adjoint of the stage-1 nodal density, U^T W A^H
        (A the QBX-FMM of qbx_drive_fmm, W the quadrature weights, U the
        stage-1 -> quad-stage-2 interpolation).  Laplace: the transpose of
        the real forward map (pass real adjoints, take the real part)."""
        from ._lib import call
        from .geometry import _p, _stream
        from . import fmm as _fmm
        if self.ambient_dim != 2 and not (kernel == "helmholtz" and source_kind == "S"):
            raise NotImplementedError("the adjoint QBX-FMM is for the 2D kernels and the 3D "
                                      "Helmholtz single layer")
        if self._direct:
            raise NotImplementedError("the adjoint is that of the QBX-FMM (fmm_order given)")
        sides = [(t, 0 if sd is None else sd) for t, sd in targets_and_sides]
        geo = self.qbx_fmm_geometry_data(tuple(sides))
        w = self.make_wrangler(geo, kernel, k, source_kind, 1)
        d = self.density_discr
        ga = torch.cat([as_dev(g, torch.complex128).reshape(-1) for g in pot_adj])
        nuser = w.tree.ntargets - geo.ncenters
        if ga.numel() != nuser:
            raise ValueError(f"adjoint has {ga.numel()} values, the targets {nuser}")
        nrm = None
        if source_kind == "D":
            zero = torch.zeros(d.nsources, dtype=torch.complex128, device="cuda")
            _, nrm = w.reorder_sources(zero, self.dev["quad_weights"], self.dev["quad_normals"])
        sa = _fmm.drive_fmm_adjoint(w, ga, nrm)
        # tree -> user source order, the quadrature weights, the interpolation
        uid = w.tree.user_source_ids.long()
        qa = torch.empty_like(sa)
        qa[uid] = sa
        qa = (qa * self.dev["quad_weights"]).contiguous()
        out = torch.empty(d.ndofs, dtype=torch.complex128, device="cuda")
        call("qbx_upsample_adjoint", self.ctx.h, _stream(), d.nelements,
             d.nodes_per_element, d.fine_nodes_per_element, 1, _p(self.dev["upsample"]),
             _p(qa), _p(out))
        return out
