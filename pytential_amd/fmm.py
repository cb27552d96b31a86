"""QBX expansion wrangler and FMM driver (mirror of pytential/qbx/fmm.py).

Each wrangler method is one C-ABI call into libqbx_mi355x; ``drive_fmm``
calls them in the reference's stage order (pytential/qbx/fmm.py:414-594).
``drive_fmm_fused`` runs the same sequence from a single library entry point.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np
import torch

from . import _lib
from ._lib import KernelDesc, call
from .geometry import _p, _stream


class QBXExpansionWrangler:
    """Stages of the QBX FMM on one tree (pytential/qbx/fmm.py:118-396)."""

    def __init__(self, geo_data, kernel, k, source_kind, qbx_order, level_orders):
        self.geo_data = geo_data
        self.tree = geo_data.tree()
        self.traversal = geo_data.traversal()
        self.geo = geo_data._geo
        self.ctx = geo_data.ctx
        self.kernel = kernel
        lo = np.ascontiguousarray(level_orders, dtype=np.int32)
        if len(lo) < self.tree.nlevels:
            raise ValueError(f"need {self.tree.nlevels} level orders, got {len(lo)}")
        self._lo = lo
        self.kd = KernelDesc(_lib.LAPLACE2D if kernel == "laplace" else _lib.HELMHOLTZ2D,
                             float(k), _lib.SOURCE_S if source_kind == "S" else _lib.SOURCE_D,
                             qbx_order, len(lo), lo.ctypes.data_as(ct.POINTER(ct.c_int)))
        self.source_kind = source_kind
        self.qbx_order = qbx_order
        self.ncoeff_fmm = _lib.lib().qbx_ncoeff_fmm(ct.byref(self.kd))
        self.ncoeff_qbx = _lib.lib().qbx_ncoeff_qbx(ct.byref(self.kd))

    # ---- data vector utilities
    def multipole_expansion_zeros(self):
        return torch.zeros((self.tree.nboxes, self.ncoeff_fmm), dtype=torch.complex128,
                           device="cuda")

    local_expansion_zeros = multipole_expansion_zeros

    def qbx_local_expansion_zeros(self):
        return torch.zeros((self.geo_data.ncenters, self.ncoeff_qbx), dtype=torch.complex128,
                           device="cuda")

    def full_output_zeros(self):
        return torch.zeros(self.tree.ntargets, dtype=torch.complex128, device="cuda")

    output_zeros = full_output_zeros

    def _a(self, name, *args):
        call(name, self.ctx.h, _stream(), *args)

    # ---- sources
    def reorder_sources(self, quad_density, weights, normals):
        """strengths = weights * density in tree order (+ normals for D)."""
        dens = quad_density.to(torch.complex128).contiguous()
        out = torch.empty(self.tree.nsources, dtype=torch.complex128, device="cuda")
        nout = None
        if normals is not None:
            nout = torch.empty_like(normals)
        self._a("qbx_reorder_sources", self.tree.h, _p(dens), _p(weights), _p(normals),
                _p(out), _p(nout))
        return out, nout

    # ---- point FMM stages
    def form_multipoles(self, strengths, normals):
        m = self.multipole_expansion_zeros()
        self._a("qbx_form_multipoles", self.tree.h, self.traversal.h, ct.byref(self.kd),
                _p(strengths), _p(normals), _p(m))
        return m

    def coarsen_multipoles(self, mpole):
        self._a("qbx_coarsen_multipoles", self.tree.h, self.traversal.h, ct.byref(self.kd),
                _p(mpole))
        return mpole

    def eval_direct(self, strengths, normals, pot=None):
        acc = pot is not None
        pot = pot if acc else self.full_output_zeros()
        self._a("qbx_eval_direct", self.tree.h, self.traversal.h, self.geo.h,
                ct.byref(self.kd), _p(strengths), _p(normals), _p(pot), int(acc))
        return pot

    def multipole_to_local(self, mpole):
        loc = self.local_expansion_zeros()
        self._a("qbx_multipole_to_local", self.tree.h, self.traversal.h, ct.byref(self.kd),
                _p(mpole), _p(loc))
        return loc

    def eval_multipoles(self, mpole, pot=None):
        acc = pot is not None
        pot = pot if acc else self.full_output_zeros()
        self._a("qbx_eval_multipoles", self.tree.h, self.traversal.h, self.geo.h,
                ct.byref(self.kd), _p(mpole), _p(pot), int(acc))
        return pot

    def form_locals(self, strengths, normals, loc=None):
        acc = loc is not None
        loc = loc if acc else self.local_expansion_zeros()
        self._a("qbx_form_locals", self.tree.h, self.traversal.h, ct.byref(self.kd),
                _p(strengths), _p(normals), _p(loc), int(acc))
        return loc

    def refine_locals(self, loc):
        self._a("qbx_refine_locals", self.tree.h, self.traversal.h, ct.byref(self.kd), _p(loc))
        return loc

    def eval_locals(self, loc, pot=None):
        acc = pot is not None
        pot = pot if acc else self.full_output_zeros()
        self._a("qbx_eval_locals", self.tree.h, self.traversal.h, self.geo.h,
                ct.byref(self.kd), _p(loc), _p(pot), int(acc))
        return pot

    # ---- QBX stages
    def form_global_qbx_locals(self, strengths, normals):
        q = self.qbx_local_expansion_zeros()
        self._a("qbx_form_global_qbx_locals", self.tree.h, self.traversal.h, self.geo.h,
                ct.byref(self.kd), _p(strengths), _p(normals), _p(q), 0)
        return q

    def translate_box_multipoles_to_qbx_local(self, mpole, qexp=None):
        acc = qexp is not None
        qexp = qexp if acc else self.qbx_local_expansion_zeros()
        self._a("qbx_translate_box_multipoles_to_qbx_local", self.tree.h, self.traversal.h,
                self.geo.h, ct.byref(self.kd), _p(mpole), _p(qexp), int(acc))
        return qexp

    def translate_box_local_to_qbx_local(self, loc, qexp=None):
        acc = qexp is not None
        qexp = qexp if acc else self.qbx_local_expansion_zeros()
        self._a("qbx_translate_box_local_to_qbx_local", self.tree.h, self.traversal.h,
                self.geo.h, ct.byref(self.kd), _p(loc), _p(qexp), int(acc))
        return qexp

    def eval_qbx_expansions(self, qexp, pot=None):
        pot = pot if pot is not None else self.full_output_zeros()
        self._a("qbx_eval_qbx_expansions", self.tree.h, self.geo.h, ct.byref(self.kd),
                _p(qexp), _p(pot))
        return pot

    def eval_target_specific_qbx_locals(self, strengths):
        # TSQBX is 3D-only in the reference (qbx/fmmlib.py:141-148); zero here
        return self.full_output_zeros()

    def finalize_potentials(self, pot_tree):
        nuser = self.tree.ntargets - self.geo_data.ncenters
        out = torch.empty(nuser, dtype=torch.complex128, device="cuda")
        self._a("qbx_finalize_potentials", self.tree.h, self.geo.h, _p(pot_tree), _p(out))
        return out


def drive_fmm(wrangler: QBXExpansionWrangler, strengths, normals):
    """Stage-by-stage QBX FMM (pytential/qbx/fmm.py:414-594).  Returns the
    potential at the user targets, in user order."""
    w = wrangler
    mpole = w.form_multipoles(strengths, normals)
    mpole = w.coarsen_multipoles(mpole)
    pot = w.eval_direct(strengths, normals)
    local = w.multipole_to_local(mpole)
    w.eval_multipoles(mpole, pot)
    w.form_locals(strengths, normals, local)
    w.refine_locals(local)
    w.eval_locals(local, pot)
    qexp = w.form_global_qbx_locals(strengths, normals)
    w.translate_box_multipoles_to_qbx_local(mpole, qexp)
    w.translate_box_local_to_qbx_local(local, qexp)
    w.eval_qbx_expansions(qexp, pot)
    return w.finalize_potentials(pot)


def drive_fmm_fused(wrangler: QBXExpansionWrangler, strengths, normals, out=None):
    """Same stage sequence from the single ``qbx_drive_fmm`` entry point."""
    w = wrangler
    nuser = w.tree.ntargets - w.geo_data.ncenters
    if out is None:
        out = torch.empty(nuser, dtype=torch.complex128, device="cuda")
    call("qbx_drive_fmm", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
         ct.byref(w.kd), _p(strengths), _p(normals), _p(out))
    return out
