"""QBX expansion wrangler and FMM driver (mirror of pytential/qbx/fmm.py).

Each wrangler method is one C-ABI call into libqbx_mi355x; ``drive_fmm``
calls them in the reference's stage order (pytential/qbx/fmm.py:414-594).
``drive_fmm_fused`` runs the same sequence from a single library entry point.

A wrangler built with ``nrhs > 1`` evaluates that many densities at once:
strengths are ``[nsources, nrhs]``, potentials ``[ntargets, nrhs]`` and every
expansion array carries the right-hand sides innermost.  With ``nrhs == 1``
arrays keep the reference's one-density shapes.
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

    def __init__(self, geo_data, kernel, k, source_kind, qbx_order, level_orders,
                 output="pot", use_target_specific_qbx=False, nrhs=1, ctx=None,
                 target_dirs=None):
        """``target_dirs`` ([dim, n_user_targets], device): evaluate the
        derivative along these directions at the user targets instead of the
        potential (sumpy's DirectionalTargetDerivative; Sp and Dp take the
        target normals), 2D only; the expansions are unchanged."""
        self.geo_data = geo_data
        self.tree = geo_data.tree()
        self.traversal = geo_data.traversal()
        self.geo = geo_data._geo
        # the context whose workspaces and tables the stages use (the
        # geometry's shared one unless the caller gives the wrangler its own)
        self.ctx = geo_data.ctx if ctx is None else ctx
        self.kernel = kernel
        self.dim = self.tree.dimensions
        self.k = complex(k)
        if output not in ("pot", "grad"):
            raise ValueError(f"unknown output {output!r}")
        if output == "grad" and source_kind == "D":
            # impl.pyx:637-638
            raise ValueError("Does not support computing gradient of dipole sources")
        self.output = output
        lo = np.ascontiguousarray(level_orders, dtype=np.int32)
        if len(lo) < self.tree.nlevels:
            raise ValueError(f"need {self.tree.nlevels} level orders, got {len(lo)}")
        self._lo = lo
        code = {(2, "laplace"): _lib.LAPLACE2D, (2, "helmholtz"): _lib.HELMHOLTZ2D,
                (3, "laplace"): _lib.LAPLACE3D, (3, "helmholtz"): _lib.HELMHOLTZ3D}
        self.kernel_code = code[(self.dim, kernel)]
        if nrhs < 1:
            raise ValueError("nrhs must be >= 1")
        self.nrhs = int(nrhs)
        self.kd = KernelDesc(self.kernel_code, self.k.real,
                             _lib.SOURCE_S if source_kind == "S" else _lib.SOURCE_D,
                             qbx_order, len(lo), lo.ctypes.data_as(ct.POINTER(ct.c_int)),
                             self.nrhs)
        self.target_dirs_tree = None
        if target_dirs is not None:
            if self.dim != 2:
                raise NotImplementedError("target derivatives through the expansion stages "
                                          "are 2D only")
            nc = geo_data.ncenters
            td = torch.as_tensor(target_dirs, dtype=torch.float64, device="cuda")
            if td.shape != (2, self.tree.ntargets - nc):
                raise ValueError(f"target_dirs must be (2, {self.tree.ntargets - nc}), "
                                 f"got {tuple(td.shape)}")
            allv = torch.cat([torch.zeros((2, nc), dtype=torch.float64, device="cuda"), td],
                             dim=1)
            order = self.tree.sorted_target_ids.long()
            self.target_dirs_tree = allv[:, order].contiguous()
            self.kd.target_deriv = 1
            self.kd.target_dir = self.target_dirs_tree.data_ptr()
        self.source_kind = source_kind
        self.qbx_order = qbx_order
        # TSQBX replaces P2QBXL's list-1 part (qbx/fmmlib.py:73-108, 295-298)
        if use_target_specific_qbx and (self.dim != 3 or output != "pot"):
            raise ValueError("target-specific QBX needs a 3D geometry and potential output")
        if use_target_specific_qbx and self.nrhs != 1 and self.k != 0:
            raise NotImplementedError("3D Helmholtz target-specific QBX evaluates one density "
                                      "at a time")
        self.using_tsqbx = bool(use_target_specific_qbx)
        # the fused driver runs the sum in place of P2QBXL (3D Laplace)
        self.kd.target_specific_qbx = int(self.using_tsqbx and self.k == 0)
        self.expansion_size = _lib.lib().qbx_expansion_size(ct.byref(self.kd), self.tree.h)
        if self.expansion_size < 0:
            raise _lib.QBXError(1, _lib.lib().qbx_last_error().decode())
        offs = (ct.c_long * (self.tree.nlevels + 1))()
        call("qbx_expansion_level_offsets", ct.byref(self.kd), self.tree.h, offs)
        self.level_offsets = np.array(offs[:], dtype=np.int64)
        self.ncoeff_qbx = _lib.lib().qbx_ncoeff_qbx(ct.byref(self.kd))

    # ---- data vector utilities
    def multipole_expansion_zeros(self):
        """Level-compact box expansions (see qbx_expansion_size)."""
        return torch.zeros(max(self.expansion_size, 1), dtype=torch.complex128,
                           device="cuda")

    def expansions_to_dense(self, e):
        """[nboxes][ncoeff(pmax)] (and [nrhs] when nrhs > 1) view of
        level-compact expansions (2D Helmholtz coefficient n at column
        n + pmax; 3D (n, m) at n*n + n + m), for inspection and tests."""
        e = e.detach().cpu().numpy()
        lo = self._lo[:self.tree.nlevels]
        pmax = int(max(lo))
        helm = self.kernel != "laplace"
        nr = self.nrhs
        ls = self.tree.level_start_box_nrs.cpu().numpy()
        if self.dim == 3:
            ncf = (pmax + 1) ** 2
        else:
            ncf = 2 * pmax + 1 if helm else pmax + 1
        out = np.zeros((self.tree.nboxes, ncf, nr), np.complex128)
        for lev in range(self.tree.nlevels):
            p = int(lo[lev])
            if self.dim == 3:
                nc, c0 = (p + 1) ** 2, 0
            else:
                nc = 2 * p + 1 if helm else p + 1
                c0 = pmax - p if helm else 0
            blk = e[self.level_offsets[lev]:self.level_offsets[lev + 1]].reshape(-1, nc, nr)
            out[ls[lev]:ls[lev + 1], c0:c0 + nc] = blk
        return out[:, :, 0] if nr == 1 else out

    local_expansion_zeros = multipole_expansion_zeros

    def _rhs_shape(self, *lead):
        return lead if self.nrhs == 1 else lead + (self.nrhs,)

    def qbx_local_expansion_zeros(self):
        return torch.zeros(self._rhs_shape(self.geo_data.ncenters, self.ncoeff_qbx),
                           dtype=torch.complex128, device="cuda")

    def full_output_zeros(self):
        return torch.zeros(self._rhs_shape(self.tree.ntargets), dtype=torch.complex128,
                           device="cuda")

    output_zeros = full_output_zeros

    def _a(self, name, *args):
        call(name, self.ctx.h, _stream(), *args)

    # ---- sources
    def reorder_sources(self, quad_density, weights, normals):
        """strengths = weights * density in tree order (+ normals for D);
        quad_density is [nsources] or [nsources, nrhs]."""
        dens = quad_density.to(torch.complex128).contiguous()
        if dens.numel() != self.tree.nsources * self.nrhs:
            raise ValueError(f"density has {dens.numel()} values, expected "
                             f"{self.tree.nsources} x {self.nrhs}")
        out = torch.empty(self._rhs_shape(self.tree.nsources), dtype=torch.complex128,
                          device="cuda")
        nout = None
        if normals is not None:
            nout = torch.empty_like(normals)
        self._a("qbx_reorder_sources", self.tree.h, self.nrhs, _p(dens), _p(weights),
                _p(normals), _p(out), _p(nout))
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
        if self.using_tsqbx:
            return q
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

    def eval_target_specific_qbx_locals(self, strengths, normals=None):
        """TSQBX (qbx/fmmlib.py:587-639, impl.pyx:566-785): potential (or its
        gradient, ``output="grad"``) at every QBX target from the list-1
        sources of its center's box, by the target-specific expansion, in the
        wrangler's units (G = 1/(4 pi r); the C entry point returns the
        unscaled 4 pi G form like impl.pyx).  3D only; zeros otherwise, like
        the reference (qbx/fmmlib.py:141-148, 590-591).
        Returns [ntargets] complex, or [3, ntargets] for the gradient."""
        if self.dim != 3:
            return self.full_output_zeros()
        kernel = _lib.LAPLACE3D if self.k == 0 else _lib.HELMHOLTZ3D
        if self.nrhs != 1:
            # many densities: each pair's kernel value once, applied to all of
            # them on the matrix cores (3D Laplace potential)
            if kernel != _lib.LAPLACE3D or self.output != "pot":
                raise NotImplementedError("many-density target-specific QBX: 3D Laplace "
                                          "potential only")
            op = _lib.TSQBX_D if self.source_kind == "D" else _lib.TSQBX_S
            if op == _lib.TSQBX_D and normals is None:
                raise ValueError("Missing dipvec")
            pot = self.full_output_zeros()
            self._a("qbx_eval_target_specific_qbx_locals_nrhs", self.tree.h,
                    self.traversal.h, self.geo.h, kernel, op, self.qbx_order, self.nrhs,
                    _p(strengths), _p(normals), _p(pot), 0)
            return pot.mul_(1.0 / (4 * np.pi))
        nt = self.tree.ntargets
        grad = pot = None
        if self.output == "grad":
            op = _lib.TSQBX_SP
            grad = torch.zeros((3, nt), dtype=torch.complex128, device="cuda")
        else:
            op = _lib.TSQBX_D if self.source_kind == "D" else _lib.TSQBX_S
            pot = self.full_output_zeros()
        if op == _lib.TSQBX_D and normals is None:
            raise ValueError("Missing dipvec")
        self._a("qbx_eval_target_specific_qbx_locals", self.tree.h, self.traversal.h,
                self.geo.h, kernel, op, self.k.real, self.k.imag, self.qbx_order,
                _p(strengths), _p(normals), _p(pot), _p(grad))
        out = grad if op == _lib.TSQBX_SP else pot
        return out.mul_(1.0 / (4 * np.pi))

    def finalize_potentials(self, pot_tree):
        nuser = self.tree.ntargets - self.geo_data.ncenters
        out = torch.empty(self._rhs_shape(nuser), dtype=torch.complex128, device="cuda")
        self._a("qbx_finalize_potentials", self.tree.h, self.geo.h, self.nrhs, _p(pot_tree),
                _p(out))
        return out


# the transposed stages of qbx_adjoint_stage (include/qbx_mi355x.h)
ADJOINT_STAGES = {"eval_qbx_expansions": 0, "translate_box_local_to_qbx_local": 1,
                  "translate_box_multipoles_to_qbx_local": 2, "form_global_qbx_locals": 3,
                  "eval_locals": 4, "eval_multipoles": 5, "eval_direct": 6,
                  "refine_locals": 7, "multipole_to_local": 8, "form_locals": 9,
                  "coarsen_multipoles": 10, "form_multipoles": 11}


def adjoint_stage(wrangler: QBXExpansionWrangler, stage, adj_out, adj_in=None, normals=None):
    """The conjugate transpose of one wrangler stage (named as the forward
    method): ``adj_out`` is the adjoint of the stage's output, the result
    (added into ``adj_in`` when given) that of its input.  refine_locals and
    coarsen_multipoles update ``adj_out`` in place and return it; the others
    return the input adjoint (eval_qbx_expansions overwrites the QBX
    expansions' adjoint).  The 2D kernels and the 3D Helmholtz single layer,
    one density."""
    w = wrangler
    sid = ADJOINT_STAGES[stage]
    adj_out = adj_out.to(torch.complex128).contiguous()
    if stage in ("refine_locals", "coarsen_multipoles"):
        call("qbx_adjoint_stage", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
             ct.byref(w.kd), sid, _p(adj_out), None, _p(normals))
        return adj_out
    if adj_in is None:
        if stage in ("eval_qbx_expansions",):
            adj_in = w.qbx_local_expansion_zeros()
        elif stage in ("translate_box_local_to_qbx_local", "eval_locals"):
            adj_in = w.local_expansion_zeros()
        elif stage in ("translate_box_multipoles_to_qbx_local", "eval_multipoles",
                       "multipole_to_local"):
            adj_in = w.multipole_expansion_zeros()
        else:
            adj_in = torch.zeros(w.tree.nsources, dtype=torch.complex128, device="cuda")
    call("qbx_adjoint_stage", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
         ct.byref(w.kd), sid, _p(adj_out), _p(adj_in), _p(normals))
    return adj_in


def drive_fmm_adjoint(wrangler: QBXExpansionWrangler, pot_adj, normals):
    """The conjugate transpose of ``drive_fmm_fused``: for the forward map
    strengths (tree order) -> potentials at the user targets, the adjoint of
    the strengths from that of the potentials (``pot_adj`` [n_user_targets]
    complex; Laplace: only its real part enters).  The 2D kernels and the 3D
    Helmholtz single layer, one density."""
    w = wrangler
    out = torch.empty(w.tree.nsources, dtype=torch.complex128, device="cuda")
    pa = pot_adj.to(torch.complex128).contiguous()
    call("qbx_drive_fmm_adjoint", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
         ct.byref(w.kd), _p(pa), _p(normals), _p(out))
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
    if w.using_tsqbx:
        # QBX targets: QBXL2P overwrote them; add the list-1 part (fmm.py:559-565)
        pot.add_(w.eval_target_specific_qbx_locals(strengths, normals))
    return w.finalize_potentials(pot)


def direct_qbx(wrangler: QBXExpansionWrangler, strengths, normals):
    """Direct QBX, no FMM (pytential/qbx/__init__.py:866-1068, the evaluation
    of ``fmm_order=False``): every QBX center's expansion from all sources,
    every other target summed over all sources.  Returns the potential at the
    user targets in user order.  3D kernels take one density at a time."""
    w = wrangler
    if w.dim == 3 and w.nrhs != 1:
        raise NotImplementedError("direct 3D QBX evaluates one density at a time")
    pot = w.full_output_zeros()
    call("qbx_direct_qbx", w.ctx.h, _stream(), w.tree.h, w.geo.h, ct.byref(w.kd),
         _p(strengths), _p(normals), _p(pot))
    return w.finalize_potentials(pot)


def drive_fmm_fused(wrangler: QBXExpansionWrangler, strengths, normals, out=None):
    """Same stage sequence from the single ``qbx_drive_fmm`` entry point
    (with target-specific QBX: 3D Laplace)."""
    w = wrangler
    if w.using_tsqbx and not w.kd.target_specific_qbx:
        raise ValueError("the fused driver's target-specific QBX stage is 3D Laplace; "
                         "use drive_fmm")
    nuser = w.tree.ntargets - w.geo_data.ncenters
    if out is None:
        out = torch.empty(w._rhs_shape(nuser), dtype=torch.complex128, device="cuda")
    call("qbx_drive_fmm", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
         ct.byref(w.kd), _p(strengths), _p(normals), _p(out))
    return out


def drive_fmm_noalloc(wrangler: QBXExpansionWrangler, strengths, normals, out):
    """The same evaluation through ``qbx_drive_fmm_noalloc``: the library
    allocates nothing (its workspace is a fixed device array shared by every
    such call on the device); ``out`` ([n_user_targets] or
    [n_user_targets, nrhs] complex128) is the caller's.  2D kernels only."""
    w = wrangler
    if w.using_tsqbx:
        raise ValueError("the fused driver has no target-specific QBX stage; use drive_fmm")
    call("qbx_drive_fmm_noalloc", w.ctx.h, _stream(), w.tree.h, w.traversal.h, w.geo.h,
         ct.byref(w.kd), _p(strengths), _p(normals), _p(out))
    return out
