"""Dirichlet boundary integral operator for the QBX-FMM path (mirror of
pytential/symbolic/pde/scalar.py:202-362, ``DirichletOperator``).

    operator(u)       = -1/2 loc_sign u + sqrt_w (alpha S(u/sqrt_w) - D(u/sqrt_w) [+ ones])
    representation(u) = alpha S(u/sqrt_w) - D(u/sqrt_w)

with ``S`` taken at ``qbx_forced_limit=+1`` and ``D`` at ``"avg"`` (the mean of
the two one-sided limits, as the reference's QBXPreprocessor splits it,
mappers.py:876-879), alpha = 0 for Laplace and 1j for Helmholtz (the
combined-field choice of the reference), and for the exterior Laplace problem
the mean of the density added to remove the constant nullspace
(scalar.py:327-340).  Densities are nodal values on the stage-1 nodes; the
weights ``sqrt_w`` are the square roots of the stage-1 quadrature weights when
``use_l2_weighting``.

One geometry evaluates both one-sided limits in a single QBX-FMM (targets
"self" at side +1 and side -1), so a matvec is one library call; ``matvec``
can be captured into a HIP graph (``pytential_amd.linalg.MatVecGraph``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import fmm as _fmm
from .geometry import as_dev


class DirichletOperator:
    def __init__(self, lpot_source, loc_sign, kernel="laplace", k=0.0, alpha=None,
                 use_l2_weighting=False):
        if loc_sign not in (-1, 1):
            raise ValueError("loc_sign must be -1 or +1")
        if kernel not in ("laplace", "helmholtz"):
            raise ValueError(f"unknown kernel {kernel!r}")
        self.src = lpot_source
        self.loc_sign = loc_sign
        self.kernel = kernel
        self.k = k
        self.alpha = (1j if kernel == "helmholtz" else 0) if alpha is None else alpha
        d = lpot_source.density_discr
        self.n = d.ndofs
        w = np.asarray(d.weights, np.float64)
        self.sqrt_w = as_dev(np.sqrt(w) if use_l2_weighting else np.ones_like(w))
        self.weights = as_dev(w)
        self.area = float(w.sum())
        self.ones = kernel == "laplace" and loc_sign > 0
        # D at both one-sided limits in one evaluation
        self.geo_d = lpot_source.qbx_fmm_geometry_data((("self", +1), ("self", -1)))
        # each wrangler owns its workspaces and tables, so matvecs of several
        # operators can be captured into HIP graphs and replayed in any order
        self.w_d = lpot_source.make_wrangler(self.geo_d, kernel, k, "D", private_context=True)
        self.w_s = None
        if self.alpha != 0:
            self.geo_s = lpot_source.qbx_fmm_geometry_data((("self", +1),))
            self.w_s = lpot_source.make_wrangler(self.geo_s, kernel, k, "S",
                                                 private_context=True)
        self.dtype = torch.float64 if (kernel == "laplace" and self.alpha == 0) \
            else torch.complex128

    def _layer(self, w, kind, dens):
        src = self.src
        qd = src.quad_density(dens)
        normals = src.dev["quad_normals"] if kind == "D" else None
        st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], normals)
        return _fmm.drive_fmm_fused(w, st, nrm)

    def matvec(self, u):
        """operator(u) for nodal u [ndofs] on the device."""
        sig = u / self.sqrt_w
        dd = self._layer(self.w_d, "D", sig)
        d_avg = 0.5 * (dd[:self.n] + dd[self.n:])
        inner = -d_avg
        if self.w_s is not None:
            inner = inner + self.alpha * self._layer(self.w_s, "S", sig)
        if self.ones:
            inner = inner + torch.sum(sig * self.weights) / self.area
        out = -0.5 * self.loc_sign * u + self.sqrt_w * inner
        return out.real.to(self.dtype) if self.dtype == torch.float64 else out

    __call__ = matvec

    def library_contexts(self):
        """The library contexts of this operator's evaluations (its own, see
        private_context), e.g. to set their QBX_TUNE_* knobs."""
        return [w.ctx for w in (self.w_d, self.w_s) if w is not None]

    @property
    def shape(self):
        return (self.n, self.n)

    def representation(self, u, targets, qbx_forced_limit=None):
        """alpha S - D of u/sqrt_w at off-surface targets (dim, m)."""
        sig = u / self.sqrt_w
        side = 0 if qbx_forced_limit is None else qbx_forced_limit
        pot_d, = self.src.exec_compute_potential(self.kernel, sig, [(targets, side)], k=self.k,
                                                 source_kind="D")
        res = -pot_d
        if self.alpha != 0:
            pot_s, = self.src.exec_compute_potential(self.kernel, sig, [(targets, side)],
                                                     k=self.k, source_kind="S")
            res = res + self.alpha * pot_s
        return res
