"""Differentiable layer potentials: ``torch.autograd.Function``s whose
backward applies hand-written adjoint kernels.

* ``layer_potential`` (QBX-FMM): S or D of a curve density (2D Laplace and
  Helmholtz), or the 3D Helmholtz single layer of a surface density, at on-
  and off-surface targets; the backward runs the transposed QBX-FMM
  (QBXLayerPotential below).
* ``single_layer_far`` (3D Helmholtz, far targets): the single layer of a
  surface density at smooth targets, described next.

The forward is what pytential computes for ``bind(places, S_k(sigma))``
evaluated at targets far from the surface (``qbx_forced_limit=None``, the
targets need no QBX center, so drive_fmm's point-FMM part evaluates them:
pytential/qbx/fmm.py:216-225, pytential/qbx/__init__.py:694-793):

    phi = G_k W U sigma,
    U  the stage-1 -> quad-stage-2 interpolation (qbx_upsample),
    W  the quad-stage-2 weights (area element times rule weight,
       pytential/qbx/__init__.py:1075-1096),
    G_k[t, s] = e^{ik|t-s|} / (4 pi |t-s|)  (sumpy HelmholtzKernel(3)).

Backward (torch's convention for complex tensors: for a real loss L the
incoming gradient is g = dL/d(conj phi) times 2):

    d/dsigma = U^T W G_k^H g                  (qbx_helmholtz3d_p2p_adjoint,
                                               qbx_upsample_adjoint)
    d/dk     = Re sum_t conj(g_t) dphi_t/dk,  dphi/dk = sum_s i e^{ik r}/(4 pi) q_s
                                              (a by-product of the forward).

The sums over sources are exact (no FMM truncation); the reference offers
derivatives only by differencing or through its dense matrix builder
(pytential/symbolic/matrix.py:376-522), whose transpose is what the adjoint
applies.
"""
from __future__ import annotations

import torch

from ._lib import call
from .geometry import _p, _stream, as_dev


def _check_far(src, targets):
    """Targets must be well separated from the surface: outside the sources'
    bounding box grown by the largest element extent."""
    q = src.dev["quad_nodes"]
    lo, hi = q.min(dim=1).values, q.max(dim=1).values
    h = float((hi - lo).max()) / max(src.density_discr.nelements, 1) ** 0.5
    inside = ((targets >= (lo - h)[:, None]) & (targets <= (hi + h)[:, None])).all(dim=0)
    if bool(inside.any()):
        raise ValueError("targets near the surface need QBX: use exec_compute_potential")


def helmholtz3d_point_sum(src, strengths, targets, k, want_dk=False):
    """phi[t] = sum_s G_k(t, s) strengths[s] over the quad-stage-2 nodes
    (and dphi/dk when want_dk)."""
    nt = targets.shape[1]
    pot = torch.empty(nt, dtype=torch.complex128, device="cuda")
    dk = torch.empty_like(pot) if want_dk else None
    call("qbx_helmholtz3d_p2p", src.ctx.h, _stream(), src.density_discr.nsources,
         _p(src.dev["quad_nodes"]), _p(strengths), nt, _p(targets), float(k), _p(pot), _p(dk))
    return pot, dk


class HelmholtzSingleLayerFar(torch.autograd.Function):
    """phi = S_k sigma at smooth targets; sigma complex [ndofs] (stage-1
    nodal values), k a real 0-dim tensor, targets (3, nt) on the device."""

    @staticmethod
    def forward(ctx, sigma, k, src, targets):
        q = src.quad_density(sigma) * src.dev["quad_weights"]
        pot, dk = helmholtz3d_point_sum(src, q, targets, float(k),
                                        want_dk=ctx.needs_input_grad[1])
        ctx.src, ctx.targets, ctx.k = src, targets, float(k)
        ctx.k_device = k.device
        ctx.dk = dk
        return pot

    @staticmethod
    def backward(ctx, g):
        src, targets, k = ctx.src, ctx.targets, ctx.k
        d = src.density_discr
        g = g.to(torch.complex128).contiguous()
        grad_sigma = grad_k = None
        if ctx.needs_input_grad[0]:
            v = torch.empty(d.nsources, dtype=torch.complex128, device="cuda")
            call("qbx_helmholtz3d_p2p_adjoint", src.ctx.h, _stream(), d.nsources,
                 _p(src.dev["quad_nodes"]), targets.shape[1], _p(targets), k, _p(g), _p(v))
            v = (v * src.dev["quad_weights"]).contiguous()
            grad_sigma = torch.empty(d.ndofs, dtype=torch.complex128, device="cuda")
            call("qbx_upsample_adjoint", src.ctx.h, _stream(), d.nelements,
                 d.nodes_per_element, d.fine_nodes_per_element, 1, _p(src.dev["upsample"]),
                 _p(v), _p(grad_sigma))
        if ctx.needs_input_grad[1]:
            grad_k = torch.real(torch.sum(torch.conj(g) * ctx.dk)).to(
                device=ctx.k_device, dtype=torch.float64)
        return grad_sigma, grad_k, None, None


def single_layer_far(src, sigma, k, targets):
    """S_k sigma at smooth targets (3, nt), differentiable in sigma and k."""
    if src.ambient_dim != 3:
        raise ValueError("single_layer_far needs a 3D surface source")
    tg = as_dev(targets)
    _check_far(src, tg)
    if not isinstance(k, torch.Tensor):
        k = torch.tensor(float(k), dtype=torch.float64)
    sig = sigma if isinstance(sigma, torch.Tensor) else torch.as_tensor(sigma)
    sig = sig.to(device="cuda", dtype=torch.complex128)
    return HelmholtzSingleLayerFar.apply(sig, k, src, tg)


class QBXLayerPotential(torch.autograd.Function):
    """S or D of a stage-1 nodal density on a curve (or S_k on a surface, 3D
    Helmholtz) at (targets, side) pairs through the QBX-FMM, differentiable
    in the density and (Helmholtz) the wavenumber.  Backward:

        d/dsigma = U^T W A(k)^H g   (the transposed QBX-FMM,
                                    qbx_drive_fmm_adjoint: QBXL2P^T, L2QBXL^T,
                                    M2QBXL^T, P2QBXL^T and the point-FMM
                                    transposes, then the weights and the
                                    interpolation transposed)
        d/dk     = Re sum_t conj(g_t) (dA/dk sigma)_t, dA/dk sigma by the
                   fourth-order central difference of the forward in k
                   (step 1e-4 k; same tree, lists and orders)

    Laplace densities and potentials are real."""

    @staticmethod
    def forward(ctx, sigma, k, src, kernel, source_kind, targets_and_sides):
        kv = float(k)
        outs = src.exec_compute_potential(kernel, sigma, list(targets_and_sides), k=kv,
                                          source_kind=source_kind)
        ctx.src, ctx.kernel, ctx.k = src, kernel, kv
        ctx.source_kind, ctx.pairs = source_kind, list(targets_and_sides)
        ctx.sizes = [o.shape[0] for o in outs]
        ctx.real = not torch.is_complex(sigma)
        ctx.k_device = k.device
        ctx.save_for_backward(sigma)
        return torch.cat(outs)

    @staticmethod
    def backward(ctx, g):
        sigma, = ctx.saved_tensors
        grad_sigma = grad_k = None
        if ctx.needs_input_grad[0]:
            parts = list(torch.split(g.to(torch.complex128), ctx.sizes))
            ga = ctx.src.exec_compute_potential_adjoint(ctx.kernel, parts, ctx.pairs,
                                                        k=ctx.k, source_kind=ctx.source_kind)
            grad_sigma = ga.real if ctx.real else ga
        if ctx.needs_input_grad[1] and ctx.kernel == "helmholtz":
            h = 1e-4 * ctx.k

            def at(kk):
                return torch.cat(ctx.src.exec_compute_potential(
                    ctx.kernel, sigma, ctx.pairs, k=kk, source_kind=ctx.source_kind))
            dk = (-at(ctx.k + 2 * h) + 8 * at(ctx.k + h) - 8 * at(ctx.k - h)
                  + at(ctx.k - 2 * h)) / (12 * h)
            grad_k = torch.real(torch.sum(torch.conj(g.to(torch.complex128)) * dk)).to(
                device=ctx.k_device, dtype=torch.float64)
        return grad_sigma, grad_k, None, None, None, None


QBXLayerPotential2D = QBXLayerPotential


def layer_potential(src, kernel, sigma, targets_and_sides, k=0.0, source_kind="S"):
    """``exec_compute_potential`` of one density, concatenated over the
    target pairs, differentiable in ``sigma`` and (a tensor) ``k`` (the
    QBX-FMM of the 2D kernels or of the 3D Helmholtz single layer; see
    QBXLayerPotential)."""
    if src.ambient_dim == 3 and not (kernel == "helmholtz" and source_kind == "S"):
        raise ValueError("layer_potential differentiates the 2D QBX-FMM and the 3D "
                         "Helmholtz single layer")
    sig = sigma if isinstance(sigma, torch.Tensor) else torch.as_tensor(sigma)
    sig = sig.to(device="cuda")
    if kernel == "laplace":
        sig = sig.to(torch.float64)
    else:
        sig = sig.to(torch.complex128)
    if not isinstance(k, torch.Tensor):
        k = torch.tensor(float(k), dtype=torch.float64)
    return QBXLayerPotential.apply(sig, k, src, kernel, source_kind,
                                     tuple(targets_and_sides))
