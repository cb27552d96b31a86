"""Restarted GMRES on device tensors (mirror of pytential/linalg/gmres.py).

The interface follows the reference's ``gmres`` (gmres.py:352-413): the same
arguments and defaults (restart = min(n, 20), tol = 1e-5, maxiter = 2n,
hard_failure, stall detection after ``stall_iterations`` iterations whose
residual fell by less than ``no_progress_factor``, monotonicity check) and the
same ``GMRESResult`` fields and outcome states ("success", "non-monotonic
residuals", "stalled", "max iterations"), raising ``GMRESError`` on failure
when ``hard_failure``.

The iteration itself is written here: Arnoldi with modified Gram-Schmidt and
Givens rotations on the Hessenberg matrix, the basis kept on the device as one
[restart + 1, n] tensor, the small least-squares problem on the host.  Each
Arnoldi step costs one operator application and one host synchronisation (the
new column's norms); the operator is any callable on device vectors,
typically a captured-graph QBX-FMM matvec (``MatVecGraph``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np
import torch


class GMRESError(RuntimeError):
    pass


@dataclass(frozen=True)
class GMRESResult:
    solution: torch.Tensor
    residual_norms: Sequence[float]
    iteration_count: int
    success: bool
    state: str


def _default_dot(a, b):
    return torch.vdot(a.reshape(-1), b.reshape(-1))


def gmres(op, rhs, restart=None, tol=None, x0=None, inner_product=None, maxiter=None,
          hard_failure=None, no_progress_factor=None, stall_iterations=None,
          callback: Callable | None = None, progress=False,
          require_monotonicity=True) -> GMRESResult:
    """Solve ``op(x) = rhs``; ``op`` is a callable (or has ``matvec``) on
    tensors shaped like ``rhs``.  ``inner_product(a, b)`` returns conj(a).b."""
    apply = op.matvec if hasattr(op, "matvec") else op
    dot = inner_product or _default_dot
    b = rhs
    n = b.numel()
    restart = min(n, 20) if restart is None else restart
    tol = 1e-5 if tol is None else tol
    maxiter = 2 * n if maxiter is None else maxiter
    hard_failure = True if hard_failure is None else hard_failure
    stall_iterations = 10 if stall_iterations is None else stall_iterations
    no_progress_factor = 1.25 if no_progress_factor is None else no_progress_factor

    def norm(v):
        return float(abs(dot(v, v)) ** 0.5)

    norm_b = norm(b)
    x = torch.zeros_like(b) if x0 is None else x0.clone()
    if norm_b == 0:
        return GMRESResult(x, [0.0], 0, True, "success")
    residual_norms: list[float] = []
    iteration = 0

    def finish(state, success):
        if not success and hard_failure:
            raise GMRESError(state)
        return GMRESResult(x, residual_norms, iteration, success, state)

    while True:
        r = b - apply(x) if (x0 is not None or iteration > 0) else b.clone()
        beta = norm(r)
        if residual_norms:
            residual_norms[-1] = beta / norm_b      # the true residual at a restart
        else:
            residual_norms.append(beta / norm_b)
        if beta / norm_b < tol:
            return finish("success", True)
        V = torch.empty((restart + 1,) + tuple(b.shape), dtype=b.dtype, device=b.device)
        V[0] = r / beta
        H = np.zeros((restart + 1, restart), dtype=np.complex128)
        cs = np.zeros(restart, dtype=np.complex128)
        sn = np.zeros(restart, dtype=np.complex128)
        g = np.zeros(restart + 1, dtype=np.complex128)
        g[0] = beta
        j_done = 0
        for j in range(restart):
            w = apply(V[j])
            for i in range(j + 1):
                hij = dot(V[i], w)
                w = w - hij * V[i]
                H[i, j] = complex(hij)
            hn = norm(w)
            H[j + 1, j] = hn
            if hn > 0:
                V[j + 1] = w / hn
            # apply the previous rotations, then form this one
            for i in range(j):
                a, c = H[i, j], H[i + 1, j]
                H[i, j] = np.conj(cs[i]) * a + np.conj(sn[i]) * c
                H[i + 1, j] = -sn[i] * a + cs[i] * c
            a, c = H[j, j], H[j + 1, j]
            rho = np.hypot(abs(a), abs(c))
            if rho == 0:
                cs[j], sn[j] = 1.0, 0.0
            else:
                cs[j] = a / rho
                sn[j] = c / rho
            H[j, j] = np.conj(cs[j]) * a + np.conj(sn[j]) * c
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = np.conj(cs[j]) * g[j]
            iteration += 1
            j_done = j + 1
            res = abs(g[j + 1]) / norm_b
            residual_norms.append(float(res))
            if callback is not None:
                callback(res)
            if progress:
                print(f"GMRES {iteration}: residual {res:.3e}")
            if res < tol or hn == 0 or iteration >= maxiter:
                break
            if require_monotonicity and len(residual_norms) >= 2 and \
                    residual_norms[-1] > residual_norms[-2] * (1 + 1e-8):
                break
            if stall_iterations and len(residual_norms) > stall_iterations and \
                    residual_norms[-stall_iterations] < no_progress_factor * residual_norms[-1]:
                break
        # solve the triangular system and update x
        y = np.zeros(j_done, dtype=np.complex128)
        for i in range(j_done - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:j_done] @ y[i + 1:]) / H[i, i]
        yt = torch.as_tensor(y, dtype=torch.complex128, device=b.device)
        upd = torch.tensordot(yt.to(b.dtype) if b.is_complex() else yt.real.to(b.dtype),
                              V[:j_done], dims=1)
        x = x + upd
        last = residual_norms[-1]
        if last < tol:
            # confirm with the true residual
            true_res = norm(b - apply(x)) / norm_b
            residual_norms[-1] = true_res
            if true_res < tol:
                return finish("success", True)
        if require_monotonicity and len(residual_norms) >= 2 and \
                residual_norms[-1] > residual_norms[-2] * (1 + 1e-8):
            return finish("non-monotonic residuals", False)
        if stall_iterations and len(residual_norms) > stall_iterations and \
                residual_norms[-stall_iterations] < no_progress_factor * residual_norms[-1]:
            return finish("stalled", False)
        if iteration >= maxiter:
            return finish("max iterations", False)


class MatVecGraph:
    """An operator on device vectors captured once into a HIP graph and
    replayed per application (the matvec of an iterative solve).  ``fn``
    must be capturable: the stage entry points allocate nothing after their
    first call and clear memory with kernels, so one eager warm-up call
    precedes the capture."""

    def __init__(self, fn, example: torch.Tensor, use_graph=True):
        self.fn = fn
        self.use_graph = use_graph
        self.shape = (example.numel(), example.numel())
        self._in = example.clone()
        if use_graph:
            fn(self._in)                        # warm-up: tables, workspaces
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._out = fn(self._in)
            torch.cuda.synchronize()

    def matvec(self, x):
        if not self.use_graph:
            return self.fn(x)
        self._in.copy_(x)
        self.graph.replay()
        return self._out.clone()

    __call__ = matvec
