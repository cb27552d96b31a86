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

import gc
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
            # Unreachable cycles holding library handles (a source and its
            # cached geometry reference each other) free device memory in
            # their finalizers; a free during the capture would invalidate
            # it.  Collect them now and keep the collector off while capturing.
            gc.collect()
            was_enabled = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(self.graph):
                    self._out = fn(self._in)
            finally:
                if was_enabled:
                    gc.enable()
            torch.cuda.synchronize()

    def matvec(self, x):
        if not self.use_graph:
            return self.fn(x)
        self._in.copy_(x)
        self.graph.replay()
        return self._out.clone()

    __call__ = matvec


class BatchedMatVecGraph:
    """Several operators captured into ONE HIP graph, on parallel branches:
    the capturing stream forks to ``nstreams`` side streams, operator i runs
    on side stream i % nstreams, and the branches join again, so one replay
    applies every operator and independent evaluations of small geometries
    share the chip instead of running one after another.  Each operator must
    be capturable (see MatVecGraph) and own its library contexts; those
    contexts are set to run qbx_drive_fmm on their caller's stream alone
    (QBX_TUNE_DRIVE_STREAMS = 1), so no two branches meet on the driver's
    shared side streams.  ``apply_many([(i, x_i), ...])`` copies the given
    inputs in, replays once and returns the outputs of those operators
    (operators not named run on their previous input; their outputs are not
    returned)."""

    def __init__(self, fns, examples, contexts=(), nstreams=8):
        from . import _lib
        for ctx in contexts:
            _lib.call("qbx_ctx_set_tuning", ctx.h, _lib.TUNE_DRIVE_STREAMS, 1)
        self.fns = list(fns)
        self._in = [e.clone() for e in examples]
        for f, x in zip(self.fns, self._in):    # warm-up: tables, workspaces
            f(x)
        torch.cuda.synchronize()
        self.streams = [torch.cuda.Stream() for _ in range(max(1, nstreams))]
        self.graph = torch.cuda.CUDAGraph()
        gc.collect()
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            with torch.cuda.graph(self.graph):
                cap = torch.cuda.current_stream()
                for st in self.streams:
                    st.wait_stream(cap)
                outs = []
                for i, (f, x) in enumerate(zip(self.fns, self._in)):
                    with torch.cuda.stream(self.streams[i % len(self.streams)]):
                        outs.append(f(x))
                for st in self.streams:
                    cap.wait_stream(st)
        finally:
            if was_enabled:
                gc.enable()
        self._out = outs
        torch.cuda.synchronize()

    def apply_many(self, pairs):
        for i, x in pairs:
            self._in[i].copy_(x)
        self.graph.replay()
        return [self._out[i].clone() for i, _ in pairs]

    def matvec(self, i, x):
        return self.apply_many([(i, x)])[0]


class GroupedMatVecGraphs:
    """Operators split into groups of ``group`` consecutive ones, each group a
    BatchedMatVecGraph: ``apply_many`` replays only the groups holding a
    requested operator, so systems that have converged stop costing their
    group's replay once the whole group has."""

    def __init__(self, fns, examples, contexts_of=None, group=8, nstreams=4):
        fns, examples = list(fns), list(examples)
        self.group = max(1, group)
        self.batches = []
        for g0 in range(0, len(fns), self.group):
            idx = range(g0, min(len(fns), g0 + self.group))
            ctxs = [c for i in idx for c in (contexts_of(i) if contexts_of else ())]
            self.batches.append(BatchedMatVecGraph([fns[i] for i in idx],
                                                   [examples[i] for i in idx], ctxs, nstreams))

    def apply_many(self, pairs):
        by = {}
        for k, (i, x) in enumerate(pairs):
            by.setdefault(i // self.group, []).append((k, i % self.group, x))
        out = [None] * len(pairs)
        for g, items in by.items():
            ys = self.batches[g].apply_many([(j, x) for _, j, x in items])
            for (k, _, _), y in zip(items, ys):
                out[k] = y
        return out


def gmres_many(ops, rhss, restart=None, tol=None, maxiter=None, hard_failure=False,
               no_progress_factor=None, stall_iterations=None,
               require_monotonicity=True, apply_many=None) -> list[GMRESResult]:
    """Solve independent systems ``ops[i](x) = rhss[i]`` in lockstep: every
    iteration advances each unfinished system by one Arnoldi step (its
    operator application, classical Gram-Schmidt applied twice, all on the
    device) and then transfers every system's new Hessenberg column to the
    host in one synchronisation, so the host round trip is paid once per
    iteration for the whole batch instead of once per system.  Restart,
    tolerance, stall and monotonicity rules and the outcome states are those
    of ``gmres`` (the reference's, gmres.py:352-413), applied per system;
    ``hard_failure`` raises on the first failed system.  ``apply_many``
    (e.g. ``BatchedMatVecGraph.apply_many``) applies the operators of a list
    of ``(index, vector)`` pairs at once; given, each iteration's operator
    applications are one call."""
    m = len(ops)
    if len(rhss) != m:
        raise ValueError("need one right-hand side per operator")
    tol = 1e-5 if tol is None else tol
    stall_iterations = 10 if stall_iterations is None else stall_iterations
    no_progress_factor = 1.25 if no_progress_factor is None else no_progress_factor
    applies = [op.matvec if hasattr(op, "matvec") else op for op in ops]

    class _S:
        pass

    sys_ = []
    for i in range(m):
        s = _S()
        s.b = rhss[i]
        s.n = s.b.numel()
        s.restart = min(s.n, 20) if restart is None else restart
        s.maxiter = 2 * s.n if maxiter is None else maxiter
        s.x = torch.zeros_like(s.b)
        s.res, s.it, s.state, s.done = [], 0, None, False
        sys_.append(s)
    nb = torch.stack([torch.linalg.vector_norm(s.b) for s in sys_]).cpu().numpy()
    for s, v in zip(sys_, nb):
        s.norm_b = float(v)
        if s.norm_b == 0:
            s.done, s.state = True, "success"
            s.res = [0.0]

    def apply(pairs):
        if apply_many is not None:
            return apply_many(pairs) if pairs else []
        return [applies[i](x) for i, x in pairs]

    def start_cycle(active):
        need = [(i, s.x) for i, s in active if s.it > 0]
        ax = dict(zip([i for i, _ in need], apply(need)))
        rs = [s.b - ax[i] if s.it > 0 else s.b.clone() for i, s in active]
        betas = torch.stack([torch.linalg.vector_norm(r) for r in rs]).cpu().numpy()
        out = []
        for (i, s), r, beta in zip(active, rs, betas):
            beta = float(beta)
            if s.res:
                s.res[-1] = beta / s.norm_b
            else:
                s.res.append(beta / s.norm_b)
            if beta / s.norm_b < tol:
                s.done, s.state = True, "success"
                continue
            s.V = torch.empty((s.restart + 1,) + tuple(s.b.shape), dtype=s.b.dtype,
                              device=s.b.device)
            s.V[0] = r / beta
            s.H = np.zeros((s.restart + 1, s.restart), dtype=np.complex128)
            s.cs = np.zeros(s.restart, dtype=np.complex128)
            s.sn = np.zeros(s.restart, dtype=np.complex128)
            s.g = np.zeros(s.restart + 1, dtype=np.complex128)
            s.g[0] = beta
            s.j = 0
            out.append((i, s))
        return out

    def finish_cycle(s, i):
        jd = s.j
        y = np.zeros(jd, dtype=np.complex128)
        for a in range(jd - 1, -1, -1):
            y[a] = (s.g[a] - s.H[a, a + 1:jd] @ y[a + 1:]) / s.H[a, a]
        yt = torch.as_tensor(y, dtype=torch.complex128, device=s.b.device)
        coef = yt.to(s.b.dtype) if s.b.is_complex() else yt.real.to(s.b.dtype)
        s.x = s.x + torch.tensordot(coef, s.V[:jd], dims=1)

    active = start_cycle([(i, s) for i, s in enumerate(sys_) if not s.done])
    while active:
        # one Arnoldi step per active system, device only
        cols = []
        ws = apply([(i, s.V[s.j]) for i, s in active])
        for (i, s), w in zip(active, ws):
            Vj = s.V[:s.j + 1].reshape(s.j + 1, -1)
            wf = w.reshape(-1)
            h = torch.conj(Vj) @ wf
            wf = wf - h @ Vj
            h2 = torch.conj(Vj) @ wf
            wf = wf - h2 @ Vj
            hn = torch.linalg.vector_norm(wf)
            s.w = wf.reshape(s.b.shape)
            cols.append(torch.cat([(h + h2).to(torch.complex128),
                                   hn.to(torch.complex128).reshape(1)]))
        host = torch.cat(cols).cpu().numpy()          # the one synchronisation
        off = 0
        next_active, cycle_over = [], []
        for (i, s), c in zip(active, cols):
            j = s.j
            hcol = host[off:off + j + 2]
            off += j + 2
            hn = float(hcol[-1].real)
            s.H[:j + 1, j] = hcol[:j + 1]
            s.H[j + 1, j] = hn
            if hn > 0:
                s.V[j + 1] = s.w / hn
            for a in range(j):
                x0, x1 = s.H[a, j], s.H[a + 1, j]
                s.H[a, j] = np.conj(s.cs[a]) * x0 + np.conj(s.sn[a]) * x1
                s.H[a + 1, j] = -s.sn[a] * x0 + s.cs[a] * x1
            x0, x1 = s.H[j, j], s.H[j + 1, j]
            rho = np.hypot(abs(x0), abs(x1))
            s.cs[j], s.sn[j] = (1.0, 0.0) if rho == 0 else (x0 / rho, x1 / rho)
            s.H[j, j] = np.conj(s.cs[j]) * x0 + np.conj(s.sn[j]) * x1
            s.H[j + 1, j] = 0.0
            s.g[j + 1] = -s.sn[j] * s.g[j]
            s.g[j] = np.conj(s.cs[j]) * s.g[j]
            s.it += 1
            s.j = j + 1
            s.res.append(float(abs(s.g[j + 1]) / s.norm_b))
            r = s.res
            stop = (r[-1] < tol or hn == 0 or s.it >= s.maxiter or s.j == s.restart
                    or (require_monotonicity and len(r) >= 2 and r[-1] > r[-2] * (1 + 1e-8))
                    or (stall_iterations and len(r) > stall_iterations and
                        r[-stall_iterations] < no_progress_factor * r[-1]))
            (cycle_over if stop else next_active).append((i, s))
        restart_list = []
        for i, s in cycle_over:
            finish_cycle(s, i)
            r = s.res
            if r[-1] < tol:
                restart_list.append((i, s))       # confirm with the true residual
                continue
            if require_monotonicity and len(r) >= 2 and r[-1] > r[-2] * (1 + 1e-8):
                s.done, s.state = True, "non-monotonic residuals"
            elif stall_iterations and len(r) > stall_iterations and \
                    r[-stall_iterations] < no_progress_factor * r[-1]:
                s.done, s.state = True, "stalled"
            elif s.it >= s.maxiter:
                s.done, s.state = True, "max iterations"
            else:
                restart_list.append((i, s))
        if restart_list:
            next_active += start_cycle(restart_list)
        active = next_active
    results = []
    for s in sys_:
        ok = s.state == "success"
        if not ok and hard_failure:
            raise GMRESError(s.state)
        results.append(GMRESResult(s.x, s.res, s.it, ok, s.state))
    return results
