"""A thin ``bind`` / ``sym`` front end over the QBX layer potentials.

The part of pytential's symbolic layer that the FMM-backed evaluation path is
reached through (pytential/symbolic/primitives.py:2271-2465,
pytential/symbolic/execution.py:933-964, pytential/symbolic/mappers.py:829-881):

* kernels ``LaplaceKernel(dim)`` and ``HelmholtzKernel(dim, helmholtz_k_name="k")``
  (sumpy's names);
* ``var(name)`` and linear combinations: ``+``, ``-``, scalar ``*`` and ``/``,
  and products of an expression with a variable (``sym.var("w") * u``);
* layer potentials ``S``, ``D``, ``Sp``, ``Dp`` with the reference's default
  ``qbx_forced_limit`` (``S`` +1, ``D`` and ``Sp`` "avg", ``Dp`` +1; omitting it
  warns, as there) and kernel arguments (``k=...`` for Helmholtz, a number or
  a variable);
* ``bind(places, expr)`` with ``places`` a layer-potential source (targets: its
  own nodes) or a ``(source, targets)`` pair, ``targets`` a ``PointsTarget`` or
  a (dim, n) array; the bound expression is called with the variables'
  values: ``bind(qbx, expr)(sigma=sigma, k=20)``.

Preprocessing follows the reference's QBXPreprocessor: on-surface ("self")
evaluation requires ``qbx_forced_limit`` in {+1, -1, "avg"} and "avg" becomes
``0.5 * (limit +1 + limit -1)``; off-surface evaluation takes ``None`` (the
side chosen by target association) or +-2 (a forced side), never +-1 or
"avg"; 0 is rejected.  Sp and Dp are the target normal derivatives of S and D
(DirectionalTargetDerivative along the target normals), 2D; on points targets
they need explicit directions (``PointsTarget(points, normals=...)``).
Evaluation calls ``QBXLayerPotentialSource.exec_compute_potential`` per layer
potential; the results are device tensors.  The rest of pytential's symbolic
compiler (operator algebra, derivatives of arbitrary expressions, matrices)
is not part of this front end.
"""
from __future__ import annotations

import numbers
import warnings
from dataclasses import dataclass, replace

import numpy as np


# ---- kernels ----------------------------------------------------------------

@dataclass(frozen=True)
class LaplaceKernel:
    dim: int
    name = "laplace"


@dataclass(frozen=True)
class HelmholtzKernel:
    dim: int
    helmholtz_k_name: str = "k"
    name = "helmholtz"


# ---- expressions --------------------------------------------------------------

class Expression:
    def __add__(self, other):
        return Sum((self, _as_expr(other)))

    def __radd__(self, other):
        return Sum((_as_expr(other), self))

    def __sub__(self, other):
        return Sum((self, Product(-1.0, _as_expr(other))))

    def __rsub__(self, other):
        return Sum((_as_expr(other), Product(-1.0, self)))

    def __neg__(self):
        return Product(-1.0, self)

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return Product(other, self)
        return Pointwise(self, _as_expr(other))

    def __rmul__(self, other):
        if isinstance(other, numbers.Number):
            return Product(other, self)
        return Pointwise(_as_expr(other), self)

    def __truediv__(self, other):
        if not isinstance(other, numbers.Number):
            raise TypeError("only division by numbers is supported")
        return Product(1.0 / other, self)


def _as_expr(x):
    if isinstance(x, Expression):
        return x
    if isinstance(x, numbers.Number):
        return Constant(x)
    raise TypeError(f"cannot use {type(x).__name__} in an expression")


@dataclass(frozen=True, eq=False)
class Constant(Expression):
    value: complex


@dataclass(frozen=True, eq=False)
class Variable(Expression):
    name: str


@dataclass(frozen=True, eq=False)
class Sum(Expression):
    children: tuple


@dataclass(frozen=True, eq=False)
class Product(Expression):
    factor: complex
    child: Expression


@dataclass(frozen=True, eq=False)
class Pointwise(Expression):
    """elementwise product of two expressions (e.g. a weight variable and a
    density)"""
    a: Expression
    b: Expression


@dataclass(frozen=True, eq=False)
class IntG(Expression):
    """One layer potential: ``source_kind`` "S" (single layer) or "D" (double
    layer, the source normal derivative of the kernel), ``target_deriv``
    whether the target normal derivative is taken (Sp, Dp)."""
    kernel: object
    density: Expression
    source_kind: str
    target_deriv: bool
    qbx_forced_limit: object
    kernel_arguments: dict


def var(name: str) -> Variable:
    return Variable(name)


_unspecified = object()


def _int_g(kernel, density, kind, deriv, qbx_forced_limit, default, opname,
           kernel_arguments=None, **kwargs):
    if qbx_forced_limit is _unspecified:
        warnings.warn(f"Not specifying 'qbx_forced_limit' on call to '{opname}' is "
                      f"deprecated. Choosing default {default!r}.", stacklevel=3)
        qbx_forced_limit = default
    args = dict(kernel_arguments or {})
    args.update(kwargs)
    return IntG(kernel, _as_expr(density), kind, deriv, qbx_forced_limit, args)


def S(kernel, density, qbx_forced_limit=_unspecified, kernel_arguments=None, **kwargs):
    """Single layer (primitives.py:2271-2305); default limit +1."""
    return _int_g(kernel, density, "S", False, qbx_forced_limit, +1, "S",
                  kernel_arguments, **kwargs)


def D(kernel, density, qbx_forced_limit=_unspecified, kernel_arguments=None, **kwargs):
    """Double layer (primitives.py:2394-2426); default limit "avg"."""
    return _int_g(kernel, density, "D", False, qbx_forced_limit, "avg", "D",
                  kernel_arguments, **kwargs)


def Sp(kernel, density, qbx_forced_limit=_unspecified, kernel_arguments=None, **kwargs):
    """Target normal derivative of the single layer (primitives.py:2334-2363);
    default limit "avg"."""
    return _int_g(kernel, density, "S", True, qbx_forced_limit, "avg", "Sp",
                  kernel_arguments, **kwargs)


def Dp(kernel, density, qbx_forced_limit=_unspecified, kernel_arguments=None, **kwargs):
    """Target normal derivative of the double layer (primitives.py:2428-2465);
    default limit +1."""
    return _int_g(kernel, density, "D", True, qbx_forced_limit, +1, "Dp",
                  kernel_arguments, **kwargs)


# ---- targets and binding ------------------------------------------------------

class PointsTarget:
    """Evaluation points (dim, n), optionally with unit directions for target
    derivatives (pytential.target.PointsTarget)."""

    def __init__(self, points, normals=None):
        self.points = np.ascontiguousarray(points, np.float64)
        self.normals = None if normals is None else np.ascontiguousarray(normals, np.float64)
        self.ambient_dim = self.points.shape[0]

    @property
    def ndofs(self):
        return self.points.shape[1]


def _preprocess(expr, is_self):
    """The QBXPreprocessor of mappers.py:829-881 on our tree: validates the
    limits and splits "avg" into the mean of the two one-sided limits."""
    if isinstance(expr, (Constant, Variable)):
        return expr
    if isinstance(expr, Sum):
        return Sum(tuple(_preprocess(c, is_self) for c in expr.children))
    if isinstance(expr, Product):
        return Product(expr.factor, _preprocess(expr.child, is_self))
    if isinstance(expr, Pointwise):
        return Pointwise(_preprocess(expr.a, is_self), _preprocess(expr.b, is_self))
    if isinstance(expr, IntG):
        lim = expr.qbx_forced_limit
        expr = replace(expr, density=_preprocess(expr.density, is_self))
        if isinstance(lim, numbers.Number) and not isinstance(lim, bool) and lim == 0:
            raise ValueError("qbx_forced_limit == 0 was a bad idea and is no longer "
                             "supported. Use qbx_forced_limit == 'avg' to request "
                             "two-sided averaging explicitly if needed.")
        if not is_self:
            if lim in ("avg", 1, -1):
                raise ValueError("May not specify +/-1 or 'avg' for qbx_forced_limit for "
                                 "non-self evaluation. Specify 'None' for automatic choice "
                                 "or +/-2 to force a QBX side in the near-evaluation regime.")
            return expr
        if lim is None:
            raise ValueError("qbx_forced_limit == None is not supported for self "
                             "evaluation--must pick evaluation side")
        if isinstance(lim, int) and abs(lim) == 2:
            raise ValueError("May not specify qbx_forced_limit == +/-2 for self-evaluation. "
                             "Specify +/-1 or 'avg' instead.")
        if lim == "avg":
            return Product(0.5, Sum((replace(expr, qbx_forced_limit=+1),
                                     replace(expr, qbx_forced_limit=-1))))
        return expr
    raise TypeError(f"unknown expression node {type(expr).__name__}")


class BoundExpression:
    """An expression bound to its source and targets (execution.py:933-964);
    calling it evaluates with the given variable values."""

    def __init__(self, source, target, expr):
        self.source = source
        self.target = target
        self.is_self = target is None
        self.expr = _preprocess(expr, self.is_self)

    def __call__(self, actx=None, **context):
        import torch
        self._context = context
        try:
            out = self._eval(self.expr)
        finally:
            self._context = None
        if isinstance(out, numbers.Number):
            n = self.source.density_discr.ndofs if self.is_self else self.target.ndofs
            out = torch.full((n,), out, dtype=torch.complex128, device="cuda")
        return out

    def _value(self, name):
        import torch
        if name not in self._context:
            raise ValueError(f"no value given for variable '{name}'")
        v = self._context[name]
        if isinstance(v, numbers.Number):
            return v
        if isinstance(v, torch.Tensor):
            return v.to("cuda")
        return torch.as_tensor(np.asarray(v), device="cuda")

    def _eval(self, e):
        if isinstance(e, Constant):
            return e.value
        if isinstance(e, Variable):
            return self._value(e.name)
        if isinstance(e, Sum):
            acc = 0
            for c in e.children:
                acc = acc + self._eval(c)
            return acc
        if isinstance(e, Product):
            return e.factor * self._eval(e.child)
        if isinstance(e, Pointwise):
            return self._eval(e.a) * self._eval(e.b)
        if isinstance(e, IntG):
            return self._eval_int_g(e)
        raise TypeError(f"unknown expression node {type(e).__name__}")

    def _eval_int_g(self, e):
        import torch
        src = self.source
        dens = self._eval(e.density)
        n = src.density_discr.ndofs
        if isinstance(dens, numbers.Number):
            dens = torch.full((n,), dens, dtype=torch.complex128, device="cuda")
        kernel = e.kernel
        if kernel.dim != src.ambient_dim:
            raise ValueError(f"kernel dimension {kernel.dim} does not match the source "
                             f"({src.ambient_dim})")
        k = 0.0
        if kernel.name == "helmholtz":
            kname = kernel.helmholtz_k_name
            if kname not in e.kernel_arguments:
                raise ValueError(f"Helmholtz kernel needs the argument '{kname}'")
            kv = e.kernel_arguments[kname]
            k = self._eval(kv) if isinstance(kv, Expression) else kv
            if isinstance(k, torch.Tensor):
                k = k.item()
            k = float(np.real(k))
        if self.is_self:
            tgt, side, dirs = "self", e.qbx_forced_limit, None
        else:
            tgt = self.target.points
            lim = e.qbx_forced_limit
            side = 0 if lim is None else int(np.sign(lim))
            dirs = [self.target.normals] if e.target_deriv else None
            if e.target_deriv and self.target.normals is None:
                raise ValueError("a target derivative at points targets needs their "
                                 "directions (PointsTarget(points, normals=...))")
        out, = src.exec_compute_potential(kernel.name, dens, [(tgt, side)], k=k,
                                          source_kind=e.source_kind,
                                          target_derivative=e.target_deriv,
                                          target_dirs=dirs)
        return out


def bind(places, expr, auto_where=None):
    """Bind ``expr`` to a layer-potential source (self evaluation) or to a
    ``(source, target)`` pair (execution.py:933-964)."""
    if auto_where is not None:
        places = auto_where
    if isinstance(places, tuple):
        source, target = places
        if target is source:
            target = None
        elif not isinstance(target, PointsTarget):
            target = PointsTarget(target)
    else:
        source, target = places, None
    return BoundExpression(source, target, expr)


class _SymNamespace:
    """``sym``: the expression constructors under pytential's names."""
    LaplaceKernel = LaplaceKernel
    HelmholtzKernel = HelmholtzKernel
    var = staticmethod(var)
    S = staticmethod(S)
    D = staticmethod(D)
    Sp = staticmethod(Sp)
    Dp = staticmethod(Dp)
    PointsTarget = PointsTarget


sym = _SymNamespace()
