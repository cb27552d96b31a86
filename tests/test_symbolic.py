"""The thin bind/sym front end on the CPU: the reference's QBXPreprocessor
rules (pytential/symbolic/mappers.py:829-881) -- "avg" split into the mean
of the two one-sided limits, the limits allowed for self and non-self
evaluation, qbx_forced_limit == 0 rejected -- and the operator defaults with
their deprecation warnings (primitives.py:2271-2465)."""
import warnings

import numpy as np
import pytest

from pytential_amd import symbolic as sym


K2 = sym.LaplaceKernel(2)


def _int_gs(e, out=None):
    out = [] if out is None else out
    if isinstance(e, sym.IntG):
        out.append(e)
    elif isinstance(e, sym.Sum):
        for c in e.children:
            _int_gs(c, out)
    elif isinstance(e, sym.Product):
        _int_gs(e.child, out)
    elif isinstance(e, sym.Pointwise):
        _int_gs(e.a, out)
        _int_gs(e.b, out)
    return out


def test_avg_split_is_mean_of_one_sided_limits():
    e = sym.D(K2, sym.var("sigma"), qbx_forced_limit="avg")
    p = sym._preprocess(e, is_self=True)
    assert isinstance(p, sym.Product) and p.factor == 0.5
    lims = sorted(g.qbx_forced_limit for g in _int_gs(p))
    assert lims == [-1, 1]


def test_one_sided_and_nested_expressions_kept():
    s = sym.var("sigma")
    e = 0.5 * s - sym.D(K2, s, qbx_forced_limit=-1) + 2 * sym.S(K2, s, qbx_forced_limit=+1)
    p = sym._preprocess(e, is_self=True)
    assert sorted(g.qbx_forced_limit for g in _int_gs(p)) == [-1, 1]


@pytest.mark.parametrize("lim,is_self,msg", [
    (0, True, "no longer supported"), (0, False, "no longer supported"),
    (None, True, "must pick evaluation side"), (2, True, "May not specify"),
    ("avg", False, "non-self"), (1, False, "non-self"), (-1, False, "non-self")])
def test_forced_limit_rules(lim, is_self, msg):
    e = sym.S(K2, sym.var("sigma"), qbx_forced_limit=lim)
    with pytest.raises(ValueError, match=msg):
        sym._preprocess(e, is_self)


@pytest.mark.parametrize("lim", [None, 2, -2])
def test_non_self_limits_accepted(lim):
    e = sym.S(K2, sym.var("sigma"), qbx_forced_limit=lim)
    assert _int_gs(sym._preprocess(e, False))[0].qbx_forced_limit == lim


@pytest.mark.parametrize("op,default", [(sym.S, 1), (sym.D, "avg"), (sym.Sp, "avg"),
                                        (sym.Dp, 1)])
def test_defaults_warn(op, default):
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        e = op(K2, sym.var("sigma"))
    assert e.qbx_forced_limit == default
    assert any("qbx_forced_limit" in str(x.message) for x in w)


def test_operator_kinds_and_kernel_arguments():
    h = sym.HelmholtzKernel(2)
    e = sym.Dp(h, sym.var("sigma"), qbx_forced_limit=+1, k=sym.var("k"))
    assert (e.source_kind, e.target_deriv) == ("D", True)
    assert isinstance(e.kernel_arguments["k"], sym.Variable)
    assert (sym.Sp(K2, 1.0, qbx_forced_limit=1).source_kind, sym.S(K2, 1.0, qbx_forced_limit=1)
            .target_deriv) == ("S", False)


def test_bind_places():
    class Src:
        ambient_dim = 2
    src = Src()
    b = sym.bind(src, sym.S(K2, sym.var("s"), qbx_forced_limit=+1))
    assert b.is_self
    pts = np.zeros((2, 5))
    b2 = sym.bind((src, pts), sym.S(K2, sym.var("s"), qbx_forced_limit=None))
    assert not b2.is_self and b2.target.ndofs == 5
    with pytest.raises(ValueError):
        sym.bind((src, pts), sym.S(K2, sym.var("s"), qbx_forced_limit=+1))
