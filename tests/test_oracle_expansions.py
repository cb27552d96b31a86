"""Pin the oracle's kernels and expansion operators (CPU).

Independent checks: scipy Bessel/Hankel values, a numpy restatement of the
Green's functions, a finite difference of the single layer for the double
layer, and convergence of every translation to the direct sum.  A wrong
sign, index or transposed shift fails one of these.
"""
import numpy as np
import pytest
from scipy import special as sp

import oracle as O

KERNELS = [("laplace", 0.0), ("helmholtz", 5.0), ("helmholtz", 20.0)]


def green(kernel, k, t, s):
    r = np.hypot(t[0][:, None] - s[0][None, :], t[1][:, None] - s[1][None, :])
    if kernel == "laplace":
        return -np.log(r) / (2 * np.pi)
    return 0.25j * sp.hankel1(0, k * r)


def test_bessel_against_scipy(oracle_lib):
    import ctypes as ct
    L = O.lib()
    for x in [1e-3, 0.37, 2.0, 11.5, 40.0]:
        j = np.zeros(31)
        y = np.zeros(31)
        L.orc_bessel_jy(30, x, j.ctypes.data_as(ct.c_void_p), y.ctypes.data_as(ct.c_void_p))
        n = np.arange(31)
        jr, yr = sp.jv(n, x), sp.yv(n, x)
        assert np.allclose(j, jr, rtol=1e-12, atol=1e-300)
        finite = np.isfinite(yr)
        assert np.allclose(y[finite], yr[finite], rtol=1e-11)


@pytest.mark.parametrize("kernel,k", KERNELS)
def test_p2p_single_layer_matches_green(oracle_lib, kernel, k):
    rng = np.random.default_rng(0)
    s = rng.standard_normal((2, 7))
    t = rng.standard_normal((2, 5)) + 3
    q = rng.standard_normal(7) + (1j * rng.standard_normal(7) if kernel != "laplace" else 0)
    kp = O.kparams(kernel, k, "S")
    got = O.p2p(kp, s, np.zeros_like(s), q, t)
    ref = green(kernel, k, t, s) @ q
    assert np.allclose(got, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("kernel,k", KERNELS)
def test_double_layer_is_normal_derivative_of_single(oracle_lib, kernel, k):
    rng = np.random.default_rng(1)
    s = rng.standard_normal((2, 4))
    nu = rng.standard_normal((2, 4))
    nu /= np.linalg.norm(nu, axis=0)
    t = rng.standard_normal((2, 3)) + 2.5
    q = rng.standard_normal(4)
    kpD = O.kparams(kernel, k, "D")
    got = O.p2p(kpD, s, nu, q, t)
    h = 1e-5
    fd = (green(kernel, k, t, s + h * nu) - green(kernel, k, t, s - h * nu)) / (2 * h) @ q
    assert np.allclose(got, fd, rtol=1e-8, atol=1e-12)


def _setup(kernel, rng):
    c = np.array([0.1, -0.2])
    src = c[:, None] + 0.1 * rng.standard_normal((2, 25))
    nrm = rng.standard_normal((2, 25))
    nrm /= np.linalg.norm(nrm, axis=0)
    st = rng.standard_normal(25) + (1j * rng.standard_normal(25) if kernel != "laplace" else 0)
    tg = c[:, None] + np.array([[1.5, -1.2, 0.3, 0.9], [0.4, 1.1, -1.6, -0.9]])
    return c, src, nrm, st, tg


@pytest.mark.parametrize("kernel,k", KERNELS)
@pytest.mark.parametrize("kind", ["S", "D"])
def test_translation_chain_converges(oracle_lib, kernel, k, kind):
    """P2M->M2P, P2M->M2M->M2P, P2M->M2L->L2P, P2L->L2L->L2P all converge
    geometrically to the direct sum and reach round-off."""
    rng = np.random.default_rng(2)
    c, src, nrm, st, tg = _setup(kernel, rng)
    kp = O.kparams(kernel, k, kind)
    exact = O.p2p(kp, src, nrm, st, tg)
    scale = np.max(np.abs(exact))
    errs = []
    for p in [8, 16, 24, 34]:
        m = O.p2e(kp, kernel, 0, p, c, 0.3, src, nrm, st)
        e1 = O.e2p(kp, 0, p, c, 0.3, m, tg)
        c2 = c + np.array([0.05, 0.07])
        m2 = O.e2e(kp, kernel, "m2m", p, p, c, 0.3, c2, 0.5, m)
        e2 = O.e2p(kp, 0, p, c2, 0.5, m2, tg)
        c3 = tg[:, 0] + np.array([0.02, -0.01])
        l3 = O.e2e(kp, kernel, "m2l", p, p, c, 0.3, c3, 0.2, m)
        e3 = O.e2p(kp, 1, p, c3, 0.2, l3, tg[:, :1])
        l4 = O.p2e(kp, kernel, 1, p, c3, 0.2, src, nrm, st)
        c5 = c3 + np.array([0.03, 0.0])
        l5 = O.e2e(kp, kernel, "l2l", p, p, c3, 0.2, c5, 0.1, l4)
        e5 = O.e2p(kp, 1, p, c5, 0.1, l5, tg[:, :1])
        errs.append([np.max(np.abs(e1 - exact)), np.max(np.abs(e2 - exact)),
                     np.max(np.abs(e3 - exact[:1])), np.max(np.abs(e5 - exact[:1]))])
    errs = np.array(errs) / scale
    assert np.all(errs[-1] < 1e-12), errs
    assert np.all(errs[0][:3] > errs[-1][:3])


def _circle(n_el=32, order=8, fine=16):
    from pytential_amd.discretization import discretize_curve, ellipse
    return discretize_curve(ellipse(1.0, 1.0), n_el, order, fine)


@pytest.mark.parametrize("m", [1, 3, 6])
def test_laplace_single_layer_circle_eigenvalue(oracle_lib, m):
    """S[cos m t] = cos(m t) / (2m) on the unit circle (the eigenvalue the
    reference checks for ellipses, test_layer_pot_eigenvalues.py:162)."""
    d = _circle()
    sigma = np.cos(m * d.t_nodes)
    ttc, status = O.target_assoc(d.nodes, d.tunnel_radii(), d.centers(),
                                 d.center_radii() * (1 + 1e-13), d.nodes,
                                 np.full(d.ndofs, 1, np.int32))
    assert np.all(ttc >= 0)
    kp = O.kparams("laplace", 0, "S", qbx_order=8)
    got = O.direct_qbx(kp, d.quad_nodes, d.quad_normals, d.strengths(sigma), d.centers(),
                       d.center_radii(), d.nodes, ttc).real
    assert np.max(np.abs(got - sigma / (2 * m))) < 1e-6


@pytest.mark.parametrize("m", [0, 2, 5])
def test_helmholtz_single_layer_circle_eigenvalue(oracle_lib, m):
    """S_k[e^{imt}] = (i pi / 2) J_m(k) H_m(k) e^{imt} on the unit circle."""
    k = 3.0
    d = _circle(48)
    sigma = np.exp(1j * m * d.t_nodes)
    ttc, status = O.target_assoc(d.nodes, d.tunnel_radii(), d.centers(),
                                 d.center_radii() * (1 + 1e-13), d.nodes,
                                 np.full(d.ndofs, -1, np.int32))
    kp = O.kparams("helmholtz", k, "S", qbx_order=8)
    got = O.direct_qbx(kp, d.quad_nodes, d.quad_normals, d.strengths(sigma), d.centers(),
                       d.center_radii(), d.nodes, ttc)
    lam = 0.5j * np.pi * sp.jv(m, k) * sp.hankel1(m, k)
    assert np.max(np.abs(got - lam * sigma)) < 1e-6


def test_laplace_double_layer_gauss_law(oracle_lib):
    """D[1] = -1 inside, 0 outside (test_layer_pot.py:129-146), FMM and direct."""
    from pytential_amd import problems
    cfg = problems.config1(nelements=64, grid_n=2)
    d = cfg.discr
    inside = np.array([[0.1, -0.3, 0.5], [0.2, 0.1, -0.4]])
    outside = np.array([[2.5, -2.0, 0.0], [0.0, 1.9, -2.2]])
    tg = np.concatenate([inside, outside], axis=1)
    sides = np.zeros(6, np.int32)
    g = O.GeometryData(d, tg, sides)
    st = d.strengths(np.ones(d.ndofs))
    fmm, _ = g.fmm("laplace", 0, "D", 4, [16] * g.tree.nlevels, st)
    direct = g.direct("laplace", 0, "D", 4, st)
    expect = np.array([-1, -1, -1, 0, 0, 0.0])
    assert np.max(np.abs(direct.real - expect)) < 1e-10
    assert np.max(np.abs(fmm.real - expect)) < 1e-10


@pytest.mark.parametrize("kernel,k,kind", [("laplace", 0, "S"), ("helmholtz", 10.0, "D")])
def test_oracle_fmm_converges_to_direct_qbx(oracle_lib, kernel, k, kind):
    from pytential_amd import problems
    cfg = problems.config1(nelements=48, grid_n=40)
    cfg = problems.refine_until_associated(cfg, O.assoc_status_and_flags)
    d = cfg.discr
    rng = np.random.default_rng(3)
    dens = cfg.density + (1j * rng.standard_normal(d.ndofs) if kernel != "laplace" else 0)
    g = O.GeometryData(d, cfg.targets, cfg.target_sides)
    st = d.strengths(dens)
    ref = g.direct(kernel, k, kind, 4, st)
    errs = []
    for p in [8, 16, 28]:
        pot, _ = g.fmm(kernel, k, kind, 4, [p] * g.tree.nlevels, st)
        errs.append(np.max(np.abs(pot - ref)) / np.max(np.abs(ref)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-10, errs
