"""The CPU oracle of the differentiable 3D Helmholtz path (smooth targets):
the dense operator A = G_k W U, pinned by the sphere's exterior single-layer
closed form S_k Y_lm = i k a^2 j_l(ka) h_l(kr) Y_lm (the addition theorem of
e^{ikr}/(4 pi r)), and the k-derivative the forward kernel emits,
dphi/dk = sum_s i e^{ikr}/(4 pi) q_s, pinned by central differences of A."""
import numpy as np
import pytest
from scipy.special import sph_harm_y

import oracle
from pytential_amd.surface import ellipsoid, icosphere


def _targets(n, seed, rmin=2.0, rmax=3.0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    r = rng.uniform(rmin, rmax, n)
    return v * r, r


def _angles(x):
    r = np.linalg.norm(x, axis=0)
    return np.arccos(x[2] / r), np.arctan2(x[1], x[0])


@pytest.mark.parametrize("l,m,tol", [(0, 0, 1e-10), (2, 1, 1e-5), (4, -3, 1e-4)])
def test_sphere_exterior_single_layer(l, m, tol):
    d = icosphere(2, order=4, fine_order=6)
    tg, r = _targets(60, 5)
    th, ph = _angles(d.nodes)
    sig = sph_harm_y(l, m, th, ph)
    k = 5.0
    got = oracle.helmholtz3d_far_operator(d, tg, k) @ sig
    tth, tph = _angles(tg)
    ex = oracle.sphere_single_layer_exterior(l, k, 1.0, r) * sph_harm_y(l, m, tth, tph)
    err = np.max(np.abs(got - ex)) / np.max(np.abs(ex))
    assert err < tol, err


def test_sphere_closed_form_converges():
    """Y_4^-3 on nsub = 1, 2: the error falls with refinement (a wrong
    constant in the closed form would not)."""
    errs = []
    tg, r = _targets(40, 6)
    for nsub in (1, 2):
        d = icosphere(nsub, order=4, fine_order=6)
        th, ph = _angles(d.nodes)
        sig = sph_harm_y(4, -3, th, ph)
        got = oracle.helmholtz3d_far_operator(d, tg, 5.0) @ sig
        tth, tph = _angles(tg)
        ex = oracle.sphere_single_layer_exterior(4, 5.0, 1.0, r) * sph_harm_y(4, -3, tth, tph)
        errs.append(np.max(np.abs(got - ex)) / np.max(np.abs(ex)))
    assert errs[1] < 0.1 * errs[0], errs


def test_k_derivative_formula():
    d = ellipsoid(1, order=4, fine_order=6)
    tg, _ = _targets(30, 7)
    rng = np.random.default_rng(8)
    sig = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    k, h = 5.0, 1e-5
    fd = (oracle.helmholtz3d_far_operator(d, tg, k + h) @ sig -
          oracle.helmholtz3d_far_operator(d, tg, k - h) @ sig) / (2 * h)
    q = d.strengths(sig)
    dd = tg[:, :, None] - d.quad_nodes[:, None, :]
    r = np.sqrt(np.sum(dd * dd, axis=0))
    an = (1j * np.exp(1j * k * r) / (4 * np.pi)) @ q
    assert np.max(np.abs(fd - an)) / np.max(np.abs(an)) < 1e-8


def test_chunked_apply_and_adjoint_equal_dense_operator():
    """helmholtz3d_far_apply / _adjoint (rows built in chunks) are A sigma and
    A^H y of the dense operator G_k W U."""
    import oracle
    from pytential_amd.surface import ellipsoid
    d = ellipsoid(1, order=4, fine_order=5)
    rng = np.random.default_rng(9)
    tg = rng.normal(size=(3, 70)) * 3.0
    A = oracle.helmholtz3d_far_operator(d, tg, 2.5)
    s = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    y = rng.normal(size=70) + 1j * rng.normal(size=70)
    a = oracle.helmholtz3d_far_apply(d, tg, 2.5, s, chunk=32)
    b = oracle.helmholtz3d_far_adjoint(d, tg, 2.5, y, chunk=32)
    assert np.max(np.abs(a - A @ s)) < 1e-13 * np.max(np.abs(A @ s))
    assert np.max(np.abs(b - A.conj().T @ y)) < 1e-13 * np.max(np.abs(A.conj().T @ y))
