"""Pins the 3D Helmholtz part of the CPU oracle (oracle/fmmh3d_oracle.c).

* the Gaunt coefficients against an independent quadrature of scipy's
  spherical harmonics (including the top degree l = n + nu, where an
  undersized quadrature fails);
* P2M/M2P, M2L/L2P and P2L/L2P against the direct e^{ikr}/(4 pi r) sum
  (S and D) with the geometric convergence rate (a dropped term, wrong sign
  or transposed index stalls it);
* M2M and L2L against forming the expansion directly about the new center
  (equal up to the truncation of the input expansion, which a high input
  order makes negligible);
* the target-specific Helmholtz expansion (impl.pyx:222-358, pinned by the
  reference's own golden vectors elsewhere) equals P2QBXL followed by QBXL2P
  to round-off;
* the whole QBX-FMM against direct QBX on a triangulated sphere (S and D),
  converging with the FMM order, and direct QBX against the sphere's
  single-layer eigenvalue S_k Y_l^m = i k j_l(k) h_l(k) Y_l^m.
"""
import numpy as np
import pytest
from scipy.special import sph_harm_y, spherical_jn, spherical_yn

import oracle
from oracle import e2e, e2p, kparams, p2e, p2p

FOURPI = 4 * np.pi
K = 1.7


def _kp(kind="S", p=3, lo=(0,), k=K):
    return kparams("helmholtz", k, kind, p, list(lo), 0, 3)


def test_gaunt_vs_scipy_quadrature(oracle_lib):
    x, w = np.polynomial.legendre.leggauss(80)
    th = np.arccos(x)

    def ref(n, m, nu, mu, l):
        f = sph_harm_y(n, m, th, 0) * np.conj(sph_harm_y(nu, mu, th, 0)) * \
            np.conj(sph_harm_y(l, m - mu, th, 0))
        return float(np.real(2 * np.pi * np.sum(w * f)))
    rng = np.random.default_rng(0)
    cases = [(12, 0, 12, 0, 24), (10, 4, 7, -5, 11), (2, 1, 3, -1, 3), (20, -7, 18, 3, 38)]
    for _ in range(30):
        n, nu = rng.integers(0, 16, size=2)
        m, mu = rng.integers(-n, n + 1), rng.integers(-nu, nu + 1)
        l = abs(n - nu) + 2 * rng.integers(0, min(n, nu) + 1)
        cases.append((int(n), int(m), int(nu), int(mu), int(l)))
    for c in cases:
        a, b = oracle.gaunt3(*c), ref(*c)
        assert abs(a - b) <= 1e-12, (c, a, b)
    # selection rules: odd n + nu + l and |m - mu| > l vanish
    assert oracle.gaunt3(3, 1, 2, 0, 2) == 0.0
    assert oracle.gaunt3(4, 3, 4, -3, 2) == 0.0


def _cloud(seed, n=9, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=(3, n)) * scale, rng.normal(size=n) + 1j * rng.normal(size=n),
            rng.normal(size=(3, n)))


@pytest.mark.parametrize("kind", ["S", "D"])
def test_far_field_convergence(oracle_lib, kind):
    src, q, nu = _cloud(1)
    kp = _kp(kind)
    c = np.zeros(3)
    c2 = np.array([2.1, 0.4, -0.9])
    x = (c2 + np.array([0.1, -0.15, 0.12]))[:, None]
    d = p2p(kp, src, nu, q, x)[0]
    errs = {}
    for p in (4, 8, 16):
        M = p2e(kp, "helmholtz", 0, p, c, 1.0, src, nu, q)
        L = e2e(kp, "helmholtz", "m2l", p, p, c, 1.0, c2, 1.0, M)
        Ld = p2e(kp, "helmholtz", 1, p, c2, 1.0, src, nu, q)
        errs[p] = [abs(e2p(kp, 0, p, c, 1.0, M, x)[0] - d) / abs(d),
                   abs(e2p(kp, 1, p, c2, 1.0, L, x)[0] - d) / abs(d),
                   abs(e2p(kp, 1, p, c2, 1.0, Ld, x)[0] - d) / abs(d)]
    for i in range(3):
        assert errs[16][i] < 1e-6, errs
        assert errs[16][i] < 1e-2 * errs[8][i] or errs[16][i] < 1e-12, errs
        assert errs[8][i] < errs[4][i], errs


@pytest.mark.parametrize("kind", ["S", "D"])
def test_m2m_matches_direct_multipole(oracle_lib, kind):
    src, q, nu = _cloud(2, scale=0.1)
    kp = _kp(kind)
    c, cp = np.array([0.05, -0.02, 0.01]), np.array([0.3, -0.2, 0.15])
    for p in (3, 8):
        M = p2e(kp, "helmholtz", 0, 22, c, 1.0, src, nu, q)
        Mp = e2e(kp, "helmholtz", "m2m", 22, p, c, 1.0, cp, 1.0, M)
        Md = p2e(kp, "helmholtz", 0, p, cp, 1.0, src, nu, q)
        assert np.max(np.abs(Mp - Md)) <= 1e-12 * np.max(np.abs(Md))


@pytest.mark.parametrize("kind", ["S", "D"])
def test_l2l_matches_direct_local(oracle_lib, kind):
    src, q, nu = _cloud(3)
    src = src + np.array([[3.0], [1.0], [-2.0]])
    kp = _kp(kind)
    c, c2 = np.zeros(3), np.array([0.07, -0.05, 0.03])
    for p in (3, 8):
        L = p2e(kp, "helmholtz", 1, 22, c, 1.0, src, nu, q)
        L2 = e2e(kp, "helmholtz", "l2l", 22, p, c, 1.0, c2, 1.0, L)
        Ld = p2e(kp, "helmholtz", 1, p, c2, 1.0, src, nu, q)
        assert np.max(np.abs(L2 - Ld)) <= 1e-12 * np.max(np.abs(Ld))


@pytest.mark.parametrize("kind", ["S", "D"])
@pytest.mark.parametrize("p", [0, 1, 3, 6])
def test_tsqbx_equals_p2qbxl_qbxl2p(oracle_lib, kind, p):
    rng = np.random.default_rng(50 + p)
    c = rng.normal(size=3)
    src = c[:, None] + rng.normal(size=(3, 12)) * 0.5 + 0.6 * np.sign(rng.normal(size=(3, 12)))
    q = rng.normal(size=12) + 1j * rng.normal(size=12)
    nu = rng.normal(size=(3, 12))
    tg = c[:, None] + rng.normal(size=(3, 5)) * 0.05
    kp = _kp(kind, p)
    L = p2e(kp, "helmholtz", 1, p, c, 1.0, src, nu, q)
    pot = e2p(kp, 1, p, c, 1.0, L, tg) * FOURPI
    csr = dict(qbx_centers=[0], c2tb=[0], ctt_starts=[0, 5], ctt_lists=np.arange(5),
               l1_starts=[0, 1], l1_lists=[0], box_src_starts=[0], box_src_counts=[12])
    ts = oracle.tsqbx("helmholtz", kind, K, p, src, tg, c[:, None], charge=q, dipvec=nu, **csr)
    assert np.max(np.abs(pot - ts)) <= 1e-13 * np.max(np.abs(ts))


@pytest.fixture(scope="module")
def sphere_geo():
    from pytential_amd.surface import icosphere
    discr = icosphere(1, order=4, fine_order=5)
    rng = np.random.default_rng(1003)
    v = rng.normal(size=(3, 200))
    v /= np.linalg.norm(v, axis=0)
    tgt = v * np.where(rng.uniform(size=200) < 0.5, rng.uniform(0.3, 0.8, 200),
                       rng.uniform(1.2, 2.0, 200))
    ut = np.concatenate([discr.nodes, tgt], axis=1)
    sides = np.concatenate([np.ones(discr.ndofs, np.int32), np.zeros(200, np.int32)])
    geo = oracle.GeometryData(discr, ut, sides, max_leaf_refine_weight=32)
    assert (geo.target_status != 1).all()
    return discr, geo


@pytest.mark.parametrize("kind", ["S", "D"])
def test_qbx_fmmh3d_vs_direct_qbx(oracle_lib, sphere_geo, kind):
    discr, geo = sphere_geo
    assert geo.tree.nlevels >= 3
    rng = np.random.default_rng(6)
    sigma = rng.normal(size=discr.ndofs) + 1j * rng.normal(size=discr.ndofs)
    st = discr.strengths(sigma)
    ref = geo.direct("helmholtz", K, kind, 3, st)
    errs = []
    for pf in (6, 12, 18):
        out, _ = geo.fmm("helmholtz", K, kind, 3, [pf] * geo.tree.nlevels, st)
        errs.append(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    assert errs[2] < 1e-9, errs
    assert errs[2] < errs[1] < errs[0], errs


def test_sphere_single_layer_eigenvalue(oracle_lib):
    """S_k Y_l^m = i k j_l(k) h_l(k) Y_l^m on the unit sphere (G = e^{ikr} /
    (4 pi r)), on the surface through direct QBX; the error (relative to
    max |sigma|, as the Laplace eigenvalue test) falls with refinement."""
    from pytential_amd.surface import icosphere
    l, m, k = 3, 2, 1.2
    lam = 1j * k * spherical_jn(l, k) * (spherical_jn(l, k) + 1j * spherical_yn(l, k))
    errs = []
    for nsub in (1, 2):
        discr = icosphere(nsub, order=4, fine_order=6)
        r = np.linalg.norm(discr.nodes, axis=0)
        th, ph = np.arccos(discr.nodes[2] / r), np.arctan2(discr.nodes[1], discr.nodes[0])
        sigma = sph_harm_y(l, m, th, ph)
        st = discr.strengths(sigma)
        geo = oracle.GeometryData(discr, discr.nodes, np.ones(discr.ndofs, np.int32),
                                  max_leaf_refine_weight=64)
        out = geo.direct("helmholtz", k, "S", 3, st)
        errs.append(np.max(np.abs(out - lam * sigma)) / np.max(np.abs(sigma)))
    assert errs[1] < 2e-3, errs
    assert errs[1] < 0.5 * errs[0], errs
