"""Pins the 3D Laplace part of the CPU oracle (oracle/fmm3d_oracle.c).

* solid harmonics against scipy's associated Legendre functions;
* P2M/M2P, M2L/L2P and P2L/L2P against the direct 1/(4 pi r) sum with the
  geometric convergence rate (a dropped term or wrong sign stalls it);
* M2M is exact (the parent multipole of a child's multipole equals the
  parent's own P2M) and L2L is exact on the truncated local (polynomial
  shift), both to round-off;
* the target-specific expansion (impl.pyx:222-358, pinned separately) equals
  P2QBXL followed by QBXL2P to round-off -- the identity the reference tests
  at 1e-13 (test/test_target_specific_qbx.py:204-214);
* the whole QBX-FMM against direct QBX on a triangulated sphere, and direct
  QBX against the sphere's single-layer eigenvalue S Y_l^m = Y_l^m / (2l+1).
"""
import numpy as np
import pytest
from scipy.special import factorial, lpmv, sph_harm_y

import oracle
from oracle import e2e, e2p, kparams, p2e, p2p

FOURPI = 4 * np.pi


def _kp(kind="S", p=3, lo=(0,)):
    return kparams("laplace", 0.0, kind, p, list(lo), 0, 3)


def test_solid_harmonics_vs_scipy(oracle_lib):
    rng = np.random.default_rng(0)
    for _ in range(4):
        x = rng.normal(size=3)
        Y, T = oracle.solid_harmonics(8, x)
        r = np.linalg.norm(x)
        th, ph = np.arccos(x[2] / r), np.arctan2(x[1], x[0])
        for n in range(9):
            for m in range(-n, n + 1):
                am = abs(m)
                P = lpmv(am, n, np.cos(th))
                if m < 0:
                    P *= (-1) ** am * factorial(n - am) / factorial(n + am)
                y = r ** n * P * np.exp(1j * m * ph) / factorial(n + m)
                t = factorial(n - m) * P * np.exp(1j * m * ph) / r ** (n + 1)
                assert abs(Y[n * n + n + m] - y) <= 1e-13 * max(abs(y), 1e-300)
                assert abs(T[n * n + n + m] - t) <= 1e-13 * abs(t)


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
        M = p2e(kp, "laplace", 0, p, c, 1.0, src, nu, q)
        L = e2e(kp, "laplace", "m2l", p, p, c, 1.0, c2, 1.0, M)
        Ld = p2e(kp, "laplace", 1, p, c2, 1.0, src, nu, q)
        errs[p] = [abs(e2p(kp, 0, p, c, 1.0, M, x)[0] - d) / abs(d),
                   abs(e2p(kp, 1, p, c2, 1.0, L, x)[0] - d) / abs(d),
                   abs(e2p(kp, 1, p, c2, 1.0, Ld, x)[0] - d) / abs(d)]
    # |src| <= ~0.9 against |x - c| ~ 2.4 (M2P), and ~0.25/2.3 (P2L): rates < 0.45
    for i in range(3):
        assert errs[16][i] < 1e-6
        assert errs[16][i] < 1e-2 * errs[8][i] or errs[16][i] < 1e-12
        assert errs[8][i] < errs[4][i]


@pytest.mark.parametrize("kind", ["S", "D"])
def test_m2m_exact(oracle_lib, kind):
    src, q, nu = _cloud(2)
    kp = _kp(kind)
    c, cp = np.array([0.05, -0.02, 0.01]), np.array([0.4, -0.3, 0.25])
    for p in (3, 9):
        M = p2e(kp, "laplace", 0, p, c, 1.0, src, nu, q)
        Mp = e2e(kp, "laplace", "m2m", p, p, c, 1.0, cp, 1.0, M)
        Md = p2e(kp, "laplace", 0, p, cp, 1.0, src, nu, q)
        assert np.max(np.abs(Mp - Md)) <= 1e-13 * np.max(np.abs(Md))


@pytest.mark.parametrize("kind", ["S", "D"])
def test_l2l_exact_on_truncated_local(oracle_lib, kind):
    src, q, nu = _cloud(3)
    src = src + np.array([[3.0], [1.0], [-2.0]])
    kp = _kp(kind)
    c, c2 = np.zeros(3), np.array([0.07, -0.05, 0.03])
    x = np.array([[0.1, -0.02], [0.03, 0.08], [-0.05, 0.02]])
    for p in (3, 8):
        L = p2e(kp, "laplace", 1, p, c, 1.0, src, nu, q)
        L2 = e2e(kp, "laplace", "l2l", p, p, c, 1.0, c2, 1.0, L)
        a = e2p(kp, 1, p, c, 1.0, L, x)
        b = e2p(kp, 1, p, c2, 1.0, L2, x)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(a))
        # lower target order keeps the leading coefficients exactly
        L3 = e2e(kp, "laplace", "l2l", p, p - 1, c, 1.0, c2, 1.0, L)
        assert np.allclose(L3, L2[:p * p], rtol=0, atol=1e-14 * np.max(np.abs(L2)))


@pytest.mark.parametrize("kind", ["S", "D"])
@pytest.mark.parametrize("p", [0, 1, 3, 5])
def test_tsqbx_equals_p2qbxl_qbxl2p(oracle_lib, kind, p):
    """impl.pyx's Legendre form and the solid-harmonic QBX local agree."""
    rng = np.random.default_rng(40 + p)
    c = rng.normal(size=3)
    src = c[:, None] + rng.normal(size=(3, 12)) * 0.5 + 0.6 * np.sign(rng.normal(size=(3, 12)))
    q = rng.normal(size=12) + 1j * rng.normal(size=12)
    nu = rng.normal(size=(3, 12))
    tg = c[:, None] + rng.normal(size=(3, 5)) * 0.05
    kp = _kp(kind, p)
    L = p2e(kp, "laplace", 1, p, c, 1.0, src, nu, q)
    pot = e2p(kp, 1, p, c, 1.0, L, tg) * FOURPI
    csr = dict(qbx_centers=[0], c2tb=[0], ctt_starts=[0, 5], ctt_lists=np.arange(5),
               l1_starts=[0, 1], l1_lists=[0], box_src_starts=[0], box_src_counts=[12])
    ts = oracle.tsqbx("laplace", kind, 0, p, src, tg, c[:, None], charge=q, dipvec=nu, **csr)
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
def test_qbx_fmm3d_vs_direct_qbx(oracle_lib, sphere_geo, kind):
    discr, geo = sphere_geo
    assert geo.tree.nlevels >= 3
    rng = np.random.default_rng(5)
    sigma = rng.normal(size=discr.ndofs)
    st = discr.strengths(sigma)
    ref = geo.direct("laplace", 0, kind, 3, st)
    errs = []
    for pf in (6, 12, 20):
        out, _ = geo.fmm("laplace", 0, kind, 3, [pf] * geo.tree.nlevels, st)
        errs.append(np.max(np.abs(out - ref)) / np.max(np.abs(ref)))
    assert errs[2] < 1e-9, errs
    assert errs[2] < errs[1] < errs[0], errs


def test_sphere_single_layer_eigenvalue(oracle_lib):
    """S Y_l^m = Y_l^m / (2l+1) on the unit sphere (G = 1/(4 pi r)), on the
    surface through direct QBX; the error falls with refinement."""
    from pytential_amd.surface import icosphere
    errs = []
    for nsub in (1, 2):
        discr = icosphere(nsub, order=4, fine_order=6)
        r, th, ph = (np.linalg.norm(discr.nodes, axis=0),
                     np.arccos(discr.nodes[2] / np.linalg.norm(discr.nodes, axis=0)),
                     np.arctan2(discr.nodes[1], discr.nodes[0]))
        l, m = 3, 2
        sigma = sph_harm_y(l, m, th, ph).real
        st = discr.strengths(sigma)
        ttc_in = np.ones(discr.ndofs, np.int32)
        geo = oracle.GeometryData(discr, discr.nodes, ttc_in, max_leaf_refine_weight=64)
        out = geo.direct("laplace", 0, "S", 3, st).real
        errs.append(np.max(np.abs(out - sigma / (2 * l + 1))) / np.max(np.abs(sigma)))
    assert errs[1] < 2e-3, errs
    assert errs[1] < 0.5 * errs[0], errs
