"""Pins the CPU oracle of target-specific QBX (3D) and its special functions.

* spherical Bessel / Hankel: the reference's golden vectors
  (test/test_target_specific_qbx.py:47-128, fixture tests/golden/tsqbx_bessel.json),
  scipy.special, and the reference's own helmholtz_utils.c when it was built
  into oracle/_ref;
* TSQBX sums (impl.pyx:222-561): an independent spherical-harmonic expansion
  (scipy's Y_n^m, j_n, y_n) of the same truncated local expansion, which fixes
  the truncation order, powers, Bessel arguments and kernel scaling; the
  addition theorem's limit (1/|t-s|, e^{ik|t-s|}/|t-s|) at high order; and
  central finite differences for the target gradient (Sp) and the source
  dipole derivative (D), which fix their signs and directions.
"""
import ctypes as ct
import json
import os

import numpy as np
import pytest
import scipy.special as sp

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _golden():
    with open(os.path.join(HERE, "golden", "tsqbx_bessel.json")) as f:
        g = json.load(f)

    def c(a):
        return np.array([complex(x, y) for x, y in a])
    return g, c


def test_sph_jn_golden(oracle_lib):
    g, c = _golden()
    z = complex(*g["jn"]["z"])
    j, jd = oracle.sph_jn(g["jn"]["nterms"], z)
    np.testing.assert_allclose(j, c(g["jn"]["j"]), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(jd, c(g["jn"]["jder"]), rtol=1e-13, atol=1e-15)


def test_sph_hn_golden(oracle_lib):
    g, c = _golden()
    z = complex(*g["hn"]["z"])
    h, hd = oracle.sph_hn(g["hn"]["nterms"], z)
    np.testing.assert_allclose(h, c(g["hn"]["h"]), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(hd, c(g["hn"]["hder"]), rtol=1e-13, atol=1e-15)


ZS = [0.3, 2.0, 7.5, 25.0, 0.5j, 3j, 2 + 3j, 12 + 1.2j, 0.05 + 0.01j, 30 + 2j, 1e-3]


@pytest.mark.parametrize("z", ZS)
def test_sph_bessel_vs_scipy(oracle_lib, z):
    n = np.arange(13)
    j, jd = oracle.sph_jn(12, z)
    h, hd = oracle.sph_hn(12, z)
    jr = sp.spherical_jn(n, z)
    jdr = sp.spherical_jn(n, z, derivative=True)
    hr = jr + 1j * sp.spherical_yn(n, z)
    hdr = jdr + 1j * sp.spherical_yn(n, z, derivative=True)
    np.testing.assert_allclose(j, jr, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(jd, jdr, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(h, hr, rtol=1e-12)
    np.testing.assert_allclose(hd, hdr, rtol=1e-12)


def _ref_helmholtz_utils():
    p = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libhelmholtz_utils.so")
    if not os.path.exists(p):
        pytest.skip("oracle/_ref/libhelmholtz_utils.so not built (needs the reference tree)")
    return ct.CDLL(p)


@pytest.mark.parametrize("z", ZS)
def test_sph_bessel_vs_reference_c(oracle_lib, z):
    """Against jfuns3d_ / h3dall_ (helmholtz_utils.c:70-285, 400-515) with
    scale 1, the reference's own C built from its sources into oracle/_ref."""
    L = _ref_helmholtz_utils()
    nterms = 12
    # jfuns3d_ runs its backward recurrence inside the caller's buffers and
    # returns ier = 8 when they are too short (|z| = 25 needs more than the
    # reference's BUFSIZE of 64), so give it room
    buf = 1024
    zc = np.array([z], np.complex128)
    fj = np.zeros(buf, np.complex128)
    fjd = np.zeros(buf, np.complex128)
    iscale = np.zeros(buf, np.int32)
    ier, nt, ifder, lw, ntop = (ct.c_int(0), ct.c_int(nterms), ct.c_int(1), ct.c_int(buf),
                                ct.c_int(0))
    scale = ct.c_double(1.0)
    P = ct.c_void_p
    L.jfuns3d_(ct.byref(ier), ct.byref(nt), P(zc.ctypes.data), ct.byref(scale),
               P(fj.ctypes.data), ct.byref(ifder), P(fjd.ctypes.data), ct.byref(lw),
               P(iscale.ctypes.data), ct.byref(ntop))
    assert ier.value == 0
    h = np.zeros(nterms + 1, np.complex128)
    hd = np.zeros(nterms + 1, np.complex128)
    L.h3dall_(ct.byref(nt), P(zc.ctypes.data), ct.byref(scale), P(h.ctypes.data),
              ct.byref(ifder), P(hd.ctypes.data))
    j, jd = oracle.sph_jn(nterms, z)
    ho, hdo = oracle.sph_hn(nterms, z)
    np.testing.assert_allclose(j, fj[:nterms + 1], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(jd, fjd[:nterms + 1], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(ho, h, rtol=1e-12)
    np.testing.assert_allclose(hdo, hd, rtol=1e-12)


# ---- TSQBX sums on a one-box "tree"

def _one_box_case(seed, ns=7, nt=5, rt=0.3, rs=1.0):
    """One center at a random point; targets at distance rt*(0.5..1), sources
    at distance rs*(1..1.5); everything in one box whose list 1 is itself."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=3)

    def sphere_pts(n, rmin, rmax):
        v = rng.normal(size=(3, n))
        v /= np.linalg.norm(v, axis=0)
        return c[:, None] + v * rng.uniform(rmin, rmax, n)
    src = sphere_pts(ns, rs, 1.5 * rs)
    tgt = sphere_pts(nt, 0.5 * rt, rt)
    q = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    dv = rng.normal(size=(3, ns))
    csr = dict(qbx_centers=[0], c2tb=[0], ctt_starts=[0, nt], ctt_lists=np.arange(nt),
               l1_starts=[0, 1], l1_lists=[0], box_src_starts=[0], box_src_counts=[ns])
    return c, src, tgt, q, dv, csr


def _ts(kernel, op, k, order, c, src, tgt, q, dv, csr):
    return oracle.tsqbx(kernel, op, k, order, src, tgt, c[:, None], charge=q,
                        dipvec=dv, **csr)


def _sph(v):
    r = np.linalg.norm(v, axis=0)
    return r, np.arccos(np.clip(v[2] / r, -1, 1)), np.arctan2(v[1], v[0])


def _brute_S(kernel, k, order, c, src, tgt, q):
    """sum_s q_s sum_{n<=p} sum_m R_n(t) S_n(s) Y_n^m(t) conj(Y_n^m(s)) with
    Laplace: 4pi/(2n+1) rt^n / rs^(n+1);  Helmholtz: 4pi i k j_n(k rt) h_n(k rs)."""
    rt, tht, pht = _sph(tgt - c[:, None])
    rs, ths, phs = _sph(src - c[:, None])
    out = np.zeros(tgt.shape[1], np.complex128)
    for n in range(order + 1):
        if kernel == "laplace":
            rad = 4 * np.pi / (2 * n + 1) * np.outer(rt ** n, rs ** (-n - 1.0))
        else:
            jn = sp.spherical_jn(n, k * rt)
            hn = sp.spherical_jn(n, k * rs) + 1j * sp.spherical_yn(n, k * rs)
            rad = 4 * np.pi * 1j * k * np.outer(jn, hn)
        ang = np.zeros((len(rt), len(rs)), np.complex128)
        for m in range(-n, n + 1):
            ang += np.outer(sp.sph_harm_y(n, m, tht, pht),
                            np.conj(sp.sph_harm_y(n, m, ths, phs)))
        out += (rad * ang) @ q
    return out


KS = [0, 1.2, 12 + 1.2j]


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("order", [0, 1, 5, 12])
def test_tsqbx_S_vs_spherical_harmonics(oracle_lib, k, order):
    kern = "laplace" if k == 0 else "helmholtz"
    c, src, tgt, q, dv, csr = _one_box_case(10 + order)
    pot = _ts(kern, "S", k, order, c, src, tgt, q, dv, csr)
    ref = _brute_S(kern, k, order, c, src, tgt, q)
    np.testing.assert_allclose(pot, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("k", KS)
def test_tsqbx_S_addition_theorem_limit(oracle_lib, k):
    """At order 12 with rt/rs <= 0.1 the truncation error is ~0.1^13: the sum
    equals the fmmlib-scaled kernel 1/|t-s| or e^{ik|t-s|}/|t-s|."""
    kern = "laplace" if k == 0 else "helmholtz"
    c, src, tgt, q, dv, csr = _one_box_case(3, rt=0.1, rs=1.0)
    pot = _ts(kern, "S", k, 12, c, src, tgt, q, dv, csr)
    r = np.linalg.norm(tgt[:, :, None] - src[:, None, :], axis=0)
    G = 1 / r if k == 0 else np.exp(1j * k * r) / r
    ref = G @ q
    # Helmholtz with Im k > 0: terms grow like |j_n(k rt) h_n(k rs)|; allow 1e-9
    np.testing.assert_allclose(pot, ref, rtol=1e-9 if k != 0 else 1e-11)


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("order", [1, 5])
def test_tsqbx_Sp_is_target_gradient(oracle_lib, k, order):
    kern = "laplace" if k == 0 else "helmholtz"
    c, src, tgt, q, dv, csr = _one_box_case(20 + order)
    grad = _ts(kern, "Sp", k, order, c, src, tgt, q, dv, csr)
    h = 1e-5
    for ax in range(3):
        e = np.zeros((3, 1))
        e[ax] = h
        fp = _ts(kern, "S", k, order, c, src, tgt + e, q, dv, csr)
        fm = _ts(kern, "S", k, order, c, src, tgt - e, q, dv, csr)
        fd = (fp - fm) / (2 * h)
        np.testing.assert_allclose(grad[ax], fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())


@pytest.mark.parametrize("k", KS)
@pytest.mark.parametrize("order", [0, 1, 5])
def test_tsqbx_D_is_source_dipole_derivative(oracle_lib, k, order):
    """D with strength q and direction nu equals q nu . grad_s of S's kernel."""
    kern = "laplace" if k == 0 else "helmholtz"
    c, src, tgt, q, dv, csr = _one_box_case(30 + order)
    pot = _ts(kern, "D", k, order, c, src, tgt, q, dv, csr)
    h = 1e-5
    fd = np.zeros_like(pot)
    for s in range(src.shape[1]):
        qs = np.zeros_like(q)
        qs[s] = q[s]
        e = (dv[:, s] * h)[:, None]
        sp_, sm_ = src.copy(), src.copy()
        sp_[:, s:s + 1] += e
        sm_[:, s:s + 1] -= e
        fd += (_ts(kern, "S", k, order, c, sp_, tgt, qs, dv, csr) -
               _ts(kern, "S", k, order, c, sm_, tgt, qs, dv, csr)) / (2 * h)
    np.testing.assert_allclose(pot, fd, rtol=1e-6, atol=1e-7 * np.abs(fd).max())


def test_tsqbx_csr_walk(oracle_lib):
    """Two centers, two target boxes with distinct list-1 sets: each target
    sums exactly its center's list-1 boxes (checked against per-box calls)."""
    c, src, tgt, q, dv, _ = _one_box_case(5, ns=8, nt=4)
    ctr = np.concatenate([c[:, None], c[:, None] + 0.01], axis=1)
    csr = dict(qbx_centers=[0, 1], c2tb=[1, 0], ctt_starts=[0, 2, 4],
               ctt_lists=[3, 0, 1, 2], l1_starts=[0, 1, 3], l1_lists=[2, 0, 1],
               box_src_starts=[0, 3, 5], box_src_counts=[3, 2, 3])
    pot = oracle.tsqbx("laplace", "S", 0, 4, src, tgt, ctr, charge=q, **csr)

    def part(ci, boxes, t):
        s = np.concatenate([np.arange(csr["box_src_starts"][b],
                                      csr["box_src_starts"][b] + csr["box_src_counts"][b])
                            for b in boxes])
        one = dict(qbx_centers=[0], c2tb=[0], ctt_starts=[0, 1], ctt_lists=[0],
                   l1_starts=[0, 1], l1_lists=[0], box_src_starts=[0],
                   box_src_counts=[len(s)])
        return oracle.tsqbx("laplace", "S", 0, 4, src[:, s], tgt[:, t:t + 1],
                            ctr[:, ci:ci + 1], charge=q[s], **one)[0]
    # center 0 -> target box 1 -> list-1 boxes {0, 1}; center 1 -> box 0 -> {2}
    np.testing.assert_allclose(pot[3], part(0, [0, 1], 3), rtol=1e-14)
    np.testing.assert_allclose(pot[0], part(0, [0, 1], 0), rtol=1e-14)
    for t in (1, 2):
        np.testing.assert_allclose(pot[t], part(1, [2], t), rtol=1e-14)
