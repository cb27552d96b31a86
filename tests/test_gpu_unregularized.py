"""GPU unregularized layer potentials (pytential/unregularized.py:66-322):
element-based Nystrom sums with the plain kernel at point targets.

* direct (``fmm_order=False``, sumpy P2P, :154-198) through ``qbx_p2p_all``
  against the oracle's point sums (2D Laplace and Helmholtz, 3D Laplace:
  ``oracle.p2p``; 3D Helmholtz: ``oracle.helmholtz3d_dense`` and its
  normal derivative), to 1e-12 relative to the largest value;
* point FMM (:214-289): tree and traversal bit-exact against the oracle's
  (sources weight 1, targets weight 0, 32 per leaf, no target extent), the
  potential against the oracle's FMM on that tree to 1e-12;
* the reference's own check (test/test_layer_pot.py:386-432): Laplace D of
  sigma = 1 on a 300-element curve, FMM order 4 against direct on a 100 x 100
  grid of extent 5, max error below 5e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def curve():
    from pytential_amd.discretization import discretize_curve, starfish
    d = discretize_curve(starfish(), 300, 8, 16)
    rng = np.random.default_rng(1101)
    tg = rng.uniform(-2.5, 2.5, (2, 3000))
    return d, tg


@pytest.fixture(scope="module")
def surface():
    from pytential_amd.surface import icosphere
    d = icosphere(1, order=4, fine_order=4)
    rng = np.random.default_rng(1102)
    tg = rng.uniform(-1.6, 1.6, (3, 1500))
    return d, tg


CASES2 = [("laplace", 0.0, "S"), ("laplace", 0.0, "D"), ("helmholtz", 5.0, "S"),
          ("helmholtz", 5.0, "D")]


@pytest.mark.parametrize("kernel,k,kind", CASES2)
def test_direct_2d_matches_oracle(curve, kernel, k, kind):
    import oracle

    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d, tg = curve
    rng = np.random.default_rng(3)
    sig = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    src = UnregularizedLayerPotentialSource(d)
    a, b = src.exec_compute_potential(kernel, sig, [tg, "self"], k=k, source_kind=kind)
    kp = oracle.kparams(kernel, k, kind, 0, [1])
    for got, t in ((a, tg), (b, d.nodes)):
        # (Laplace: a real kernel, so the complex sum is the real and the
        # imaginary part's sums, which the host evaluates as two densities)
        exp = oracle.p2p(kp, d.nodes, d.normals, d.weights * sig, t)
        err = _rel(_np(got), exp)
        print(f"\n[unreg] direct 2D {kernel} {kind} n_t={t.shape[1]}: rel {err:.2e}", flush=True)
        assert err < 1e-12


@pytest.mark.parametrize("kernel,k,kind", [("laplace", 0.0, "S"), ("laplace", 0.0, "D"),
                                           ("helmholtz", 2.5, "S"), ("helmholtz", 2.5, "D")])
def test_direct_3d_matches_oracle(surface, kernel, k, kind):
    import oracle

    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d, tg = surface
    rng = np.random.default_rng(4)
    sig = rng.normal(size=d.ndofs)
    if kernel == "helmholtz":
        sig = sig + 1j * rng.normal(size=d.ndofs)
    q = d.weights * sig
    src = UnregularizedLayerPotentialSource(d)
    got, = src.exec_compute_potential(kernel, sig, [tg], k=k, source_kind=kind)
    if kernel == "laplace":
        kp = oracle.kparams("laplace", 0.0, kind, 0, [1], dim=3)
        exp = oracle.p2p(kp, d.nodes, d.normals, q, tg).real
    elif kind == "S":
        exp = oracle.helmholtz3d_dense(d.nodes, tg, k) @ q
    else:
        dv = tg[:, :, None] - d.nodes[:, None, :]
        r = np.sqrt(np.sum(dv * dv, axis=0))
        proj = np.sum(dv * d.normals[:, None, :], axis=0) / r ** 2
        exp = (proj * (1 - 1j * k * r) * np.exp(1j * k * r) / (4 * np.pi * r)) @ q
    err = _rel(_np(got), exp)
    print(f"\n[unreg] direct 3D {kernel} {kind}: rel {err:.2e}", flush=True)
    assert err < 1e-12


def _oracle_fmm(d, tg, kernel, k, kind, order, q):
    import oracle
    nt = tg.shape[1]
    tree = oracle.Tree(d.nodes, tg, None, None, np.zeros(nt, np.int32), 32, 0.0)
    trav = oracle.Traversal(tree, 2, False)
    kp = oracle.kparams(kernel, k, kind, 0, [order] * tree.nlevels, dim=d.ambient_dim)
    out, _ = oracle.qbx_fmm(tree, trav, kp, d.nodes, d.normals, q,
                            np.zeros((d.ambient_dim, 0)), np.zeros(0), tg,
                            -np.ones(nt, np.int32))
    return tree, trav, out


@pytest.mark.parametrize("kernel,k,kind,order", [("laplace", 0.0, "S", 10),
                                                 ("laplace", 0.0, "D", 10),
                                                 ("helmholtz", 5.0, "S", 16),
                                                 ("helmholtz", 5.0, "D", 16)])
def test_fmm_2d_matches_oracle_fmm(curve, kernel, k, kind, order):
    from test_gpu_parity import TRAV_FIELDS

    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d, tg = curve
    rng = np.random.default_rng(5)
    sig = rng.normal(size=d.ndofs)
    if kernel == "helmholtz":
        sig = sig + 1j * rng.normal(size=d.ndofs)
    src = UnregularizedLayerPotentialSource(d, fmm_order=order)
    got, = src.exec_compute_potential(kernel, sig, [tg], k=k, source_kind=kind)
    tree, trav, exp = _oracle_fmm(d, tg, kernel, k, kind, order, d.weights * sig)
    geo = src.fmm_geometry_data((tg,))
    t, r = geo.tree(), geo.traversal()
    assert (t.nboxes, t.nlevels) == (tree.nboxes, tree.nlevels)
    for f in ["box_source_starts", "box_source_counts_nonchild", "box_target_starts",
              "box_target_counts_nonchild", "level_start_box_nrs", "user_source_ids",
              "sorted_target_ids", "box_parent_ids", "box_levels"]:
        assert np.array_equal(_np(getattr(t, f)), getattr(tree, f)), f
    for f in TRAV_FIELDS:
        assert np.array_equal(_np(getattr(r, f)), getattr(trav, f)), f
    if kernel == "laplace":
        exp = exp.real
    err = _rel(_np(got), exp)
    print(f"\n[unreg] FMM 2D {kernel} {kind} p={order} nboxes={t.nboxes}: rel {err:.2e}",
          flush=True)
    assert err < 1e-12


@pytest.mark.parametrize("kind", ["S", "D"])
def test_fmm_3d_matches_oracle_fmm(surface, kind):
    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d, tg = surface
    rng = np.random.default_rng(6)
    sig = rng.normal(size=d.ndofs)
    src = UnregularizedLayerPotentialSource(d, fmm_order=6)
    got, = src.exec_compute_potential("laplace", sig, [tg], source_kind=kind)
    tree, _, exp = _oracle_fmm(d, tg, "laplace", 0.0, kind, 6, d.weights * sig)
    assert src.fmm_geometry_data((tg,)).tree().nboxes == tree.nboxes
    err = _rel(_np(got), exp.real)
    print(f"\n[unreg] FMM 3D laplace {kind} nboxes={tree.nboxes}: rel {err:.2e}", flush=True)
    assert err < 1e-12


def test_fmm_batched_matches_single(curve):
    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d, tg = curve
    rng = np.random.default_rng(7)
    sig = rng.normal(size=(d.ndofs, 5)) + 1j * rng.normal(size=(d.ndofs, 5))
    src = UnregularizedLayerPotentialSource(d, fmm_order=16)
    many, = src.exec_compute_potential("helmholtz", sig, [tg], k=5.0)
    for r in range(5):
        one, = src.exec_compute_potential("helmholtz", sig[:, r], [tg], k=5.0)
        assert _rel(_np(many[:, r]), _np(one)) < 1e-13


def test_off_surface_fmm_vs_direct():
    """test/test_layer_pot.py:386-432 with a seeded starfish in place of the
    random wobbly circle."""
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.unregularized import UnregularizedLayerPotentialSource
    d = discretize_curve(starfish(arms=8, amplitude=0.2), 300, 8, 16)
    direct = UnregularizedLayerPotentialSource(d, fmm_order=False)
    fmm = direct.copy(fmm_level_to_order=lambda kernel, kernel_args, tree, level: 4)
    x = np.linspace(-2.5, 2.5, 100)
    pts = np.stack(np.meshgrid(x, x, indexing="ij")).reshape(2, -1)
    sig = np.ones(d.ndofs)
    a, = direct.exec_compute_potential("laplace", sig, [pts], source_kind="D")
    b, = fmm.exec_compute_potential("laplace", sig, [pts], source_kind="D")
    err = float(np.max(np.abs(_np(a) - _np(b))))
    print(f"\n[unreg] FMM(4) vs direct, D of 1 on a 100x100 grid: l_inf {err:.2e}", flush=True)
    assert err < 5e-3
