"""GPU: several densities in one evaluation (nrhs > 1) give, column by
column, what one-density evaluations give, through every stage and the
right-hand-side chunking of the launches (sizes that leave a partial last
chunk); and a complex density for 2D Laplace (evaluated as a real and an
imaginary right-hand side) equals its two real evaluations combined."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REL_TOL = 1e-12


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _cfg(name):
    import oracle

    from pytential_amd import problems
    if name == "laplace2d":
        cfg = problems.config1(nelements=48, grid_n=48, fmm_order=10)
        cfg.source_kind = "D"
    else:
        cfg = problems.config2(nelements=128, grid_n=40, fmm_order=12)
    return problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)


@pytest.mark.parametrize("name,nrhs", [("laplace2d", 9), ("helmholtz2d", 5)])
def test_batched_2d_matches_single(name, nrhs):
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = _cfg(name)
    print(f"\n[batched] {name} nrhs={nrhs} ndofs={cfg.discr.ndofs}", flush=True)
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_order=cfg.fmm_order)
    rng = np.random.default_rng(3)
    n = cfg.discr.ndofs
    if name == "laplace2d":
        dens = rng.normal(size=(n, nrhs))
    else:
        dens = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    tg = [(cfg.targets, cfg.target_sides)]
    for fused in (True, False):
        multi, = src.exec_compute_potential(cfg.kernel, dens, tg, k=cfg.k,
                                            source_kind=cfg.source_kind, fused=fused)
        multi = _np(multi)
        assert multi.shape == (cfg.ntargets, nrhs)
        for r in (0, nrhs // 2, nrhs - 1):
            one, = src.exec_compute_potential(cfg.kernel, dens[:, r], tg, k=cfg.k,
                                              source_kind=cfg.source_kind, fused=fused)
            err = _rel(multi[:, r], _np(one))
            print(f"[batched] fused={fused} rhs {r}: rel diff {err:.2e}", flush=True)
            assert err < REL_TOL


def test_batched_helmholtz2d_mixed_orders_matches_single():
    """Helmholtz with 20 densities (right-hand-side chunks of 16 and 4 on the
    matrix-core M2L) and orders 30 at the coarse levels (2p + 1 > 48: those
    levels take the vector M2L) and 12 below."""
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = _cfg("helmholtz2d")
    nrhs = 20

    def orders(kernel, kernel_args, tree, level):
        return 30 if level <= 2 else 12

    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_level_to_order=orders)
    rng = np.random.default_rng(5)
    n = cfg.discr.ndofs
    dens = rng.normal(size=(n, nrhs)) + 1j * rng.normal(size=(n, nrhs))
    tg = [(cfg.targets, cfg.target_sides)]
    multi, = src.exec_compute_potential(cfg.kernel, dens, tg, k=cfg.k,
                                        source_kind=cfg.source_kind)
    multi = _np(multi)
    for r in (0, 15, 16, 19):
        one, = src.exec_compute_potential(cfg.kernel, dens[:, r], tg, k=cfg.k,
                                          source_kind=cfg.source_kind)
        err = _rel(multi[:, r], _np(one))
        print(f"[batched] mixed orders, 20 densities, rhs {r}: rel diff {err:.2e}", flush=True)
        assert err < REL_TOL


def test_laplace2d_complex_density_splits():
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = _cfg("laplace2d")
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_order=cfg.fmm_order)
    rng = np.random.default_rng(4)
    n = cfg.discr.ndofs
    dens = rng.normal(size=n) + 1j * rng.normal(size=n)
    tg = [(cfg.targets, cfg.target_sides)]
    c, = src.exec_compute_potential("laplace", dens, tg, source_kind=cfg.source_kind)
    a, = src.exec_compute_potential("laplace", dens.real.copy(), tg, source_kind=cfg.source_kind)
    b, = src.exec_compute_potential("laplace", dens.imag.copy(), tg, source_kind=cfg.source_kind)
    assert _rel(_np(c), _np(a) + 1j * _np(b)) < REL_TOL


def test_batched_3d_matches_single():
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    discr = icosphere(2, order=4, fine_order=5)
    src = QBXLayerPotentialSource(discr, fine_order=5, qbx_order=3, fmm_order=8,
                                  _max_leaf_refine_weight=64)
    rng = np.random.default_rng(6)
    nrhs = 3
    dens = rng.normal(size=(discr.ndofs, nrhs))
    print(f"\n[batched] laplace3d nrhs={nrhs} ndofs={discr.ndofs}", flush=True)
    for kind in ("S", "D"):
        multi, = src.exec_compute_potential("laplace", dens, [("self", +1)], source_kind=kind)
        multi = _np(multi)
        for r in range(nrhs):
            one, = src.exec_compute_potential("laplace", dens[:, r], [("self", +1)],
                                              source_kind=kind)
            err = _rel(multi[:, r], _np(one))
            print(f"[batched] 3D {kind} rhs {r}: rel diff {err:.2e}", flush=True)
            assert err < REL_TOL


def test_batched_3d_contraction_matches_single():
    """37 densities (a partial second block of 32 right-hand sides) through
    the batched P2QBXL contraction, S and D, against one-density runs."""
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    discr = icosphere(2, order=4, fine_order=5)
    src = QBXLayerPotentialSource(discr, fine_order=5, qbx_order=3, fmm_order=8)
    rng = np.random.default_rng(9)
    nrhs = 37
    dens = rng.normal(size=(discr.ndofs, nrhs))
    print(f"\n[batched] laplace3d contraction nrhs={nrhs} ndofs={discr.ndofs}", flush=True)
    for kind in ("S", "D"):
        multi, = src.exec_compute_potential("laplace", dens, [("self", +1)], source_kind=kind)
        multi = _np(multi)
        for r in (0, 18, 31, 32, 36):
            one, = src.exec_compute_potential("laplace", dens[:, r], [("self", +1)],
                                              source_kind=kind)
            err = _rel(multi[:, r], _np(one))
            print(f"[batched] 3D contraction {kind} rhs {r}: rel diff {err:.2e}", flush=True)
            assert err < REL_TOL
