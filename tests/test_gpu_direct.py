"""GPU direct QBX (fmm_order=False: pytential/qbx/__init__.py:866-1068) through
qbx_direct_qbx against the CPU oracle's direct QBX on the same seeded inputs,
to 1e-12 relative to the largest potential, for 2D Laplace and Helmholtz
single and double layers; plus the GPU FMM against the GPU direct QBX on
every target (the full-size form of the accuracy check)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CASES = {
    "laplace-S": dict(builder="config1", kw=dict(nelements=64, grid_n=48, fmm_order=12)),
    "laplace-D": dict(builder="config1", kw=dict(nelements=48, grid_n=40, fmm_order=12),
                      source_kind="D"),
    "helmholtz-D": dict(builder="config2", kw=dict(nelements=192, grid_n=40, fmm_order=20)),
    "helmholtz-S": dict(builder="config2", kw=dict(nelements=128, grid_n=32, fmm_order=20),
                        source_kind="S"),
}


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    import oracle

    from pytential_amd import problems
    spec = CASES[request.param]
    cfg = getattr(problems, spec["builder"])(**spec["kw"])
    if "source_kind" in spec:
        cfg.source_kind = spec["source_kind"]
    cfg = problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)
    print(f"\n[direct] {request.param}: ndofs={cfg.discr.ndofs} ntargets={cfg.ntargets}",
          flush=True)
    return request.param, cfg


def test_direct_qbx_matches_oracle(case):
    import oracle

    from pytential_amd.qbx import QBXLayerPotentialSource
    name, cfg = case
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_order=False)
    pot, = src.exec_compute_potential(cfg.kernel, cfg.density,
                                      [(cfg.targets, cfg.target_sides)], k=cfg.k,
                                      source_kind=cfg.source_kind)
    ref = oracle.GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
    d = ref.direct(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order,
                   cfg.discr.strengths(cfg.density))
    if cfg.kernel == "laplace":
        d = d.real
    err = _rel(pot.cpu().numpy(), d)
    print(f"[direct] {name}: GPU direct QBX vs oracle direct QBX rel {err:.2e}", flush=True)
    assert err < 1e-12


def test_fmm_vs_gpu_direct_every_target(case):
    from pytential_amd.qbx import QBXLayerPotentialSource, fmm_level_to_order_for_tolerance
    name, cfg = case

    def tol_orders(kernel, kernel_args, tree, level):
        return fmm_level_to_order_for_tolerance(cfg.kernel, cfg.k, 1e-10, tree.root_extent,
                                                tree.nlevels)[level]

    srcd = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                   qbx_order=cfg.qbx_order, fmm_order=False)
    srcf = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                   qbx_order=cfg.qbx_order, fmm_level_to_order=tol_orders)
    args = (cfg.kernel, cfg.density, [(cfg.targets, cfg.target_sides)])
    kw = dict(k=cfg.k, source_kind=cfg.source_kind)
    d, = srcd.exec_compute_potential(*args, **kw)
    f, = srcf.exec_compute_potential(*args, **kw)
    err = _rel(f.cpu().numpy(), d.cpu().numpy())
    print(f"[direct] {name}: FMM (orders for 1e-10) vs GPU direct QBX on all "
          f"{cfg.ntargets} targets rel {err:.2e}", flush=True)
    assert err < 1e-10
