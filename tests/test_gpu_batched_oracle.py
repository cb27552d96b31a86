"""GPU: the many-right-hand-side stages (the FP64 matrix-core kernels
k_m2l_mfma / k_m2l_mfma_rows, k_shift_mfma, k_p2qbxl_mfma, k_m2qbxl_mfma,
k_l2qbxl_mfma for 2D Helmholtz; k3_m2l_mfma, k3_p2e_mfma, k3_p2qbxl_mr for 3D
Laplace; the register-chunked vector kernels for 2D Laplace) compared column
by column with the CPU oracle's one-density stage outputs on the same tree:
multipoles after P2M+M2M, locals after M2L+P2L+L2L, QBX expansions after
P2QBXL+M2QBXL+L2QBXL, and the potentials, each to 1e-12 relative to the
array's largest entry.  Right-hand-side counts leave partial chunks (20 =
16 + 4 in 2D, 37 = 32 + 5 in 3D), and the compared columns include the first,
one inside the last chunk and the last.  The 3D cases run a second time with
the stage entry points told that the strengths are real (k3_trans_real for
M2M, M2L and L2L: half the coefficients, mirrored), orders 8, 12 and 15."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

REL_TOL = 1e-12


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale) if a.size else 0.0


def _check_columns(name, src, geo, ref, kernel, k, kind, dens, cols, strengths_user_fn,
                   qbx_order, real_strengths=False):
    from pytential_amd import fmm as F
    nrhs = dens.shape[1]
    w = src.make_wrangler(geo, kernel, k, kind, nrhs)
    # the stage entry points told that the strengths are real (the
    # conjugate-symmetric translation kernels; the fused driver finds it out)
    w.kd.real_strengths = int(real_strengths)
    lo = list(w._lo)
    qd = src.quad_density(dens)
    normals = src.dev["quad_normals"] if kind == "D" else None
    strengths, nrm = w.reorder_sources(qd, src.dev["quad_weights"], normals)
    mpole = w.coarsen_multipoles(w.form_multipoles(strengths, nrm))
    local = w.multipole_to_local(mpole)
    w.form_locals(strengths, nrm, local)
    w.refine_locals(local)
    qexp = w.form_global_qbx_locals(strengths, nrm)
    w.translate_box_multipoles_to_qbx_local(mpole, qexp)
    w.translate_box_local_to_qbx_local(local, qexp)
    pot = F.drive_fmm(w, strengths, nrm)
    fused = F.drive_fmm_fused(w, strengths, nrm)
    torch.cuda.synchronize()
    m_d = w.expansions_to_dense(mpole)
    l_d = w.expansions_to_dense(local)
    q_d, p_d, f_d = _np(qexp), _np(pot), _np(fused)
    for c in cols:
        pot_ref, st = ref.fmm(kernel, k, kind, qbx_order, lo, strengths_user_fn(dens[:, c]),
                              return_stages=True)
        errs = {"P2M+M2M": _rel(m_d[:, :, c], st["mpole"]),
                "M2L+P2L+L2L": _rel(l_d[:, :, c], st["local"]),
                "P2QBXL+M2QBXL+L2QBXL": _rel(q_d[:, :, c], st["qbx"]),
                "drive_fmm": _rel(p_d[:, c], pot_ref),
                "drive_fmm_fused": _rel(f_d[:, c], pot_ref)}
        print(f"[batched-oracle] {name} nrhs={nrhs} column {c}: "
              + ", ".join(f"{s} {e:.2e}" for s, e in errs.items()), flush=True)
        for s, e in errs.items():
            assert e < REL_TOL, (name, c, s, e)


@pytest.mark.parametrize("name,nrhs", [("helmholtz2d-D", 20), ("helmholtz2d-S", 6),
                                       ("laplace2d-D", 9)])
def test_batched_2d_stages_match_oracle(name, nrhs):
    import oracle

    from pytential_amd import problems
    from pytential_amd.qbx import QBXLayerPotentialSource
    if name.startswith("laplace"):
        cfg = problems.config1(nelements=48, grid_n=48, fmm_order=10)
    else:
        cfg = problems.config2(nelements=256, grid_n=48, fmm_order=15)
    cfg.source_kind = name[-1]
    cfg = problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_order=cfg.fmm_order)
    geo = src.qbx_fmm_geometry_data(((cfg.targets, cfg.target_sides),))
    ref = oracle.GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
    rng = np.random.default_rng(21)
    n = cfg.discr.ndofs
    dens = rng.normal(size=(n, nrhs))
    if cfg.kernel == "helmholtz":
        dens = dens + 1j * rng.normal(size=(n, nrhs))
    print(f"\n[batched-oracle] {name}: ndofs={n} ntargets={cfg.ntargets}", flush=True)
    cols = sorted({0, min(17, nrhs - 2), nrhs - 1})
    _check_columns(name, src, geo, ref, cfg.kernel, cfg.k, cfg.source_kind, dens, cols,
                   cfg.discr.strengths, cfg.qbx_order)


@pytest.mark.parametrize("kind,pf,nrhs,real", [("S", 8, 37, False), ("D", 8, 12, False),
                                               ("S", 12, 16, False), ("S", 8, 37, True),
                                               ("D", 8, 12, True), ("S", 12, 16, True),
                                               ("S", 15, 5, True)])
def test_batched_3d_stages_match_oracle(kind, pf, nrhs, real):
    import oracle

    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    discr = icosphere(2, order=4, fine_order=5)
    rng = np.random.default_rng(1003)
    v = rng.normal(size=(3, 300))
    v /= np.linalg.norm(v, axis=0)
    tgt = v * np.where(rng.uniform(size=300) < 0.5, rng.uniform(0.3, 0.85, 300),
                       rng.uniform(1.15, 2.0, 300))
    targets = (("self", +1), (tgt, 0))
    user_targets = np.concatenate([discr.nodes, tgt], axis=1)
    sides = np.concatenate([np.ones(discr.ndofs, np.int32), np.zeros(300, np.int32)])
    src = QBXLayerPotentialSource(discr, fine_order=discr.fine_order, qbx_order=3,
                                  fmm_order=pf, _max_leaf_refine_weight=64)
    geo = src.qbx_fmm_geometry_data(targets)
    ref = oracle.GeometryData(discr, user_targets, sides, max_leaf_refine_weight=64)
    dens = rng.normal(size=(discr.ndofs, nrhs))
    print(f"\n[batched-oracle] 3D {kind} p={pf}: ndofs={discr.ndofs} "
          f"nboxes={geo.tree().nboxes}", flush=True)
    cols = sorted({0, min(33, nrhs - 2), nrhs - 1})
    _check_columns(f"laplace3d-{kind}-p{pf}{'-real' if real else ''}", src, geo, ref,
                   "laplace", 0.0, kind, dens, cols, discr.strengths, 3, real_strengths=real)
