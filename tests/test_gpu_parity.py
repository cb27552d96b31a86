"""GPU parity: libqbx_mi355x (through its C ABI) against the CPU oracle on the
same seeded inputs.

Integer work (tree, traversal lists, target association, geometry lookups)
must be bit-exact; FP64 stage outputs must agree to 1e-12 relative to the
array's max magnitude (north-star tolerance for potentials: 1e-12).
"""
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


CASES = {
    "laplace-S": dict(builder="config1", kw=dict(nelements=64, grid_n=96, fmm_order=10)),
    "laplace-D": dict(builder="config1", kw=dict(nelements=48, grid_n=64, fmm_order=12),
                      source_kind="D"),
    "helmholtz-D": dict(builder="config2", kw=dict(nelements=256, grid_n=64, fmm_order=15)),
    "helmholtz-S": dict(builder="config2", kw=dict(nelements=128, grid_n=48, fmm_order=12),
                        source_kind="S"),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    from oracle import GeometryData

    from pytential_amd import problems
    from pytential_amd.qbx import QBXLayerPotentialSource
    spec = CASES[request.param]
    cfg = getattr(problems, spec["builder"])(**spec["kw"])
    if "source_kind" in spec:
        cfg.source_kind = spec["source_kind"]
    import oracle
    cfg = problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)
    print(f"\n[case] {request.param}: ndofs={cfg.discr.ndofs} ntargets={cfg.ntargets}",
          flush=True)
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_order=cfg.fmm_order)
    geo = src.qbx_fmm_geometry_data(((cfg.targets, cfg.target_sides),))
    ref = GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
    return request.param, cfg, src, geo, ref


def test_target_association_bit_exact(case):
    _, cfg, src, geo, ref = case
    ttc, status = geo._assoc
    assert np.array_equal(_np(ttc), ref.target_to_center)
    assert np.array_equal(_np(status), ref.target_status)
    assert (ref.target_to_center >= 0).sum() > 0


TREE_FIELDS = ["box_levels", "box_parent_ids", "box_child_ids", "box_coords",
               "box_source_starts", "box_source_counts_nonchild", "box_source_counts_cumul",
               "box_target_starts", "box_target_counts_nonchild", "box_target_counts_cumul",
               "level_start_box_nrs", "user_source_ids", "sorted_target_ids"]


def test_tree_bit_exact(case):
    _, cfg, src, geo, ref = case
    t, r = geo.tree(), ref.tree
    assert (t.nboxes, t.nlevels) == (r.nboxes, r.nlevels)
    assert t.root_extent == r.root_extent
    assert np.array_equal(t.bounding_box_min, r.bounding_box_min)
    for f in TREE_FIELDS:
        assert np.array_equal(_np(getattr(t, f)), getattr(r, f)), f
    # box centers: same formula, no contraction on either side -> identical bits
    assert np.array_equal(_np(t.box_centers), r.box_centers)


TRAV_FIELDS = ["source_boxes", "source_parent_boxes", "target_boxes",
               "target_or_target_parent_boxes", "level_start_source_box_nrs",
               "level_start_source_parent_box_nrs", "level_start_target_box_nrs",
               "level_start_target_or_target_parent_box_nrs",
               "neighbor_source_boxes_starts", "neighbor_source_boxes_lists",
               "from_sep_siblings_starts", "from_sep_siblings_lists",
               "from_sep_smaller_starts", "from_sep_smaller_lists",
               "from_sep_bigger_starts", "from_sep_bigger_lists"]


def test_traversal_bit_exact(case):
    _, cfg, src, geo, ref = case
    tr, rr = geo.traversal(), ref.trav
    for f in TRAV_FIELDS:
        assert np.array_equal(_np(getattr(tr, f)), getattr(rr, f)), f


def test_geometry_lookups(case):
    _, cfg, src, geo, ref = case
    r = ref.tree
    nc = ref.ncenters
    uttc = ref.user_target_to_center
    tgt_boxes = ref.trav.target_boxes
    # qbx_center_to_target_box (geometry.py:643-700)
    owner = np.empty(r.ntargets, np.int64)
    for b in range(r.nboxes):
        s, n = r.box_target_starts[b], r.box_target_counts_nonchild[b]
        owner[s:s + n] = b
    box_to_tb = -np.ones(r.nboxes, np.int64)
    box_to_tb[tgt_boxes] = np.arange(len(tgt_boxes))
    tree_of_user = np.empty(r.ntargets, np.int64)
    tree_of_user[r.sorted_target_ids] = np.arange(r.ntargets)
    c2tb = box_to_tb[owner[tree_of_user[:nc]]]
    assert np.array_equal(_np(geo.qbx_center_to_target_box()), c2tb)
    # global_qbx_centers (:752-793)
    used = np.unique(uttc[nc:][uttc[nc:] >= 0])
    assert np.array_equal(_np(geo.global_qbx_centers()), used)
    # center_to_tree_targets (:839-872): ascending tree order per center
    starts, lists = map(_np, geo.center_to_tree_targets())
    tree_ttc = uttc[r.sorted_target_ids]
    for c in used[:: max(1, len(used) // 200)]:
        expect = np.nonzero(tree_ttc == c)[0]
        assert np.array_equal(lists[starts[c]:starts[c + 1]], expect)
    assert starts[-1] == (tree_ttc >= 0).sum()
    # non-QBX target lists (:874-899)
    nq = geo.non_qbx_box_target_lists()
    flags = (r.sorted_target_ids >= nc) & (tree_ttc == -1)
    assert np.array_equal(_np(nq["unfiltered_from_filtered_target_indices"]),
                          np.nonzero(flags)[0])
    pos = np.concatenate([[0], np.cumsum(flags)])
    fs = pos[r.box_target_starts]
    fc = pos[r.box_target_starts + r.box_target_counts_nonchild] - fs
    assert np.array_equal(_np(nq["box_target_starts"]), fs)
    assert np.array_equal(_np(nq["box_target_counts_nonchild"]), fc)


def test_fmm_stages_match_oracle(case):
    from pytential_amd import fmm as F
    name, cfg, src, geo, ref = case
    w = src.make_wrangler(geo, cfg.kernel, cfg.k, cfg.source_kind)
    lo = w._lo
    strengths_user = cfg.discr.strengths(cfg.density)
    pot_ref, st = ref.fmm(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order, list(lo),
                          strengths_user, return_stages=True)
    qd = src.quad_density(cfg.density)
    normals = src.dev["quad_normals"] if cfg.source_kind == "D" else None
    strengths, nrm = w.reorder_sources(qd, src.dev["quad_weights"], normals)
    # reorder_sources + upsampling vs the oracle's tree-order strengths
    s_ref = np.asarray(strengths_user, np.complex128)[ref.tree.user_source_ids]
    assert _rel(_np(strengths), s_ref) < REL_TOL
    mpole = w.coarsen_multipoles(w.form_multipoles(strengths, nrm))
    assert _rel(w.expansions_to_dense(mpole), st["mpole"]) < REL_TOL, "P2M+M2M"
    local = w.multipole_to_local(mpole)
    w.form_locals(strengths, nrm, local)
    w.refine_locals(local)
    assert _rel(w.expansions_to_dense(local), st["local"]) < REL_TOL, "M2L+P2L+L2L"
    qexp = w.form_global_qbx_locals(strengths, nrm)
    w.translate_box_multipoles_to_qbx_local(mpole, qexp)
    w.translate_box_local_to_qbx_local(local, qexp)
    assert _rel(_np(qexp), st["qbx"]) < REL_TOL, "P2QBXL+M2QBXL+L2QBXL"
    pot = F.drive_fmm(w, strengths, nrm)
    assert _rel(_np(pot), pot_ref) < REL_TOL, "drive_fmm"
    fused = F.drive_fmm_fused(w, strengths, nrm)
    assert _rel(_np(fused), pot_ref) < REL_TOL, "drive_fmm_fused"


def test_end_to_end_vs_direct_qbx(case):
    """FMM error against the FP64 direct-QBX oracle, reported per result."""
    name, cfg, src, geo, ref = case
    pot, = src.exec_compute_potential(cfg.kernel, cfg.density,
                                      [(cfg.targets, cfg.target_sides)], k=cfg.k,
                                      source_kind=cfg.source_kind)
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(cfg.ntargets, size=min(1500, cfg.ntargets), replace=False))
    d = ref.direct(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order,
                   cfg.discr.strengths(cfg.density), idx)
    got = _np(pot)[idx]
    if cfg.kernel == "laplace":
        d = d.real
    err = _rel(got, d)
    print(f"[{name}] FMM vs direct QBX rel err = {err:.3e}")
    # orders here are the config's; errors are stated, bounds are loose
    assert err < (1e-6 if cfg.kernel == "laplace" else 5e-2)


def test_association_failure_parity():
    """Failed association: same statuses and refine flags as the oracle, and
    the reference's exception type and message."""
    import oracle

    from pytential_amd import problems
    from pytential_amd.geometry import (QBXTargetAssociationFailedError,
                                        assoc_status_and_flags,
                                        associate_targets_to_qbx_centers, as_dev)
    cfg = problems.config1(nelements=64, grid_n=96)
    st_ref, fl_ref = oracle.assoc_status_and_flags(cfg.discr, cfg.targets, cfg.target_sides)
    assert np.any(st_ref == 1)
    st, fl = assoc_status_and_flags(cfg.discr, cfg.targets, cfg.target_sides)
    assert np.array_equal(st.astype(bool), st_ref == 1)
    assert np.array_equal(fl, fl_ref)
    d = cfg.discr
    dev = {"nodes": as_dev(d.nodes), "tunnel": as_dev(d.tunnel_radii()),
           "centers": as_dev(d.centers()), "center_radii": as_dev(d.center_radii()),
           "element_to_source_starts": as_dev(d.element_to_source_starts(), torch.int32)}
    with pytest.raises(QBXTargetAssociationFailedError, match="not assigned a QBX center"):
        associate_targets_to_qbx_centers(dev, cfg.targets, cfg.target_sides)
    ref = problems.refine_until_associated(cfg, assoc_status_and_flags)
    ref2 = problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)
    assert np.array_equal(ref.discr.t_edges, ref2.discr.t_edges)


@pytest.mark.parametrize("qbx_order", [2, 8])
def test_helmholtz_qbx_locals_other_orders(qbx_order):
    """2D Helmholtz D at QBX orders other than the default 5: P2QBXL + M2QBXL
    + L2QBXL (for one density L2QBXL splits each center over four lanes) and
    the single-call driver (which skips the coarse M2M/L2L levels no stage
    reads) against the oracle."""
    import oracle
    from oracle import GeometryData

    from pytential_amd import fmm as F
    from pytential_amd import problems
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = problems.config2(nelements=128, grid_n=48, fmm_order=12, qbx_order=qbx_order)
    cfg = problems.refine_until_associated(cfg, oracle.assoc_status_and_flags)
    print(f"\n[case] helmholtz-D qbx_order={qbx_order}: ndofs={cfg.discr.ndofs}", flush=True)
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=qbx_order, fmm_order=cfg.fmm_order)
    geo = src.qbx_fmm_geometry_data(((cfg.targets, cfg.target_sides),))
    ref = GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
    w = src.make_wrangler(geo, cfg.kernel, cfg.k, cfg.source_kind)
    pot_ref, st = ref.fmm(cfg.kernel, cfg.k, cfg.source_kind, qbx_order, list(w._lo),
                          cfg.discr.strengths(cfg.density), return_stages=True)
    qd = src.quad_density(cfg.density)
    strengths, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
    mpole = w.coarsen_multipoles(w.form_multipoles(strengths, nrm))
    local = w.multipole_to_local(mpole)
    w.form_locals(strengths, nrm, local)
    w.refine_locals(local)
    qexp = w.form_global_qbx_locals(strengths, nrm)
    w.translate_box_multipoles_to_qbx_local(mpole, qexp)
    w.translate_box_local_to_qbx_local(local, qexp)
    assert _rel(_np(qexp), st["qbx"]) < REL_TOL, "P2QBXL+M2QBXL+L2QBXL"
    fused = F.drive_fmm_fused(w, strengths, nrm)
    assert _rel(_np(fused), pot_ref) < REL_TOL, "drive_fmm_fused"
