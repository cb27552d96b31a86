"""GPU direct 3D QBX (fmm_order=False, pytential/qbx/__init__.py:866-1068)
through qbx_direct_qbx, against the CPU oracle on the same tree arrays: at QBX
targets the oracle's target-specific sums (impl.pyx:566-785 restated, pinned
by its own tests) over one list holding every source, at the other targets
the direct sums of G = e^{ikr}/(4 pi r) or its source-normal derivative, to
1e-12; 3D Laplace also against the oracle's direct QBX; and the sphere's
closed form S_k Y_00 = i k j_0(k) h_0(k) Y_00 on the surface."""
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
def sphere():
    from pytential_amd.surface import icosphere
    discr = icosphere(1, order=4, fine_order=4)
    rng = np.random.default_rng(1006)
    v = rng.normal(size=(3, 200))
    v /= np.linalg.norm(v, axis=0)
    tgt = v * np.where(rng.uniform(size=200) < 0.5, rng.uniform(0.5, 0.85, 200),
                       rng.uniform(1.15, 1.5, 200))
    return discr, tgt


def _expected_tree(geo, w, kernel, kind, k, p, strengths, nrm):
    import oracle
    tree = geo.tree()
    ns = tree.nsources
    ntb = geo.traversal().ntarget_boxes
    starts, lists = geo.center_to_tree_targets()
    src, tg = _np(tree.sources), _np(tree.targets)
    q = _np(strengths)
    exp = oracle.tsqbx(kernel, kind, k, p, src, tg, _np(geo.flat_centers()),
                       _np(geo.global_qbx_centers()), _np(geo.qbx_center_to_target_box()),
                       _np(starts), _np(lists), np.arange(ntb + 1), np.zeros(ntb),
                       np.array([0]), np.array([ns]), q,
                       _np(nrm) if nrm is not None else None) / (4 * np.pi)
    nq = geo.non_qbx_box_target_lists()
    idx = _np(nq["unfiltered_from_filtered_target_indices"]).astype(np.int64)
    d = tg[:, idx][:, :, None] - src[:, None, :]
    r = np.sqrt(np.sum(d * d, axis=0))
    g = np.exp(1j * k * r) / (4 * np.pi * r)
    if kind == "D":
        nv = _np(nrm).reshape(3, ns)
        proj = np.sum(d * nv[:, None, :], axis=0) / r ** 2
        g = proj * (1 - 1j * k * r) * g
    exp[idx] = g @ q
    return exp


@pytest.mark.parametrize("kernel,k", [("laplace", 0.0), ("helmholtz", 1.2), ("helmholtz", 5.0)])
@pytest.mark.parametrize("kind", ["S", "D"])
@pytest.mark.parametrize("qbx_order", [3, 5])
def test_direct3d_matches_oracle(sphere, kernel, k, kind, qbx_order):
    from pytential_amd import fmm as F
    from pytential_amd.qbx import QBXLayerPotentialSource
    discr, tgt = sphere
    src = QBXLayerPotentialSource(discr, fine_order=discr.fine_order, qbx_order=qbx_order,
                                  fmm_order=False, _max_leaf_refine_weight=64)
    geo = src.qbx_fmm_geometry_data((("self", +1), (tgt, 0)))
    w = src.make_wrangler(geo, kernel, k, kind)
    rng = np.random.default_rng(11 + qbx_order)
    dens = rng.normal(size=discr.ndofs) + 1j * rng.normal(size=discr.ndofs)
    if kernel == "laplace":
        dens = dens.real
    qd = src.quad_density(dens)
    strengths, nrm = w.reorder_sources(qd, src.dev["quad_weights"],
                                       src.dev["quad_normals"] if kind == "D" else None)
    got = _np(F.direct_qbx(w, strengths, nrm))
    exp_tree = _expected_tree(geo, w, kernel, kind, k, qbx_order, strengths, nrm)
    tree = geo.tree()
    sti = _np(tree.sorted_target_ids).astype(np.int64)
    nc = geo.ncenters
    exp = np.zeros(tree.ntargets - nc, np.complex128)
    mask = sti >= nc
    exp[sti[mask] - nc] = exp_tree[mask]
    err = _rel(got, exp)
    print(f"\n[direct3d] {kernel} k={k} {kind} p={qbx_order}: rel {err:.2e}", flush=True)
    assert err < 1e-12


def test_direct3d_laplace_matches_oracle_direct(sphere):
    import oracle

    from pytential_amd.qbx import QBXLayerPotentialSource
    discr, tgt = sphere
    user_targets = np.concatenate([discr.nodes, tgt], axis=1)
    sides = np.concatenate([np.ones(discr.ndofs, np.int32), np.zeros(tgt.shape[1], np.int32)])
    rng = np.random.default_rng(21)
    sig = rng.normal(size=discr.ndofs)
    for kind in ("S", "D"):
        src = QBXLayerPotentialSource(discr, fine_order=discr.fine_order, qbx_order=3,
                                      fmm_order=False, _max_leaf_refine_weight=64)
        a, b = src.exec_compute_potential("laplace", sig, [("self", +1), (tgt, 0)],
                                          source_kind=kind)
        got = np.concatenate([_np(a), _np(b)])
        ref = oracle.GeometryData(discr, user_targets, sides, max_leaf_refine_weight=64)
        d = ref.direct("laplace", 0.0, kind, 3, discr.strengths(sig)).real
        err = _rel(got, d)
        print(f"[direct3d] laplace {kind} vs oracle direct QBX rel {err:.2e}", flush=True)
        assert err < 1e-12


def test_direct3d_helmholtz_sphere_closed_form():
    import oracle

    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    # discretisation error only: the oracle's Laplace direct QBX of S Y_00 = Y_00
    # at this discretisation (icosphere(2), fine order 8, qbx_order 3) is 2.4e-4
    discr = icosphere(2, order=4, fine_order=8)
    src = QBXLayerPotentialSource(discr, fine_order=8, qbx_order=3, fmm_order=False)
    k = 2.0
    sig = np.full(discr.ndofs, 0.5 / np.sqrt(np.pi))
    pot, = src.exec_compute_potential("helmholtz", sig, [("self", +1)], k=k)
    ex = oracle.sphere_single_layer_exterior(0, k, 1.0, 1.0) * 0.5 / np.sqrt(np.pi)
    err = float(np.max(np.abs(_np(pot) - ex)) / abs(ex))
    print(f"[direct3d] sphere S_k Y_00 on the surface: rel {err:.2e}", flush=True)
    assert err < 2e-3
