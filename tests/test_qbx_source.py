"""QBXLayerPotentialSource constructor defaults and ``copy`` against the
reference's (pytential/qbx/__init__.py:164-478).  Construction touches no GPU
(device arrays are made on first use), so these run on the CPU."""
import pytest

from pytential_amd.discretization import discretize_curve, starfish
from pytential_amd.qbx import QBXLayerPotentialSource


@pytest.fixture(scope="module")
def curve():
    return discretize_curve(starfish(5, 0.25), 16, 8, 16)


@pytest.fixture(scope="module")
def surface():
    from pytential_amd.surface import icosphere
    return icosphere(0, order=2, fine_order=3)


def test_defaults_by_dimension(curve, surface):
    c = QBXLayerPotentialSource(curve, fine_order=16, qbx_order=4, fmm_order=10)
    s = QBXLayerPotentialSource(surface, fine_order=6, qbx_order=3, fmm_order=8)
    # :290-304 max_leaf_refine_weight 64 (2D) / 512 (3D)
    assert c._max_leaf_refine_weight == 64 and s._max_leaf_refine_weight == 512
    # :303-310 from_sep_smaller_min_nsources_cumul 15 (curves) / 30 (surfaces)
    assert c._from_sep_smaller_min_nsources_cumul == 15
    assert s._from_sep_smaller_min_nsources_cumul == 30
    for src in (c, s):
        assert src._expansion_stick_out_factor == 0.5
        assert src._well_sep_is_n_away == 2
        assert src._box_extent_norm == "l2"
        assert src.target_association_tolerance == pytest.approx(1e3 * 2.220446049250313e-16)
        assert src._use_target_specific_qbx is False


def test_constructor_errors(curve):
    with pytest.raises(ValueError):
        QBXLayerPotentialSource(curve, qbx_order=4, fmm_order=10)
    with pytest.raises(ValueError):
        QBXLayerPotentialSource(curve, fine_order=16, fmm_order=10)
    with pytest.raises(TypeError):
        QBXLayerPotentialSource(curve, fine_order=16, qbx_order=4)
    with pytest.raises(TypeError):
        QBXLayerPotentialSource(curve, fine_order=16, qbx_order=4, fmm_order=10,
                                fmm_level_to_order=lambda *a: 3)


def test_copy(curve):
    src = QBXLayerPotentialSource(curve, fine_order=16, qbx_order=4, fmm_order=10,
                                  _max_leaf_refine_weight=32, _expansion_stick_out_factor=0.3)
    c = src.copy()
    assert type(c) is type(src)
    for attr in ("density_discr", "fine_order", "qbx_order", "fmm_level_to_order",
                 "target_association_tolerance", "_max_leaf_refine_weight",
                 "_expansion_stick_out_factor", "_well_sep_is_n_away", "_box_extent_norm",
                 "_from_sep_smaller_min_nsources_cumul", "_use_target_specific_qbx"):
        assert getattr(c, attr) == getattr(src, attr), attr
    c2 = src.copy(qbx_order=6, fmm_order=12, _expansion_stick_out_factor=0)
    assert c2.qbx_order == 6 and c2._expansion_stick_out_factor == 0
    assert c2.fmm_level_to_order("laplace", {}, None, 3) == 12
    assert c2._max_leaf_refine_weight == 32
    # fmm_order=False gives direct QBX, and copies keep it
    d = src.copy(fmm_order=False)
    assert d._direct and d.copy()._direct and not src._direct
    with pytest.raises(TypeError):
        src.copy(fmm_order=8, fmm_level_to_order=lambda *a: 8)
    # the target-association tolerance may be replaced, 0 included
    assert src.copy(target_association_tolerance=0).target_association_tolerance == 0
