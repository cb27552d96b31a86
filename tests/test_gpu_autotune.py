"""GPU: the tuning knobs and the autotuner.

* qbx_ctx_set_tuning / qbx_ctx_get_tuning round-trip every candidate and
  refuse values outside the library's range;
* every knob setting gives the default evaluation's potentials (2D Helmholtz
  D on the config-1 starfish with a grid): within 1e-13 for the unit-plan
  knobs (the order of the last additions changes), bit for bit for the
  L2QBXL lane split;
* autotune() returns a measured setting no slower than its own default
  measurement, applies it, and reuses its cache for the same workload.
"""
import functools

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@functools.lru_cache(maxsize=1)
def _config():
    import oracle

    from pytential_amd import problems
    return problems.refine_until_associated(problems.config1(nelements=48, grid_n=24),
                                            oracle.assoc_status_and_flags)


def _eval(src, pairs, dens):
    from pytential_amd.geometry import QBXFMMGeometryData
    from pytential_amd import fmm as F
    geo = QBXFMMGeometryData(src, tuple(pairs))
    w = src.make_wrangler(geo, "helmholtz", 6.0, "D", 1)
    st, nrm = w.reorder_sources(src.quad_density(dens), src.dev["quad_weights"],
                                src.dev["quad_normals"])
    return F.drive_fmm_fused(w, st, nrm).cpu().numpy()


def test_knobs_round_trip_and_range():
    from pytential_amd import _lib, autotune
    from pytential_amd.geometry import Context
    ctx = Context(0)
    for name, (_, cands) in autotune.KNOBS.items():
        for v in cands:
            autotune.apply(ctx, {name: v})
            assert autotune.get(ctx)[name] == v
        for bad in (0, 3, 4096):
            with pytest.raises(_lib.QBXError):
                autotune.apply(ctx, {name: bad})


def test_every_setting_gives_the_default_potentials():
    from pytential_amd import autotune
    from pytential_amd.geometry import Context
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = _config()
    d = cfg.discr
    src = QBXLayerPotentialSource(d, fine_order=d.fine_order, qbx_order=4, fmm_order=14)
    src._ctx = Context(0)
    pairs = [("self", +1), (cfg.targets[:, d.ndofs:], 0)]
    rng = np.random.default_rng(5)
    dens = torch.as_tensor(rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs),
                           device="cuda")
    base = autotune.get(src.ctx)
    ref = _eval(src, pairs, dens)
    for name, (_, cands) in autotune.KNOBS.items():
        for v in cands:
            autotune.apply(src.ctx, dict(base, **{name: v}))
            got = _eval(src, pairs, dens)
            err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
            print(f"[tune] {name}={v}: rel diff {err:.1e}", flush=True)
            if name == "l2qbxl_lanes":
                assert err == 0.0
            else:
                assert err < 1e-13
    autotune.apply(src.ctx, base)


def test_autotune_search_and_cache(tmp_path):
    from pytential_amd import autotune
    from pytential_amd.geometry import Context
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = _config()
    d = cfg.discr
    src = QBXLayerPotentialSource(d, fine_order=d.fine_order, qbx_order=4, fmm_order=14)
    src._ctx = Context(0)
    pairs = [("self", +1), (cfg.targets[:, d.ndofs:], 0)]
    cache = str(tmp_path / "tuning.json")
    res = autotune.autotune(src, pairs, "helmholtz", 6.0, "D", reps=3, cache=cache,
                            knobs=["p2p_unit_iter", "l2qbxl_lanes"])
    print(f"[tune] best {res['settings']} {res['ms']:.3f} ms (default "
          f"{res['default_ms']:.3f} ms, {len(res['trials'])} trials)", flush=True)
    assert not res["cached"] and res["ms"] <= res["default_ms"]
    assert autotune.get(src.ctx)["p2p_unit_iter"] == res["settings"]["p2p_unit_iter"]
    again = autotune.autotune(src, pairs, "helmholtz", 6.0, "D", reps=3, cache=cache,
                              knobs=["p2p_unit_iter", "l2qbxl_lanes"])
    assert again["cached"] and again["settings"] == res["settings"]
