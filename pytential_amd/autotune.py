"""Autotuner for the work-unit shapes of the QBX-FMM evaluation.

The library's tunable shapes (``QBX_TUNE_*`` in include/qbx_mi355x.h) are
the list items one lane walks in the work units of P2P, M2P, P2QBXL and
M2QBXL (so the lanes per target and the units per box), and the lanes per
center of the one-density 2D Helmholtz L2QBXL.  Their best values depend on
the workload: how uneven the boxes' lists are, the expansion orders, how many
targets a box holds.  ``autotune`` searches them on the caller's own
evaluation -- coordinate descent over each knob's candidate values, each
candidate timed as the median of a few fused evaluations on a geometry built
with it -- keeps the fastest, and caches the result per workload signature in
a JSON file.  ``apply`` sets a result on a context; the unit plans take it at
the next geometry build.  Every setting gives the same potentials up to the
order of the last additions.
"""
from __future__ import annotations

import ctypes as ct
import hashlib
import json
import os
import time

import numpy as np

from . import _lib

KNOBS = {
    "p2p_unit_iter": (0, (32, 64, 128, 256, 512)),
    "m2p_unit_iter": (1, (2, 4, 8, 16, 32)),
    "p2qbxl_unit_iter": (2, (8, 16, 32, 64, 128)),
    "m2qbxl_unit_iter": (3, (1, 2, 4, 8)),
    "l2qbxl_lanes": (4, (1, 2, 4, 8)),
}
# settings that are not searched (the caller picks them): the stream layout
# of qbx_drive_fmm, 4 (its stage DAG on four streams) or 1 (the caller's
# stream only, for contexts that share a graph)
SETTINGS = {
    "drive_streams": 5,
}


def get(ctx):
    """The knob values currently set on a context (``geometry.Context``)."""
    out = {}
    for name, (pid, _) in KNOBS.items():
        v = ct.c_int()
        _lib.call("qbx_ctx_get_tuning", ctx.h, pid, ct.byref(v))
        out[name] = v.value
    return out


def apply(ctx, settings):
    """Set knob values (a dict name -> value) on a context."""
    for name, value in settings.items():
        if name in KNOBS:
            pid = KNOBS[name][0]
        elif name in SETTINGS:
            pid = SETTINGS[name]
        else:
            raise ValueError(f"unknown tuning knob {name!r}")
        _lib.call("qbx_ctx_set_tuning", ctx.h, pid, int(value))


def signature(src, targets_and_sides, kernel, k, source_kind, nrhs):
    """Cache key of a workload: kernel, sizes, orders and the device."""
    import torch
    d = src.density_discr
    parts = [kernel, f"{float(k):.6g}", source_kind, str(nrhs), str(d.ndofs),
             str(getattr(d, "nsources", 0)), str(src.qbx_order),
             torch.cuda.get_device_name(src.ctx.device)]
    for t, s in targets_and_sides:
        parts.append(f"{t if isinstance(t, str) else np.asarray(t).shape}:{s}")
    return hashlib.sha1("|".join(parts).encode()).hexdigest()[:16]


def _cache_path(path):
    if path is not None:
        return path
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "pytential_amd", "tuning.json")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def coordinate_descent(knobs, start, measure, passes=1):
    """Minimise ``measure(settings)`` over ``knobs`` (name -> candidates)
    one knob at a time from ``start``; returns (best settings, best time,
    every (settings, time) tried).  ``measure`` is called once per distinct
    setting."""
    best = dict(start)
    seen = {}

    def timed(s):
        key = tuple(sorted(s.items()))
        if key not in seen:
            seen[key] = measure(dict(s))
        return seen[key]

    t_best = timed(best)
    for _ in range(passes):
        improved = False
        for name, cands in knobs.items():
            for v in cands:
                if v == best[name]:
                    continue
                trial = dict(best, **{name: v})
                t = timed(trial)
                if t < t_best:
                    best, t_best, improved = trial, t, True
        if not improved:
            break
    return best, t_best, [(dict(k), t) for k, t in seen.items()]


def autotune(src, targets_and_sides, kernel, k=0.0, source_kind="S", density=None, nrhs=1,
             reps=5, warmup=2, knobs=None, cache=None, use_cache=True):
    """Search the knobs on this workload (``src.exec_compute_potential``'s
    arguments; ``density`` defaults to random values) and set the best on
    ``src.ctx``.  Returns ``{"settings", "ms", "default_ms", "trials",
    "cached"}``; a cached result for the same signature is reused unless
    ``use_cache`` is false."""
    import torch

    from . import fmm as F
    from .geometry import QBXFMMGeometryData
    path = _cache_path(cache)
    sig = signature(src, targets_and_sides, kernel, k, source_kind, nrhs)
    db = _load(path)
    if use_cache and sig in db:
        apply(src.ctx, db[sig]["settings"])
        return dict(db[sig], cached=True)
    knobs = dict(KNOBS if knobs is None else {n: KNOBS[n] for n in knobs})
    d = src.density_discr
    if density is None:
        rng = np.random.default_rng(0)
        shape = (d.ndofs,) if nrhs == 1 else (d.ndofs, nrhs)
        density = rng.standard_normal(shape)
        if kernel == "helmholtz":
            density = density + 1j * rng.standard_normal(shape)
    dens = torch.as_tensor(density, device="cuda")
    pairs = [(t, 0 if s is None else s) for t, s in targets_and_sides]
    start = get(src.ctx)
    normals = src.dev["quad_normals"] if source_kind == "D" else None

    def measure(settings):
        apply(src.ctx, settings)
        geo = QBXFMMGeometryData(src, tuple(pairs))
        geo.build_all()
        w = src.make_wrangler(geo, kernel, k, source_kind, nrhs)
        st, nrm = w.reorder_sources(src.quad_density(dens), src.dev["quad_weights"], normals)
        for _ in range(warmup):
            F.drive_fmm_fused(w, st, nrm)
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            F.drive_fmm_fused(w, st, nrm)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    cands = {name: vals for name, (_, vals) in knobs.items()}
    best, t_best, trials = coordinate_descent(cands, {n: start[n] for n in cands}, measure)
    default_ms = [t for s, t in trials if s == {n: start[n] for n in cands}][0]
    apply(src.ctx, best)
    res = {"settings": best, "ms": t_best, "default_ms": default_ms,
           "trials": [{"settings": s, "ms": t} for s, t in trials], "cached": False}
    db[sig] = {k_: v for k_, v in res.items() if k_ != "cached"}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass
    return res
