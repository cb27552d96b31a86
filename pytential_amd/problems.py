"""Synthetic QBX workloads (geometry, density, targets) built from a seed.

Seeds follow ``numpy.random.default_rng(1000 + config_index)``.  Every
builder returns a :class:`QBXProblem`; nothing here touches the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .discretization import (CurveDiscretization, discretize_curve, grid_targets,
                             starfish)


@dataclass
class QBXProblem:
    name: str
    kernel: str                 # "laplace" | "helmholtz"
    k: float
    source_kind: str            # "S" | "D"
    qbx_order: int
    fmm_order: int
    discr: CurveDiscretization = field(repr=False)
    density: np.ndarray = field(repr=False)            # stage-1 nodal values
    targets: np.ndarray = field(repr=False)            # (2, nt) user targets
    target_sides: np.ndarray = field(repr=False)       # (nt,) int32 side flags
    seed: int = 0

    @property
    def dtype(self):
        return np.complex128 if self.kernel == "helmholtz" else np.float64

    @property
    def ntargets(self) -> int:
        return self.targets.shape[1]


def _targets_with_surface(discr, grid_n, surface_side, lo=-2.0, hi=2.0):
    grid = grid_targets(grid_n, lo, hi)
    tg = np.concatenate([discr.nodes, grid], axis=1)
    sides = np.concatenate([np.full(discr.ndofs, surface_side, np.int32),
                            np.zeros(grid.shape[1], np.int32)])
    return tg, sides


def config1(nelements=64, order=8, fine_order=16, grid_n=128, fmm_order=10,
            qbx_order=4, surface_side=+1):
    """2D Laplace S on the 5-arm starfish r = 1 + 0.25 cos 5t."""
    rng = np.random.default_rng(1000 + 0)
    discr = discretize_curve(starfish(5, 0.25), nelements, order, fine_order)
    t = discr.t_nodes
    sigma = np.sin(3 * t) + 0.1 * rng.standard_normal(discr.ndofs)
    tg, sides = _targets_with_surface(discr, grid_n, surface_side)
    return QBXProblem("laplace2d-S-starfish", "laplace", 0.0, "S", qbx_order,
                      fmm_order, discr, sigma, tg, sides, seed=1000)


def config2(nelements=4096, order=8, fine_order=16, grid_n=512, fmm_order=15,
            qbx_order=5, k=20.0, surface_side=+1):
    """2D Helmholtz k=20 D on a seeded starfish (arms 5..9, amplitude .2-.3)."""
    rng = np.random.default_rng(1000 + 1)
    arms = int(rng.integers(5, 10))
    amp = float(rng.uniform(0.2, 0.3))
    discr = discretize_curve(starfish(arms, amp), nelements, order, fine_order)
    sigma = (rng.standard_normal(discr.ndofs)
             + 1j * rng.standard_normal(discr.ndofs))
    tg, sides = _targets_with_surface(discr, grid_n, surface_side)
    p = QBXProblem(f"helmholtz2d-D-starfish{arms}", "helmholtz", k, "D", qbx_order,
                   fmm_order, discr, sigma, tg, sides, seed=1001)
    p.arms, p.amplitude = arms, amp
    return p


def refine_until_associated(problem: QBXProblem, assoc_fn, max_iter: int = 8) -> QBXProblem:
    """Split the elements that ``QBXTargetAssociationFailedError.refine_flags``
    names (pytential/qbx/target_assoc.py:751-825, 915-950) until every target
    finds a QBX center.  ``assoc_fn(discr, targets, sides) -> (status,
    refine_flags)``.  On-surface targets are the stage-1 nodes, so they move
    with the refinement; the density is carried across by the per-panel
    interpolation connection."""
    p = problem
    discr, dens = p.discr, p.density
    nsurf = discr.ndofs
    surf_side = int(p.target_sides[0]) if nsurf else 0
    vol = p.targets[:, nsurf:]
    vol_sides = p.target_sides[nsurf:]
    for _ in range(max_iter):
        tg = np.concatenate([discr.nodes, vol], axis=1)
        sides = np.concatenate([np.full(discr.ndofs, surf_side, np.int32), vol_sides])
        status, flags = assoc_fn(discr, tg, sides)
        if not np.any(status == 1):
            out = QBXProblem(p.name, p.kernel, p.k, p.source_kind, p.qbx_order,
                             p.fmm_order, discr, dens, tg, sides, seed=p.seed)
            out.__dict__.update({k: v for k, v in p.__dict__.items()
                                 if k not in out.__dict__})
            out.nrefined = discr.nelements - p.discr.nelements
            return out
        discr, conn = discr.refined(flags)
        dens = conn(dens)
    raise RuntimeError("target association still fails after refinement")


def sphere_harmonic_densities(nodes, nrhs, noise=0.01, seed=1002):
    """Config-3 densities on surface nodes (3, n): real spherical harmonics
    Y_lm (l <= 6, taken in order l = 0.., m = -l..l, cyclically) plus
    N(0, noise^2) noise from ``default_rng(seed)``.  Returns (n, nrhs) and
    the degree l of each column."""
    from scipy.special import sph_harm_y
    x = np.asarray(nodes)
    r = np.linalg.norm(x, axis=0)
    th, ph = np.arccos(np.clip(x[2] / r, -1, 1)), np.arctan2(x[1], x[0])
    lm = [(l, m) for l in range(7) for m in range(-l, l + 1)]
    rng = np.random.default_rng(seed)
    cols, degs = [], []
    for i in range(nrhs):
        l, m = lm[i % len(lm)]
        y = sph_harm_y(l, abs(m), th, ph)
        y = y.real if m >= 0 else y.imag
        cols.append(y + noise * rng.standard_normal(len(y)) if noise else y)
        degs.append(l)
    return np.stack(cols, axis=1), np.array(degs)


def dirichlet_curves(ncurves=64, seed=1003):
    """Config-4 curve specs: ``ncurves`` starfish (arms 3..9, radius 0.5..2,
    amplitude 0.25, element counts log-uniform in 256..4096, random phase)
    and one point charge outside each, from ``default_rng(seed)``.  Returns a
    list of dicts (arms, radius, nelements, phase, charge) and the generator
    (callers draw their check points from it)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(ncurves):
        arms = int(rng.integers(3, 10))
        rad = float(rng.uniform(0.5, 2.0))
        nel = int(round(np.exp(rng.uniform(np.log(256), np.log(4096)))))
        phase = float(rng.uniform(0, 2 * np.pi))
        ang = rng.uniform(0, 2 * np.pi)
        x0 = 1.6 * rad * np.array([np.cos(ang), np.sin(ang)])
        out.append({"arms": arms, "radius": rad, "nelements": nel, "phase": phase,
                    "charge": x0})
    return out, rng


def exterior_shell_targets(n, rng, rmin=2.0, rmax=3.0):
    """Config-5 targets: n points with uniform directions and radius uniform
    in [rmin, rmax]."""
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    return v * rng.uniform(rmin, rmax, n)
