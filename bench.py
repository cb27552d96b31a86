#!/usr/bin/env python3
"""QBX-FMM layer-potential throughput on one MI355X.

Workload (the metric's configuration, 2D Helmholtz): k = 20 double layer on a
seeded starfish, 4096 panels x 8 Gauss-Legendre nodes (split where target
association needs it), 16-node oversampled sources, qbx_order 5, complex128
density; targets = a 512 x 512 grid on [-2, 2]^2 plus every on-surface node.

A step is one evaluation with the geometry (tree, traversal, association)
cached, as the reference memoizes it: density upsampling, strengths, every
FMM and QBX stage through the single-call driver (qbx_drive_fmm, the
reference's fmm_driver seam; its independent stages run on three streams),
and the potential at all user targets in user order.  The per-stage times
come from a stage-by-stage run of the same evaluation (one stream, HIP
events between stages).

value   = user target points / step time, FMM orders chosen per level for a
          1e-10 relative error against the FP64 direct-QBX oracle
roofline: dominant kernel's algorithmic flops per launch / its average
          duration (HIP events on the launch stream) against FP64 peak
cpu_baseline: the oracle (C + OpenMP restatement of the same algorithm)
          timed on this host on one full evaluation; its direct-QBX leg
          supplies the accuracy reference on a sample of targets
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("QBX-FMM points/s at 1e-10 rel err vs FP64 oracle, "
          "GPU batched vs plain vs CPU")
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

# algorithmic flop model (per unit); special functions J0/J1/Y0/Y1 count as
# SPECIAL flops each (a minimal rational/asymptotic evaluation), complex
# multiply-add 8, complex multiply 6
SPECIAL = 40


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_helmholtz2d_starfish.json")


def profiled_stage(stage):
    """The committed rocprofv3 record of a stage's kernel on this workload:
    HBM bytes per evaluation from FETCH_SIZE / WRITE_SIZE (scaled by the copy
    calibration recorded there) and the hardware FP64 flop count from
    SQ_INSTS_VALU_FLOPS_FP64; None when the summary lacks the stage."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)["stages"].get(stage)
    except (OSError, ValueError, KeyError):
        return None


def measured_traffic(stage):
    rec = profiled_stage(stage)
    return None if not rec else int(rec["read_bytes"] + rec["write_bytes"])


def build_problem(grid_n, nelements):
    from pytential_amd import problems
    from pytential_amd.geometry import assoc_status_and_flags
    cfg = problems.config2(nelements=nelements, grid_n=grid_n)
    return problems.refine_until_associated(cfg, assoc_status_and_flags)


def interaction_counts(geo):
    """Per-stage interaction counts from the traversal (host copies)."""
    t, tr, g = geo.tree(), geo.traversal(), geo._geo

    def h(x):
        return x.cpu().numpy().astype(np.int64)

    def per_list(starts, lists, weight):
        """sum of weight[list box] over each CSR row"""
        if len(lists) == 0:
            return np.zeros(len(starts) - 1, np.int64)
        c = np.concatenate([[0], np.cumsum(weight[lists])])
        return c[starts[1:]] - c[starts[:-1]]

    scount = h(t.box_source_counts_nonchild)
    scumul = h(t.box_source_counts_cumul)
    levels = h(t.box_levels)
    l1s, l1l = h(tr.neighbor_source_boxes_starts), h(tr.neighbor_source_boxes_lists)
    l2s, l2l = h(tr.from_sep_siblings_starts), h(tr.from_sep_siblings_lists)
    l3s, l3l = h(tr.from_sep_smaller_starts), h(tr.from_sep_smaller_lists)
    l4s, l4l = h(tr.from_sep_bigger_starts), h(tr.from_sep_bigger_lists)
    tboxes = h(tr.target_boxes)
    ttp = h(tr.target_or_target_parent_boxes)
    sboxes = h(tr.source_boxes)
    spar = h(tr.source_parent_boxes)
    child = h(t.box_child_ids).reshape(-1, t.nboxes) if t.nboxes else np.zeros((4, 0), np.int64)
    parent = h(t.box_parent_ids)
    nq = geo.non_qbx_box_target_lists()
    nfilt_box = h(nq["box_target_counts_nonchild"])[tboxes]
    src_in_l1 = per_list(l1s, l1l, scount)
    n_l3 = np.diff(l3s)
    gc = h(geo.global_qbx_centers())
    c2tb = h(geo.qbx_center_to_target_box())[gc]
    ctt_starts = h(geo.center_to_tree_targets()[0])
    return {
        "p2p_pairs": int(np.sum(nfilt_box * src_in_l1)),
        "p2qbxl_pairs": int(np.sum(src_in_l1[c2tb])),
        # (QBX target, list-1 source of its center's box) pairs
        "tsqbx_pairs": int(np.sum(np.diff(ctt_starts)[gc] * src_in_l1[c2tb])),
        "m2qbxl_levels": levels[l3l] if len(l3l) else np.zeros(0, np.int64),
        "l2qbxl": int(len(gc)),
        "qbxl2p": int(g.nctt),
        "nglobal": int(g.nglobal), "nfiltered": int(g.nfiltered),
        "c2tb": c2tb, "l3s": l3s, "l3l": l3l, "levels": levels,
        "scount": scount, "scumul": scumul, "tboxes": tboxes, "nfilt_box": nfilt_box,
        "ttp": ttp, "l2s": l2s, "l2l": l2l, "l4s": l4s, "l4l": l4l, "sboxes": sboxes,
        "spar": spar, "child": child, "parent": parent, "n_l3": n_l3,
    }


def _miller_steps(p):
    m = max(p, 1)
    return m + 16 + int(np.sqrt(40.0 * m + 1.0))


def stage_flops(counts, orders, pq, SPECIAL=1):
    """Algorithmic flops per stage launch (Helmholtz 2D, double layer).  A
    Bessel/Hankel function value of order 0 or 1 (and atan2) counts SPECIAL
    flops (1 in the measurement plan's convention, where every special
    function is one flop), a complex multiply-add 8, a complex multiply 6."""
    o = np.asarray(orders, np.int64)
    lv = counts["levels"]
    nq = 2 * pq + 1
    # P2P D: r, proj (10), H1 = J1 + i Y1 (2 special), scale and accumulate (12)
    p2p = counts["p2p_pairs"] * (22 + 2 * SPECIAL)
    # P2QBXL D: rho/theta (6 + atan2 as special), H0,H1 (4 special), recurrence to
    # pq+1 (6 each), phases (6 each), 2*nq complex MACs for the dipole
    p2qbxl = counts["p2qbxl_pairs"] * (6 + 5 * SPECIAL + 12 * (pq + 1) + 2 * 8 * nq)
    # M2QBXL: per (center, box): H0,H1 + atan2, recurrence to p+pq (8 each),
    # 2*nq complex MACs per order
    m2q_lv = counts["m2qbxl_levels"]
    c2tb, l3s = counts["c2tb"], counts["l3s"]
    per_box = o[m2q_lv] if len(m2q_lv) else np.zeros(0, np.int64)
    m2q = 0
    if len(per_box):
        csum = np.concatenate([[0], np.cumsum(5 * SPECIAL + (per_box + pq + 1) *
                                              (8 + 2 * nq * 8))])
        m2q = int(np.sum(csum[l3s[c2tb + 1]] - csum[l3s[c2tb]]))
    # M2L (table rows): (2p+1)^2 complex MACs per (box, list-2 box)
    ttp, l2s = counts["ttp"], counts["l2s"]
    pt = o[lv[ttp]]
    m2l = int(np.sum(np.diff(l2s) * (2 * pt + 1) ** 2 * 8))
    # P2L D: per (box, list-4 source): H0,H1 (4 special), recurrence and phases
    # to p+1 (4 + 12 each), b_n from two neighbours and one MAC (6 + 6 + 8 each)
    scount = counts["scount"]
    l4s, l4l = counts["l4s"], counts["l4l"]
    c4 = np.concatenate([[0], np.cumsum(scount[l4l])]) if len(l4l) else np.zeros(1, np.int64)
    nsrc4 = c4[l4s[1:]] - c4[l4s[:-1]] if len(l4l) else np.zeros(len(ttp), np.int64)
    p2l = int(np.sum(nsrc4 * (4 * SPECIAL + 16 * (pt + 2) + 20 * (2 * pt + 1))))
    # P2M D: per leaf source: Miller's recurrence (4 per step), normalisation and
    # phases (16 each to p+1), b_n and one MAC (20 each)
    sb = counts["sboxes"]
    ps = o[lv[sb]]
    steps = np.array([_miller_steps(int(p) + 1) for p in ps], np.int64)
    p2m = int(np.sum(scount[sb] * (4 * steps + 16 * (ps + 2) + 20 * (2 * ps + 1))))
    # M2M: (2 p_parent + 1)(2 p_child + 1) complex MACs per nonempty child
    spar, child, scumul = counts["spar"], counts["child"], counts["scumul"]
    m2m = 0
    for P in spar:
        pp = o[lv[P]]
        for c in child[:, P]:
            if c >= 0 and scumul[c] > 0:
                m2m += (2 * pp + 1) * (2 * o[lv[c]] + 1) * 8
    # L2L: (2p+1)(2 p_parent + 1) complex MACs per ttp box below the root
    par = counts["parent"][ttp]
    has = par >= 0
    l2l = int(np.sum((2 * pt[has] + 1) * (2 * o[lv[par[has]]] + 1) * 8))
    # M2P: per (target, list-3 box): H0,H1 (4 special), per order a recurrence
    # step, a phase step and two complex MACs (4 + 6 + 16)
    tb, l3l, nf = counts["tboxes"], counts["l3l"], counts["nfilt_box"]
    w3 = np.zeros(len(l3s) - 1, np.int64)
    if len(l3l):
        c3 = np.concatenate([[0], np.cumsum(4 * SPECIAL + 26 * o[lv[l3l]])])
        w3 = c3[l3s[1:]] - c3[l3s[:-1]]
    m2p = int(np.sum(nf * w3))
    return {"p2p": p2p, "p2qbxl": p2qbxl, "m2qbxl": m2q, "m2l": m2l, "p2l": p2l,
            "p2m": p2m, "m2m": int(m2m), "l2l": l2l, "m2p": m2p}


def run_gpu(args, cfg, orders_fn, label, tune=False):
    import torch

    from pytential_amd import autotune
    from pytential_amd import fmm as F
    from pytential_amd.geometry import Context
    from pytential_amd.qbx import QBXLayerPotentialSource
    src = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                  qbx_order=cfg.qbx_order, fmm_level_to_order=orders_fn)
    tuned = None
    if tune and args.tuning:
        # fixed settings (a profiled run repeats the headline's tuned shapes
        # without the search's trial launches in its trace)
        src._ctx = Context(torch.cuda.current_device())
        autotune.apply(src.ctx, json.loads(args.tuning))
        tuned = {"settings": autotune.get(src.ctx), "ms": float("nan"),
                 "default_ms": float("nan"), "trials": [], "search_s": 0.0}
    elif tune:
        # the work-unit shapes searched on this very evaluation (its own
        # context, so the other workloads keep the defaults); the timed
        # evaluation below runs on a geometry built with the result
        src._ctx = Context(torch.cuda.current_device())
        t0 = time.perf_counter()
        tuned = autotune.autotune(src, [(cfg.targets, cfg.target_sides)], cfg.kernel, cfg.k,
                                  cfg.source_kind, density=cfg.density, reps=5,
                                  use_cache=False)
        tuned["search_s"] = time.perf_counter() - t0
        log(f"[{label}] autotune: {tuned['settings']} {tuned['ms']:.3f} ms (defaults "
            f"{tuned['default_ms']:.3f} ms, {len(tuned['trials'])} trials, "
            f"{tuned['search_s']:.1f}s)")
    geo = src.qbx_fmm_geometry_data(((cfg.targets, cfg.target_sides),))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    geo.build_all()
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    w = src.make_wrangler(geo, cfg.kernel, cfg.k, cfg.source_kind)
    orders = list(map(int, w._lo))
    log(f"[{label}] geometry {cold:.3f}s nboxes={geo.tree().nboxes} "
        f"nlevels={geo.tree().nlevels} orders={orders}")
    dens = torch.as_tensor(cfg.density, dtype=torch.complex128, device="cuda")

    stages = ["strengths", "p2m", "m2m", "p2p", "m2l", "m2p", "p2l", "l2l", "l2p",
              "p2qbxl", "m2qbxl", "l2qbxl", "qbxl2p", "finalize"]

    def step(events=None):
        def mark(i):
            if events is not None:
                events[i].record()
        mark(0)
        qd = src.quad_density(dens)
        st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
        mark(1)
        m = w.form_multipoles(st, nrm)
        mark(2)
        w.coarsen_multipoles(m)
        mark(3)
        pot = w.eval_direct(st, nrm)
        mark(4)
        loc = w.multipole_to_local(m)
        mark(5)
        w.eval_multipoles(m, pot)
        mark(6)
        w.form_locals(st, nrm, loc)
        mark(7)
        w.refine_locals(loc)
        mark(8)
        w.eval_locals(loc, pot)
        mark(9)
        q = w.form_global_qbx_locals(st, nrm)
        mark(10)
        w.translate_box_multipoles_to_qbx_local(m, q)
        mark(11)
        w.translate_box_local_to_qbx_local(loc, q)
        mark(12)
        w.eval_qbx_expansions(q, pot)
        mark(13)
        out = w.finalize_potentials(pot)
        mark(14)
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = step(evs[k])
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    per_stage = {s: float(np.mean([e[i].elapsed_time(e[i + 1]) for e in evs]))
                 for i, s in enumerate(stages)}
    # the single-call driver (the timed evaluation of the headline)
    def fused_step():
        qd = src.quad_density(dens)
        st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
        return F.drive_fmm_fused(w, st, nrm)

    for _ in range(max(args.warmup, 1)):
        fused_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        fused = fused_step()
    torch.cuda.synchronize()
    dt_fused = (time.perf_counter() - t1) / args.steps
    # the same evaluation (density -> potentials) captured once into a HIP
    # graph and replayed: the ~60 kernels on four streams start as soon as
    # their inputs are ready instead of when the host reaches their launch
    dt_graph, graph_diff = None, None
    if not args.no_graph:
        from pytential_amd.linalg import MatVecGraph

        def fused_of(dd):
            qd = src.quad_density(dd)
            st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
            return F.drive_fmm_fused(w, st, nrm)
        gph = MatVecGraph(fused_of, torch.as_tensor(dens, device="cuda"))
        for _ in range(max(args.warmup, 1)):
            gph.graph.replay()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        for _ in range(args.steps):
            gph.graph.replay()
        torch.cuda.synchronize()
        dt_graph = (time.perf_counter() - t3) / args.steps
        gout = gph._out.cpu().numpy()
        del gph
        torch.cuda.synchronize()
        src.ctx.release_graphs()
    # the adjoint (conjugate-transposed) QBX-FMM on the same tree and orders:
    # the backward pass of the density gradient, checked against the forward
    # by Re <y', A x> = Re <A^H y', x>
    qd = src.quad_density(dens)
    st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
    fx = F.drive_fmm_fused(w, st, nrm)
    rng = np.random.default_rng(90)
    ya = torch.as_tensor(rng.standard_normal(fx.numel()) + 1j * rng.standard_normal(fx.numel()),
                         dtype=torch.complex128, device="cuda")
    xa = F.drive_fmm_adjoint(w, ya, nrm)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    nadj = max(2, args.steps // 3)
    for _ in range(nadj):
        xa = F.drive_fmm_adjoint(w, ya, nrm)
    torch.cuda.synchronize()
    dt_adj = (time.perf_counter() - t2) / nadj
    lhs = complex(torch.sum(torch.conj(ya) * fx)).real
    rhs = complex(torch.sum(torch.conj(xa) * st)).real
    adj_rel = abs(lhs - rhs) / float(torch.linalg.norm(ya) * torch.linalg.norm(fx))
    log(f"[{label}] adjoint {dt_adj * 1e3:.3f} ms, transpose identity {adj_rel:.1e}")
    counts = interaction_counts(geo)
    flops = stage_flops(counts, orders, cfg.qbx_order)
    flops40 = stage_flops(counts, orders, cfg.qbx_order, SPECIAL=40)
    staged = out.cpu().numpy()
    fused = fused.cpu().numpy()
    if dt_graph is not None:
        graph_diff = float(np.max(np.abs(gout - fused)) / np.max(np.abs(fused)))
        log(f"[{label}] graph replay {dt_graph * 1e3:.3f} ms (eager {dt_fused * 1e3:.3f} ms), "
            f"graph vs eager {graph_diff:.1e}")
    return {"label": label, "orders": orders, "cold_s": cold, "step_s": dt,
            "step_fused_s": dt_fused, "step_graph_s": dt_graph, "graph_vs_eager": graph_diff,
            "per_stage_ms": per_stage, "counts": counts,
            "adjoint_s": dt_adj, "adjoint_identity_rel": adj_rel, "autotune": tuned,
            "flops": flops, "flops40": flops40, "result": fused, "result_stagewise": staged,
            "fused_vs_stagewise": float(np.max(np.abs(fused - staged)) /
                                        np.max(np.abs(staged))),
            "geo": geo, "src": src}


def run_batched(args, cfg, src, geo, nrhs):
    """nrhs densities per evaluation on the same geometry and orders (the
    first is the workload's own density): a step is one evaluation of all of
    them through the single-call driver."""
    import torch

    from pytential_amd import fmm as F
    rng = np.random.default_rng(77)
    n = cfg.discr.ndofs
    extra = rng.standard_normal((n, nrhs - 1)) + 1j * rng.standard_normal((n, nrhs - 1))
    dens_np = np.concatenate([np.asarray(cfg.density, np.complex128)[:, None], extra], axis=1)
    dens = torch.as_tensor(dens_np, dtype=torch.complex128, device="cuda")
    w = src.make_wrangler(geo, cfg.kernel, cfg.k, cfg.source_kind, nrhs)

    def step():
        qd = src.quad_density(dens)
        st, nrm = w.reorder_sources(qd, src.dev["quad_weights"], src.dev["quad_normals"])
        return F.drive_fmm_fused(w, st, nrm)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"[batched nrhs={nrhs}] {dt * 1e3:.3f} ms/step")
    return {"nrhs": nrhs, "step_s": dt, "result0": out[:, 0].cpu().numpy()}


def cpu_baseline(args, cfg, orders, sample_idx):
    """The oracle on the host cores: one full evaluation (timed, on the
    process's CPU share and on one thread) plus the direct-QBX accuracy
    reference on a sample of targets."""
    import oracle
    oracle.build(with_ref=False)
    nthreads = cpu_threads(args)
    g = oracle.GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
    st = cfg.discr.strengths(cfg.density)
    t0 = time.perf_counter()
    pot, _ = g.fmm(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order, orders, st,
                   nthreads=nthreads)
    t_fmm = time.perf_counter() - t0
    t0 = time.perf_counter()
    ref = g.direct(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order, st, sample_idx,
                   nthreads=nthreads)
    t_direct = time.perf_counter() - t0
    one = None
    if not args.no_cpu_1thread:
        t0 = time.perf_counter()
        g.fmm(cfg.kernel, cfg.k, cfg.source_kind, cfg.qbx_order, orders, st, nthreads=1)
        t1 = time.perf_counter() - t0
        one = {"value": round(cfg.ntargets / t1, 1), "unit": "points/s", "cores": 1,
               "seconds": round(t1, 2)}
    return {"value": cfg.ntargets / t_fmm, "unit": "points/s", "cores": nthreads,
            "kind": "port",
            "sample": (f"one full evaluation of the same workload ({cfg.ntargets} targets, "
                       f"same tree/lists/orders) through the oracle (C + OpenMP), "
                       f"{t_fmm:.2f}s; direct-QBX reference on {len(sample_idx)} sampled "
                       f"targets, {t_direct:.2f}s"),
            "host": host_cpu(), "one_thread": one,
            "fmm_result_sample": pot[sample_idx], "direct_sample": ref, "t_fmm": t_fmm}


_CEILING = None


def fp64_ceiling():
    """FP64 throughput this GPU reaches without memory traffic (the probe
    pytential_amd/fp64_ceiling built beside the library): TFLOP/s of
    v_mfma_f64_16x16x4f64 and of v_fma_f64, run once per bench process."""
    global _CEILING
    if _CEILING is None:
        exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pytential_amd",
                           "fp64_ceiling")
        try:
            r = subprocess.run([exe], capture_output=True, text=True, timeout=60, check=True)
            _CEILING = json.loads(r.stdout.strip().splitlines()[-1])
        except (OSError, subprocess.SubprocessError, ValueError) as e:
            _CEILING = {"error": str(e)}
    return _CEILING


def host_cpu():
    """The host's CPU model, logical CPU count and the CPUs this process may
    use (lscpu's "Model name", nproc)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_threads(args):
    """Threads for the oracle's OpenMP legs: --cpu-threads, else this process's
    CPU share (at most 16, the host share of one GPU)."""
    if args.cpu_threads:
        return args.cpu_threads
    share = host_cpu()["usable_cpus"] or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        share = min(share, int(env))
    return max(1, min(share, 16))


def stage_flops_3d(counts, orders, pq, nrhs):
    """Algorithmic flops per stage launch for 3D Laplace S with nrhs real
    densities, in the per-unit convention of the measurement plan: +,-,* 1,
    FMA 2, each sqrt/division/special function 1; a complex multiply-add 8, a
    complex-by-real one 4.  ncq = (pq+1)^2 QBX coefficients, ncf = (p+1)^2
    box coefficients at the box's level."""
    o = np.asarray(orders, np.int64)
    ncq = (pq + 1) ** 2
    lv = counts["levels"]
    # P2QBXL: the pair's basis (r, angles, the harmonic recurrence: ~4 ncq +
    # 20) once, then a complex-by-real MAC per coefficient and density
    p2qbxl = counts["p2qbxl_pairs"] * (4 * ncq + 20 + 4 * ncq * nrhs)
    # target-specific sum: per (QBX target, source) pair the distances and the
    # angle cosine (~20), the Legendre recurrence and the power series (~8 per
    # order), then a real MAC per density
    tsqbx = counts["tsqbx_pairs"] * (20 + 8 * (pq + 1) + 2 * nrhs)
    # M2L for real densities: the conjugate-symmetric real-linear map of the
    # m >= 0 coefficients, (p+1)(p+2) x (p+1)(p+2) real multiply-adds per
    # (box, list-2 box) and density (the dense complex square is 3.2x that)
    ttc, l2s = counts["ttp"], counts["l2s"]
    nr2 = (o[lv[ttc]] + 1) * (o[lv[ttc]] + 2)
    m2l = int(np.sum(np.diff(l2s) * nr2 * nr2 * 2 * nrhs))
    # M2QBXL: dense (ncq x ncf) complex per (center, list-3 box) and density
    m2q_lv = counts["m2qbxl_levels"]
    c2tb, l3s = counts["c2tb"], counts["l3s"]
    m2q = 0
    if len(m2q_lv):
        csum = np.concatenate([[0], np.cumsum(8 * ncq * (o[m2q_lv] + 1) ** 2 * nrhs)])
        m2q = int(np.sum(csum[l3s[c2tb + 1]] - csum[l3s[c2tb]]))
    return {"p2qbxl": int(p2qbxl), "tsqbx": int(tsqbx), "m2l": m2l, "m2qbxl": m2q}


def run_sphere3d(args):
    """Config 3: 3D Laplace single layer on the unit sphere, nrhs3d densities
    in one evaluation: icosahedron refined nsub times, order-4 triangles,
    qbx_order 3, fmm_order 8 (and the order that reaches 1e-10), on-surface
    targets (side +1).  Returns the result dict."""
    import torch

    from pytential_amd import fmm as F
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    from pytential_amd.problems import sphere_harmonic_densities
    t0 = time.perf_counter()
    discr = icosphere(args.nsub, order=4, fine_order=args.fine3d)
    nrhs = args.nrhs3d
    dens_np, degs = sphere_harmonic_densities(discr.nodes, nrhs)
    log(f"[sphere3d] {discr.nelements} triangles, ndofs={discr.ndofs} "
        f"nsources={discr.nsources} ({time.perf_counter() - t0:.1f}s)")
    dens = torch.as_tensor(dens_np, dtype=torch.float64, device="cuda")
    steps = args.steps3d

    def make(order, tsq=True):
        # the list-1 part of the QBX targets by the target-specific sum, as
        # pytential's fmmlib backend does by default (_use_target_specific_qbx
        # None: "use if possible"); the expansion pipeline is timed beside it
        src = QBXLayerPotentialSource(discr, fine_order=args.fine3d, qbx_order=3,
                                      fmm_order=order, _use_target_specific_qbx=tsq)
        geo = src.qbx_fmm_geometry_data((("self", +1),))
        return src, geo

    def stepper(src, w, d):
        def step():
            qd = src.quad_density(d)
            st, _ = w.reorder_sources(qd, src.dev["quad_weights"], None)
            return F.drive_fmm_fused(w, st, None)
        return step

    def timed(step, nsteps, warm=1):
        for _ in range(warm):
            step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(nsteps):
            out = step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / nsteps, out

    src, geo = make(8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    geo.build_all()
    torch.cuda.synchronize()
    cold = time.perf_counter() - t0
    tree = geo.tree()
    log(f"[sphere3d] geometry {cold:.3f}s nboxes={tree.nboxes} nlevels={tree.nlevels}")
    w = src.make_wrangler(geo, "laplace", 0.0, "S", nrhs)
    dt, out = timed(stepper(src, w, dens), steps)
    # the stage-by-stage run below uses the stage entry points, told that the
    # strengths are real (the fused driver decides that on the device)
    w.kd.real_strengths = 1
    res = out.real.cpu().numpy()
    # stage by stage on one stream, HIP events between stages: per-stage times
    # and the roofline of the dominant kernel
    stages = ["p2m", "m2m", "p2p", "m2l", "m2p", "p2l", "l2l", "l2p", "p2qbxl", "m2qbxl",
              "l2qbxl", "qbxl2p", "tsqbx"]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
    qd = src.quad_density(dens)
    st, _ = w.reorder_sources(qd, src.dev["quad_weights"], None)
    per_stage = {s: 0.0 for s in stages}
    for rep in range(2):
        ev[0].record()
        m = w.form_multipoles(st, None)
        ev[1].record()
        w.coarsen_multipoles(m)
        ev[2].record()
        pot = w.eval_direct(st, None)
        ev[3].record()
        loc = w.multipole_to_local(m)
        ev[4].record()
        w.eval_multipoles(m, pot)
        ev[5].record()
        w.form_locals(st, None, loc)
        ev[6].record()
        w.refine_locals(loc)
        ev[7].record()
        w.eval_locals(loc, pot)
        ev[8].record()
        q = w.form_global_qbx_locals(st, None)
        ev[9].record()
        w.translate_box_multipoles_to_qbx_local(m, q)
        ev[10].record()
        w.translate_box_local_to_qbx_local(loc, q)
        ev[11].record()
        w.eval_qbx_expansions(q, pot)
        ev[12].record()
        pot.add_(w.eval_target_specific_qbx_locals(st, None))
        ev[13].record()
        torch.cuda.synchronize()
        if rep == 1:
            per_stage = {s: ev[i].elapsed_time(ev[i + 1]) for i, s in enumerate(stages)}
    counts = interaction_counts(geo)
    lo = list(map(int, w._lo))
    flops = stage_flops_3d(counts, lo, 3, nrhs)
    # P2QBXL is the zero fill of the expansions here: the target-specific sum
    # does its work
    flops.pop("p2qbxl")
    kernel_of = {"tsqbx": "k3_tsqbx_mr", "m2l": "k3_trans_real<0,8,2>",
                 "m2qbxl": "k3_m2qbxl_mfma"}
    dom = max(flops, key=lambda k: per_stage[k])
    ach = flops[dom] / (per_stage[dom] * 1e-3) / 1e12
    roofline = {"bound": "mfma", "compute": "fp64",
                "kernel": kernel_of[dom], "stage": dom, "achieved": round(ach, 3),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "flops_per_launch": int(flops[dom]),
                "avg_launch_ms": round(per_stage[dom], 4),
                "timing": "HIP events on the launch stream, stage-by-stage run of the same "
                          "32-density evaluation"}
    stage_rl = {k: {"ms": round(per_stage[k], 4), "gflop": round(flops[k] / 1e9, 3),
                    "frac_fp64": round(flops[k] / (per_stage[k] * 1e-3) / 1e12 /
                                       FP64_PEAK_TFLOPS, 4)} for k in flops}
    log(f"[sphere3d] p=8, {nrhs} densities: {dt * 1e3:.1f} ms; stages "
        f"{ {k: round(v, 2) for k, v in per_stage.items()} }")
    # the expansion pipeline (P2QBXL + QBXL2P for the list-1 part) on the
    # same tree: same result to rounding, for comparison
    src_e, _ = make(8, tsq=False)
    we = src_e.make_wrangler(geo, "laplace", 0.0, "S", nrhs)
    dte, oute = timed(stepper(src, we, dens), steps)
    expansion_pipeline = {"ms_per_step": round(dte * 1e3, 4),
                          "rel_diff_vs_target_specific": float(
                              np.max(np.abs(oute.real.cpu().numpy() - res)) /
                              np.max(np.abs(res)))}
    log(f"[sphere3d] expansion pipeline: {dte * 1e3:.1f} ms")
    # one density per evaluation
    w1 = src.make_wrangler(geo, "laplace", 0.0, "S", 1)
    dt1, out1 = timed(stepper(src, w1, dens[:, 0].contiguous()), steps)
    # S Y_lm = Y_lm / (2l + 1) for the noise-free harmonics (untimed)
    clean, _ = sphere_harmonic_densities(discr.nodes, nrhs, noise=0.0)
    s_clean = stepper(src, w, torch.as_tensor(clean, device="cuda"))().real.cpu().numpy()
    eig = {}
    for c in range(nrhs):
        e = float(np.max(np.abs(s_clean[:, c] - clean[:, c] / (2 * degs[c] + 1))) /
                  np.max(np.abs(clean[:, c] / (2 * degs[c] + 1))))
        eig[f"l={degs[c]}"] = max(eig.get(f"l={degs[c]}", 0.0), e)
    result = {
        "value": round(discr.ndofs * nrhs / dt, 1), "unit": "points*rhs/s",
        "ms_per_step": round(dt * 1e3, 4), "steps": steps, "nrhs": nrhs,
        "dtype": "f64",
        "config": {"workload": (f"3D Laplace S, unit sphere: icosahedron refined {args.nsub}x "
                                f"({discr.nelements} triangles), order-4 nodes "
                                f"({discr.ndofs}), quad {args.fine3d}x{args.fine3d} "
                                f"collapsed Gauss ({discr.nsources} sources), qbx_order 3, "
                                f"fmm_order 8, {nrhs} densities (Y_lm l<=6 + N(0,0.01) "
                                f"noise, seed 1002), on-surface targets, target-specific "
                                f"QBX for the list-1 part"),
                   "ntargets": discr.ndofs, "nboxes": tree.nboxes, "nlevels": tree.nlevels},
        "plain_one_density": {"value": round(discr.ndofs / dt1, 1), "unit": "points/s",
                              "ms_per_step": round(dt1 * 1e3, 4),
                              "rel_diff_vs_batched_rhs0": float(
                                  np.max(np.abs(out1.real.cpu().numpy() - res[:, 0])) /
                                  np.max(np.abs(res[:, 0])))},
        "S_Ylm_vs_Ylm_over_2l_plus_1_worst_rel_err": eig,
        "cold_geometry_s": round(cold, 4),
        "expansion_pipeline": expansion_pipeline,
        "per_stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "stage_roofline": stage_rl,
        "roofline": {**roofline, "measured_ceiling": fp64_ceiling()},
        "cpu_baseline": None,
    }
    if args.no_cpu:
        return result
    import oracle
    oracle.build(with_ref=False)
    nth = cpu_threads(args)
    g = oracle.GeometryData(discr, discr.nodes, np.ones(discr.ndofs, np.int32),
                            max_leaf_refine_weight=512)
    rng = np.random.default_rng(12)
    idx = np.sort(rng.choice(discr.ndofs, min(args.sample, discr.ndofs), replace=False))
    st0 = discr.strengths(dens_np[:, 0])
    t2 = time.perf_counter()
    ref = g.direct("laplace", 0.0, "S", 3, st0, idx, nthreads=nth).real
    t_dir = time.perf_counter() - t2
    scale = np.max(np.abs(ref))
    result["rel_err_vs_direct_qbx_rhs0"] = float(np.max(np.abs(res[idx, 0] - ref)) / scale)
    # the per-level order that reaches 1e-10 against direct QBX (same sample)
    sweep = {"8": result["rel_err_vs_direct_qbx_rhs0"]}
    p_ok = None
    for p in (10, 12, 13, 14):
        sp, gp = make(p)
        wp = sp.make_wrangler(gp, "laplace", 0.0, "S", 1)
        o = stepper(sp, wp, dens[:, 0].contiguous())().real.cpu().numpy()
        sweep[str(p)] = float(np.max(np.abs(o[idx] - ref)) / scale)
        log(f"[sphere3d] fmm_order {p}: rel err vs direct QBX {sweep[str(p)]:.2e}")
        if sweep[str(p)] <= 1e-10:
            p_ok = p
            break
    result["rel_err_vs_direct_qbx_by_fmm_order"] = sweep
    if p_ok is not None:
        wq = sp.make_wrangler(gp, "laplace", 0.0, "S", nrhs)
        dtq, _ = timed(stepper(sp, wq, dens), max(2, steps // 2))
        result["at_1e-10_order"] = {"fmm_order": p_ok,
                                    "value": round(discr.ndofs * nrhs / dtq, 1),
                                    "unit": "points*rhs/s",
                                    "ms_per_step": round(dtq * 1e3, 4),
                                    "rel_err_vs_direct_qbx_rhs0": sweep[str(p_ok)]}
        log(f"[sphere3d] p={p_ok}, {nrhs} densities: {dtq * 1e3:.1f} ms")
    t2 = time.perf_counter()
    g.fmm("laplace", 0.0, "S", 3, lo, st0, nthreads=nth)
    t_fmm = time.perf_counter() - t2
    one = None
    if not args.no_cpu_1thread:
        t2 = time.perf_counter()
        g.fmm("laplace", 0.0, "S", 3, lo, st0, nthreads=1)
        t1 = time.perf_counter() - t2
        one = {"value": round(discr.ndofs / t1, 1), "unit": "points*rhs/s", "cores": 1,
               "seconds": round(t1, 2)}
    result["cpu_baseline"] = {
        "value": round(discr.ndofs / t_fmm, 1), "unit": "points*rhs/s", "cores": nth,
        "kind": "port",
        "sample": (f"one density (RHS 0) of the same workload through the oracle's QBX-FMM "
                   f"(C + OpenMP), same tree/lists/orders, {t_fmm:.2f}s; the direct-QBX "
                   f"accuracy reference on {len(idx)} targets took {t_dir:.2f}s"),
        "host": host_cpu(), "one_thread": one}
    result["speedup_vs_cpu"] = round(result["value"] / result["cpu_baseline"]["value"], 1)
    return result


def main_sphere3d(args):
    res = run_sphere3d(args)
    line = {"metric": METRIC, "value": res.pop("value"), "unit": res.pop("unit"),
            "n_gpus": 1, "steps": res.pop("steps"), "warmup": 1,
            "ms_per_step": res.pop("ms_per_step"), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
            "data": "synthetic", **res}
    print(json.dumps(line), flush=True)


def run_ellipsoid3d(args):
    """Loss and gradient of the 3D Helmholtz (k = 5) single layer on the
    ellipsoid with semi-axes (1, 0.75, 0.5) at 4096 seeded exterior targets
    (radius 2..3): L = ||S_k sigma - g||^2, g from a seeded point source inside;
    a step is the forward plus the backward (dL/dsigma and dL/dk).  Checked on
    a coarser ellipsoid (1280 triangles) against the oracle's dense operator
    (its conjugate transpose for dL/dsigma) and central differences (h = 1e-6)."""
    import torch

    import oracle
    from pytential_amd.autodiff import single_layer_far
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import ellipsoid
    k0 = 5.0
    rng = np.random.default_rng(1004)

    def targets(n):
        v = rng.normal(size=(3, n))
        v /= np.linalg.norm(v, axis=0)
        return v * rng.uniform(2.0, 3.0, n)

    x0 = rng.uniform(-0.3, 0.3, 3) * np.array([1.0, 0.75, 0.5])
    t0 = time.perf_counter()
    d = ellipsoid(args.nsub_ell, order=4, fine_order=6)
    src = QBXLayerPotentialSource(d, fine_order=6, qbx_order=3, fmm_order=10)
    tg = targets(4096)
    rp = np.linalg.norm(tg - x0[:, None], axis=0)
    g = torch.as_tensor(np.exp(1j * k0 * rp) / (4 * np.pi * rp), device="cuda")
    sig0 = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    tgd = torch.as_tensor(tg, device="cuda")
    log(f"ellipsoid: {d.nelements} triangles, ndofs={d.ndofs} nsources={d.nsources} "
        f"({time.perf_counter() - t0:.1f}s)")

    def step():
        s = torch.tensor(sig0, dtype=torch.complex128, device="cuda", requires_grad=True)
        k = torch.tensor(k0, dtype=torch.float64, requires_grad=True)
        loss = torch.sum(torch.abs(single_layer_far(src, s, k, tgd) - g) ** 2)
        loss.backward()
        return loss, s.grad, k.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    t1 = time.perf_counter()
    ev[0].record()
    for _ in range(args.steps):
        loss, gs, gk = step()
    ev[1].record()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t1) / args.steps
    npairs = 4096 * d.nsources
    # kernel times (events around one forward and one adjoint launch)
    from pytential_amd.autodiff import helmholtz3d_point_sum
    from pytential_amd._lib import call
    from pytential_amd.geometry import _p, _stream
    q = (src.quad_density(torch.as_tensor(sig0, device="cuda")) *
         src.dev["quad_weights"]).contiguous()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    helmholtz3d_point_sum(src, q, tgd, k0, want_dk=True)
    v = torch.empty(d.nsources, dtype=torch.complex128, device="cuda")
    torch.cuda.synchronize()
    e[0].record()
    helmholtz3d_point_sum(src, q, tgd, k0, want_dk=True)
    e[1].record()
    call("qbx_helmholtz3d_p2p_adjoint", src.ctx.h, _stream(), d.nsources,
         _p(src.dev["quad_nodes"]), 4096, _p(tgd), k0, _p(g), _p(v))
    e[2].record()
    torch.cuda.synchronize()
    t_fwd, t_adj = e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2])
    # per pair: |t - s|^2 (8), sqrt and 1/r (2), sincos (1, a special
    # function), the complex scale-and-add (12) and the k-derivative's complex
    # add (4); the adjoint the same without the last
    flops_fwd = npairs * (8 + 2 + 1 + 12 + 4)
    flops_adj = npairs * (8 + 2 + 1 + 12)
    dom, t_dom, f_dom = (("forward", t_fwd, flops_fwd) if t_fwd >= t_adj
                         else ("adjoint", t_adj, flops_adj))
    achieved = f_dom / (t_dom * 1e-3) / 1e12
    # the same with the sincos counted as SPECIAL flops
    achieved40 = (f_dom + npairs * (SPECIAL - 1)) / (t_dom * 1e-3) / 1e12

    # checks on the 1280-triangle coarsening against the dense operator
    dc = ellipsoid(3, order=4, fine_order=6)
    srcc = QBXLayerPotentialSource(dc, fine_order=6, qbx_order=3, fmm_order=10)
    tgc = targets(512)
    rpc = np.linalg.norm(tgc - x0[:, None], axis=0)
    gc = np.exp(1j * k0 * rpc) / (4 * np.pi * rpc)
    sc = rng.normal(size=dc.ndofs) + 1j * rng.normal(size=dc.ndofs)
    s_t = torch.tensor(sc, dtype=torch.complex128, device="cuda", requires_grad=True)
    k_t = torch.tensor(k0, dtype=torch.float64, requires_grad=True)
    gct = torch.as_tensor(gc, device="cuda")
    lossc = torch.sum(torch.abs(single_layer_far(srcc, s_t, k_t, tgc) - gct) ** 2)
    lossc.backward()
    A = oracle.helmholtz3d_far_operator(dc, tgc, k0)
    r = A @ sc - gc
    e_grad = float(np.max(np.abs(s_t.grad.cpu().numpy() - A.conj().T @ (2 * r))) /
                   np.max(np.abs(A.conj().T @ (2 * r))))

    def lossf(sv, kv):
        with torch.no_grad():
            return float(torch.sum(torch.abs(single_layer_far(srcc, sv, kv, tgc) - gct) ** 2))

    h = 1e-6
    fd_k = (lossf(sc, k0 + h) - lossf(sc, k0 - h)) / (2 * h)
    vdir = rng.normal(size=dc.ndofs) + 1j * rng.normal(size=dc.ndofs)
    fd_v = (lossf(sc + h * vdir, k0) - lossf(sc - h * vdir, k0)) / (2 * h)
    an_v = float(np.real(np.vdot(s_t.grad.cpu().numpy(), vdir)))
    e_fd_k = abs(float(k_t.grad) - fd_k) / abs(fd_k)
    e_fd_v = abs(an_v - fd_v) / abs(fd_v)
    log(f"[ellipsoid3d] {dt * 1e3:.2f} ms/step, fwd {t_fwd:.2f} ms adj {t_adj:.2f} ms; "
        f"checks: grad vs dense^H {e_grad:.2e}, FD k {e_fd_k:.2e}, FD dir {e_fd_v:.2e}")
    # the dense check at its stated size: G_k W U of the 1280-triangle
    # coarsening at all 4096 targets, built on the GPU (qbx_lpot_matrix, no
    # quad-stage-2 intermediate); its conjugate transpose against the adjoint
    # kernels' gradient, and sampled rows against the oracle's dense operator
    from pytential_amd.matrix import build_matrix
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    Ad = build_matrix(srcc, "helmholtz", k0, "S", qbx_forced_limit=None, targets=tg)
    torch.cuda.synchronize()
    t_mat = time.perf_counter() - t3
    s4 = torch.tensor(sc, dtype=torch.complex128, device="cuda", requires_grad=True)
    phi4 = single_layer_far(srcc, s4, k0, tgd)
    (g4,) = torch.autograd.grad(phi4, s4, grad_outputs=g)
    e_dense_adj = float((g4 - Ad.conj().T @ g).abs().max() / (Ad.conj().T @ g).abs().max())
    phi4 = phi4.detach()
    e_dense_fwd = float((phi4 - Ad @ s4.detach()).abs().max() / phi4.abs().max())
    rows = rng.choice(4096, 64, replace=False)
    Ao = oracle.helmholtz3d_far_operator(dc, tg[:, rows], k0)
    e_dense_rows = float(np.max(np.abs(Ad[rows].cpu().numpy() - Ao)) / np.max(np.abs(Ao)))
    dense = {"shape": list(Ad.shape), "gbytes": round(Ad.numel() * 16 / 1e9, 3),
             "build_ms": round(t_mat * 1e3, 2),
             "rows_vs_oracle_dense": e_dense_rows, "A_sigma_vs_forward": e_dense_fwd,
             "AH_g_vs_adjoint_kernels": e_dense_adj}
    log(f"[ellipsoid3d] dense G_k W U {tuple(Ad.shape)} ({dense['gbytes']} GB) built in "
        f"{t_mat * 1e3:.1f} ms; rows vs oracle {e_dense_rows:.2e}, A sigma vs forward "
        f"{e_dense_fwd:.2e}, A^H g vs adjoint {e_dense_adj:.2e}")
    del Ad
    fmm_grad = ellipsoid_fmm_gradient(src, sig0, k0, tgd, g, loss, gs, gk, args)
    qfmm = ellipsoid_qbx_fmm(d, k0, rng)
    cpu = None
    if not args.no_cpu:
        # the oracle's dense rows for a sample of targets on the host cores
        ns_t = 32
        t2 = time.perf_counter()
        Gs = oracle.helmholtz3d_dense(d.quad_nodes, tg[:, :ns_t], k0)
        _ = Gs @ (d.quad_weights * d.upsample_density(sig0))
        t_cpu = time.perf_counter() - t2
        cpu = {"value": round(ns_t / t_cpu, 1), "unit": "points/s", "cores": 1,
               "kind": "port",
               "sample": (f"forward only: {ns_t} of the 4096 targets against all "
                          f"{d.nsources} sources through the oracle's dense kernel rows "
                          f"(numpy, one thread), {t_cpu:.2f}s"),
               "host": host_cpu()}
    line = {
        "metric": METRIC,
        "value": round(4096 / dt, 1),
        "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic",
        "config": {"workload": (f"3D Helmholtz k=5 S on the ellipsoid (1, 0.75, 0.5): "
                                f"icosahedron refined {args.nsub_ell}x ({d.nelements} "
                                f"triangles), order-4 nodes ({d.ndofs}), quad 6x6 "
                                f"({d.nsources} sources); loss ||S sigma - g||^2 at 4096 "
                                f"exterior targets (radius 2..3) and its gradient in "
                                f"sigma and k per step"),
                   "ntargets": 4096, "nsources": d.nsources, "seed": 1004},
        "kernel_ms": {"forward": round(t_fwd, 4), "adjoint": round(t_adj, 4)},
        "checks_1280_triangles": {"grad_sigma_vs_dense_adjoint": e_grad,
                                  "dL_dk_vs_central_diff": e_fd_k,
                                  "dL_dsigma_dir_vs_central_diff": e_fd_v,
                                  "gpu_dense_matrix_4096_targets": dense},
        "roofline": {"bound": "valu", "compute": "fp64",
                     "kernel": "k_h3_p2p_partial" if dom == "forward" else "k_h3_p2p_adjoint",
                     "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                     "frac_special40": round(achieved40 / FP64_PEAK_TFLOPS, 4),
                     "traffic": None, "flops_per_launch": int(f_dom),
                     "avg_launch_ms": round(t_dom, 4)},
        "cpu_baseline": cpu,
        "fmm_gradient": fmm_grad,
        "qbx_fmm": qfmm,
    }
    return line


def ellipsoid_fmm_gradient(src, sig0, k0, tgd, g, loss_exact, gs_exact, gk_exact, args):
    """The same loss and gradient through the QBX-FMM (fmm_order 10) instead of
    the exact sums: forward drive_fmm, backward qbx_drive_fmm_adjoint for
    dL/dsigma and a fourth-order difference of four more forwards in k for
    dL/dk (layer_potential); timed per step and compared with the exact
    path's loss and gradient."""
    import torch

    from pytential_amd.autodiff import layer_potential
    pairs = [(tgd, 0)]

    def step():
        s = torch.tensor(sig0, dtype=torch.complex128, device="cuda", requires_grad=True)
        k = torch.tensor(k0, dtype=torch.float64, requires_grad=True)
        loss = torch.sum(torch.abs(layer_potential(src, "helmholtz", s, pairs, k=k) - g) ** 2)
        loss.backward()
        return loss, s.grad, k.grad
    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nst = max(1, min(args.steps, 3))
    for _ in range(nst):
        loss, gsf, gkf = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / nst * 1e3
    # the dL/dsigma half alone (forward plus adjoint, no k-derivative)
    t1 = time.perf_counter()
    for _ in range(nst):
        s = torch.tensor(sig0, dtype=torch.complex128, device="cuda", requires_grad=True)
        L = torch.sum(torch.abs(layer_potential(src, "helmholtz", s, pairs, k=k0) - g) ** 2)
        L.backward()
    torch.cuda.synchronize()
    ms_sig = (time.perf_counter() - t1) / nst * 1e3
    e_loss = abs(float(loss) - float(loss_exact)) / abs(float(loss_exact))
    e_gs = float((gsf - gs_exact).abs().max() / gs_exact.abs().max())
    e_gk = abs(float(gkf) - float(gk_exact)) / abs(float(gk_exact))
    log(f"[ellipsoid3d fmm-gradient] {ms:.2f} ms per step (sigma-gradient only "
        f"{ms_sig:.2f} ms); vs exact sums: loss {e_loss:.2e}, dL/dsigma {e_gs:.2e}, "
        f"dL/dk {e_gk:.2e}")
    return {"what": "loss and gradient through the QBX-FMM (fmm_order 10): forward "
                    "drive_fmm, dL/dsigma by qbx_drive_fmm_adjoint, dL/dk by a fourth-order "
                    "difference of four forwards in k",
            "ms_per_step": round(ms, 3), "ms_per_step_sigma_gradient_only": round(ms_sig, 3),
            "steps": nst, "loss_vs_exact_sums": e_loss,
            "grad_sigma_vs_exact_sums": e_gs, "grad_k_vs_exact_sums": e_gk}


def ellipsoid_qbx_fmm(d, k0, rng, steps=5):
    """The 3D Helmholtz QBX-FMM of the same ellipsoid: S_k sigma at every
    surface node (exterior limit), qbx_order 3, fmm_order 10, fused driver,
    timed over `steps` evaluations; the stage-by-stage times of one more; the
    error against the GPU direct QBX at 2000 sampled nodes."""
    import torch

    from pytential_amd import fmm as F
    from pytential_amd.qbx import QBXLayerPotentialSource
    sigma = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    src = QBXLayerPotentialSource(d, fine_order=d.fine_order, qbx_order=3, fmm_order=10)
    t0 = time.perf_counter()
    geo = src.qbx_fmm_geometry_data((("self", +1),))
    w = src.make_wrangler(geo, "helmholtz", k0, "S")
    st, nrm = w.reorder_sources(src.quad_density(torch.as_tensor(sigma, device="cuda")),
                                src.dev["quad_weights"], None)
    F.drive_fmm_fused(w, st, nrm)
    torch.cuda.synchronize()
    t_cold = time.perf_counter() - t0
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(steps):
        out = F.drive_fmm_fused(w, st, nrm)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    stages = {}

    def tm(name, f):
        a, b = (torch.cuda.Event(enable_timing=True) for _ in range(2))
        torch.cuda.synchronize()
        a.record()
        r = f()
        b.record()
        torch.cuda.synchronize()
        stages[name] = round(a.elapsed_time(b), 3)
        return r
    m = tm("p2m", lambda: w.form_multipoles(st, nrm))
    m = tm("m2m", lambda: w.coarsen_multipoles(m))
    pot = tm("p2p", lambda: w.eval_direct(st, nrm))
    loc = tm("m2l", lambda: w.multipole_to_local(m))
    tm("m2p", lambda: w.eval_multipoles(m, pot))
    tm("p2l", lambda: w.form_locals(st, nrm, loc))
    tm("l2l", lambda: w.refine_locals(loc))
    tm("l2p", lambda: w.eval_locals(loc, pot))
    q = tm("p2qbxl", lambda: w.form_global_qbx_locals(st, nrm))
    tm("m2qbxl", lambda: w.translate_box_multipoles_to_qbx_local(m, q))
    tm("l2qbxl", lambda: w.translate_box_local_to_qbx_local(loc, q))
    tm("qbxl2p", lambda: w.eval_qbx_expansions(q, pot))
    # M2L work: every list-2 pair applies one (p+1)^2 x (p+1)^2 complex table
    tr = geo.traversal()
    l2s = tr.from_sep_siblings_starts.cpu().numpy().astype(np.int64)
    n2 = int(l2s[-1])
    ncf = (10 + 1) ** 2
    m2l_flops = n2 * ncf * ncf * 8
    # accuracy at sampled nodes against the GPU direct QBX
    idx = np.sort(rng.choice(d.ndofs, 2000, replace=False))
    tgt = d.nodes[:, idx]
    got = src.exec_compute_potential("helmholtz", sigma, [(tgt, +1)], k=k0)[0]
    srcd = QBXLayerPotentialSource(d, fine_order=d.fine_order, qbx_order=3, fmm_order=False)
    ref = srcd.exec_compute_potential("helmholtz", sigma, [(tgt, +1)], k=k0)[0]
    err = float((got - ref).abs().max() / ref.abs().max())
    full_vs_sample = float((out[torch.as_tensor(idx, device="cuda")] - got).abs().max() /
                           got.abs().max())
    tree = geo.tree()
    # the transposed QBX-FMM on the same geometry (QBXL2P^T ... P2M^T): its
    # time and the identity Re <y', A x> = Re <A^H y', x>
    ya = torch.as_tensor(rng.normal(size=out.numel()) + 1j * rng.normal(size=out.numel()),
                         device="cuda")
    xa = F.drive_fmm_adjoint(w, ya, None)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(steps):
        xa = F.drive_fmm_adjoint(w, ya, None)
    ev[1].record()
    torch.cuda.synchronize()
    ms_adj = ev[0].elapsed_time(ev[1]) / steps
    lhs = complex(torch.sum(torch.conj(ya) * out)).real
    rhs = complex(torch.sum(torch.conj(xa) * st)).real
    e_tr = abs(lhs - rhs) / float(torch.linalg.norm(ya) * torch.linalg.norm(out))
    log(f"[ellipsoid3d qbx-fmm] {ms:.2f} ms per evaluation ({d.ndofs} targets, "
        f"{d.nsources} sources, {tree.nboxes} boxes, {tree.nlevels} levels), cold "
        f"{t_cold:.2f}s, vs direct QBX {err:.2e}; stages {stages}; adjoint {ms_adj:.2f} ms, "
        f"transpose identity {e_tr:.1e}")
    return {"what": "3D Helmholtz QBX-FMM (spherical waves, qbx_order 3, fmm_order 10) "
                    "of S_k sigma at every surface node (exterior limit), fused driver",
            "ms_per_step": round(ms, 4), "steps": steps,
            "value": round(d.ndofs / (ms * 1e-3), 1), "unit": "points/s",
            "ntargets": d.ndofs, "nsources": d.nsources, "nboxes": tree.nboxes,
            "nlevels": tree.nlevels, "cold_s_incl_tables": round(t_cold, 3),
            "rel_err_vs_gpu_direct_qbx_2000_nodes": err,
            "fused_vs_point_target_eval": full_vs_sample,
            "per_stage_ms": stages,
            "adjoint": {"what": "qbx_drive_fmm_adjoint (the twelve transposed stages) "
                                "of the same evaluation",
                        "ms_per_step": round(ms_adj, 4),
                        "transpose_identity_rel": e_tr},
            "m2l": {"list2_pairs": n2, "gflop": round(m2l_flops / 1e9, 2),
                    "tflops": round(m2l_flops / (stages["m2l"] * 1e-3) / 1e12, 3),
                    "frac_fp64": round(m2l_flops / (stages["m2l"] * 1e-3) / 1e12 /
                                       FP64_PEAK_TFLOPS, 4)}}


def main_ellipsoid3d(args):
    if args.ell_qbx_fmm_only:
        # the QBX-FMM of S_k sigma at the surface nodes and its adjoint alone
        # (for profiles): no loss/gradient step, dense checks or CPU leg
        from pytential_amd.surface import ellipsoid
        d = ellipsoid(args.nsub_ell, order=4, fine_order=6)
        res = ellipsoid_qbx_fmm(d, 5.0, np.random.default_rng(1004), steps=args.steps)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "data": "synthetic", **res}),
              flush=True)
        return
    print(json.dumps(run_ellipsoid3d(args)), flush=True)


def run_dirichlet64(args):
    """Config 4: ncurves independent starfish curves (arms 3..9, radius
    0.5..2, amplitude 0.25, element counts log-uniform in 256..4096, 8
    Gauss-Legendre nodes per element), the interior Laplace Dirichlet problem
    (0.5 u - D u with D at "avg") on each, boundary data from a point charge
    outside each curve, all solved by one lockstep GMRES (gmres_many) to 1e-10
    with every matvec a HIP-graph replay; the solutions' fields are checked
    against the point charges inside each curve.  A step is the whole batch of
    solves; points/s counts boundary nodes times matvecs."""
    import torch

    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.linalg import GroupedMatVecGraphs, MatVecGraph, gmres_many
    from pytential_amd.pde import DirichletOperator
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.problems import dirichlet_curves
    specs, rng = dirichlet_curves(args.ncurves, 1003)
    t0 = time.perf_counter()
    curves = []
    for sp in specs:
        rad, nel, x0 = sp["radius"], sp["nelements"], sp["charge"]
        d = discretize_curve(starfish(sp["arms"], 0.25, rad, phase=sp["phase"]), nel, 8, 16)
        src = QBXLayerPotentialSource(d, fine_order=16, qbx_order=5, fmm_order=15)
        op = DirichletOperator(src, loc_sign=-1)
        f = -np.log(np.hypot(d.nodes[0] - x0[0], d.nodes[1] - x0[1])) / (2 * np.pi)
        curves.append({"d": d, "src": src, "op": op, "x0": x0, "rad": rad, "nel": nel,
                       "f": torch.as_tensor(f, device="cuda")})
    ndofs = [c["d"].ndofs for c in curves]
    log(f"[dirichlet64] {len(curves)} curves, {sum(c['nel'] for c in curves)} elements, "
        f"{sum(ndofs)} nodes ({time.perf_counter() - t0:.1f}s)")
    # eager pass over every geometry first (geometry, workspaces at their
    # largest), then one captured matvec per geometry
    t1 = time.perf_counter()
    for c in curves:
        c["op"].matvec(c["f"])
    torch.cuda.synchronize()
    for c in curves:
        c["mv"] = MatVecGraph(c["op"].matvec, c["f"])
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t1
    counts = [0] * len(curves)

    def counted(i, mv):
        def f(x):
            counts[i] += 1
            return mv(x)
        return f

    ops = [counted(i, c["mv"]) for i, c in enumerate(curves)]
    rhss = [c["f"] for c in curves]
    # the same solves with one graph per geometry, replayed one after another
    for _ in range(args.warmup):
        gmres_many(ops, rhss, tol=1e-10, restart=30, maxiter=200)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        gmres_many(ops, rhss, tol=1e-10, restart=30, maxiter=200)
    torch.cuda.synchronize()
    dt_seq = (time.perf_counter() - t2) / args.steps
    t3 = time.perf_counter()
    for c in curves:
        c["mv"](c["f"])
    torch.cuda.synchronize()
    t_mv_seq = time.perf_counter() - t3
    # all 64 operators in ONE graph on parallel branches: each lockstep
    # iteration applies every pending operator with one replay
    t4 = time.perf_counter()
    bat = GroupedMatVecGraphs([c["op"].matvec for c in curves], [c["f"] for c in curves],
                              contexts_of=lambda i: curves[i]["op"].library_contexts(),
                              group=args.batch_group, nstreams=args.batch_streams)
    torch.cuda.synchronize()
    t_setup += time.perf_counter() - t4

    def apply_many(pairs):
        for i, _ in pairs:
            counts[i] += 1
        return bat.apply_many(pairs)
    for _ in range(args.warmup):
        gmres_many(ops, rhss, tol=1e-10, restart=30, maxiter=200, apply_many=apply_many)
    torch.cuda.synchronize()
    counts[:] = [0] * len(curves)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        res = gmres_many(ops, rhss, tol=1e-10, restart=30, maxiter=200, apply_many=apply_many)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t2) / args.steps
    nmv = sum(counts) / args.steps
    node_mv = sum(n * c for n, c in zip(ndofs, counts)) / args.steps
    # one replay of the batched graph: every geometry's matvec at once
    t3 = time.perf_counter()
    bat.apply_many([(i, c["f"]) for i, c in enumerate(curves)])
    torch.cuda.synchronize()
    t_mv_all = time.perf_counter() - t3
    errs = []
    for c, r in zip(curves, res):
        th = rng.uniform(0, 2 * np.pi, 64)
        rr = 0.5 * 0.75 * c["rad"] * np.sqrt(rng.uniform(0, 1, 64))
        pts = np.stack([rr * np.cos(th), rr * np.sin(th)])
        u = c["op"].representation(r.solution, pts).real.cpu().numpy()
        ex = -np.log(np.hypot(pts[0] - c["x0"][0], pts[1] - c["x0"][1])) / (2 * np.pi)
        errs.append(float(np.max(np.abs(u - ex)) / np.max(np.abs(ex))))
    its = [r.iteration_count for r in res]
    states = sorted(set(r.state for r in res))
    log(f"[dirichlet64] {dt * 1e3:.1f} ms for all solves, {nmv:.0f} matvecs, iterations "
        f"{min(its)}..{max(its)}, states {states}, interior error max {max(errs):.2e}")
    line = {
        "metric": METRIC,
        "value": round(node_mv / dt, 1),
        "unit": "points*matvecs/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"{len(curves)} starfish curves (arms 3..9, radius 0.5..2, "
                                f"amplitude 0.25, 256..4096 elements log-uniform, 8 nodes "
                                f"each, fine 16, qbx_order 5, fmm_order 15): interior Laplace "
                                f"Dirichlet, 0.5u - D u, GMRES(30) to 1e-10 in lockstep, "
                                f"every iteration's matvecs replays of HIP graphs holding "
                                f"{args.batch_group} geometries each on {args.batch_streams} "
                                f"parallel branches (only groups with unconverged systems)"),
                   "ncurves": len(curves), "nodes": int(sum(ndofs)),
                   "elements": int(sum(c["nel"] for c in curves)), "seed": 1003},
        "iterations": {"min": int(min(its)), "max": int(max(its)),
                       "mean": round(float(np.mean(its)), 2)},
        "states": states,
        "matvecs_per_step": nmv,
        "all_geometries_one_matvec_ms": round(t_mv_all * 1e3, 3),
        "one_graph_per_geometry": {
            "ms_per_step": round(dt_seq * 1e3, 4),
            "all_geometries_one_matvec_ms": round(t_mv_seq * 1e3, 3),
            "what": "the same solves with one HIP graph per geometry, replayed one after "
                    "another"},
        "interior_rel_err_max": max(errs),
        "setup_s": round(t_setup, 3),
        "roofline": None,
        "cpu_baseline": None,
    }
    return line


def main_dirichlet64(args):
    print(json.dumps(run_dirichlet64(args)), flush=True)


def run_config1(args):
    """Config 1, the reference's CPU-runnable case: 2D Laplace S on the 5-arm
    starfish, 64 panels x 8 nodes, qbx_order 4, fmm_order 10, on-surface plus
    a 128^2 grid; compared with direct QBX (fmm_order=False) on every target.
    Returns the result dict."""
    import torch

    from pytential_amd import fmm as F
    from pytential_amd import problems
    from pytential_amd.geometry import assoc_status_and_flags
    from pytential_amd.qbx import QBXLayerPotentialSource
    cfg = problems.refine_until_associated(problems.config1(), assoc_status_and_flags)
    tg = ((cfg.targets, cfg.target_sides),)
    src = QBXLayerPotentialSource(cfg.discr, fine_order=16, qbx_order=4, fmm_order=10)
    geo = src.qbx_fmm_geometry_data(tg)
    geo.build_all()
    w = src.make_wrangler(geo, "laplace", 0.0, "S")
    dens = torch.as_tensor(cfg.density, dtype=torch.float64, device="cuda")

    def step():
        qd = src.quad_density(dens)
        st, _ = w.reorder_sources(qd, src.dev["quad_weights"], None)
        return F.drive_fmm_fused(w, st, None)

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    n = max(args.steps, 20)
    t0 = time.perf_counter()
    for _ in range(n):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / n
    got = out.real.cpu().numpy()
    srcd = QBXLayerPotentialSource(cfg.discr, fine_order=16, qbx_order=4, fmm_order=False)
    dref, = srcd.exec_compute_potential("laplace", cfg.density, list(tg))
    dref = dref.cpu().numpy()
    res = {"value": round(cfg.ntargets / dt, 1), "unit": "points/s",
           "ms_per_step": round(dt * 1e3, 4), "steps": n, "dtype": "f64",
           "config": {"workload": (f"2D Laplace S, starfish 5 arms / 0.25, "
                                   f"{cfg.discr.nelements} panels x 8 nodes, fine 16, "
                                   f"qbx_order 4, fmm_order 10, 128^2 grid on [-2,2]^2 + "
                                   f"on-surface targets, density sin 3t + 0.1 N(0,1)"),
                      "ntargets": cfg.ntargets, "seed": cfg.seed},
           "rel_err_vs_gpu_direct_qbx_all_targets": float(
               np.max(np.abs(got - dref)) / np.max(np.abs(dref))),
           "cpu_baseline": None}
    if not args.no_cpu:
        import oracle
        oracle.build(with_ref=False)
        g = oracle.GeometryData(cfg.discr, cfg.targets, cfg.target_sides)
        st = cfg.discr.strengths(cfg.density)
        lo = [10] * g.tree.nlevels
        nth = cpu_threads(args)
        times = {}
        for th in (nth, 1):
            reps = 0
            t0 = time.perf_counter()
            while reps < 3 or time.perf_counter() - t0 < 1.0:
                f_ref, _ = g.fmm("laplace", 0.0, "S", 4, lo, st, nthreads=th)
                reps += 1
            times[th] = (time.perf_counter() - t0) / reps
        d_ref = g.direct("laplace", 0.0, "S", 4, st, nthreads=nth).real
        res["rel_err_vs_oracle_fmm_all_targets"] = float(
            np.max(np.abs(got - f_ref.real)) / np.max(np.abs(f_ref)))
        res["rel_err_vs_oracle_direct_qbx_all_targets"] = float(
            np.max(np.abs(got - d_ref)) / np.max(np.abs(d_ref)))
        res["cpu_baseline"] = {
            "value": round(cfg.ntargets / times[nth], 1), "unit": "points/s", "cores": nth,
            "kind": "port",
            "sample": (f"the whole evaluation through the oracle (C + OpenMP), same "
                       f"tree/lists/orders, {times[nth] * 1e3:.1f} ms per evaluation"),
            "host": host_cpu(),
            "one_thread": {"value": round(cfg.ntargets / times[1], 1), "unit": "points/s",
                           "cores": 1, "seconds": round(times[1], 4)}}
        res["speedup_vs_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--nelements", type=int, default=4096)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the headline with the default work-unit shapes")
    ap.add_argument("--tuning", default=None,
                    help="JSON dict of work-unit shapes (pytential_amd.autotune.KNOBS) "
                         "for the headline instead of searching them")
    ap.add_argument("--no-cpu-1thread", action="store_true",
                    help="skip the one-thread oracle timings")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--sample", type=int, default=1000)
    ap.add_argument("--nrhs", type=int, default=8,
                    help="densities per evaluation for the batched figure")
    ap.add_argument("--workload", choices=["helmholtz2d", "sphere3d", "ellipsoid3d",
                                           "dirichlet64", "laplace2d"],
                    default="helmholtz2d")
    ap.add_argument("--headline-only", action="store_true",
                    help="helmholtz2d without the other workloads' nested results")
    ap.add_argument("--ncurves", type=int, default=64, help="dirichlet64: curves")
    ap.add_argument("--nsub-ell", type=int, default=5,
                    help="ellipsoid3d: icosahedron refinements (5: 20480 triangles)")
    ap.add_argument("--no-graph", action="store_true",
                    help="helmholtz2d: skip the HIP-graph replay of the evaluation")
    ap.add_argument("--ell-qbx-fmm-only", action="store_true",
                    help="ellipsoid3d: only the surface QBX-FMM and its adjoint")
    ap.add_argument("--nsub", type=int, default=5, help="sphere3d: icosahedron refinements")
    ap.add_argument("--fine3d", type=int, default=6, help="sphere3d: quad rule n x n")
    ap.add_argument("--nrhs3d", type=int, default=32, help="sphere3d: densities")
    ap.add_argument("--steps3d", type=int, default=5, help="sphere3d: timed steps")
    ap.add_argument("--batch-streams", type=int, default=4,
                    help="dirichlet64: parallel branches of each batched matvec graph")
    ap.add_argument("--batch-group", type=int, default=4,
                    help="dirichlet64: geometries per batched matvec graph")
    ap.add_argument("--profile", action="store_true",
                    help="one configuration only (tolerance orders), no CPU leg: "
                         "the run profiled under rocprofv3")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 512 MiB device buffer 5 times (HBM counter scale)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the QBX-FMM "
              "path here is single-GPU.", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration:
        n = 512 << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / 5
        print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n,
                          "GBps_read_plus_write": round(2 * n / dt / 1e9, 1)}), flush=True)
        return
    if args.workload == "ellipsoid3d":
        main_ellipsoid3d(args)
        return
    if args.workload == "dirichlet64":
        main_dirichlet64(args)
        return
    if args.workload == "sphere3d":
        main_sphere3d(args)
        return
    if args.workload == "laplace2d":
        res = run_config1(args)
        print(json.dumps({"metric": METRIC, "n_gpus": 1, "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "data": "synthetic",
                          **res}), flush=True)
        return
    from pytential_amd.qbx import fmm_level_to_order_for_tolerance

    t0 = time.perf_counter()
    cfg = build_problem(args.grid, args.nelements)
    log(f"problem: {cfg.name} panels={cfg.discr.nelements} ndofs={cfg.discr.ndofs} "
        f"nsources={cfg.discr.nsources} ntargets={cfg.ntargets} "
        f"({time.perf_counter() - t0:.2f}s)")

    def tol_orders(kernel, kernel_args, tree, level):
        return fmm_level_to_order_for_tolerance(cfg.kernel, cfg.k, args.tol,
                                                tree.root_extent, tree.nlevels)[level]

    def config_orders(kernel, kernel_args, tree, level):
        return cfg.fmm_order

    main_run = run_gpu(args, cfg, tol_orders, "tol", tune=not args.no_autotune)
    if args.profile:
        log(f"profile run: {main_run['step_s'] * 1e3:.3f} ms/step, per stage "
            f"{ {k: round(v, 4) for k, v in main_run['per_stage_ms'].items()} }")
        return
    cfg_run = run_gpu(args, cfg, config_orders, f"p{cfg.fmm_order}")
    # accuracy on every target: the GPU direct QBX (fmm_order=False) of the
    # same discretization, untimed
    from pytential_amd.qbx import QBXLayerPotentialSource
    srcd = QBXLayerPotentialSource(cfg.discr, fine_order=cfg.discr.fine_order,
                                   qbx_order=cfg.qbx_order, fmm_order=False)
    t_d = time.perf_counter()
    dref, = srcd.exec_compute_potential(cfg.kernel, cfg.density,
                                        [(cfg.targets, cfg.target_sides)], k=cfg.k,
                                        source_kind=cfg.source_kind)
    torch.cuda.synchronize()
    t_d = time.perf_counter() - t_d
    dref = dref.cpu().numpy()
    err_all = {key: float(np.max(np.abs(run["result"] - dref)) / np.max(np.abs(dref)))
               for key, run in (("tol", main_run), ("config_order", cfg_run))}
    log(f"GPU direct QBX on all {cfg.ntargets} targets ({t_d:.2f}s incl. geometry): "
        f"FMM rel err {err_all}")
    del srcd
    batched = run_batched(args, cfg, main_run["src"], main_run["geo"], args.nrhs)

    rng = np.random.default_rng(11)
    ttc, _ = main_run["geo"]._assoc
    ttc = ttc.cpu().numpy()
    q_idx, n_idx = np.nonzero(ttc >= 0)[0], np.nonzero(ttc < 0)[0]
    half = args.sample // 2
    sample_idx = np.sort(np.concatenate([rng.choice(q_idx, min(half, len(q_idx)), False),
                                         rng.choice(n_idx, min(half, len(n_idx)), False)]))
    cpu = None
    err = {"tol": None, "config_order": None}
    if not args.no_cpu:
        cpu = cpu_baseline(args, cfg, main_run["orders"], sample_idx)
        ref = cpu["direct_sample"]
        scale = np.max(np.abs(ref))
        for key, run in (("tol", main_run), ("config_order", cfg_run)):
            err[key] = float(np.max(np.abs(run["result"][sample_idx] - ref)) / scale)
        err["gpu_vs_cpu_fmm"] = float(np.max(np.abs(main_run["result"][sample_idx]
                                                    - cpu["fmm_result_sample"])) / scale)
        log(f"errors: {err}")

    ps = main_run["per_stage_ms"]
    flops, flops40 = main_run["flops"], main_run["flops40"]
    # the dominant kernel: the slowest stage of the stage-by-stage run (every
    # stage is one kernel launch, or one per level for the level-by-level
    # translations), timed by HIP events on its launch stream
    dominant = max(ps, key=ps.get)
    kernel_name = {"p2p": "k_p2p", "p2qbxl": "k_p2qbxl", "m2l": "k_m2l_blk",
                   "m2m": "k_m2m_h", "p2l": "k_p2e_h", "p2m": "k_p2e_h",
                   "l2l": "k_l2l_h", "m2p": "k_m2p", "m2qbxl": "k_m2qbxl_mfma"}
    rl = None
    if dominant in flops:
        achieved = flops[dominant] / (ps[dominant] * 1e-3) / 1e12
        achieved40 = flops40[dominant] / (ps[dominant] * 1e-3) / 1e12
        # FP64 compute; MI355X's FP64 vector and FP64 matrix peaks are the
        # same 78.6 TFLOP/s, the figure this fraction is taken against
        rl = {"bound": "mfma" if dominant == "m2qbxl" else "valu", "compute": "fp64",
              "kernel": kernel_name.get(dominant, dominant), "stage": dominant,
              "achieved": round(achieved, 3),
              "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
              "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
              "frac_special40": round(achieved40 / FP64_PEAK_TFLOPS, 4),
              "traffic": measured_traffic(dominant),
              "flops_per_launch": int(flops[dominant]),
              "flops_per_launch_special40": int(flops40[dominant]),
              "avg_launch_ms": round(ps[dominant], 4),
              "timing": "HIP events on the launch stream, stage-by-stage run of the same "
                        "evaluation on one stream"}
        rec = profiled_stage(dominant)
        if rec and rec.get("hw_fp64_flops"):
            # the same launch time against the FP64 flops the hardware counted
            hw = rec["hw_fp64_flops"] / (ps[dominant] * 1e-3) / 1e12
            rl["hw_fp64_flops_per_launch"] = int(rec["hw_fp64_flops"])
            rl["hw_achieved"] = round(hw, 3)
            rl["hw_frac"] = round(hw / FP64_PEAK_TFLOPS, 4)
    stage_rl = {k: {"ms": round(ps[k], 4), "gflop": round(flops[k] / 1e9, 3),
                    "gflop_special40": round(flops40[k] / 1e9, 3),
                    "frac_fp64": round(flops[k] / (ps[k] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                    "frac_fp64_special40": round(flops40[k] / (ps[k] * 1e-3) / 1e12 /
                                                 FP64_PEAK_TFLOPS, 4)}
                for k in flops if ps.get(k, 0) > 0}
    line = {
        "metric": METRIC,
        "value": round(cfg.ntargets / main_run["step_fused_s"], 1),
        "unit": "points/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(main_run["step_fused_s"] * 1e3, 4),
        "graph_replay": None if main_run["step_graph_s"] is None else {
            "what": "the same evaluation captured into one HIP graph (four stream "
                    "branches) and replayed",
            "ms_per_step": round(main_run["step_graph_s"] * 1e3, 4),
            "rel_diff_vs_eager": main_run["graph_vs_eager"]},
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128", "data": "synthetic",
        "config": {
            "workload": (f"2D Helmholtz k={cfg.k:g} double layer, starfish arms="
                         f"{cfg.arms} amp={cfg.amplitude:.3f}, {cfg.discr.nelements} panels x "
                         f"{cfg.discr.order} nodes, fine {cfg.discr.fine_order}, qbx_order "
                         f"{cfg.qbx_order}, {args.grid}^2 grid + on-surface targets"),
            "ntargets": cfg.ntargets, "nsources": cfg.discr.nsources,
            "ncenters": 2 * cfg.discr.ndofs, "fmm_orders_per_level": main_run["orders"],
            "tol": args.tol, "seed": cfg.seed},
        "rel_err_vs_direct_qbx": err["tol"],
        "rel_err_vs_gpu_direct_qbx_all_targets": err_all["tol"],
        "fused_vs_stagewise_rel_diff": main_run["fused_vs_stagewise"],
        "gpu_direct_qbx_s": round(t_d, 3),
        "at_config_order": {
            "fmm_order": cfg.fmm_order,
            "value": round(cfg.ntargets / cfg_run["step_fused_s"], 1),
            "ms_per_step": round(cfg_run["step_fused_s"] * 1e3, 4),
            "rel_err_vs_direct_qbx": err["config_order"],
            "rel_err_vs_gpu_direct_qbx_all_targets": err_all["config_order"]},
        "batched": {
            "nrhs": batched["nrhs"],
            "value": round(cfg.ntargets * batched["nrhs"] / batched["step_s"], 1),
            "unit": "points*rhs/s",
            "ms_per_step": round(batched["step_s"] * 1e3, 4),
            "rel_diff_rhs0_vs_plain": float(
                np.max(np.abs(batched["result0"] - main_run["result"])) /
                np.max(np.abs(main_run["result"])))},
        "autotune": None if main_run["autotune"] is None else {
            "settings": main_run["autotune"]["settings"],
            "tuned_ms": round(main_run["autotune"]["ms"], 4),
            "default_ms": round(main_run["autotune"]["default_ms"], 4),
            "trials": len(main_run["autotune"]["trials"]),
            "search_s": round(main_run["autotune"]["search_s"], 2),
            "what": "pytential_amd.autotune: coordinate descent over the work-unit shapes "
                    "(QBX_TUNE_*), each candidate the median of 5 fused evaluations on a "
                    "geometry built with it; the headline is timed with the result"},
        "adjoint": {"ms_per_step": round(main_run["adjoint_s"] * 1e3, 4),
                    "what": "qbx_drive_fmm_adjoint: the conjugate-transposed QBX-FMM "
                            "(density gradient backward pass), same tree and orders",
                    "transpose_identity_rel": main_run["adjoint_identity_rel"]},
        "stagewise_ms_per_step": round(main_run["step_s"] * 1e3, 4),
        "cold_geometry_s": round(main_run["cold_s"], 4),
        "per_stage_ms": {k: round(v, 4) for k, v in ps.items()},
        "stage_roofline": stage_rl,
        "roofline": {**rl, "measured_ceiling": fp64_ceiling()},
        "cpu_baseline": None if cpu is None else {
            k: (round(v, 1) if isinstance(v, float) else v) for k, v in cpu.items()
            if k in ("value", "unit", "cores", "kind", "sample", "host", "one_thread")},
    }
    if cpu is not None:
        line["gpu_vs_cpu_fmm_rel_diff"] = err["gpu_vs_cpu_fmm"]
        line["speedup_vs_cpu"] = round(line["value"] / cpu["value"], 2)
    if not args.headline_only:
        # the other workloads, each measured in this same run
        others = {}
        for name, fn, steps, warmup in (
                ("laplace2d_starfish_c1", run_config1, args.steps, args.warmup),
                ("laplace3d_sphere_32rhs_c3", run_sphere3d, args.steps, args.warmup),
                ("dirichlet64_gmres_c4", run_dirichlet64, 2, 1),
                ("helmholtz3d_ellipsoid_adjoint_c5", run_ellipsoid3d, 10, 2)):
            t0 = time.perf_counter()
            r = fn(argparse.Namespace(**{**vars(args), "steps": steps, "warmup": warmup}))
            for k in ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline",
                      "data"):
                r.pop(k, None)
            r["wall_s"] = round(time.perf_counter() - t0, 1)
            others[name] = r
            log(f"[{name}] done in {r['wall_s']}s: {r.get('value')} {r.get('unit')}")
        line["other_workloads"] = others
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
