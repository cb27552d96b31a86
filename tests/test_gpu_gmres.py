"""GPU: interior Laplace Dirichlet problem (pytential/symbolic/pde/scalar.py
DirichletOperator, loc_sign = -1) solved by GMRES with a HIP-graph captured
QBX-FMM matvec, as in the reference's test_scalar_int_eq.py:186-262: boundary
data from a point charge outside the curve, whose field the solution must
reproduce inside.  The captured graph is replayed several times, with eager
work and synchronisations between replays, and each replay is checked
against the eager matvec."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _setup(nelements=128):
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.pde import DirichletOperator
    from pytential_amd.qbx import QBXLayerPotentialSource
    discr = discretize_curve(starfish(5, 0.25), nelements, 8, 16)
    src = QBXLayerPotentialSource(discr, fine_order=16, qbx_order=5, fmm_order=15)
    op = DirichletOperator(src, loc_sign=-1)
    return discr, src, op


def _field(x, x0):
    return -np.log(np.hypot(x[0] - x0[0], x[1] - x0[1])) / (2 * np.pi)


def test_graph_replay_matches_eager():
    from pytential_amd.linalg import MatVecGraph
    discr, src, op = _setup(96)
    rng = np.random.default_rng(0)
    ex = torch.as_tensor(rng.normal(size=discr.ndofs), device="cuda")
    mv = MatVecGraph(op.matvec, ex)
    for i in range(4):
        x = torch.as_tensor(rng.normal(size=discr.ndofs), device="cuda")
        y_graph = mv(x)
        torch.cuda.synchronize()
        junk = (x * 2.0).sum()          # eager work between replays
        y_eager = op.matvec(x)
        torch.cuda.synchronize()
        err = float((y_graph - y_eager).abs().max() / y_eager.abs().max())
        print(f"[graph] replay {i}: rel diff {err:.2e} ({float(junk):.3f})", flush=True)
        assert err < 1e-13


def test_interior_dirichlet_point_charge():
    from pytential_amd.linalg import MatVecGraph, gmres
    discr, src, op = _setup(128)
    x0 = np.array([3.0, 1.0])
    f = torch.as_tensor(_field(discr.nodes, x0), device="cuda")
    mv = MatVecGraph(op.matvec, f)
    res = gmres(mv, f, tol=1e-10, restart=60, maxiter=200)
    print(f"[gmres] {res.state} after {res.iteration_count} iterations, "
          f"residual {res.residual_norms[-1]:.2e}", flush=True)
    assert res.success
    rng = np.random.default_rng(1)
    th = rng.uniform(0, 2 * np.pi, 200)
    rr = 0.5 * np.sqrt(rng.uniform(0, 1, 200))
    pts = np.stack([rr * np.cos(th), rr * np.sin(th)])
    u = op.representation(res.solution, pts).cpu().numpy()
    exact = _field(pts, x0)
    err = np.max(np.abs(u - exact)) / np.max(np.abs(exact))
    print(f"[gmres] interior field rel err {err:.2e}", flush=True)
    assert err < 1e-6


def test_two_captured_helmholtz_operators_interleaved():
    """Two Helmholtz Dirichlet operators (different curves, sizes and k, each
    with a D and an S wrangler on their own trees) captured into HIP graphs
    and replayed interleaved: every replay equals the eager matvec, so no
    graph reads workspaces or translation tables another operator replaced."""
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.linalg import MatVecGraph
    from pytential_amd.pde import DirichletOperator
    from pytential_amd.qbx import QBXLayerPotentialSource
    ops = []
    for arms, nel, k in ((5, 64, 3.0), (7, 160, 11.0)):
        d = discretize_curve(starfish(arms, 0.25), nel, 8, 16)
        s = QBXLayerPotentialSource(d, fine_order=16, qbx_order=5, fmm_order=15)
        ops.append((d, DirichletOperator(s, loc_sign=-1, kernel="helmholtz", k=k)))
    rng = np.random.default_rng(3)
    graphs = [MatVecGraph(op.matvec, torch.as_tensor(
        rng.normal(size=d.ndofs) + 0j, device="cuda")) for d, op in ops]
    for rep in range(3):
        for (d, op), mv in zip(ops, graphs):
            x = torch.as_tensor(rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs),
                                device="cuda")
            y_graph = mv(x)
            torch.cuda.synchronize()
            y_eager = op.matvec(x)
            torch.cuda.synchronize()
            err = float((y_graph - y_eager).abs().max() / y_eager.abs().max())
            print(f"[graph2] replay {rep} n={d.ndofs}: rel diff {err:.2e}", flush=True)
            assert err < 1e-13


def test_shared_context_refuses_to_replace_captured_workspaces():
    """A graph captured through the shared context pins its workspaces: a later
    evaluation on that context that needs larger ones raises instead of
    freeing memory the graph replays; after release_graphs it runs."""
    from pytential_amd import fmm as F
    from pytential_amd._lib import QBXError
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.geometry import Context
    from pytential_amd.linalg import MatVecGraph
    from pytential_amd.qbx import QBXLayerPotentialSource
    ctx = Context(torch.cuda.current_device())

    def evaluator(nel):
        d = discretize_curve(starfish(5, 0.25), nel, 8, 16)
        s = QBXLayerPotentialSource(d, fine_order=16, qbx_order=5, fmm_order=10)
        geo = s.qbx_fmm_geometry_data((("self", +1),))
        w = F.QBXExpansionWrangler(geo, "laplace", 0.0, "S", 5,
                                   s.level_orders("laplace", 0.0, geo.tree()), ctx=ctx)

        def fn(x):
            st, _ = w.reorder_sources(s.quad_density(x), s.dev["quad_weights"], None)
            return F.drive_fmm_fused(w, st, None)
        return d, fn

    d1, f1 = evaluator(32)
    d2, f2 = evaluator(256)
    x1 = torch.ones(d1.ndofs, dtype=torch.float64, device="cuda")
    x2 = torch.ones(d2.ndofs, dtype=torch.float64, device="cuda")
    mv = MatVecGraph(f1, x1)
    with pytest.raises(QBXError, match="captured"):
        f2(x2)
    torch.cuda.synchronize()
    y1 = mv(x1)
    ref1 = f1(x1)                       # same sizes: no replacement needed
    torch.cuda.synchronize()
    assert float((y1 - ref1).abs().max() / ref1.abs().max()) < 1e-13
    del mv
    ctx.release_graphs()
    y2 = f2(x2)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(y2.real).all())


def test_batched_graph_matches_each_operator():
    """Three operators (different curves) captured into one graph on parallel
    branches: each replay's outputs equal every operator's own eager matvec,
    over several replays with eager work between them, for all operators at
    once and for a subset."""
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.linalg import BatchedMatVecGraph
    from pytential_amd.pde import DirichletOperator
    from pytential_amd.qbx import QBXLayerPotentialSource
    rng = np.random.default_rng(3)
    ops, xs = [], []
    for arms, nel in ((3, 64), (5, 96), (7, 128)):
        d = discretize_curve(starfish(arms, 0.25), nel, 8, 16)
        src = QBXLayerPotentialSource(d, fine_order=16, qbx_order=5, fmm_order=15)
        op = DirichletOperator(src, loc_sign=-1)
        ops.append(op)
        xs.append(torch.as_tensor(rng.normal(size=d.ndofs), device="cuda"))
    bat = BatchedMatVecGraph([op.matvec for op in ops], xs,
                             contexts=[c for op in ops for c in op.library_contexts()],
                             nstreams=2)
    for rep in range(3):
        ins = [torch.as_tensor(rng.normal(size=x.numel()), device="cuda") for x in xs]
        pick = [0, 1, 2] if rep != 1 else [2, 0]
        outs = bat.apply_many([(i, ins[i]) for i in pick])
        torch.cuda.synchronize()
        junk = float(sum(x.sum() for x in ins))
        for i, y in zip(pick, outs):
            ref = ops[i].matvec(ins[i])
            torch.cuda.synchronize()
            err = float((y - ref).abs().max() / ref.abs().max())
            print(f"[batched graph] replay {rep} operator {i}: rel diff {err:.2e} "
                  f"({junk:.2f})", flush=True)
            assert err < 1e-13


def test_grouped_graphs_apply_only_requested_groups():
    """Operators in groups of two (each group one graph): outputs of a mixed
    request equal the eager matvecs, in the order asked."""
    from pytential_amd.discretization import discretize_curve, starfish
    from pytential_amd.linalg import GroupedMatVecGraphs
    from pytential_amd.pde import DirichletOperator
    from pytential_amd.qbx import QBXLayerPotentialSource
    rng = np.random.default_rng(4)
    ops, xs = [], []
    for arms, nel in ((3, 64), (5, 80), (7, 96)):
        d = discretize_curve(starfish(arms, 0.25), nel, 8, 16)
        src = QBXLayerPotentialSource(d, fine_order=16, qbx_order=5, fmm_order=15)
        ops.append(DirichletOperator(src, loc_sign=-1))
        xs.append(torch.as_tensor(rng.normal(size=d.ndofs), device="cuda"))
    grp = GroupedMatVecGraphs([op.matvec for op in ops], xs,
                              contexts_of=lambda i: ops[i].library_contexts(), group=2,
                              nstreams=2)
    assert len(grp.batches) == 2
    ins = [torch.as_tensor(rng.normal(size=x.numel()), device="cuda") for x in xs]
    order = [2, 0, 1]
    outs = grp.apply_many([(i, ins[i]) for i in order])
    torch.cuda.synchronize()
    for i, y in zip(order, outs):
        ref = ops[i].matvec(ins[i])
        torch.cuda.synchronize()
        err = float((y - ref).abs().max() / ref.abs().max())
        print(f"[grouped graphs] operator {i}: rel diff {err:.2e}", flush=True)
        assert err < 1e-13

