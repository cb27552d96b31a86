"""GMRES (pytential_amd/linalg.py) on CPU tensors: solutions against a direct
solve, restarts, and the reference's outcome states and error behaviour
(pytential/linalg/gmres.py:144-413)."""
import numpy as np
import pytest
import torch

from pytential_amd.linalg import GMRESError, gmres


def _system(n, seed, complex_=True):
    rng = np.random.default_rng(seed)
    R = rng.normal(size=(n, n))
    if complex_:
        R = R + 1j * rng.normal(size=(n, n))
    A = np.eye(n) + 0.3 * R / np.sqrt(n)
    b = rng.normal(size=n) + (1j * rng.normal(size=n) if complex_ else 0)
    At = torch.as_tensor(A)
    return A, b, (lambda x: At @ x)


@pytest.mark.parametrize("complex_", [True, False])
@pytest.mark.parametrize("restart", [5, 50])
def test_gmres_solves(complex_, restart):
    n = 60
    A, b, op = _system(n, 1, complex_)
    res = gmres(op, torch.as_tensor(b), restart=restart, tol=1e-12, maxiter=500)
    assert res.success and res.state == "success"
    x = res.solution.numpy()
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-11)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-11
    assert res.residual_norms[-1] < 1e-12


def test_gmres_x0_and_zero_rhs():
    n = 30
    A, b, op = _system(n, 2)
    x_true = np.linalg.solve(A, b)
    res = gmres(op, torch.as_tensor(b), tol=1e-12,
                x0=torch.as_tensor(x_true + 1e-3), maxiter=200)
    np.testing.assert_allclose(res.solution.numpy(), x_true, atol=1e-11)
    z = gmres(op, torch.zeros(n, dtype=torch.complex128))
    assert z.success and float(z.solution.abs().max()) == 0


def test_gmres_failure_states():
    n = 40
    A, b, op = _system(n, 3)
    with pytest.raises(GMRESError, match="max iterations"):
        gmres(op, torch.as_tensor(b), tol=1e-14, maxiter=3, restart=2)
    res = gmres(op, torch.as_tensor(b), tol=1e-14, maxiter=3, restart=2, hard_failure=False)
    assert not res.success and res.state == "max iterations"
    assert res.iteration_count == 3
    # a rotation: GMRES stalls on it for restart < n
    P = np.roll(np.eye(n), 1, axis=0)
    Pt = torch.as_tensor(P.astype(complex))
    res = gmres(lambda x: Pt @ x, torch.as_tensor(np.eye(n)[0].astype(complex)),
                restart=5, tol=1e-10, hard_failure=False)
    assert not res.success and res.state in ("stalled", "max iterations")


def test_gmres_many_matches_single_solves():
    """Lockstep batch: each system converges like its own gmres call (same
    iteration count) to the numpy solution; a zero right-hand side and a
    complex system ride along."""
    from pytential_amd.linalg import gmres, gmres_many
    rng = np.random.default_rng(31)
    ops, rhss, mats = [], [], []
    for n, cplx in ((40, False), (17, True), (63, False), (5, False)):
        A = 3 * np.eye(n) + rng.normal(size=(n, n)) / np.sqrt(n)
        if cplx:
            A = A + 0.5j * rng.normal(size=(n, n)) / np.sqrt(n)
        At = torch.as_tensor(A)
        mats.append(A)
        ops.append(lambda x, At=At: At @ x)
        b = rng.normal(size=n) + (1j * rng.normal(size=n) if cplx else 0)
        rhss.append(torch.as_tensor(b))
    rhss[3] = torch.zeros(5, dtype=torch.float64)
    res = gmres_many(ops, rhss, tol=1e-11, restart=7, maxiter=300)
    for A, b, op, r in zip(mats, rhss, ops, res):
        assert r.success and r.state == "success"
        np.testing.assert_allclose(r.solution.numpy(), np.linalg.solve(A, b.numpy()),
                                   rtol=1e-9, atol=1e-10)
        one = gmres(op, b, tol=1e-11, restart=7, maxiter=300)
        assert one.iteration_count == r.iteration_count


def test_gmres_many_apply_many_matches_per_system_calls():
    """The batched application path (one call applying every pending
    operator, as BatchedMatVecGraph does) gives the per-call path's
    iterations and solutions exactly."""
    from pytential_amd.linalg import gmres_many
    rng = np.random.default_rng(32)
    ops, rhss = [], []
    for n in (30, 12, 45):
        A = torch.as_tensor(2 * np.eye(n) + rng.normal(size=(n, n)) / np.sqrt(n))
        ops.append(lambda x, A=A: A @ x)
        rhss.append(torch.as_tensor(rng.normal(size=n)))
    calls = []

    def apply_many(pairs):
        calls.append(len(pairs))
        return [ops[i](x) for i, x in pairs]
    a = gmres_many(ops, rhss, tol=1e-11, restart=5, maxiter=200)
    b = gmres_many(ops, rhss, tol=1e-11, restart=5, maxiter=200, apply_many=apply_many)
    for ra, rb in zip(a, b):
        assert ra.state == rb.state == "success"
        assert ra.iteration_count == rb.iteration_count
        assert torch.equal(ra.solution, rb.solution)
    # one call per lockstep iteration or restart, each for several systems
    assert len(calls) < sum(r.iteration_count for r in a)
    assert max(calls) == 3


def test_gmres_many_failure_states():
    from pytential_amd.linalg import GMRESError, gmres_many
    rng = np.random.default_rng(32)
    n = 30
    good = 3 * np.eye(n) + 0.2 * rng.normal(size=(n, n))
    # a rotation-like operator on which restarted GMRES(1) makes no progress
    bad = np.roll(np.eye(n), 1, axis=0)
    ops = [lambda x, A=torch.as_tensor(good): A @ x, lambda x, A=torch.as_tensor(bad): A @ x]
    b = [torch.as_tensor(rng.normal(size=n)) for _ in range(2)]
    res = gmres_many(ops, b, tol=1e-10, restart=1, maxiter=60)
    assert res[0].success
    assert not res[1].success and res[1].state in ("stalled", "max iterations",
                                                   "non-monotonic residuals")
    with pytest.raises(GMRESError):
        gmres_many(ops, b, tol=1e-10, restart=1, maxiter=60, hard_failure=True)
