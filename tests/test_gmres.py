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
