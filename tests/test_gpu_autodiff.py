"""GPU: the differentiable 3D Helmholtz single layer at smooth targets
(pytential_amd.autodiff) through its C-ABI kernels, against the CPU oracle's
dense operator A = G_k W U on a coarse ellipsoid (semi-axes 1, 0.75, 0.5):

* forward phi = A sigma and the k-derivative, to 1e-12;
* backward of L = ||phi - g||^2 (g from a point source inside): dL/dsigma =
  A^H (2 r) and dL/dk = 2 Re sum conj(r) dA/dk sigma, to 1e-12;
* both against central finite differences of the GPU forward (h = 1e-6),
  to 1e-7 relative;
* torch.autograd.gradcheck on a tiny surface;
* the sphere's exterior closed form S_k Y_00 = i k j_0(k) h_0(k r) Y_00."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

K = 5.0


def _targets(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    r = rng.uniform(2.0, 3.0, n)
    return v * r, r


def _point_source(tg, x0, k):
    r = np.linalg.norm(tg - x0[:, None], axis=0)
    return np.exp(1j * k * r) / (4 * np.pi * r)


@pytest.fixture(scope="module")
def setup():
    import oracle
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import ellipsoid
    d = ellipsoid(2, order=4, fine_order=6)
    src = QBXLayerPotentialSource(d, fine_order=6, qbx_order=3, fmm_order=10)
    tg, _ = _targets(256, 4000)
    rng = np.random.default_rng(4001)
    sig = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    g = _point_source(tg, np.array([0.1, -0.2, 0.05]), K)
    A = oracle.helmholtz3d_far_operator(d, tg, K)
    print(f"\n[autodiff] ellipsoid nsub=2: {d.nelements} triangles, ndofs={d.ndofs}, "
          f"nsources={d.nsources}, ntargets={tg.shape[1]}", flush=True)
    return d, src, tg, sig, g, A


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def test_forward_matches_dense(setup):
    from pytential_amd.autodiff import single_layer_far
    d, src, tg, sig, g, A = setup
    phi = single_layer_far(src, sig, K, tg).cpu().numpy()
    err = _rel(phi, A @ sig)
    print(f"[autodiff] forward vs dense rel {err:.2e}", flush=True)
    assert err < 1e-12


def test_backward_matches_dense_adjoint(setup):
    import oracle
    from pytential_amd.autodiff import single_layer_far
    d, src, tg, sig, g, A = setup
    s = torch.tensor(sig, dtype=torch.complex128, device="cuda", requires_grad=True)
    k = torch.tensor(K, dtype=torch.float64, requires_grad=True)
    phi = single_layer_far(src, s, k, tg)
    gt = torch.as_tensor(g, device="cuda")
    loss = torch.sum(torch.abs(phi - gt) ** 2)
    loss.backward()
    r = A @ sig - g
    grad_ref = A.conj().T @ (2 * r)
    e_s = _rel(s.grad.cpu().numpy(), grad_ref)
    h = 1e-5
    dA = (oracle.helmholtz3d_far_operator(d, tg, K + h) -
          oracle.helmholtz3d_far_operator(d, tg, K - h)) / (2 * h)
    gk_ref = 2 * np.real(np.vdot(r, dA @ sig))
    e_k = abs(float(k.grad) - gk_ref) / abs(gk_ref)
    print(f"[autodiff] dL/dsigma vs A^H(2r) rel {e_s:.2e}; dL/dk {float(k.grad):.10e} "
          f"vs dense-FD {gk_ref:.10e} rel {e_k:.2e}", flush=True)
    assert e_s < 1e-12
    assert e_k < 1e-8


def test_gradients_match_central_differences(setup):
    from pytential_amd.autodiff import single_layer_far
    d, src, tg, sig, g, A = setup
    gt = torch.as_tensor(g, device="cuda")

    def loss_of(sv, kv):
        with torch.no_grad():
            phi = single_layer_far(src, sv, kv, tg)
            return float(torch.sum(torch.abs(phi - gt) ** 2))

    s = torch.tensor(sig, dtype=torch.complex128, device="cuda", requires_grad=True)
    k = torch.tensor(K, dtype=torch.float64, requires_grad=True)
    loss = torch.sum(torch.abs(single_layer_far(src, s, k, tg) - gt) ** 2)
    loss.backward()
    h = 1e-6
    fd_k = (loss_of(sig, K + h) - loss_of(sig, K - h)) / (2 * h)
    e_k = abs(float(k.grad) - fd_k) / abs(fd_k)
    # directional derivative along a random complex direction: for real L,
    # dL(sigma + t v)/dt = Re <grad, v> with torch's complex gradient
    rng = np.random.default_rng(4002)
    v = rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs)
    fd_v = (loss_of(sig + h * v, K) - loss_of(sig - h * v, K)) / (2 * h)
    an_v = float(np.real(np.vdot(s.grad.cpu().numpy(), v)))
    e_v = abs(an_v - fd_v) / abs(fd_v)
    print(f"[autodiff] central differences h=1e-6: dL/dk rel {e_k:.2e}, "
          f"directional dL/dsigma rel {e_v:.2e}", flush=True)
    assert e_k < 1e-7
    assert e_v < 1e-7


def test_gradcheck_tiny():
    from pytential_amd.autodiff import single_layer_far
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import ellipsoid
    d = ellipsoid(0, order=2, fine_order=3)
    src = QBXLayerPotentialSource(d, fine_order=3, qbx_order=3, fmm_order=10)
    tg, _ = _targets(8, 4003)
    rng = np.random.default_rng(4004)
    s = torch.tensor(rng.normal(size=d.ndofs) + 1j * rng.normal(size=d.ndofs),
                     dtype=torch.complex128, device="cuda", requires_grad=True)
    k = torch.tensor(3.0, dtype=torch.float64, requires_grad=True)
    ok = torch.autograd.gradcheck(lambda a, b: single_layer_far(src, a, b, tg), (s, k),
                                  eps=1e-6, atol=1e-8, rtol=1e-6)
    assert ok


def test_sphere_exterior_closed_form():
    import oracle
    from scipy.special import sph_harm_y

    from pytential_amd.autodiff import single_layer_far
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import icosphere
    d = icosphere(2, order=4, fine_order=6)
    src = QBXLayerPotentialSource(d, fine_order=6, qbx_order=3, fmm_order=10)
    tg, r = _targets(128, 4005)
    sig = np.full(d.ndofs, sph_harm_y(0, 0, 0.0, 0.0))
    phi = single_layer_far(src, sig, K, tg).cpu().numpy()
    ex = oracle.sphere_single_layer_exterior(0, K, 1.0, r) * sph_harm_y(0, 0, 0.0, 0.0)
    err = _rel(phi, ex)
    print(f"[autodiff] sphere S_k Y_00 exterior closed form rel {err:.2e}", flush=True)
    assert err < 1e-10


def test_near_targets_rejected():
    from pytential_amd.autodiff import single_layer_far
    from pytential_amd.qbx import QBXLayerPotentialSource
    from pytential_amd.surface import ellipsoid
    d = ellipsoid(1, order=4, fine_order=6)
    src = QBXLayerPotentialSource(d, fine_order=6, qbx_order=3, fmm_order=10)
    with pytest.raises(ValueError):
        single_layer_far(src, np.ones(d.ndofs), K, np.array([[0.9], [0.0], [0.0]]))


def test_point_kernel_phase_accuracy():
    """The pair kernel's e^{ikr} / (4 pi r) (own sin/cos: Cody-Waite
    reduction and Taylor series, a refined reciprocal square root) against
    numpy for single source-target pairs with k r from 1e-3 to 1e6, through
    qbx_helmholtz3d_p2p and its adjoint: within 4e-16 (1 + k r) relative
    (the distance itself carries a few ulp, which the phase multiplies by
    k r; the sin/cos reduction stays below 2 ulp up to |k r| = 1.6e6, beyond
    which the device library takes over)."""
    from pytential_amd._lib import call
    from pytential_amd.geometry import Context, _p, _stream
    ctx = Context.get()
    rng = np.random.default_rng(7)
    kr = np.concatenate([np.geomspace(1e-3, 1e6, 400), rng.uniform(0, 50, 200)])
    k = 3.7
    r = kr / k
    d = rng.normal(size=(3, kr.size))
    d /= np.linalg.norm(d, axis=0)
    worst = 0.0
    for i in range(kr.size):
        s = torch.as_tensor(np.array([[0.1], [-0.2], [0.3]]), device="cuda")
        t = torch.as_tensor(np.array([[0.1], [-0.2], [0.3]]) + r[i] * d[:, i:i + 1],
                            device="cuda")
        q = torch.ones(1, dtype=torch.complex128, device="cuda")
        pot = torch.empty(1, dtype=torch.complex128, device="cuda")
        call("qbx_helmholtz3d_p2p", ctx.h, _stream(), 1, _p(s), _p(q), 1, _p(t), k, _p(pot),
             None)
        adj = torch.empty(1, dtype=torch.complex128, device="cuda")
        call("qbx_helmholtz3d_p2p_adjoint", ctx.h, _stream(), 1, _p(s), 1, _p(t), k, _p(q),
             _p(adj))
        rr = float(np.linalg.norm(t.cpu().numpy() - s.cpu().numpy()))
        ref = np.exp(1j * k * rr) / (4 * np.pi * rr)
        e1 = abs(complex(pot.cpu()[0]) - ref) / abs(ref)
        e2 = abs(complex(adj.cpu()[0]) - np.conj(ref)) / abs(ref)
        worst = max(worst, max(e1, e2) / (1.0 + k * rr))
    print(f"\n[helm3d] pair kernel vs numpy over k r in [1e-3, 1e6]: worst rel / (1 + k r) "
          f"{worst:.2e}", flush=True)
    assert worst < 4e-16
