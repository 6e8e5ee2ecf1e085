"""The operator pair behind torch.autograd (backward = the adjoint kernel),
checked on the config-3 loss 0.5 ||A gamma - u||^2 + lambda ||Psi gamma||_1
(Psi: orthogonal Daubechies-4, lambda = 1e-3 ||A^T u||_inf, u from a point
phantom plus complex Gaussian noise at 20 dB SNR) against FP64 central
finite differences (relative step 1e-3) on 64 random directions and against
the gradient assembled from the FP64 oracle's adjoint."""
import numpy as np
import pytest

from comp_ui_toolbox_amd import configs as C
from comp_ui_toolbox_amd import OperatorBorn, OperatorBatch
from comp_ui_toolbox_amd import wavelet as W
from oracle import born_oracle as O
from tests.test_oracle import tiny_spec

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def problem(spec, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    A = O.assemble_matrix(O.build_tables(spec))
    g_true = C.phantom_points(spec, n_points=12, seed=seed)
    s = A @ g_true
    noise = rng.standard_normal(s.shape) + 1j * rng.standard_normal(s.shape)
    noise *= 0.1 * np.linalg.norm(s) / np.linalg.norm(noise)            # 20 dB SNR
    u = s + noise
    lam = 1e-3 * np.max(np.abs((A.conj().T @ u).real))
    return A, u, lam, rng


def test_gradient_against_fp64_finite_differences(torch_cuda):
    torch = torch_cuda
    from comp_ui_toolbox_amd.autograd import born_apply
    spec = tiny_spec(N_e=4, Nx=16, Nz=16, npp=4)
    op = OperatorBorn.from_sequence(spec)
    A, u, lam, rng = problem(spec, 61)
    Nz, Nx = 16, 16
    lv = 2
    # a point whose wavelet coefficients are all at least 0.5 in magnitude, so
    # no step of the finite differences crosses the kink of |.|
    c = np.sign(rng.standard_normal((Nz, Nx))) * (0.5 + np.abs(rng.standard_normal((Nz, Nx))))
    g0 = W.idwt2(torch.tensor(c), lv).numpy().reshape(-1)

    def psi64(x):
        return W.dwt2(torch.tensor(x.reshape(Nz, Nx)), lv).numpy()

    def loss64(x):
        r = A @ x - u
        return 0.5 * np.vdot(r, r).real + lam * np.abs(psi64(x)).sum()

    # GPU gradient through the HIP forward/adjoint
    x = torch.tensor(g0.astype(np.float32), device="cuda", requires_grad=True)
    ut = torch.tensor(u.astype(np.complex64), device="cuda")
    up = born_apply(op, x)
    loss = 0.5 * (up - ut).abs().pow(2).sum() + lam * W.dwt2(x.view(Nz, Nx), lv).abs().sum()
    loss.backward()
    loss = loss.detach()
    grad = x.grad.double().cpu().numpy()

    # against FP64 central differences
    D = rng.standard_normal((64, g0.size))
    fd = np.empty(64)
    for k in range(64):
        h = 1e-3 * np.linalg.norm(g0) / np.linalg.norm(D[k])
        fd[k] = (loss64(g0 + h * D[k]) - loss64(g0 - h * D[k])) / (2 * h)
    gd = D @ grad
    e_fd = np.linalg.norm(gd - fd) / np.linalg.norm(fd)
    # against the FP64 oracle's adjoint
    g64 = (A.conj().T @ (A @ g0 - u)).real + lam * W.idwt2(torch.tensor(np.sign(psi64(g0))), lv).numpy().reshape(-1)
    e_adj = np.linalg.norm(grad - g64) / np.linalg.norm(g64)
    log("autograd: 64 directions vs FP64 central differences rel %.2e; gradient vs FP64 adjoint rel %.2e"
        % (e_fd, e_adj))
    assert e_fd < 1e-5 and e_adj < 1e-5
    assert abs(loss.item() - loss64(g0)) / loss64(g0) < 1e-5


def test_complex_input_gradient_is_adjoint(torch_cuda):
    torch = torch_cuda
    from comp_ui_toolbox_amd.autograd import born_apply
    spec = tiny_spec(N_e=4, Nx=16, Nz=6, npp=4)
    op = OperatorBorn.from_sequence(spec)
    A = O.assemble_matrix(O.build_tables(spec))
    rng = np.random.Generator(np.random.PCG64(62))
    g0 = rng.standard_normal(A.shape[1]) + 1j * rng.standard_normal(A.shape[1])
    w = rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0])
    x = torch.tensor(g0.astype(np.complex64), device="cuda", requires_grad=True)
    wt = torch.tensor(w.astype(np.complex64), device="cuda")
    # L = Re <w, A x>  ->  dL/dx (torch convention, conj Wirtinger x2) = A^H w
    L = torch.real(torch.vdot(wt, born_apply(op, x)))
    L.backward()
    ref = A.conj().T @ w
    e = np.linalg.norm(x.grad.cpu().numpy() - ref) / np.linalg.norm(ref)
    log("complex gradient vs A^H w: rel %.2e" % e)
    assert e < 1e-5


def test_batch_autograd_matches_single(torch_cuda):
    torch = torch_cuda
    from comp_ui_toolbox_amd.autograd import born_apply, born_batch_apply
    specs = [tiny_spec(N_e=16, Nx=n, Nz=3, npp=4) for n in (24, 64, 100)]
    ops = [OperatorBorn.from_sequence(s) for s in specs]
    B = OperatorBatch(ops)
    rng = np.random.Generator(np.random.PCG64(63))
    gs = [torch.tensor(rng.standard_normal(op.N_points).astype(np.float32), device="cuda", requires_grad=True)
          for op in ops]
    us = [torch.tensor((rng.standard_normal(op.N_observations) + 1j * rng.standard_normal(op.N_observations))
                       .astype(np.complex64), device="cuda") for op in ops]
    outs = born_batch_apply(B, gs)
    loss = sum(0.5 * (o - u).abs().pow(2).sum() for o, u in zip(outs, us))
    loss.backward()
    for op, g, u in zip(ops, gs, us):
        x = g.detach().clone().requires_grad_(True)
        l1 = 0.5 * (born_apply(op, x) - u).abs().pow(2).sum()
        l1.backward()
        e = float((g.grad - x.grad).norm() / x.grad.norm())
        assert e < 1e-6, e
    log("batched autograd: gradients equal the per-instance ones")
