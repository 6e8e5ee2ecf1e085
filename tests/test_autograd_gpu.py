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


# ------------------------------------------------------------ config-3 scale
@pytest.fixture(scope="module")
def c3_loss(torch_cuda):
    """The config-3 loss over its 24 uneven instances through one batch
    (born_batch_apply): per instance 0.5 ||A gamma - u||^2 + lambda
    ||Psi gamma||_1 with Psi the D4 wavelet at full depth, u = A gamma_true +
    complex Gaussian noise at 20 dB SNR (gamma_true: the config-1 speckle
    phantom on the instance's FOV), lambda = 1e-3 ||Re A^H u||_inf.  The
    gradient is taken at gamma_0 = Psi^T c with every |c| >= 0.5, so no
    finite-difference step crosses the kink of |.|."""
    torch = torch_cuda
    from comp_ui_toolbox_amd.autograd import born_batch_apply
    specs = C.config3_instances()
    ops = [OperatorBorn.from_sequence(s) for s in specs]
    B = OperatorBatch(ops)
    rng = np.random.Generator(np.random.PCG64(404))
    clean = B.forward_dev([torch.tensor(C.phantom(s).astype(np.complex64), device="cuda") for s in specs])
    inst = []
    for s, op, c in zip(specs, ops, clean):
        Nx, Nz = s.fov.N_points_axis
        lv = W.max_levels((Nz, Nx))
        c64 = c.cpu().numpy().astype(np.complex128)
        noise = rng.standard_normal(c64.shape) + 1j * rng.standard_normal(c64.shape)
        noise *= 0.1 * np.linalg.norm(c64) / np.linalg.norm(noise)
        u = c64 + noise
        lam = 1e-3 * float(op.adjoint_dev(torch.tensor(u.astype(np.complex64), device="cuda")).real.abs().max())
        coef = np.sign(rng.standard_normal((Nz, Nx))) * (0.5 + np.abs(rng.standard_normal((Nz, Nx))))
        g0 = W.idwt2(torch.tensor(coef), lv).numpy().reshape(-1)
        inst.append(dict(spec=s, op=op, u=u, lam=lam, lv=lv, g0=g0, Nx=Nx, Nz=Nz))
    xs = [torch.tensor(d["g0"].astype(np.float32), device="cuda", requires_grad=True) for d in inst]
    outs = born_batch_apply(B, xs)
    loss = 0
    for d, x, o in zip(inst, xs, outs):
        ut = torch.tensor(d["u"].astype(np.complex64), device="cuda")
        loss = loss + 0.5 * (o - ut).abs().pow(2).sum() + d["lam"] * W.dwt2(x.view(d["Nz"], d["Nx"]), d["lv"]).abs().sum()
    loss.backward()
    for d, x, o in zip(inst, xs, outs):
        d["grad"] = x.grad.double().cpu().numpy()
        d["r_gpu"] = o.detach().cpu().numpy().astype(np.complex128) - d["u"]
    log("config3 loss over 24 instances: %.6e" % loss.item())
    return inst


def l1_gradient(d, x):
    """lambda Psi^T sign(Psi x) in FP64."""
    torch = __import__("torch")
    Nz, Nx, lv = d["Nz"], d["Nx"], d["lv"]
    s = np.sign(W.dwt2(torch.tensor(x.reshape(Nz, Nx)), lv).numpy())
    return d["lam"] * W.idwt2(torch.tensor(s), lv).numpy().reshape(-1)


@pytest.fixture(scope="module")
def c3_fp64_tables(c3_loss):
    """FP64 tables of the full operator of each 128 x 128 instance (every
    frequency; the FOV is narrower than the aperture, so they index the
    laterally extended grid)."""
    out = {}
    for d in c3_loss:
        if d["Nx"] == 128:
            s = d["spec"]
            T, rows = O.frequency_tables(s, np.arange(len(s.frequencies)))
            assert np.array_equal(rows, np.arange(d["op"].N_observations))
            out[s.name] = T
    return out


def test_config3_gradient_against_fp64_at_128(c3_loss, c3_fp64_tables):
    """The three 128 x 128 instances: the whole gradient against the FP64
    oracle's (residual and adjoint of the full operator in FP64)."""
    for d in c3_loss:
        if d["Nx"] != 128:
            continue
        s = d["spec"]
        T = c3_fp64_tables[s.name]
        r64 = O.forward_by_frequency(T, d["g0"]) - d["u"]
        g64 = O.adjoint_by_frequency(T, r64).real + l1_gradient(d, d["g0"])
        e = np.linalg.norm(d["grad"] - g64) / np.linalg.norm(g64)
        log("%s (%d f, %d waves): gradient vs FP64 rel %.2e" % (s.name, len(s.frequencies), s.N_measurements, e))
        assert e < 1e-5


def test_config3_gradient_against_fp64_finite_differences(c3_loss, c3_fp64_tables):
    """The smallest 128 x 128 instance: the gradient against FP64 central
    differences of the loss (relative step 1e-3) on 16 random directions.
    The data term's central difference is evaluated exactly as
    Re <A gamma_0 - u, A v> (the term is quadratic), with A in FP64."""
    torch = __import__("torch")
    d = min((d for d in c3_loss if d["Nx"] == 128),
            key=lambda d: len(d["spec"].frequencies) * d["spec"].N_measurements)
    s = d["spec"]
    T = c3_fp64_tables[s.name]
    rng = np.random.Generator(np.random.PCG64(405))
    V = rng.standard_normal((d["g0"].size, 16))
    AV = O.forward_by_frequency(T, np.concatenate([d["g0"][:, None], V], axis=1))
    r64 = AV[:, 0] - d["u"]
    Nz, Nx, lv = d["Nz"], d["Nx"], d["lv"]

    def l1(x):
        return d["lam"] * np.abs(W.dwt2(torch.tensor(x.reshape(Nz, Nx)), lv).numpy()).sum()

    fd = np.empty(16)
    for k in range(16):
        h = 1e-3 * np.linalg.norm(d["g0"]) / np.linalg.norm(V[:, k])
        fd[k] = np.vdot(r64, AV[:, k + 1]).real + (l1(d["g0"] + h * V[:, k]) - l1(d["g0"] - h * V[:, k])) / (2 * h)
    gd = V.T @ d["grad"]
    e = np.linalg.norm(gd - fd) / np.linalg.norm(fd)
    log("%s: 16 directions vs FP64 central differences rel %.2e" % (s.name, e))
    assert e < 1e-5


def test_config3_gradient_against_reference_adjoint_all_instances(c3_loss):
    """Every one of the 24 instances: the data term of the gradient (the
    gradient minus the FP64 l1 term) at 4 sampled points against the FP64
    oracle's adjoint of the residual, with tables restricted exactly to
    those points."""
    rng = np.random.Generator(np.random.PCG64(406))
    worst = 0.0
    for d in c3_loss:
        s = d["spec"]
        pts = np.sort(rng.choice(s.fov.N_points, 4, replace=False))
        R = O.restricted_tables(s, pts)
        ref = O.adjoint(R, d["r_gpu"]).real
        got = (d["grad"] - l1_gradient(d, d["g0"]))[pts]
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        worst = max(worst, e)
        assert e < 1e-5, (s.name, e)
    log("config3: data-term gradient at 4 points of each of the 24 instances vs FP64 adjoint, worst rel %.2e" % worst)
