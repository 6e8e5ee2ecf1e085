"""Time gain compensation on the device against the oracle's DFT-based
convolution (fft.m), and the operator with TGC (op_A with tgc=True)
against the oracle's operator followed by that convolution.

Tolerances: the FP64 host entry point 1e-12 relative L2; the complex64
device entry point 1e-6 (FP32 accumulation over at most 2 N_f - 1 taps);
op_A with TGC 1e-5, the plain FP32 operator's bar."""
import numpy as np
import pytest

from comp_ui_toolbox_amd import BornError, OperatorBorn
from comp_ui_toolbox_amd import configs as C
from comp_ui_toolbox_amd import tgc
from oracle import born_oracle as O
from tests.test_oracle import mixed_tiny_tables, tiny_spec

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-12
TOL_DEV = 1e-6
TOL_OP = 1e-5


def rnd(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def log(*a):
    print(*a, flush=True)


def random_kernels(sizes, halves, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return [rng.standard_normal(2 * M + 1) + 1j * rng.standard_normal(2 * M + 1) for M in halves]


def check_both_entry_points(op, kernels, sizes, n_rhs, seed):
    import torch
    op.set_tgc(kernels)
    u = rnd((op.N_observations, n_rhs), seed)
    ud = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    errs = []
    for mode in (1, 2):
        ref = O.tgc_apply(kernels, sizes, u, mode)
        e_h = O.rel_l2(op._tgc_host(u, mode), ref)
        ref32 = O.tgc_apply(kernels, sizes, u.astype(np.complex64), mode)
        e_d = O.rel_l2(op.tgc_dev(ud, mode).cpu().numpy().T, ref32)
        errs.append((e_h, e_d))
        assert e_h < TOL_F64 and e_d < TOL_DEV, (mode, e_h, e_d)
    return errs


@pytest.mark.parametrize("n_rhs", [1, 3])
def test_tgc_tiny_mixes(torch_cuda, n_rhs):
    """Mixes of 3 and 4 frequencies; kernels shorter than, equal to and
    longer than the mix (truncated); non-Hermitian taps."""
    _, T = mixed_tiny_tables(distinct=True)
    op = OperatorBorn.from_tables(T)
    sizes = op.meta["mix_n_obs"]
    halves = [1, 0, 7, 2][: len(sizes)]
    kernels = random_kernels(sizes, halves, 11)
    errs = check_both_entry_points(op, kernels, sizes, n_rhs, 12)
    log("tgc tiny mixes, %d rhs: %s" % (n_rhs, " ".join("%.1e/%.1e" % e for e in errs)))


def test_tgc_long_mixes_several_tiles(torch_cuda):
    """600 frequencies per mix: three tiles of 256 rows, and tap counts that
    take one, two and five chunks of 256 (M = 100, 300, 599 and one past the
    mix, truncated)."""
    freqs = 2.0e6 + 5e3 * np.arange(600)
    spec = tiny_spec(N_e=2, Nx=9, Nz=3, npp=4, freqs=freqs)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    sizes = op.meta["mix_n_obs"]
    assert sizes == [600] * 4
    kernels = random_kernels(sizes, [100, 300, 599, 700], 21)
    errs = check_both_entry_points(op, kernels, sizes, 2, 22)
    log("tgc long mixes: %s" % " ".join("%.1e/%.1e" % e for e in errs))


def test_tgc_exponential_config1(torch_cuda):
    """The exponential TGC of the 128-element, 257-frequency, 11-wave
    sequence (1408 mixes) on the device, against the oracle, forward and
    adjoint, two right-hand sides."""
    spec = C.config1()
    op = OperatorBorn.from_sequence(spec)
    sizes = op.meta["mix_n_obs"]
    n_mix = len(sizes)
    df = float(spec.frequencies[1] - spec.frequencies[0])
    iv = np.tile([[2.0e-6, 30.0e-6]], (n_mix, 1))
    a = np.log(10 ** (40 / 20)) / 28e-6          # 40 dB over the recording
    kernels = tgc.exponential_kernels(a, iv, df, -60.0)
    errs = check_both_entry_points(op, kernels, sizes, 2, 31)
    log("tgc config 1 (M=%d, N_f=%d): %s" % ((len(kernels[0]) - 1) // 2, sizes[0],
                                             " ".join("%.1e/%.1e" % e for e in errs)))


def test_op_A_with_tgc(torch_cuda):
    """op_A(x, 1, tgc) = TGC(A x) and op_A(y, 2, tgc) = A^H TGC^H y
    (forward_scalar / adjoint_scalar), and the pair stays adjoint."""
    _, T = mixed_tiny_tables(distinct=True)
    op = OperatorBorn.from_tables(T)
    sizes = op.meta["mix_n_obs"]
    kernels = [tgc.kernel_from_coefficients(tgc.exponential_coefficients(2e5, (1e-6, 9e-6), 10e-6, -40.0))
               for _ in sizes]
    op.set_tgc(kernels)
    x = rnd((T.N_points, 2), 41)
    y = rnd((T.N_observations(), 2), 42)
    e_f = O.rel_l2(op.op_A(x, 1, tgc=True), O.tgc_apply(kernels, sizes, O.forward(T, x), 1))
    e_a = O.rel_l2(op.op_A(y, 2, tgc=True), O.adjoint(T, O.tgc_apply(kernels, sizes, y, 2)))
    lhs = np.vdot(y[:, 0], op.op_A(x[:, 0], 1, tgc=True))
    rhs = np.vdot(op.op_A(y[:, 0], 2, tgc=True), x[:, 0])
    log("op_A with TGC: forward %.2e adjoint %.2e adjointness %.2e" % (e_f, e_a, abs(lhs - rhs) / abs(lhs)))
    assert e_f < TOL_OP and e_a < TOL_OP
    assert abs(lhs - rhs) < TOL_OP * abs(lhs)


def test_tgc_errors(torch_cuda):
    import torch
    _, T = mixed_tiny_tables(distinct=True)
    op = OperatorBorn.from_tables(T)
    u = rnd(T.N_observations(), 51)
    with pytest.raises(BornError):
        op.tgc_forward(u)                                   # nothing set
    sizes = op.meta["mix_n_obs"]
    with pytest.raises(BornError):
        op.set_tgc([np.ones(3)] * (len(sizes) - 1))         # one kernel short
    with pytest.raises(BornError):
        op.set_tgc([np.ones(2)] * len(sizes))               # even length
    op.set_tgc([np.ones(1)] * len(sizes))
    assert np.allclose(op.tgc_forward(u), u, atol=1e-15)    # identity kernel
    ud = torch.from_numpy(u.astype(np.complex64)).cuda()
    with pytest.raises(BornError):
        op.tgc_dev(ud, 1, out=ud)                           # in place
    op.set_tgc(None)
    with pytest.raises(BornError):
        op.tgc_forward(u)


def _kernels_of_measurement(op, T, kernels, m):
    first = sum(len(ms.mixes) for ms in T.meas[:m])
    return kernels[first:first + len(T.meas[m].mixes)]


def test_energy_rx_tgc_mixed_tables(torch_cuda):
    """Received energies under TGC (energy_rx_canonical with a diagonal
    LT_tgc) on mixed tables, host and device entry points, both
    measurements, against the oracle."""
    _, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    sizes = op.meta["mix_n_obs"]
    kernels = random_kernels(sizes, [1, 0, 5, 2], 71)
    op.set_tgc(kernels)
    for m in range(len(T.meas)):
        ref = O.energy_rx(T, m, _kernels_of_measurement(op, T, kernels, m))
        e_h = O.rel_l2(op.energy_rx(m, tgc=True), ref)
        e_d = O.rel_l2(op.energy_rx_dev(m, tgc=True).cpu().numpy(), ref)
        log("energy_rx with TGC measurement %d: host %.2e device %.2e" % (m, e_h, e_d))
        assert e_h < TOL_OP and e_d < TOL_OP
    op.set_tgc(None)
    with pytest.raises(BornError):
        op.energy_rx(0, tgc=True)


def test_energy_rx_tgc_long_mixes(torch_cuda):
    """Mixes of 600 frequencies, longer than the LDS holds: the rows go to
    the spill buffer."""
    freqs = 2.0e6 + 5e3 * np.arange(600)
    spec = tiny_spec(N_e=2, Nx=9, Nz=3, npp=4, freqs=freqs)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    kernels = random_kernels(op.meta["mix_n_obs"], [40, 0, 350, 7], 72)
    op.set_tgc(kernels)
    for m in range(2):
        ref = O.energy_rx(T, m, _kernels_of_measurement(op, T, kernels, m))
        e = O.rel_l2(op.energy_rx(m, tgc=True), ref)
        log("energy_rx with TGC, 600-frequency mixes, measurement %d: %.2e" % (m, e))
        assert e < TOL_OP


def test_energy_rx_tgc_config1(torch_cuda):
    """Config 1, first measurement (128 mixes of 257 frequencies), exponential
    TGC, at 24 sampled points against tables restricted exactly to them."""
    spec = C.config1()
    op = OperatorBorn.from_sequence(spec)
    sizes = op.meta["mix_n_obs"]
    df = float(spec.frequencies[1] - spec.frequencies[0])
    a = np.log(100.0) / 28e-6
    kernels = tgc.exponential_kernels(a, np.tile([[2.0e-6, 30.0e-6]], (len(sizes), 1)), df, -60.0)
    op.set_tgc(kernels)
    rng = np.random.Generator(np.random.PCG64(73))
    pts = np.sort(rng.choice(spec.fov.N_points, 24, replace=False))
    R = O.restricted_tables(spec, pts)
    E = op.energy_rx(0, tgc=True)
    e = O.rel_l2(E[pts], O.energy_rx(R, 0, kernels[: len(R.meas[0].mixes)]))
    log("energy_rx with TGC, config 1 measurement 0 at 24 points: %.2e" % e)
    assert e < TOL_OP
