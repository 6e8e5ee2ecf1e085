"""The allocation-free apply born_apply_dev_noalloc, called through the C ABI:
its output equals that of the entry points it mirrors (born_forward_dev,
born_adjoint_dev) on two calls in a row, the second on a smaller plan and
input, while every device allocation of the library is made to fail; the
small plan also against the FP64 oracle.  A plan beyond the static
scratch array is rejected with the sizes in the message."""
import ctypes as C

import numpy as np
import pytest

from comp_ui_toolbox_amd import BornError, OperatorBorn, _lib
from comp_ui_toolbox_amd import configs as Cf
from comp_ui_toolbox_amd._lib import check
from oracle import born_oracle as O
from tests.test_oracle import tiny_spec

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "split_bf16": 1e-4, "split_fp16": 1e-4}
# Against the allocating entry points: the same arithmetic in FP32.  In the
# split forms the allocating entry points pack several right-hand sides into
# wider chunks, and a one-wave plan's one-column chunks (the layout of the
# allocation-free apply) run on the FP32 vector kernels, the wider ones on
# the split MFMA kernels: the two then differ by the split form's own error
# (about 5e-6 in BF16), well inside its 1e-4 bar against FP64.
AGREE = {"fp32": 1e-6, "split_bf16": 2e-5, "split_fp16": 2e-5}


def rnd(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def noalloc(op, mode, x, n_rhs, out):
    import torch
    s = torch.cuda.current_stream().cuda_stream
    check(_lib.lib().born_apply_dev_noalloc(op._plan, int(mode), C.c_void_p(x.data_ptr()), int(n_rhs),
                                            C.c_void_p(out.data_ptr()), C.c_void_p(s)))


def rel(a, b):
    import torch
    a = a.to(torch.complex128)
    b = b.to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


@pytest.mark.parametrize("precision", ["fp32", "split_bf16", "split_fp16"])
def test_noalloc_matches_the_allocating_entry_points(torch_cuda, precision):
    torch = torch_cuda
    L = _lib.lib()
    small_spec = tiny_spec(N_e=20, Nx=80, Nz=4)
    big = OperatorBorn.from_sequence(Cf.config0(), precision=precision)
    small = OperatorBorn.from_sequence(small_spec, precision=precision)
    T = O.build_tables(small_spec)
    for call, (op, n_rhs, seed) in enumerate(((big, 2, 1), (small, 1, 2))):
        print(f"no-alloc call {call}: {precision}, {op.N_points} points, {n_rhs} rhs", flush=True)
        x = torch.tensor(rnd((n_rhs, op.N_points), seed), device="cuda")
        y = torch.tensor(rnd((n_rhs, op.N_observations), seed + 10), device="cuda")
        u_ref = op.forward_dev(x)
        g_ref = op.adjoint_dev(y)
        u = torch.full_like(u_ref, float("nan"))
        g = torch.full_like(g_ref, float("nan"))
        torch.cuda.synchronize()
        live = L.born_debug_live_allocations()
        L.born_debug_fail_allocation(0)          # the library's next device allocation would fail
        try:
            noalloc(op, 1, x, n_rhs, u)
            noalloc(op, 2, y, n_rhs, g)
            torch.cuda.synchronize()
        finally:
            L.born_debug_fail_allocation(-1)
        assert L.born_debug_live_allocations() == live
        e_u, e_g = rel(u, u_ref), rel(g, g_ref)
        print(f"  against born_forward_dev {e_u:.2e}, born_adjoint_dev {e_g:.2e}", flush=True)
        assert e_u < AGREE[precision] and e_g < AGREE[precision]
        if op is small:
            x64 = x.cpu().numpy().T.astype(np.complex128)
            y64 = y.cpu().numpy().T.astype(np.complex128)
            assert O.rel_l2(u.cpu().numpy().T, O.forward(T, x64)) < TOL[precision]
            assert O.rel_l2(g.cpu().numpy().T, O.adjoint(T, y64)) < TOL[precision]
    big.close()
    small.close()


def test_noalloc_rejects_what_it_cannot_serve(torch_cuda):
    """A plan on the general path, and a plan whose one-rhs scratch exceeds
    BORN_NOALLOC_SCRATCH_BYTES (64 MiB: here 2^23 grid points), get
    born:Unsupported; the allocating entry points still serve them."""
    torch = torch_cuda
    spec = tiny_spec(N_e=20, Nx=80, Nz=4)
    op = OperatorBorn.from_sequence(spec)
    op.force_general(True)
    x = torch.tensor(rnd((1, op.N_points), 3), device="cuda")
    u = torch.empty((1, op.N_observations), dtype=torch.complex64, device="cuda")
    with pytest.raises(BornError, match="Unsupported"):
        noalloc(op, 1, x, 1, u)
    op.close()
    big = tiny_spec(N_e=20, Nx=1024, Nz=8192, npp=4, nfp=1, freqs=np.array([3e6]))
    big.waves = big.waves[:1]
    big.mixes = big.mixes[:1]
    opb = OperatorBorn.from_sequence(big)
    assert opb.path == "structured"
    xb = torch.zeros((1, opb.N_points), dtype=torch.complex64, device="cuda")
    ub = torch.empty((1, opb.N_observations), dtype=torch.complex64, device="cuda")
    with pytest.raises(BornError, match="bytes of scratch"):
        noalloc(opb, 1, xb, 1, ub)
    opb.forward_dev(xb, out=ub)
    torch.cuda.synchronize()
    opb.close()
