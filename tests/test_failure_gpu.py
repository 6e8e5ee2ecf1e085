"""Forced device-allocation failures (born_debug_fail_allocation): at every
allocation a call can make, the call reports born:OutOfMemory, releases what
it had allocated (born_debug_live_allocations returns to its value before
the call), and the library keeps giving correct results afterwards.  The
failures are forced on the host, before hipMalloc; no kernel is affected."""
import ctypes as C

import numpy as np
import pytest

from comp_ui_toolbox_amd import BornError, OperatorBorn, _lib
from comp_ui_toolbox_amd import configs as Cf
from oracle import born_oracle as O
from tests.test_oracle import mixed_tiny_tables, tiny_spec

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rnd(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def live():
    return _lib.lib().born_debug_live_allocations()


def fail_at(k):
    _lib.lib().born_debug_fail_allocation(k)


def sweep(make, max_allocs=400, strict=True):
    """Fail allocation k = 0, 1, ... of make() until it succeeds; every
    failure must be an OutOfMemory and, when `strict`, leave the live count
    unchanged (apply calls may keep the staging buffers they got before the
    failure: the plan owns and reuses them)."""
    failures = 0
    try:
        for k in range(max_allocs):
            before = live()
            fail_at(k)
            try:
                out = make()
            except BornError as e:
                assert e.identifier == "born:OutOfMemory", e
                if strict:
                    assert live() == before, f"allocation {k}: {live() - before} device allocations leaked"
                failures += 1
                continue
            finally:
                fail_at(-1)
            return out, failures
    finally:
        fail_at(-1)
    raise AssertionError("never succeeded")


def test_plan_creation_from_tables(torch_cuda):
    _, T = mixed_tiny_tables()
    base = live()
    op, n = sweep(lambda: OperatorBorn.from_tables(T))
    assert n >= 10
    g = rnd((T.N_points, 2), 1)
    assert O.rel_l2(op.forward(g), O.forward(T, g)) < TOL
    op.close()
    assert live() == base
    print(f"from_tables: {n} allocation points, each failed cleanly", flush=True)


def test_plan_creation_from_geometry(torch_cuda):
    spec = tiny_spec(N_e=5, Nx=17, Nz=7, npp=3)
    T = O.build_tables(spec)
    base = live()
    op, n = sweep(lambda: OperatorBorn.from_sequence(spec))
    assert n >= 10
    u = rnd((T.N_observations(), 1), 2)
    assert O.rel_l2(op.adjoint(u), O.adjoint(T, u)) < TOL
    op.close()
    assert live() == base
    print(f"from_sequence: {n} allocation points, each failed cleanly", flush=True)


@pytest.mark.parametrize("precision", ["fp32", "split_bf16"])
def test_first_apply_workspace(torch_cuda, precision):
    """The first forward / adjoint with a new right-hand-side count builds
    the plan's workspace; a failure there must leave the plan usable."""
    spec = tiny_spec(N_e=40, Nx=160, Nz=3, npp=4)
    T = O.build_tables(spec)
    base = live()
    op = OperatorBorn.from_tables(T, precision=precision)
    g = rnd((T.N_points, 3), 3)
    u = rnd((T.N_observations(), 3), 4)
    out, n = sweep(lambda: op.forward(g), strict=False)
    assert n >= 3
    assert O.rel_l2(out, O.forward(T, g)) < (TOL if precision == "fp32" else 1e-4)
    out, _ = sweep(lambda: op.adjoint(u[:, :2]), strict=False)   # another rhs count: workspace rebuilt
    assert O.rel_l2(out, O.adjoint(T, u[:, :2])) < (TOL if precision == "fp32" else 1e-4)
    op.close()
    assert live() == base


def test_tgc_and_energy_paths(torch_cuda):
    _, T = mixed_tiny_tables()
    base = live()
    op = OperatorBorn.from_tables(T)
    sizes = op.meta["mix_n_obs"]
    kernels = [np.array([0.2, 1.0 + 0.1j, 0.2])] * len(sizes)
    _, n = sweep(lambda: op.set_tgc(kernels))
    assert n >= 1
    u = rnd(T.N_observations(), 5)
    out, _ = sweep(lambda: op.tgc_forward(u))
    assert O.rel_l2(out, O.tgc_apply(kernels, sizes, u, 1)) < 1e-12
    E, _ = sweep(lambda: op.energy_rx(0, tgc=True))
    first = len(T.meas[0].mixes)
    assert O.rel_l2(E, O.energy_rx(T, 0, kernels[:first])) < TOL
    op.close()
    assert live() == base


def test_batch_creation(torch_cuda):
    """Batch creation over two uneven plans, then an eager forward through it."""
    import torch
    from comp_ui_toolbox_amd import OperatorBatch
    s1 = tiny_spec(N_e=5, Nx=17, Nz=7, npp=3)
    s2 = tiny_spec(N_e=5, Nx=21, Nz=4, npp=3)
    T1, T2 = O.build_tables(s1), O.build_tables(s2)
    base = live()
    ops = [OperatorBorn.from_tables(T1), OperatorBorn.from_tables(T2)]
    B, n = sweep(lambda: OperatorBatch(ops), strict=False)   # plans keep the workspaces built for the batch
    assert n >= 1
    g1, g2 = rnd(T1.N_points, 6), rnd(T2.N_points, 7)
    xs = [torch.tensor(g.astype(np.complex64), device="cuda") for g in (g1, g2)]
    us = B.forward_dev(xs)
    torch.cuda.synchronize()
    assert O.rel_l2(us[0].cpu().numpy(), O.forward(T1, g1.astype(np.complex64))) < TOL
    assert O.rel_l2(us[1].cpu().numpy(), O.forward(T2, g2.astype(np.complex64))) < TOL
    B.close()
    for op in ops:
        op.close()
    assert live() == base
