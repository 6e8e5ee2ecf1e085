"""Batched operator instances (born_batch_*): one persistent contraction
launch per direction over plans of uneven FOV, frequency and wave counts.
Checked against the plans applied one by one, against the FP64 oracle, under
hipGraph replay, and after autotuning."""
import numpy as np
import pytest

from comp_ui_toolbox_amd import configs as C
from comp_ui_toolbox_amd import OperatorBorn, OperatorBatch
from oracle import born_oracle as O
from tests.test_oracle import tiny_spec

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5
TOL_SPLIT = 1e-4


def log(*a):
    print(*a, flush=True)


def uneven_specs(N_e):
    med = C.MediumSpec()

    def waves(arr, n):
        return [C.qpw_wave(arr, a, med.c_avg) for a in np.linspace(-0.2, 0.2, n)]

    out = []
    for Nx, Nz, nf, nw in [(64, 4, 3, 2), (24, 3, 4, 1), (100, 2, 5, 3), (300, 2, 2, 2), (40, 5, 3, 1)]:
        f = np.linspace(2.0e6, 5.0e6, nf)
        s = tiny_spec(N_e=N_e, Nx=Nx, Nz=Nz, npp=4, freqs=f)
        s.waves = waves(s.array, nw)
        s.mixes = [C.identity_mixes(N_e) for _ in s.waves]
        out.append(s)
    return out


def inputs(torch, ops, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    g = [torch.tensor((rng.standard_normal(op.N_points) + 1j * rng.standard_normal(op.N_points)).astype(np.complex64),
                      device="cuda") for op in ops]
    u = [torch.tensor((rng.standard_normal(op.N_observations)
                       + 1j * rng.standard_normal(op.N_observations)).astype(np.complex64), device="cuda")
         for op in ops]
    return g, u


def rel(a, b):
    return float((a - b).norm() / b.norm())


@pytest.fixture(scope="module")
def fp32_set(torch_cuda):
    specs = uneven_specs(16)
    ops = [OperatorBorn.from_sequence(s) for s in specs]
    assert all(op.path == "structured" for op in ops)
    return specs, ops


def test_batch_matches_single_plans(torch_cuda, fp32_set):
    torch = torch_cuda
    specs, ops = fp32_set
    B = OperatorBatch(ops)
    g, u = inputs(torch, ops, 11)
    fb = B.forward_dev(g)
    ab = B.adjoint_dev(u)
    for i, op in enumerate(ops):
        ef = rel(fb[i], op.forward_dev(g[i]))
        ea = rel(ab[i], op.adjoint_dev(u[i]))
        log("batch instance %d: forward %.2e adjoint %.2e vs single plan" % (i, ef, ea))
        assert ef < 1e-6 and ea < 1e-6


def test_batch_against_oracle(torch_cuda, fp32_set):
    torch = torch_cuda
    specs, ops = fp32_set
    B = OperatorBatch(ops)
    g, u = inputs(torch, ops, 12)
    fb = B.forward_dev(g)
    ab = B.adjoint_dev(u)
    for i, spec in enumerate(specs):
        T = O.direct_tables(spec)
        ef = O.rel_l2(fb[i].cpu().numpy(), O.forward(T, g[i].cpu().numpy().astype(np.complex128)))
        ea = O.rel_l2(ab[i].cpu().numpy(), O.adjoint(T, u[i].cpu().numpy().astype(np.complex128)))
        log("batch instance %d vs FP64 oracle: forward %.2e adjoint %.2e" % (i, ef, ea))
        assert ef < TOL_FP32 and ea < TOL_FP32


def test_batch_graph_replay(torch_cuda, fp32_set):
    torch = torch_cuda
    specs, ops = fp32_set
    B = OperatorBatch(ops)
    g, u = inputs(torch, ops, 13)
    ref_f = [x.clone() for x in B.forward_dev(g)]
    ref_a = [x.clone() for x in B.adjoint_dev(u)]
    B.set_graph(True)
    outs_f = [torch.empty_like(x) for x in ref_f]
    outs_a = [torch.empty_like(x) for x in ref_a]
    for rep in range(3):
        B.forward_dev(g, outs=outs_f)
        B.adjoint_dev(u, outs=outs_a)
        torch.cuda.synchronize()
        for i in range(len(ops)):
            assert torch.equal(outs_f[i], ref_f[i]), (rep, i)
            assert torch.equal(outs_a[i], ref_a[i]), (rep, i)
        # eager work between replays
        ops[rep % len(ops)].forward_dev(g[rep % len(ops)])
        for x in outs_f + outs_a:
            x.zero_()
        torch.cuda.synchronize()
    log("graph replay: 3 replays equal to the eager launches")


def test_batch_autotune_keeps_results(torch_cuda, fp32_set):
    torch = torch_cuda
    specs, ops = fp32_set
    B = OperatorBatch(ops)
    g, u = inputs(torch, ops, 14)
    f0 = [x.clone() for x in B.forward_dev(g)]
    a0 = [x.clone() for x in B.adjoint_dev(u)]
    tf, ta = B.autotune()
    log("autotune: forward %.3f ms adjoint %.3f ms" % (tf, ta))
    f1 = B.forward_dev(g)
    a1 = B.adjoint_dev(u)
    for i in range(len(ops)):
        assert rel(f1[i], f0[i]) < 1e-6 and rel(a1[i], a0[i]) < 1e-6


def test_plan_autotune_keeps_results(torch_cuda, fp32_set):
    torch = torch_cuda
    specs, ops = fp32_set
    op = ops[3]
    g, u = inputs(torch, [op], 15)
    f0 = op.forward_dev(g[0]).clone()
    a0 = op.adjoint_dev(u[0]).clone()
    op.autotune()
    assert rel(op.forward_dev(g[0]), f0) < 1e-6 and rel(op.adjoint_dev(u[0]), a0) < 1e-6


@pytest.mark.parametrize("form", ["split_bf16", "split_fp16"])
def test_batch_split(torch_cuda, form):
    torch = torch_cuda
    specs = uneven_specs(20)
    ops = [OperatorBorn.from_sequence(s, precision=form) for s in specs]
    B = OperatorBatch(ops)
    g, u = inputs(torch, ops, 16)
    fb = B.forward_dev(g)
    ab = B.adjoint_dev(u)
    for i, spec in enumerate(specs):
        T = O.direct_tables(spec)
        ef = O.rel_l2(fb[i].cpu().numpy(), O.forward(T, g[i].cpu().numpy().astype(np.complex128)))
        ea = O.rel_l2(ab[i].cpu().numpy(), O.adjoint(T, u[i].cpu().numpy().astype(np.complex128)))
        log("%s batch instance %d vs FP64 oracle: forward %.2e adjoint %.2e" % (form, i, ef, ea))
        assert ef < TOL_SPLIT and ea < TOL_SPLIT
    # graph replays (the FP16 operand scales are computed inside the graph)
    ref_f = [x.clone() for x in fb]
    ref_a = [x.clone() for x in ab]
    B.set_graph(True)
    f2 = [torch.empty_like(x) for x in ref_f]
    a2 = [torch.empty_like(x) for x in ref_a]
    for rep in range(3):
        B.forward_dev(g, outs=f2)
        B.adjoint_dev(u, outs=a2)
        torch.cuda.synchronize()
        for i in range(len(ops)):
            assert torch.equal(f2[i], ref_f[i]) and torch.equal(a2[i], ref_a[i]), (rep, i)
        ops[rep % len(ops)].forward_dev(g[rep % len(ops)] * 3.0)   # eager work with other scales between replays
        torch.cuda.synchronize()


def test_batch_rejects_mixed_precision(torch_cuda, fp32_set):
    specs, ops = fp32_set
    other = OperatorBorn.from_sequence(uneven_specs(16)[0], precision="fp32")
    other.set_precision("fp32")
    from comp_ui_toolbox_amd import BornError
    s20 = OperatorBorn.from_sequence(uneven_specs(20)[0], precision="split_bf16")
    with pytest.raises(BornError):
        OperatorBatch([ops[0], s20])


def test_batch_follows_plan_workspace_changes(torch_cuda):
    """A plan of the batch applied on its own with two right-hand sides gets
    a new workspace; the batch then rebuilds its records (and its captured
    graph) instead of reading the freed one.  A plan whose precision changed
    is refused."""
    torch = torch_cuda
    from comp_ui_toolbox_amd import BornError
    specs = uneven_specs(16)[:3]
    ops = [OperatorBorn.from_sequence(s) for s in specs]
    B = OperatorBatch(ops, graph=True)
    g, u = inputs(torch, ops, 19)
    fb0 = [t.clone() for t in B.forward_dev(g)]
    rng = np.random.Generator(np.random.PCG64(20))
    two = rng.standard_normal((ops[1].N_points, 2)) + 1j * rng.standard_normal((ops[1].N_points, 2))
    cols = np.stack([ops[1].forward(two[:, r]) for r in range(2)], axis=1)
    assert O.rel_l2(ops[1].forward(two), cols) < 1e-6   # plan 1 now holds a 2-rhs workspace
    fb = B.forward_dev(g)
    ab = B.adjoint_dev(u)
    torch.cuda.synchronize()
    for i, op in enumerate(ops):
        assert rel(fb[i], fb0[i]) < 1e-6
        assert rel(fb[i], op.forward_dev(g[i])) < 1e-6
        assert rel(ab[i], op.adjoint_dev(u[i])) < 1e-6
    ops[2].set_precision("split_bf16")
    with pytest.raises(BornError):
        B.forward_dev(g)
    ops[2].set_precision("fp32")
    fb = B.forward_dev(g)
    torch.cuda.synchronize()
    assert all(rel(fb[i], fb0[i]) < 1e-6 for i in range(len(ops)))
    B.close()


def test_config3_batch(torch_cuda):
    """The 24 uneven config-3 instances (most with a FOV narrower than the
    aperture) as one batch: every instance equals its plan applied alone, and
    three instances match the FP64 oracle at sampled points."""
    torch = torch_cuda
    specs = C.config3_instances()
    ops = [OperatorBorn.from_sequence(s) for s in specs]
    B = OperatorBatch(ops, graph=True)
    g, u = inputs(torch, ops, 17)
    fb = B.forward_dev(g)
    ab = B.adjoint_dev(u)
    torch.cuda.synchronize()
    for i, op in enumerate(ops):
        assert rel(fb[i], op.forward_dev(g[i])) < 1e-6
        assert rel(ab[i], op.adjoint_dev(u[i])) < 1e-6
    sides = [s.fov.N_points_axis[0] for s in specs]
    # the smallest and largest rows, the last instance and the first with 5-8
    # waves (the stacked 8-column MFMA form)
    mid = [k for k, s in enumerate(specs) if 5 <= s.N_measurements <= 8][:1]
    pick = sorted({sides.index(v) for v in (128, 512) if v in sides} | {len(specs) - 1} | set(mid))
    rng = np.random.Generator(np.random.PCG64(18))
    for i in pick:
        spec, op = specs[i], ops[i]
        pts = np.sort(rng.choice(spec.fov.N_points, 6, replace=False))
        T = O.direct_tables(spec, pts)
        gi = np.zeros(spec.fov.N_points, dtype=np.complex64)
        gi[pts] = rng.standard_normal(len(pts))
        xs = [torch.tensor(gi, device="cuda") if k == i else g[k] for k in range(len(ops))]
        fi = B.forward_dev(xs)[i].cpu().numpy()
        ef = O.rel_l2(fi, O.forward(T, gi[pts].astype(np.complex128)))
        ea = O.rel_l2(ab[i].cpu().numpy()[pts], O.adjoint(T, u[i].cpu().numpy().astype(np.complex128)))
        log("config3 instance %d (side %d, %d f, %d waves): forward %.2e adjoint %.2e vs FP64 oracle"
            % (i, sides[i], len(spec.frequencies), spec.N_measurements, ef, ea))
        assert ef < TOL_FP32 and ea < TOL_FP32
