"""Parity of the HIP operator pair with the FP64 oracle, through the C ABI.

Tolerance of the plain FP32 path: relative L2 error <= 1e-5 against FP64
(the operator's stated bar); device-built tables <= 1e-6 (FP32 storage of
an FP64 evaluation)."""
import math

import numpy as np
import pytest

from comp_ui_toolbox_amd import configs as C
from comp_ui_toolbox_amd import OperatorBorn, combined_quick_gpu, BornError
from oracle import born_oracle as O
from tests.test_oracle import GOLD, mixed_tiny_tables, tiny_spec

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5
TOL_TABLE = 1e-6


def rnd(shape, seed, cplx=True):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x


def log(*a):
    print(*a, flush=True)


# ------------------------------------------------------------- tiny tables
@pytest.mark.parametrize("general", [False, True])
def test_tiny_mixed_tables_forward_adjoint(torch_cuda, general):
    spec, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    log("tiny mixed tables, path", op.path, "general" if general else "")
    assert op.path == "structured"
    if general:
        op.force_general(True)
        assert op.path == "general"
    g = rnd((T.N_points, 3), 1)
    u = rnd((T.N_observations(), 3), 2)
    assert op.size == (T.N_observations(), T.N_points)
    assert O.rel_l2(op.forward(g), O.forward(T, g)) < TOL_FP32
    assert O.rel_l2(op.adjoint(u), O.adjoint(T, u)) < TOL_FP32
    assert O.rel_l2(op.forward(g[:, 1]), O.forward(T, g[:, 1])) < TOL_FP32


@pytest.mark.parametrize("N_e,Nx,Nz,npp", [(5, 17, 7, 3), (2, 9, 3, 2), (19, 80, 6, 4), (33, 131, 3, 4)])
def test_tiny_odd_geometries(torch_cuda, N_e, Nx, Nz, npp):
    spec = tiny_spec(N_e=N_e, Nx=Nx, Nz=Nz, npp=npp)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    log("geometry", N_e, Nx, Nz, npp, "path", op.path)
    assert op.path == "structured"
    g = rnd((T.N_points, 2), 3)
    u = rnd((T.N_observations(), 2), 4)
    assert O.rel_l2(op.forward(g), O.forward(T, g)) < TOL_FP32
    assert O.rel_l2(op.adjoint(u), O.adjoint(T, u)) < TOL_FP32


def test_unstructured_shift_map_uses_general_path(torch_cuda):
    spec, T = mixed_tiny_tables()
    rng = np.random.Generator(np.random.PCG64(9))
    T.shift = rng.integers(1, T.N_points + 1, size=T.shift.shape)
    op = OperatorBorn.from_tables(T)
    assert op.path == "general"
    g = rnd(T.N_points, 5)
    u = rnd(T.N_observations(), 6)
    assert O.rel_l2(op.forward(g), O.forward(T, g)) < TOL_FP32
    assert O.rel_l2(op.adjoint(u), O.adjoint(T, u)) < TOL_FP32


@pytest.mark.parametrize("general", [False, True])
def test_3d_tables(torch_cuda, general):
    """3-D grid from the reference's tables: the separable lateral maps are
    found in the shift map (structured path, one x-Toeplitz product per y
    slice) or the shift map is gathered (general path)."""
    arr = C.ArraySpec((3, 2), (3e-4, 3e-4), (2.5e-4, 2.5e-4), (2, 2))
    fov = C.FovSpec((8, 6, 4), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
    spec = C.SequenceSpec("t3", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5e6]),
                          [C.qsw_wave(arr, np.array([0, 0, -1e-2]), 1500.0),
                           C.qsw_wave(arr, np.array([3e-4, 0, -1e-2]), 1500.0)],
                          [C.identity_mixes(6), C.identity_mixes(6)])
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    assert op.path == "structured"
    if general:
        op.force_general(True)
    g = rnd((T.N_points, 2), 7)
    u = rnd((T.N_observations(), 2), 8)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("3-D tables (%s): forward %.2e adjoint %.2e" % ("general" if general else "structured", e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


@pytest.mark.parametrize("precision,tol,nex", [("fp32", TOL_FP32, 8), ("fp32", TOL_FP32, 32),
                                               ("split_bf16", 1e-4, 32), ("split_fp16", 1e-4, 32)])
@pytest.mark.parametrize("shape,div", [((20, 12, 5), 2), ((10, 7, 3), 2), ((40, 12, 3), 4), ((100, 9, 2), 4)])
def test_3d_geometry_structured(torch_cuda, shape, div, precision, tol, nex):
    """3-D plans built from geometry on the structured path: an nex x 4
    matrix array at pitch/div over FOVs as wide as the aperture (mirrored
    tables) or narrower (tables on the extended grid), against every
    element's own transfer function.  The split form needs 17 or more
    elements per x row; with 32 elements per x row and 4 y rows the
    forwards take the four y slices in one work item."""
    arr = C.ArraySpec((nex, 4), (3e-4, 3e-4), (2.7e-4, 2.7e-4), (2, 2))
    fov = C.FovSpec(shape, (3e-4 / div,) * 3, 2e-3)
    waves = [C.qsw_wave(arr, np.array([x, 0.0, -1e-2]), 1500.0) for x in (0.0, 6e-4)]
    spec = C.SequenceSpec("t3g", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5.5e6]), waves,
                          [C.identity_mixes(4 * nex) for _ in waves])
    op = OperatorBorn.from_sequence(spec, precision=precision)
    assert op.path == "structured"
    T = O.direct_tables(spec)
    g = rnd((T.N_points, 2), 9)
    u = rnd((T.N_observations(), 2), 10)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("3-D geometry %s pitch/%d %s %d: forward %.2e adjoint %.2e" % (shape, div, precision, nex, e_f, e_a))
    assert e_f < tol and e_a < tol


# -------------------------------------------------------------- config 0
@pytest.fixture(scope="module")
def c0_tables():
    return O.build_tables(C.config0())


def test_config0_golden(torch_cuda, c0_tables):
    d = np.load(GOLD / "config0.npz")
    op = OperatorBorn.from_tables(c0_tables)
    assert op.path == "structured"
    u = op.forward(d["gamma"])
    e_f = O.rel_l2(u, d["u"])
    gh = op.adjoint(d["u"])
    e_a = O.rel_l2(gh[d["points"]], d["gamma_hat"])
    log("config0 golden: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


def test_config0_dense_input_full_adjoint(torch_cuda, c0_tables):
    T = c0_tables
    op = OperatorBorn.from_tables(T)
    spec = C.config0()
    g = C.phantom_speckle(spec, n_disks=2, n_bright=5)
    rows = np.arange(0, T.N_observations(), 37)
    e_f = O.rel_l2(op.forward(g)[rows], O.forward(T, g, rows=rows))
    u = rnd(T.N_observations(), 10)
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("config0 dense: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


def test_config0_device_tables(torch_cuda, c0_tables):
    spec = C.config0()
    op = OperatorBorn.from_sequence(spec)
    assert op.path == "structured"
    h = op.table("h_ref")
    p = op.table("p_incident", 0)
    e_h = O.rel_l2(h, c0_tables.h_ref)
    e_p = O.rel_l2(p, c0_tables.meas[0].p_incident)
    log("config0 device tables: h_ref rel-L2 %.2e p_inc rel-L2 %.2e" % (e_h, e_p))
    assert e_h < TOL_TABLE and e_p < TOL_TABLE
    d = np.load(GOLD / "config0.npz")
    assert O.rel_l2(op.forward(d["gamma"]), d["u"]) < TOL_FP32


@pytest.mark.parametrize("filt", ["off", "logistic", "boxcar"])
def test_device_tables_filters_tiny(torch_cuda, filt):
    spec = tiny_spec(N_e=5, Nx=20, Nz=6, npp=4, filt=filt)
    spec.filter_tx = C.FilterSpec("raised_cosine", 0.3)
    if filt == "logistic":
        spec.filter_rx = C.FilterSpec("logistic", 4.0)
    T = O.build_tables(spec)
    op = OperatorBorn.from_sequence(spec)
    assert O.rel_l2(op.table("h_ref"), T.h_ref) < TOL_TABLE
    for m in range(spec.N_measurements):
        assert O.rel_l2(op.table("p_incident", m), T.meas[m].p_incident) < TOL_TABLE


def test_device_tables_3d_tiny(torch_cuda):
    arr = C.ArraySpec((3, 2), (3e-4, 3e-4), (2.5e-4, 2.5e-4), (2, 2))
    fov = C.FovSpec((8, 6, 4), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
    spec = C.SequenceSpec("t3", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5e6]),
                          [C.qsw_wave(arr, np.array([0, 0, -1e-2]), 1500.0)], [C.identity_mixes(6)])
    T = O.build_tables(spec)
    op = OperatorBorn.from_sequence(spec)
    assert O.rel_l2(op.table("h_ref"), T.h_ref) < TOL_TABLE
    assert O.rel_l2(op.table("p_incident", 0), T.meas[0].p_incident) < TOL_TABLE


# ----------------------------------------------------------- full sizes
@pytest.fixture(scope="module")
def c1():
    spec = C.config1()
    op = OperatorBorn.from_sequence(spec)
    return spec, op


@pytest.mark.parametrize("sparse", [0, -1])
def test_config1_forward_points_against_oracle(torch_cuda, c1, sparse):
    """Host forward of gamma on 24 points: through the grid contraction
    (sparse forward off) and through the gather path that the plan picks for
    so few occupied points (the MEX's occupancy compaction)."""
    spec, op = c1
    rng = np.random.Generator(np.random.PCG64(21))
    pts = np.sort(rng.choice(spec.fov.N_points, 24, replace=False))
    R = O.restricted_tables(spec, pts)
    g = np.zeros(spec.fov.N_points)
    g[pts] = rng.standard_normal(len(pts))
    op.set_sparse_forward(sparse)
    try:
        u = op.forward(g)
    finally:
        op.set_sparse_forward(-1)
    e = O.rel_l2(u, O.forward(R, g[pts]))
    log("config1 forward (24 points, %s) rel-L2 %.2e" % ("contraction" if sparse == 0 else "gather", e))
    assert e < TOL_FP32


def test_config1_adjoint_points_against_oracle(torch_cuda, c1):
    spec, op = c1
    rng = np.random.Generator(np.random.PCG64(22))
    pts = np.sort(rng.choice(spec.fov.N_points, 24, replace=False))
    R = O.restricted_tables(spec, pts)
    u = rnd(op.N_observations, 23)
    gh = op.adjoint(u)
    e = O.rel_l2(gh[pts], O.adjoint(R, u))
    log("config1 adjoint (24 points) rel-L2 %.2e" % e)
    assert e < TOL_FP32


def test_config1_adjoint_identity_and_linearity(torch_cuda, c1):
    torch = torch_cuda
    spec, op = c1
    x = torch.tensor(C.phantom_speckle(spec).astype(np.complex64), device="cuda")
    y = torch.tensor(rnd(op.N_observations, 31).astype(np.complex64), device="cuda")
    Ax = op.forward_dev(x).to(torch.complex128)
    Ahy = op.adjoint_dev(y).to(torch.complex128)
    lhs = torch.vdot(y.to(torch.complex128), Ax).item()
    rhs = torch.vdot(Ahy, x.to(torch.complex128)).item()
    e = abs(lhs - rhs) / abs(lhs)
    log("config1 <Ax,y> vs <x,A^H y> rel diff %.2e" % e)
    assert e < TOL_FP32
    x2 = torch.tensor(rnd(op.N_points, 32).astype(np.complex64), device="cuda")
    A12 = op.forward_dev(x + 2 * x2).to(torch.complex128)
    lin = (A12 - Ax - 2 * op.forward_dev(x2).to(torch.complex128)).norm() / A12.norm()
    assert lin.item() < TOL_FP32


def test_config1_multi_rhs_matches_single(torch_cuda, c1):
    torch = torch_cuda
    spec, op = c1
    X = torch.tensor(rnd((3, op.N_points), 41).astype(np.complex64), device="cuda")
    U = op.forward_dev(X)
    for r in range(3):
        ur = op.forward_dev(X[r].contiguous())
        assert ((U[r] - ur).norm() / ur.norm()).item() < 1e-6
    Y = torch.tensor(rnd((2, op.N_observations), 42).astype(np.complex64), device="cuda")
    G = op.adjoint_dev(Y)
    for r in range(2):
        gr = op.adjoint_dev(Y[r].contiguous())
        assert ((G[r] - gr).norm() / gr.norm()).item() < 1e-6


def test_repeatability_bitwise(torch_cuda, c1):
    torch = torch_cuda
    spec, op = c1
    x = torch.tensor(rnd(op.N_points, 51).astype(np.complex64), device="cuda")
    y = torch.tensor(rnd(op.N_observations, 52).astype(np.complex64), device="cuda")
    a1, a2 = op.forward_dev(x).clone(), op.forward_dev(x).clone()
    b1, b2 = op.adjoint_dev(y).clone(), op.adjoint_dev(y).clone()
    assert torch.equal(a1, a2) and torch.equal(b1, b2)


# ---------------------------------------------------------- MEX semantics
def test_zero_and_subnormal_input(torch_cuda):
    spec, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    assert np.all(op.forward(np.zeros(T.N_points)) == 0)
    g = np.zeros(T.N_points)
    g[4] = 1e-17           # below DBL_EPSILON: dropped as in the MEX
    assert np.all(op.forward(g) == 0)
    g[4] = 1.0
    g[7] = 1e-17
    ref = np.zeros(T.N_points)
    ref[4] = 1.0
    assert O.rel_l2(op.forward(g), O.forward(T, ref)) < TOL_FP32


def test_combined_quick_gpu_contract(torch_cuda):
    spec, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    g = rnd((T.N_points, 2), 61, cplx=False)           # real input is promoted
    u = combined_quick_gpu(op, 1, g)
    assert u.shape == (T.N_observations(), 2) and u.dtype == np.complex128
    assert O.rel_l2(u, O.forward(T, g)) < TOL_FP32
    gh = combined_quick_gpu(op, 2.0, u)
    assert gh.shape == (T.N_points, 2)
    assert O.rel_l2(gh, O.adjoint(T, u)) < TOL_FP32
    with pytest.raises(BornError, match="NoValidDoubleMatrix"):
        combined_quick_gpu(op, 1, g[:-1])
    with pytest.raises(BornError, match="NoValidDoubleMatrix"):
        combined_quick_gpu(op, 1, g.astype(np.float32))
    with pytest.raises(BornError, match="NoNumericScalarMode"):
        combined_quick_gpu(op, [1, 2], g)
    with pytest.raises(BornError, match="NoOperatorBorn"):
        combined_quick_gpu("op", 1, g)
    assert op.op_A(None, 0) == [T.N_observations(), T.N_points]


# ------------------------------------------------------- split precision
TOL_SPLIT = 1e-4


@pytest.mark.parametrize("form", ["split_bf16", "split_fp16"])
@pytest.mark.parametrize("N_e,Nx,Nz,npp", [(20, 80, 4, 4), (40, 300, 3, 4), (128, 520, 2, 4),
                                           (20, 40, 4, 2), (64, 300, 3, 2), (128, 260, 2, 2)])
def test_split_small_tables(torch_cuda, N_e, Nx, Nz, npp, form):
    spec = tiny_spec(N_e=N_e, Nx=Nx, Nz=Nz, npp=npp)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T, precision=form)
    g = rnd((T.N_points, 2), 71)
    u = rnd((T.N_observations(), 2), 72)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    assert op.path == "structured"
    log("%s N_e=%d Nx=%d npp=%d: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (form, N_e, Nx, npp, e_f, e_a))
    assert e_f < TOL_SPLIT and e_a < TOL_SPLIT


@pytest.mark.parametrize("form", ["split_bf16", "split_fp16"])
def test_split_config1(torch_cuda, c1, form):
    spec, op = c1
    rng = np.random.Generator(np.random.PCG64(81))
    pts = np.sort(rng.choice(spec.fov.N_points, 24, replace=False))
    R = O.restricted_tables(spec, pts)
    g = np.zeros(spec.fov.N_points)
    g[pts] = rng.standard_normal(len(pts))
    u = rnd(op.N_observations, 82)
    op.set_precision(form)
    op.set_sparse_forward(0)          # the contraction of this form, not the gather path
    try:
        e_f = O.rel_l2(op.forward(g), O.forward(R, g[pts]))
        e_a = O.rel_l2(op.adjoint(u)[pts], O.adjoint(R, u))
        torch = torch_cuda
        x = torch.tensor(C.phantom_speckle(spec).astype(np.complex64), device="cuda")
        us = op.forward_dev(x)
        op.set_precision("fp32")
        u32 = op.forward_dev(x)
        e_d = ((us - u32).norm() / u32.norm()).item()
    finally:
        op.set_precision("fp32")
        op.set_sparse_forward(-1)
    log("%s config1: forward rel-L2 %.2e adjoint rel-L2 %.2e dense vs fp32 %.2e" % (form, e_f, e_a, e_d))
    assert e_f < TOL_SPLIT and e_a < TOL_SPLIT and e_d < TOL_SPLIT


# ------------------------------------------- FOV narrower than the aperture
@pytest.mark.parametrize("general", [False, True])
def test_narrow_fov_extended_tables(torch_cuda, general):
    """FOV of 24 points (6 pitches) under a 16-element aperture: the mirror map
    is undefined there, the plan keeps h_ref on the laterally extended grid.
    Checked against every element's own transfer function (no shift map)."""
    spec = tiny_spec(N_e=16, Nx=24, Nz=3, npp=4)
    op = OperatorBorn.from_sequence(spec)
    assert op.path == "structured"
    if general:
        op.force_general(True)
    T = O.direct_tables(spec)
    g = rnd((T.N_points, 2), 91)
    u = rnd((T.N_observations(), 2), 92)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("narrow FOV (%s): forward rel-L2 %.2e adjoint rel-L2 %.2e" % ("general" if general else "structured", e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


def test_narrow_fov_extended_table_values(torch_cuda):
    """The device h_ref on the extended grid equals the first element's transfer
    function at the extended lateral positions."""
    spec = tiny_spec(N_e=8, Nx=10, Nz=2, npp=2)
    op = OperatorBorn.from_sequence(spec)
    h = op.table("h_ref")
    ox = (8 - 1) * 2
    Nxt = 10 + ox
    dx = spec.fov.delta_axis[0]
    xs = (np.arange(Nxt) - ox - 0.5 * (10 - 1)) * dx
    zs = spec.fov.z0 + np.arange(2) * spec.fov.delta_axis[1]
    pos = np.stack([np.tile(xs, 2), np.repeat(zs, Nxt)], axis=1)
    ref = O.transfer_function(spec, spec.frequencies, 0, spec.filter_rx, positions=pos)
    assert h.shape == ref.shape
    e = O.rel_l2(h, ref)
    log("extended h_ref rel-L2 %.2e" % e)
    assert e < TOL_TABLE


def test_narrow_fov_split_bf16(torch_cuda):
    spec = tiny_spec(N_e=20, Nx=40, Nz=2, npp=4)
    op = OperatorBorn.from_sequence(spec, precision="split_bf16")
    T = O.direct_tables(spec)
    g = rnd((T.N_points, 1), 93)
    u = rnd((T.N_observations(), 1), 94)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("narrow FOV split bf16: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (e_f, e_a))
    assert e_f < TOL_SPLIT and e_a < TOL_SPLIT


# ----------------------------------------------- config 2: random waves, mixing
@pytest.fixture(scope="module")
def c2(torch_cuda):
    """The plan and the oracle's answers at 16 sampled points (computed once)."""
    spec = C.config2()
    op = OperatorBorn.from_sequence(spec)
    rng = np.random.Generator(np.random.PCG64(101))
    pts = np.sort(rng.choice(spec.fov.N_points, 16, replace=False))
    R = O.restricted_tables(spec, pts)
    g = np.zeros(spec.fov.N_points)
    g[pts] = rng.standard_normal(len(pts))
    u = rnd(op.N_observations, 102)
    return spec, op, pts, g, u, O.forward(R, g[pts]), O.adjoint(R, u)


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
def test_config2_against_oracle(torch_cuda, c2, precision, tol):
    """32 random waves (+-1 apodization, delays U[0, 5 us]) compounded into 16
    mixes of all 128 elements with +-1 weights: sampled points against tables
    the oracle restricts exactly to them, through the grid contraction."""
    spec, op, pts, g, u, ref_f, ref_a = c2
    op.set_precision(precision)
    op.set_sparse_forward(0)
    try:
        e_f = O.rel_l2(op.forward(g), ref_f)
        e_a = O.rel_l2(op.adjoint(u)[pts], ref_a)
    finally:
        op.set_precision("fp32")
        op.set_sparse_forward(-1)
    log("config2 %s: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (precision, e_f, e_a))
    assert e_f < tol and e_a < tol


# ------------------------------- dense inputs over the full FOV, against FP64
C12_FREQS = np.array([0, 128, 256])      # first, middle and last of the 257 bins


def dense_case(spec, seed):
    """The FP64 oracle of a dense input over the whole FOV at 3 of the 257
    frequencies: the forward of the speckle phantom at those frequencies'
    observations, and the adjoint of random u supported on them."""
    T, rows = O.frequency_tables(spec, C12_FREQS)
    g = C.phantom_speckle(spec)
    u_sub = rnd(len(rows), seed)
    return dict(rows=rows, g=g, u_sub=u_sub, u_ref=O.forward_by_frequency(T, g),
                g_ref=O.adjoint_by_frequency(T, u_sub))


def dense_errors(op, d, precision):
    """Forward and adjoint of the dense case through the device entry points
    (the contraction kernels the benchmark times) in one numeric form:
    relative L2 and max-abs error per output, max-abs also relative to the
    largest reference entry."""
    import torch
    op.set_precision(precision)
    try:
        x = torch.tensor(d["g"].astype(np.complex64), device="cuda")
        u = op.forward_dev(x).cpu().numpy()[d["rows"]]
        uu = torch.zeros(op.N_observations, dtype=torch.complex64, device="cuda")
        uu[torch.from_numpy(d["rows"]).cuda()] = torch.from_numpy(d["u_sub"].astype(np.complex64)).cuda()
        gh = op.adjoint_dev(uu).cpu().numpy()
    finally:
        op.set_precision("fp32")
    out = {}
    for name, x, ref in (("forward", u, d["u_ref"]), ("adjoint", gh, d["g_ref"])):
        m = O.max_abs(x, ref)
        out[name] = (O.rel_l2(x, ref), m, m / np.max(np.abs(ref)))
    return out


@pytest.fixture(scope="module")
def c1_dense(c1):
    log("config1: dense oracle at 3 frequencies over 262144 points")
    return dense_case(c1[0], 111)


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
def test_config1_dense_against_fp64(torch_cuda, c1, c1_dense, precision, tol):
    """Config 1 at full size: the speckle phantom over all 262144 points (and
    u on every observation of 3 frequencies) through the forward and adjoint
    contraction kernels, against the FP64 oracle."""
    errs = dense_errors(c1[1], c1_dense, precision)
    for name, (e, m, mr) in errs.items():
        log("config1 dense %s %s: rel-L2 %.2e max-abs %.3e (%.2e of max |ref|)" % (precision, name, e, m, mr))
        assert e < tol and mr < tol


@pytest.fixture(scope="module")
def c2_dense(c2):
    log("config2: dense oracle at 3 frequencies over 262144 points")
    return dense_case(c2[0], 112)


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
def test_config2_dense_against_fp64(torch_cuda, c2, c2_dense, precision, tol):
    """Config 2 at full size (32 random waves, 16 +-1 mixes of 128 elements):
    dense forward and adjoint against the FP64 oracle at 3 frequencies."""
    errs = dense_errors(c2[1], c2_dense, precision)
    for name, (e, m, mr) in errs.items():
        log("config2 dense %s %s: rel-L2 %.2e max-abs %.3e (%.2e of max |ref|)" % (precision, name, e, m, mr))
        assert e < tol and mr < tol


# ------------------------------------------------- face weights (apodization, lens)
def lens_specs():
    """A 2-D array whose faces carry a symmetric apodization and a lens
    focusing along x, and a 3-D matrix array with an elevation lens (the
    L14-5/38 kind, focus along y)."""
    from tests.test_oracle import tiny_spec
    s2 = tiny_spec(N_e=20, Nx=80, Nz=4, npp=4, nfp=4)
    s2.array.apodization = lambda r: 1.0 - 0.4 * r[:, 0] ** 2
    s2.array.lens = C.LensSpec((3e-3,), 1200.0)
    arr = C.ArraySpec((20, 2), (3e-4, 3e-4), (2.5e-4, 2.5e-4), (2, 4), lens=C.LensSpec((math.inf, 4e-3), 1000.0))
    fov = C.FovSpec((40, 4, 3), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
    s3 = C.SequenceSpec("lens3", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5e6]),
                        [C.qsw_wave(arr, np.array([0, 0, -1e-2]), 1500.0),
                         C.qsw_wave(arr, np.array([3e-4, 0, -1e-2]), 1500.0)],
                        [C.identity_mixes(40), C.identity_mixes(40)])
    return {"2d": s2, "3d": s3}


@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_device_tables_with_face_weights(torch_cuda, dims):
    """Complex, non-unit face weights (face.m:359-376) passed to the device
    table build: h_ref and every p_incident against the oracle's mirrored
    tables, and the operator pair in plain FP32 and split BF16 against every
    element's own weighted transfer function (no shift map)."""
    spec = lens_specs()[dims]
    T = O.build_tables(spec)
    D = O.direct_tables(spec)
    ref_plain = O.transfer_function(spec, spec.frequencies, 0, spec.filter_rx,
                                    face_weights_fq=np.ones((len(spec.array.face_offsets()[0]),
                                                             len(spec.frequencies))))
    assert O.rel_l2(T.h_ref, ref_plain) > 1e-2          # the weights do change the table
    op = OperatorBorn.from_sequence(spec)
    e_h = O.rel_l2(op.table("h_ref"), T.h_ref)
    e_p = max(O.rel_l2(op.table("p_incident", m), T.meas[m].p_incident) for m in range(spec.N_measurements))
    log("face weights %s: device h_ref rel-L2 %.2e p_inc rel-L2 %.2e" % (dims, e_h, e_p))
    assert e_h < TOL_TABLE and e_p < TOL_TABLE
    g = rnd((T.N_points, 2), 121)
    u = rnd((T.N_observations(), 2), 122)
    for form, tol in (("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT)):
        op.set_precision(form)
        op.set_sparse_forward(0)
        e_f = O.rel_l2(op.forward(g), O.forward(D, g))
        e_a = O.rel_l2(op.adjoint(u), O.adjoint(D, u))
        log("face weights %s %s: forward rel-L2 %.2e adjoint rel-L2 %.2e" % (dims, form, e_f, e_a))
        assert e_f < tol and e_a < tol


# ---------------------------------------------------------------- config 4 (3-D)
C4_FREQS = np.array([0, 32, 64])


@pytest.fixture(scope="module")
def config4_case():
    """The 3-D workload (32 x 32 matrix array, 128 x 128 x 192 FOV at pitch/2,
    65 bins, 5 QSWs) with 4 gamma right-hand sides, and the FP64 oracle of
    its block at a 32^3 sub-FOV (x, y in [48, 80), z in [80, 112)) and 3 of
    the 65 frequencies (first, middle, last): the forward of gamma supported
    on the sub-FOV at those frequencies' observations, and the adjoint at
    the sub-FOV points of u supported on those observations."""
    from oracle import ref_loop
    spec = C.config4()
    Nx, Ny, Nz = spec.fov.N_points_axis
    ix, iy, iz = np.meshgrid(np.arange(48, 80), np.arange(48, 80), np.arange(80, 112), indexing="ij")
    pts = np.sort((ix + Nx * (iy + Ny * iz)).reshape(-1))
    log("config4: sampled oracle tables (%d points, %d frequencies)" % (len(pts), len(C4_FREQS)))
    T, rows = O.sampled_tables(spec, pts, C4_FREQS)
    nrhs = 4
    rng = np.random.Generator(np.random.PCG64(5))
    g_sub = rng.standard_normal((len(pts), nrhs)) + 1j * rng.standard_normal((len(pts), nrhs))
    u_sub = rng.standard_normal((len(rows), nrhs)) + 1j * rng.standard_normal((len(rows), nrhs))
    preps = [ref_loop.Prepared(T, m) for m in range(len(T.meas))]
    u_ref = np.stack([np.concatenate([ref_loop.apply_measurement(P, 1, g_sub[:, r], fused=True) for P in preps])
                      for r in range(nrhs)], axis=1)
    g_ref = np.zeros((len(pts), nrhs), dtype=np.complex128)
    for r in range(nrhs):
        o = 0
        for P in preps:
            g_ref[:, r] += ref_loop.apply_measurement(P, 2, u_sub[o:o + P.n_obs, r], fused=True)
            o += P.n_obs
    log("config4: oracle done")
    return dict(spec=spec, pts=pts, rows=rows, g_sub=g_sub, u_sub=u_sub, u_ref=u_ref, g_ref=g_ref)


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
def test_config4_subfov_oracle(torch_cuda, config4_case, precision, tol):
    import torch
    d = config4_case
    op = OperatorBorn.from_sequence(d["spec"], precision=precision)
    assert op.path == "structured"
    nrhs = d["g_sub"].shape[1]
    g = torch.zeros((nrhs, op.N_points), dtype=torch.complex64, device="cuda")
    g[:, torch.from_numpy(d["pts"]).cuda()] = torch.from_numpy(d["g_sub"].T.astype(np.complex64)).cuda()
    u = op.forward_dev(g)
    u_gpu = u.cpu().numpy().T[d["rows"]]
    e_f = O.rel_l2(u_gpu, d["u_ref"])
    uu = torch.zeros((nrhs, op.N_observations), dtype=torch.complex64, device="cuda")
    uu[:, torch.from_numpy(d["rows"]).cuda()] = torch.from_numpy(d["u_sub"].T.astype(np.complex64)).cuda()
    gh = op.adjoint_dev(uu)
    g_gpu = gh.cpu().numpy().T[d["pts"]]
    e_a = O.rel_l2(g_gpu, d["g_ref"])
    log("config4 %s (4 rhs, 32^3 sub-FOV, 3 frequencies): forward rel-L2 %.2e adjoint rel-L2 %.2e" % (precision, e_f, e_a))
    assert e_f < tol and e_a < tol
    op.close()


# ---------------------------------------------------------- received energy
def test_energy_rx_mixed_tables(torch_cuda):
    """Received energies per measurement (energy_rx_canonical) on mixed tables
    with repeated frequencies and several active elements per mix, host and
    device entry points, against the FP64 oracle."""
    import torch
    _, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    for m in range(len(T.meas)):
        ref = O.energy_rx(T, m)
        e_h = O.rel_l2(op.energy_rx(m), ref)
        e_d = O.rel_l2(op.energy_rx_dev(m).cpu().numpy(), ref)
        log("energy_rx measurement %d: host rel-L2 %.2e device rel-L2 %.2e" % (m, e_h, e_d))
        assert e_h < TOL_FP32 and e_d < TOL_FP32
    with pytest.raises(BornError):
        op.energy_rx(len(T.meas))


def test_energy_rx_3d(torch_cuda):
    arr = C.ArraySpec((8, 4), (3e-4, 3e-4), (2.7e-4, 2.7e-4), (2, 2))
    fov = C.FovSpec((10, 7, 3), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
    waves = [C.qsw_wave(arr, np.array([x, 0.0, -1e-2]), 1500.0) for x in (0.0, 6e-4)]
    spec = C.SequenceSpec("t3e", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5.5e6]), waves,
                          [C.identity_mixes(32) for _ in waves])
    op = OperatorBorn.from_sequence(spec)
    T = O.direct_tables(spec)
    for m in range(2):
        e = O.rel_l2(op.energy_rx(m), O.energy_rx(T, m))
        log("energy_rx 3-D measurement %d: rel-L2 %.2e" % (m, e))
        assert e < TOL_FP32


def test_energy_rx_config1(torch_cuda, c1):
    """Config 1, first and last measurement, at 24 sampled points against
    tables restricted exactly to those points."""
    spec, op = c1
    rng = np.random.Generator(np.random.PCG64(83))
    pts = np.sort(rng.choice(spec.fov.N_points, 24, replace=False))
    R = O.restricted_tables(spec, pts)
    for m in (0, spec.N_measurements - 1):
        E = op.energy_rx(m)
        e = O.rel_l2(E[pts], O.energy_rx(R, m))
        log("energy_rx config1 measurement %d at 24 points: rel-L2 %.2e" % (m, e))
        assert e < TOL_FP32


# ---------------------------------------------------------- time-domain signals
@pytest.mark.parametrize("delta_t", [0.0, 1.0 / (0.5e6 * 32)])
def test_time_signals(torch_cuda, delta_t):
    """Mixed voltage signals in the time domain (format_voltages + signal)
    for two right-hand sides, against the oracle's real inverse FFT."""
    import torch
    freqs = np.array([2.0e6, 2.5e6, 3.0e6, 4.5e6, 5.0e6])
    _, T = mixed_tiny_tables(freqs=freqs, distinct=True)
    op = OperatorBorn.from_tables(T)
    u = rnd((T.N_observations(), 2), 95)
    ut = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    rf, n_t = op.time_signals(ut, freqs, delta_t=delta_t)
    assert n_t == (20 if delta_t == 0.0 else 32)
    ref = O.time_signals(T, freqs, u, 0.5e6, n_t)
    e = O.rel_l2(rf.cpu().numpy(), ref)
    log("time signals (N_t=%d): rel-L2 %.2e" % (n_t, e))
    assert e < TOL_FP32


def test_time_signals_rejects_unquantized_axis_and_repeated_frequencies(torch_cuda):
    import torch
    _, T = mixed_tiny_tables(distinct=True)   # 2.0, 2.6, 3.1, 4.5, 5.0 MHz: no common delta_f at the smallest gap
    op = OperatorBorn.from_tables(T)
    ut = torch.zeros((1, T.N_observations()), dtype=torch.complex64, device="cuda")
    with pytest.raises(BornError):
        op.time_signals(ut, np.array([2.0e6, 2.6e6, 3.1e6, 4.5e6, 5.0e6]))
    freqs = np.array([2.0e6, 2.5e6, 3.0e6, 4.5e6, 5.0e6])
    _, T = mixed_tiny_tables(freqs=freqs)   # one mix repeats a frequency: no frequency axis
    op = OperatorBorn.from_tables(T)
    with pytest.raises(BornError):
        op.time_signals(ut, freqs)


# ---------------------------------------------------- several right-hand sides
@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
@pytest.mark.parametrize("n_rhs", [1, 2, 3, 5, 16, 17])
def test_many_rhs(torch_cuda, precision, tol, n_rhs):
    """Right-hand-side counts that pack differently into the 16 MFMA columns
    (2 waves per frequency: 1 chunk per rhs, or rhs packed per frequency with
    k = 2, 4, 8, 16 lanes per pair, or 17 > 16 one chunk per rhs again),
    through the host and the device entry points, against the oracle."""
    import torch
    spec = tiny_spec(N_e=40, Nx=160, Nz=3, npp=4)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T, precision=precision)
    g = rnd((T.N_points, n_rhs), 101)
    u = rnd((T.N_observations(), n_rhs), 102)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    gd = torch.from_numpy(g.T.astype(np.complex64)).cuda().contiguous()
    ud = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    e_fd = O.rel_l2(op.forward_dev(gd).cpu().numpy().T, O.forward(T, g.astype(np.complex64)))
    e_ad = O.rel_l2(op.adjoint_dev(ud).cpu().numpy().T, O.adjoint(T, u.astype(np.complex64)))
    log("%d rhs %s: forward %.2e adjoint %.2e (device %.2e %.2e)" % (n_rhs, precision, e_f, e_a, e_fd, e_ad))
    assert max(e_f, e_a, e_fd, e_ad) < tol


@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT), ("split_fp16", TOL_SPLIT)])
@pytest.mark.parametrize("n_rhs", [3, 4])
def test_many_rhs_many_waves(torch_cuda, precision, tol, n_rhs):
    """11 waves per frequency, as config 1.  With 3 right-hand sides packing
    saves no chunk (9 either way); with 4 it saves 3 of 12, too few for the
    split forms, whose forward then keeps one chunk set per rhs beside the
    packed adjoint, while FP32 packs both directions."""
    import torch
    med = C.MediumSpec()
    spec = tiny_spec(N_e=40, Nx=160, Nz=3, npp=4)
    spec.waves = [C.qpw_wave(spec.array, a, med.c_avg) for a in np.linspace(-0.3, 0.3, 11)]
    spec.mixes = [C.identity_mixes(40) for _ in spec.waves]
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T, precision=precision)
    g = rnd((T.N_points, n_rhs), 103)
    u = rnd((T.N_observations(), n_rhs), 104)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    gd = torch.from_numpy(g.T.astype(np.complex64)).cuda().contiguous()
    ud = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    e_fd = O.rel_l2(op.forward_dev(gd).cpu().numpy().T, O.forward(T, g.astype(np.complex64)))
    e_ad = O.rel_l2(op.adjoint_dev(ud).cpu().numpy().T, O.adjoint(T, u.astype(np.complex64)))
    log("11 waves, %d rhs %s: forward %.2e adjoint %.2e (device %.2e %.2e)" % (n_rhs, precision, e_f, e_a, e_fd, e_ad))
    assert max(e_f, e_a, e_fd, e_ad) < tol


# ---------------------------------------------------------------- columns
def test_columns_tiny(torch_cuda):
    """Columns of A (born_columns): against the assembled matrix, between
    the host and device entry points, and against the dense forward of the
    unit vectors; a repeated point is allowed, out-of-range points are not."""
    _, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    A = O.assemble_matrix(T)
    pts = np.array([0, 5, T.N_points - 1, 5])
    e_h = O.rel_l2(op.columns(pts), A[:, pts])
    e_d = O.rel_l2(op.columns_dev(pts).cpu().numpy().T, A[:, pts])
    E = np.zeros((T.N_points, len(pts)))
    E[pts, np.arange(len(pts))] = 1.0
    e_f = O.rel_l2(op.columns(pts), op.forward(E))
    log("columns (tiny): host %.2e device %.2e vs dense forward %.2e" % (e_h, e_d, e_f))
    assert max(e_h, e_d, e_f) < TOL_FP32
    with pytest.raises(BornError):
        op.columns([T.N_points])
    with pytest.raises(BornError):
        op.columns([])


def test_columns_config1(torch_cuda, c1):
    """Config 1: 16 columns against tables restricted exactly to those
    points; 1024 columns (a CoSaMP batch) timed on the device."""
    import torch
    spec, op = c1
    rng = np.random.Generator(np.random.PCG64(91))
    pts = np.sort(rng.choice(spec.fov.N_points, 16, replace=False))
    R = O.restricted_tables(spec, pts)
    e = O.rel_l2(op.columns(pts), O.forward(R, np.eye(len(pts))))
    big = rng.choice(spec.fov.N_points, 1024, replace=False)
    out = torch.empty((1024, op.N_observations), dtype=torch.complex64, device="cuda")
    op.columns_dev(big, out=out)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    op.columns_dev(big, out=out)
    t1.record()
    torch.cuda.synchronize()
    log("columns (config 1): 16 columns rel-L2 %.2e; 1024 columns in %.2f ms" % (e, t0.elapsed_time(t1)))
    assert e < TOL_FP32
    assert O.rel_l2(out[3].cpu().numpy(), op.columns(big[3:4])[:, 0]) < 1e-6


# ---------------------------------------------------------------- sparse forward
def test_sparse_forward_paths_agree(torch_cuda):
    """Host forward with few occupied points: the gather path (forced) and
    the contraction (forced off) against the oracle, two right-hand sides
    with different occupied points (the union is occupied, cu:466-488), an
    entry below DBL_EPSILON that does not occupy its point, and zero input."""
    _, T = mixed_tiny_tables()
    op = OperatorBorn.from_tables(T)
    g = np.zeros((T.N_points, 2), dtype=np.complex128)
    g[[3, 17], 0] = [1.5 - 0.5j, -2.0]
    g[[17, 40], 1] = [0.25j, 1.0]
    g[50, 1] = 1e-17                      # not occupied
    ref = O.forward(T, g)
    op.set_sparse_forward(1000)
    u_s = op.forward(g)
    op.set_sparse_forward(0)
    u_d = op.forward(g)
    op.set_sparse_forward(-1)
    e_s, e_d = O.rel_l2(u_s, ref), O.rel_l2(u_d, ref)
    log("sparse forward: gather %.2e contraction %.2e" % (e_s, e_d))
    assert e_s < 1e-6 and e_d < TOL_FP32
    op.set_sparse_forward(1000)
    assert np.all(op.forward(np.zeros((T.N_points, 1))) == 0)


def test_sparse_forward_config0(torch_cuda, c0_tables):
    """Config 0's input (30 occupied points) takes the gather path by the
    cost rule; it agrees with the contraction and the oracle, and is faster."""
    import time
    spec = C.config0()
    T = c0_tables
    op = OperatorBorn.from_tables(T)
    g = C.phantom_points(spec)
    ref = O.forward(T, g)
    times = {}
    outs = {}
    for mode, lim in (("gather", -1), ("contraction", 0)):
        op.set_sparse_forward(lim)
        op.forward(g)
        t0 = time.perf_counter()
        for _ in range(5):
            outs[mode] = op.forward(g)
        times[mode] = (time.perf_counter() - t0) / 5 * 1e3
    e_g, e_c = O.rel_l2(outs["gather"], ref), O.rel_l2(outs["contraction"], ref)
    log("config0 forward, 30 points: gather %.2e in %.2f ms, contraction %.2e in %.2f ms (host entry)"
        % (e_g, times["gather"], e_c, times["contraction"]))
    assert e_g < 1e-6 and e_c < TOL_FP32


def test_columns_and_sparse_forward_3d(torch_cuda):
    """3-D plan: columns and the gather forward against the direct oracle tables."""
    arr = C.ArraySpec((8, 4), (3e-4, 3e-4), (2.7e-4, 2.7e-4), (2, 2))
    fov = C.FovSpec((10, 7, 3), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
    waves = [C.qsw_wave(arr, np.array([x, 0.0, -1e-2]), 1500.0) for x in (0.0, 6e-4)]
    spec = C.SequenceSpec("t3c", arr, C.MediumSpec(), fov, np.array([2.5e6, 4e6, 5.5e6]), waves,
                          [C.identity_mixes(32) for _ in waves])
    op = OperatorBorn.from_sequence(spec)
    T = O.direct_tables(spec)
    pts = np.array([0, 37, 123, spec.fov.N_points - 1])
    E = np.zeros((T.N_points, len(pts)))
    E[pts, np.arange(len(pts))] = 1.0
    e_c = O.rel_l2(op.columns(pts), O.forward(T, E))
    g = np.zeros(T.N_points, dtype=np.complex128)
    g[pts] = [1.0, -0.5j, 2.0, 0.25]
    op.set_sparse_forward(100)
    e_s = O.rel_l2(op.forward(g), O.forward(T, g))
    log("3-D columns %.2e, gather forward %.2e" % (e_c, e_s))
    assert e_c < TOL_FP32 and e_s < TOL_FP32


def test_random_syntheses_plan(torch_cuda):
    """Plans built on the device from ra / rd / rad waves (syntheses/random)
    against the oracle's tables."""
    spec = tiny_spec(N_e=6, Nx=24, Nz=5, npp=4)
    arr, c = spec.array, spec.medium.c_avg
    spec.waves = [C.ra_wave(arr, np.random.default_rng(11)),
                  C.rd_wave(arr, np.random.default_rng(12), [math.sin(0.25)], c),
                  C.rad_wave(arr, np.random.default_rng(13), np.random.default_rng(14), [math.sin(-0.1)], c)]
    spec.mixes = [C.identity_mixes(arr.N_elements) for _ in spec.waves]
    T = O.build_tables(spec)
    op = OperatorBorn.from_sequence(spec)
    g = rnd((T.N_points, 2), 121)
    u = rnd((T.N_observations(), 2), 122)
    e_f, e_a = O.rel_l2(op.forward(g), O.forward(T, g)), O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    log("ra/rd/rad plan: forward %.2e adjoint %.2e" % (e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


def test_op_A_with_dictionary_and_tgc(torch_cuda):
    """op_A(theta, 1, dictionary=D4) = A Psi^T theta and op_A(y, 2) =
    Psi A^H y (forward_scalar / adjoint_scalar), with and without TGC; the
    pair stays adjoint."""
    from comp_ui_toolbox_amd.wavelet import D4Dictionary
    spec = tiny_spec(N_e=4, Nx=16, Nz=8, npp=4)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    D = D4Dictionary((16, 8))
    theta = rnd(T.N_points, 131)
    y = rnd(T.N_observations(), 132)
    e_f = O.rel_l2(op.op_A(theta, 1, dictionary=D), O.forward(T, D.adjoint_transform(theta)))
    e_a = O.rel_l2(op.op_A(y, 2, dictionary=D), D.forward_transform(O.adjoint(T, y)))
    op.set_tgc([np.array([0.1, 1.0, 0.1])] * len(op.meta["mix_n_obs"]))
    for tgc in (False, True):
        lhs = np.vdot(y, op.op_A(theta, 1, tgc=tgc, dictionary=D))
        rhs = np.vdot(op.op_A(y, 2, tgc=tgc, dictionary=D), theta)
        assert abs(lhs - rhs) < TOL_FP32 * abs(lhs)
    log("op_A with D4 dictionary: forward %.2e adjoint %.2e" % (e_f, e_a))
    assert e_f < TOL_FP32 and e_a < TOL_FP32


class _Identity:
    """The canonical basis as a dictionary (linear_transforms.identity)."""

    def adjoint_transform(self, x):
        return x


def test_energy_rx_arbitrary(torch_cuda):
    """Received energies of dictionary coefficients (energy_rx_arbitrary,
    operator_born.m:606-713): with the D4 wavelet, the squared column norms
    of the measurement's rows of A Psi^T (A assembled entry by entry in
    FP64), per measurement and summed over both (energy_rx_scalar), with and
    without TGC; with the canonical basis, the canonical energies
    (energy_rx_canonical).  Batches of 48 atoms leave a partial last batch."""
    from comp_ui_toolbox_amd.wavelet import D4Dictionary
    spec = tiny_spec(N_e=4, Nx=16, Nz=8, npp=4)
    T = O.build_tables(spec)
    op = OperatorBorn.from_tables(T)
    D = D4Dictionary((16, 8))
    A = O.assemble_matrix(T)
    Psi_T = D.adjoint_transform(np.eye(T.N_points))          # columns: the atoms
    APt = A @ Psi_T
    kernels = [np.array([0.1 - 0.05j, 1.0, 0.1 + 0.05j])] * len(op.meta["mix_n_obs"])
    APt_tgc = O.tgc_apply(kernels, op.meta["mix_n_obs"], APt, 1)
    for m in (0, 1):
        rows = np.asarray(op.measurement_rows(m))
        ref = np.sum(np.abs(APt[rows]) ** 2, axis=0)
        e = O.rel_l2(op.energy_rx_arbitrary(D, [m], batch=48), ref)
        e_c = O.rel_l2(op.energy_rx_arbitrary(_Identity(), [m], batch=48), O.energy_rx(T, m))
        log("energy_rx_arbitrary measurement %d: D4 rel-L2 %.2e, canonical %.2e" % (m, e, e_c))
        assert e < TOL_FP32 and e_c < TOL_FP32
    e_s = O.rel_l2(op.energy_rx_arbitrary(D), np.sum(np.abs(APt) ** 2, axis=0))
    op.set_tgc(kernels)
    e_t = O.rel_l2(op.energy_rx_arbitrary(D, tgc=True), np.sum(np.abs(APt_tgc) ** 2, axis=0))
    log("energy_rx_arbitrary summed over measurements %.2e, with TGC %.2e" % (e_s, e_t))
    assert e_s < TOL_FP32 and e_t < TOL_FP32


# ------------------------------------------------------ few-column chunks
@pytest.mark.parametrize("precision,tol", [("fp32", TOL_FP32), ("split_bf16", TOL_SPLIT)])
@pytest.mark.parametrize("n_waves,n_rhs", [(1, 1), (1, 2), (1, 4), (3, 1), (4, 1), (2, 2), (1, 3), (1, 20)])
@pytest.mark.parametrize("dims", ["2d", "3d"])
def test_few_column_kernels(torch_cuda, dims, n_waves, n_rhs, precision, tol):
    """Chunks of at most 4 columns (1-4 waves per frequency; with several
    right-hand sides packed per column) run on the FP32 vector kernels: one
    or two columns in both forms, 3-4 in the FP32 form.  Rows whose points per phase
    are not a multiple of the task length, 2-D at pitch/4 and 3-D at pitch/2
    (20 x 2 elements, 130-point rows), host and device entry points, against
    the oracle.  20 right-hand sides are too many to pack per column: each
    chunk then holds one rhs, and a chunk group walks chunks of several rhs
    (in 3-D slice by slice)."""
    import torch
    med = C.MediumSpec()
    if dims == "2d":
        spec = tiny_spec(N_e=40, Nx=170, Nz=5, npp=4)
        spec.waves = [C.qpw_wave(spec.array, a, med.c_avg) for a in np.linspace(-0.25, 0.25, n_waves)]
        spec.mixes = [C.identity_mixes(40) for _ in spec.waves]
        T = O.build_tables(spec)
    else:
        arr = C.ArraySpec((20, 2), (3e-4, 3e-4), (2.7e-4, 2.7e-4), (2, 2))
        fov = C.FovSpec((130, 6, 2), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
        waves = [C.qsw_wave(arr, np.array([x, 0.0, -1e-2]), med.c_avg) for x in np.linspace(-6e-4, 6e-4, n_waves)]
        spec = C.SequenceSpec("fc3", arr, med, fov, np.array([2.5e6, 4e6, 5.5e6]), waves,
                              [C.identity_mixes(40) for _ in waves])
        T = O.direct_tables(spec)
    if dims == "2d":
        op = OperatorBorn.from_tables(T, precision=precision)
    else:
        op = OperatorBorn.from_sequence(spec, precision=precision)
        assert op.path == "structured"
    g = rnd((T.N_points, n_rhs), 111)
    u = rnd((T.N_observations(), n_rhs), 112)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    gd = torch.from_numpy(g.T.astype(np.complex64)).cuda().contiguous()
    ud = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    e_fd = O.rel_l2(op.forward_dev(gd).cpu().numpy().T, O.forward(T, g.astype(np.complex64)))
    e_ad = O.rel_l2(op.adjoint_dev(ud).cpu().numpy().T, O.adjoint(T, u.astype(np.complex64)))
    log("few columns %s, %d waves, %d rhs, %s: forward %.2e adjoint %.2e (device %.2e %.2e)"
        % (dims, n_waves, n_rhs, precision, e_f, e_a, e_fd, e_ad))
    assert max(e_f, e_a, e_fd, e_ad) < tol
    op.close()


# ------------------------------------------------- stacked 8-column chunks
@pytest.mark.parametrize("n_waves,n_rhs", [(5, 1), (8, 1), (9, 1), (3, 2), (1, 8)])
@pytest.mark.parametrize("dims", ["2d", "2d_dual", "2d_mt1", "3d"])
def test_stacked_mfma_kernels(torch_cuda, dims, n_waves, n_rhs):
    """FP32 chunks of 5-8 columns (5-8 waves; 3 waves x 2 and 1 wave x 8
    right-hand sides packed per column) run on the plain MFMA kernels in the
    stacked form (real parts of the 8 columns in one half of the B operand,
    imaginary parts in the other); 9 waves run in the 16-column form.  Rows
    of 170 points (40 elements), of 100 points (24 elements, two rows per
    forward step), of 150 points with 12 elements (one element tile) and 3-D
    at pitch/2, host and device entry points, against the oracle."""
    import torch
    med = C.MediumSpec()
    if dims != "3d":
        N_e, Nx = {"2d": (40, 170), "2d_dual": (24, 100), "2d_mt1": (12, 150)}[dims]
        spec = tiny_spec(N_e=N_e, Nx=Nx, Nz=5, npp=4)
        spec.waves = [C.qpw_wave(spec.array, a, med.c_avg) for a in np.linspace(-0.25, 0.25, n_waves)]
        spec.mixes = [C.identity_mixes(N_e) for _ in spec.waves]
        T = O.build_tables(spec)
        op = OperatorBorn.from_tables(T)
    else:
        arr = C.ArraySpec((20, 2), (3e-4, 3e-4), (2.7e-4, 2.7e-4), (2, 2))
        fov = C.FovSpec((130, 6, 2), (1.5e-4, 1.5e-4, 1.5e-4), 2e-3)
        waves = [C.qsw_wave(arr, np.array([x, 0.0, -1e-2]), med.c_avg) for x in np.linspace(-6e-4, 6e-4, n_waves)]
        spec = C.SequenceSpec("stk3", arr, med, fov, np.array([2.5e6, 4e6, 5.5e6]), waves,
                              [C.identity_mixes(40) for _ in waves])
        T = O.direct_tables(spec)
        op = OperatorBorn.from_sequence(spec)
        assert op.path == "structured"
    g = rnd((T.N_points, n_rhs), 121)
    u = rnd((T.N_observations(), n_rhs), 122)
    e_f = O.rel_l2(op.forward(g), O.forward(T, g))
    e_a = O.rel_l2(op.adjoint(u), O.adjoint(T, u))
    gd = torch.from_numpy(g.T.astype(np.complex64)).cuda().contiguous()
    ud = torch.from_numpy(u.T.astype(np.complex64)).cuda().contiguous()
    e_fd = O.rel_l2(op.forward_dev(gd).cpu().numpy().T, O.forward(T, g.astype(np.complex64)))
    e_ad = O.rel_l2(op.adjoint_dev(ud).cpu().numpy().T, O.adjoint(T, u.astype(np.complex64)))
    log("stacked %s, %d waves, %d rhs: forward %.2e adjoint %.2e (device %.2e %.2e)"
        % (dims, n_waves, n_rhs, e_f, e_a, e_fd, e_ad))
    assert max(e_f, e_a, e_fd, e_ad) < TOL_FP32
    op.close()
