"""CPU checks of the boundary: the C ABI library loads and exports every
symbol include/born.h declares, status identifiers match the reference's
error IDs, early argument errors are reported without touching a GPU, and
the host-side table producers agree with the oracle."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from comp_ui_toolbox_amd import _lib, physics, configs
from comp_ui_toolbox_amd.operator_born import combined_quick_gpu
from oracle import born_oracle as O

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "born.h").read_text()
    return sorted(set(re.findall(r"\b(born_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTED)


def test_status_identifiers():
    L = _lib.lib()
    ids = {s: L.born_status_id(s).decode() for s in range(0, 13)}
    assert ids[1] == "combined_quick_gpu:NumberArguments"
    assert ids[2] == "combined_quick_gpu:NoSymmetricGridsSetup"
    assert ids[3] == "combined_quick_gpu:NoScatteringOptionsGPUActive"
    assert ids[4] == "combined_quick_gpu:NoComplexDoubles"
    assert ids[5] == "combined_quick_gpu:NoOperatorBorn"
    assert ids[6] == "combined_quick_gpu:NoNumericScalarMode"
    assert ids[7] == "combined_quick_gpu:NoValidDoubleMatrix"
    assert ids[8] == "combined_quick_gpu:ConversionFailed"
    assert ids[9] == "adjoint_quick_gpu:ErrorCUDA"


def test_early_argument_errors_without_device():
    L = _lib.lib()
    plan = C.c_void_p()
    assert L.born_plan_create_tables(None, C.byref(plan)) == 1
    d = _lib.TablesDesc()
    assert L.born_plan_create_tables(C.byref(d), C.byref(plan)) == 4          # h_ref missing
    buf = np.zeros(4)
    d.h_ref = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert L.born_plan_create_tables(C.byref(d), C.byref(plan)) == 2          # no shift map
    assert L.born_forward(None, None, 1, None) == 5
    assert "operator_born" in L.born_last_error().decode()
    assert L.born_plan_path(None) == -1
    with pytest.raises(Exception, match="NoOperatorBorn"):
        combined_quick_gpu(object(), 1, np.zeros((3, 1)))


def test_host_physics_matches_oracle():
    spec = configs.config2()
    f = spec.frequencies
    assert np.max(np.abs(physics.wavenumbers(f, spec.medium) - O.wavenumbers(f, spec.medium))) < 1e-9
    m = 5
    assert np.allclose(physics.normal_velocities(spec, m), O.normal_velocities(spec, m), rtol=1e-14, atol=0)
    T = O.build_tables(spec, h_ref=np.zeros((len(f), 1)), with_p_inc=False) if False else None
    pref = physics.prefactors_mix(spec, m, 3)
    k = O.wavenumbers(f, spec.medium)
    expect = (-spec.fov.cell_volume * k ** 2)[:, None] * spec.mixes[m][3].weights[None, :]
    assert np.allclose(pref, expect, rtol=1e-14, atol=0)


def test_config_shapes():
    c1 = configs.config1()
    assert c1.N_observations() == 361856 and c1.fov.N_points == 262144
    assert c1.N_points_per_pitch_axis == (4,)
    c2 = configs.config2()
    assert c2.N_observations() == 131584
    c0 = configs.config0()
    assert c0.N_observations() == 4160
    g = configs.phantom_points(c0, 30)
    assert np.count_nonzero(g) == 30
