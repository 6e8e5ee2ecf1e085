"""Small host-side table producers of the operator (all O(N_f x N_e)).

The large tables (transfer function, incident field) are built on the
device by libborn; these supply the per-frequency and per-element inputs.
"""
from __future__ import annotations

import math

import numpy as np

from .configs import T_CLK, SequenceSpec

DBL_EPS = float(np.finfo(np.float64).eps)


def wavenumbers(f: np.ndarray, medium) -> np.ndarray:
    """k-tilde(f) of the time-causal power-law model: real part 2 pi f/c_ref
    plus causal dispersion (log term for odd integer exponents, tan term
    otherwise), imaginary part -(alpha_0 + alpha_1 |f|^y), k = 0 real part at
    f = 0 (time_causal.m:82-83, :118-153)."""
    f = np.asarray(f, dtype=np.float64)
    y = float(medium.exponent)
    a0 = medium.absorption_constant * math.log(10.0) / 0.2
    a1 = medium.absorption * math.log(10.0) / (0.2 * 1e6 ** y)
    kr = 2.0 * math.pi * f / medium.c_ref
    if medium.dispersion:
        with np.errstate(divide="ignore", invalid="ignore"):
            if y % 2 == 0 or math.floor(y) != y:
                kr = kr + a1 * math.tan(y * math.pi / 2.0) * f * (np.abs(f) ** (y - 1.0) - abs(medium.f_ref) ** (y - 1.0))
            else:
                kr = kr - 2.0 * a1 * f ** y * np.log(np.abs(f / medium.f_ref)) / math.pi
    kr = np.where(np.abs(f) < DBL_EPS, 0.0, kr)
    return kr - 1j * (a0 + a1 * np.abs(f) ** y)


def prefactors_mix(spec: SequenceSpec, m: int, k: int) -> np.ndarray:
    """-dV k(f)^2 times the receive impulse-response weights of mix k of
    measurement m: complex (N_f_mix, N_active) (setup.m:759, sequence.m:623)."""
    fidx = spec.measurement_frequencies(m)
    mix = spec.mixes[m][k]
    fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
    kk = wavenumbers(spec.frequencies[fidx[fk]], spec.medium)
    return (-spec.fov.cell_volume * kk ** 2)[:, None] * np.asarray(mix.weights, dtype=np.complex128)[None, :]


def normal_velocities(spec: SequenceSpec, m: int) -> np.ndarray:
    """v_d(f, a) = apod_a U(f) exp(-j 2 pi f round(tau_a/T_clk) T_clk) for the
    active transmit elements of measurement m, unit transmit impulse
    response: complex (N_f_meas, N_tx) (setup.m:625, :635-646; tx.m:264-286)."""
    wave = spec.waves[m]
    f = spec.frequencies[spec.measurement_frequencies(m)]
    tau_q = np.round(np.asarray(wave.delays) / T_CLK) * T_CLK
    U = spec.excitation_spectrum(f)
    return U[:, None] * np.asarray(wave.apodization)[None, :] * np.exp(-2j * math.pi * f[:, None] * tau_q[None, :])


def face_weights(spec: SequenceSpec, f: np.ndarray):
    """Complex weights of the face integration points, N_face x N_f
    (face.m:359-376): apodization(r) exp(-j k_lens(f) t(r)), r the points'
    positions relative to the face centre normalized by the half widths
    (face.m:219-226), t the lens thickness of +thickness/focus_axial.m
    (sum over focused axes of (sqrt(1 + F^2) - sqrt(r^2 + F^2)) w/2 with
    F = 2 focus / w), k_lens = 2 pi f / c_lens (lossless lens, none.m:83).
    Returns None for uniform apodization without a lens (all weights 1).
    The face must be mirror symmetric (face.m:290-347), as the shared
    reference transfer function requires; otherwise ValueError."""
    arr = spec.array
    if arr.lens is None and arr.apodization is None:
        return None
    offs, _ = arr.face_offsets()
    width = np.asarray(arr.width_axis, dtype=np.float64)
    rel = 2.0 * offs / width[None, :]
    apod = np.ones(rel.shape[0]) if arr.apodization is None else \
        np.asarray(arr.apodization(rel), dtype=np.float64).reshape(-1)
    thick = np.zeros(rel.shape[0])
    k_lens = np.zeros(len(np.atleast_1d(f)))
    if arr.lens is not None:
        foc = np.asarray(arr.lens.axial_focus_axis, dtype=np.float64)
        if foc.shape != width.shape:
            raise ValueError("the lens needs one axial focus per face axis")
        for d in np.nonzero(np.isfinite(foc))[0]:
            Fn = 2.0 * foc[d] / width[d]
            thick += (math.sqrt(1.0 + Fn * Fn) - np.sqrt(rel[:, d] ** 2 + Fn * Fn)) * width[d] / 2.0
        k_lens = 2.0 * math.pi * np.asarray(f, dtype=np.float64).reshape(-1) / arr.lens.c_lens
    # mirror symmetry of apodization and thickness along every face axis
    grid = tuple(arr.N_face_points_axis)
    for vals in (apod, thick):
        v = vals.reshape(grid, order="F")
        for ax in range(v.ndim):
            if np.max(np.abs(v - np.flip(v, axis=ax)), initial=0.0) > 1e-12 * max(1.0, np.max(np.abs(v))):
                raise ValueError("combined_quick_gpu:NoSymmetricGridsSetup: face apodization and lens must be "
                                 "mirror symmetric")
    return apod[:, None] * np.exp(-1j * thick[:, None] * k_lens[None, :])
