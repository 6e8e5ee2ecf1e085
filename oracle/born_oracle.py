"""FP64 CPU restatement of the Born operator pair -- TEST INFRASTRUCTURE ONLY.

This module is the oracle: tests, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` use it as the checker.  The product path
(`comp_ui_toolbox_amd`) never imports it.

It restates, in NumPy/SciPy double precision:

* the table producers the operator consumes -- complex wavenumbers
  (time_causal.m:99-155), the Green's function (greens_function.m:67, :88,
  :103), the face-integrated spatial transfer function (setup.m:1216-1281),
  the anti-aliasing taper (setup.m:916-1032, on.m:49-72, raised_cosine.m:92-121,
  logistic.m:92-106, boxcar.m:72-86), the lateral shift/mirror map
  (setup_grid_symmetric.m:95-241), the prefactors (setup.m:709-768,
  sequence.m:590-634), the normal velocities and the incident field
  (setup.m:531-653, tx.m:264-286, sequence.m:768-861);
* the operator itself -- forward_quick_cpu / adjoint_quick_cpu
  (operator_born.m:332-596) with the MEX's occupancy rule
  (combined_quick_gpu.cu:454-499) and observation ordering (cu:953-991);
* a dense assembly of A for tiny cases (brute force).

Every function is pinned by a check in tests/test_oracle.py (closed forms,
scipy/mpmath special functions, brute force, invariants).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Sequence

import numpy as np
import scipy.special

DBL_EPS = np.finfo(np.float64).eps
T_CLK = 1.0 / 80e6


# ---------------------------------------------------------------------------
# physics
# ---------------------------------------------------------------------------
def wavenumbers(f: np.ndarray, medium) -> np.ndarray:
    """Complex wavenumbers of the time-causal power-law model
    (time_causal.m:82-83 for alpha_0/alpha_1, :118-153 for k-tilde)."""
    f = np.asarray(f, dtype=np.float64)
    y = float(medium.exponent)
    alpha_0 = medium.absorption_constant * math.log(10) / (20 * 0.01)
    alpha_1 = medium.absorption * math.log(10) / (20 * 0.01 * (1e6) ** y)
    k_ref = 2 * math.pi * f / medium.c_ref
    k_imag = -alpha_0 - alpha_1 * np.abs(f) ** y
    if medium.dispersion:
        if (y % 2 == 0) or (math.floor(y) != y):
            with np.errstate(divide="ignore", invalid="ignore"):
                k_real = k_ref + alpha_1 * math.tan(y * math.pi / 2) * f * (
                    np.abs(f) ** (y - 1) - abs(medium.f_ref) ** (y - 1))
        else:
            with np.errstate(divide="ignore", invalid="ignore"):
                k_real = k_ref - 2 * alpha_1 * f ** y * np.log(np.abs(f / medium.f_ref)) / math.pi
    else:
        k_real = k_ref.copy()
    k_real = np.where(np.abs(f) < DBL_EPS, 0.0, k_real)
    return k_real + 1j * k_imag


def greens_function(k: np.ndarray, D: np.ndarray, n_dims: int) -> np.ndarray:
    """Free-space Green's function (greens_function.m:67, :88, :103):
    2-D (j/4) H0^(2)(k D), 3-D -exp(-j k D)/(4 pi D).  Broadcasts k vs D."""
    kD = k * D
    if n_dims == 2:
        return 0.25j * scipy.special.hankel2(0, kD)
    if n_dims == 3:
        return -np.exp(-1j * kD) / (4 * math.pi * D)
    raise ValueError("greens_function:UnknownDimensions")


def hankel2_0_series(z: np.ndarray) -> np.ndarray:
    """H0^(2)(z) written out from its published expansions: the ascending
    series for |z| < 12 (DLMF 10.2.2, 10.8.2) and the Hankel asymptotic
    expansion truncated at its smallest term for |z| >= 12 (DLMF 10.17.4,
    10.17.6).  The device table build evaluates the same formulas; this
    CPU copy exists so tests can pin them against scipy and mpmath."""
    z = np.asarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    flat_z = z.reshape(-1)
    flat_o = out.reshape(-1)
    for i, zz in enumerate(flat_z):
        flat_o[i] = _h20_scalar(complex(zz))
    return out


def _h20_scalar(z: complex) -> complex:
    az = abs(z)
    if az >= 12.0:
        # sum_k (-i)^k a_k(0) / z^k with a_0 = 1, a_k = a_{k-1} * (-(2k-1)^2) / (8k)
        s = 1.0 + 0.0j
        term = 1.0 + 0.0j
        prev = math.inf
        for k in range(1, 60):
            term = term * (-(2 * k - 1) ** 2) / (8.0 * k) * (-1j) / z
            at = abs(term)
            if at >= prev:
                break
            s += term
            prev = at
            if at < 1e-17 * abs(s):
                break
        # sqrt(2/(pi z)) * exp(-i (z - pi/4))
        return np.sqrt(2.0 / (math.pi * z)) * np.exp(-1j * (z - math.pi / 4)) * s
    q = 0.25 * z * z
    term = 1.0 + 0.0j
    j0 = 1.0 + 0.0j
    ysum = 0.0 + 0.0j
    hk = 0.0
    for k in range(1, 80):
        term = term * (-q) / (k * k)
        hk += 1.0 / k
        j0 += term
        ysum += -term * hk          # (-1)^{k+1} H_k (z^2/4)^k / (k!)^2
        if abs(term) < 1e-18 * max(1.0, abs(j0)):
            break
    y0 = (2.0 / math.pi) * ((np.log(z / 2.0) + np.euler_gamma) * j0 + ysum)
    return j0 - 1j * y0


# ---------------------------------------------------------------------------
# anti-aliasing taper
# ---------------------------------------------------------------------------
def filter_samples(flags: np.ndarray, filt) -> np.ndarray:
    """Taper of the local phase shift per pitch, product over the last axis
    (raised_cosine.m:92-121, logistic.m:92-106, boxcar.m:72-86, off = 1)."""
    kind = filt.kind
    if kind == "off":
        return np.ones(flags.shape[:-1])
    if kind == "raised_cosine":
        lb = math.pi * (1 - filt.param)
        ub = math.pi
        out = np.where(flags <= lb, 1.0, 0.0)
        taper = (flags > lb) & (flags < ub)
        out = np.where(taper, 0.5 * (1 + np.cos(math.pi * (flags - lb) / (ub - lb))), out)
        return np.prod(out, axis=-1)
    if kind == "logistic":
        return np.prod(1.0 / (1.0 + np.exp(filt.param * (flags - math.pi))), axis=-1)
    if kind == "boxcar":
        return np.all(flags < math.pi, axis=-1).astype(np.float64)
    raise ValueError(kind)


def filter_flags(spec, f: np.ndarray, element: int, positions: np.ndarray) -> np.ndarray:
    """Local phase shift per pitch real(k) * pitch * |e_lateral| for the unit
    vectors from the element centre to the grid points, lateral dimensions
    with more than one element only (setup.m:980-1017).  Shape
    (N_f, N_pts, N_lat_relevant)."""
    k = wavenumbers(f, spec.medium)
    ctr = spec.array.positions_ctr()[element]
    nd = spec.fov.N_dimensions
    ctr_full = np.concatenate([ctr, [0.0]])
    vec = positions - ctr_full[None, :]
    dist = np.linalg.norm(vec, axis=1)
    e = np.abs(vec[:, : nd - 1] / dist[:, None])
    rel = [d for d in range(nd - 1) if spec.array.N_elements_axis[d] > 1]
    e = e[:, rel]
    pitch = np.array([spec.array.pitch_axis[d] for d in rel])
    return np.real(k)[:, None, None] * pitch[None, None, :] * e[None, :, :]


# ---------------------------------------------------------------------------
# face weights (apodization and acoustic lens)
# ---------------------------------------------------------------------------
def lens_thickness(focus_axis, width_axis, rel_norm: np.ndarray) -> np.ndarray:
    """Thickness profile of a lens focusing laterally at the axial distances
    `focus_axis` (+thickness/focus_axial.m:21-36): for each face point with
    normalized relative position rel_norm (rows, one column per face axis),
    sum over the finite foci F of (sqrt(1 + Fn^2) - sqrt(r^2 + Fn^2)) w / 2,
    Fn = 2 F / w.  All foci infinite gives zeros."""
    rel_norm = np.atleast_2d(np.asarray(rel_norm, dtype=np.float64))
    out = np.zeros(rel_norm.shape[0])
    for d, (F, w) in enumerate(zip(focus_axis, width_axis)):
        if math.isinf(F):
            continue
        Fn = 2.0 * F / w
        peak = math.sqrt(1.0 + Fn ** 2)
        for i in range(rel_norm.shape[0]):
            out[i] += (peak - math.sqrt(rel_norm[i, d] ** 2 + Fn ** 2)) * w / 2.0
    return out


def face_weights(spec, f: np.ndarray) -> Optional[np.ndarray]:
    """Complex apodization weights of the face integration points
    (face.compute_weights_scalar, face.m:359-376): apodization(rel_norm) .*
    exp(-1j * thickness * k_lens(f)), with the normalized positions
    2 (r - r_ctr) / widths (face.m:219-226), the thickness of
    lens_thickness and the wavenumbers of the lossless lens material,
    2 pi f / c_lens (absorption_models/none.m:83).  (N_face, N_f); None for
    a uniform face without a lens (every weight 1)."""
    arr = spec.array
    if getattr(arr, "lens", None) is None and getattr(arr, "apodization", None) is None:
        return None
    offs, _ = arr.face_offsets()
    rel_norm = offs / (np.asarray(arr.width_axis, dtype=np.float64)[None, :] / 2.0)
    f = np.atleast_1d(np.asarray(f, dtype=np.float64))
    apod = np.ones(offs.shape[0]) if arr.apodization is None else np.asarray(arr.apodization(rel_norm), float)
    W = np.empty((offs.shape[0], len(f)), dtype=np.complex128)
    if arr.lens is None:
        W[:] = apod.reshape(-1, 1)
        return W
    thick = lens_thickness(arr.lens.axial_focus_axis, arr.width_axis, rel_norm)
    k_lens = 2.0 * math.pi * f / arr.lens.c_lens
    for q in range(offs.shape[0]):
        W[q] = apod.reshape(-1)[q] * np.exp(-1j * thick[q] * k_lens)
    return W


# ---------------------------------------------------------------------------
# spatial transfer function, shift map
# ---------------------------------------------------------------------------
def transfer_function(spec, f: np.ndarray, element: int = 0, filt=None,
                      positions: Optional[np.ndarray] = None,
                      face_weights_fq: Optional[np.ndarray] = None) -> np.ndarray:
    """h[f, p] = -2 dA sum_face w G(k_f, |r_p - r_face|) (setup.m:1216-1281),
    then the anti-aliasing taper of the same element (setup.m:902,
    on.m:49-72).  `element` is 0-based; `positions` defaults to the FOV grid.
    The face weights w (N_face, N_f) default to face_weights(spec, f)
    (setup.m:1241, :1262)."""
    if positions is None:
        positions = spec.fov.positions()
    nd = spec.fov.N_dimensions
    k = wavenumbers(f, spec.medium)
    offs, dA = spec.array.face_offsets()
    ctr = spec.array.positions_ctr()[element]
    if face_weights_fq is None:
        face_weights_fq = face_weights(spec, f)
    h = np.zeros((len(f), positions.shape[0]), dtype=np.complex128)
    for q in range(offs.shape[0]):
        fp = np.concatenate([ctr + offs[q], [0.0]])
        D = np.linalg.norm(positions - fp[None, :], axis=1)
        G = greens_function(k[:, None], D[None, :], nd)
        w = 1.0 if face_weights_fq is None else face_weights_fq[q][:, None]
        h -= 2.0 * dA * w * G
    if filt is not None and filt.kind != "off":
        flags = filter_flags(spec, f, element, positions)
        h *= filter_samples(flags, filt)
    return h


def shift_lateral(spec) -> np.ndarray:
    """1-based N_pts x N_elements map: element e's transfer function at p is
    h_ref[:, shift[p, e]] (setup_grid_symmetric.m:95-241)."""
    npp = spec.N_points_per_pitch_axis
    N_axis = spec.fov.N_points_axis
    nd = len(N_axis)
    nlat = nd - 1
    sub = np.stack(np.unravel_index(np.arange(spec.fov.N_points), N_axis, order="F"), axis=1) + 1
    e_sub = np.stack(np.unravel_index(np.arange(spec.array.N_elements),
                                      spec.array.N_elements_axis, order="F"), axis=1) + 1
    out = np.empty((spec.fov.N_points, spec.array.N_elements), dtype=np.int64)
    for e in range(spec.array.N_elements):
        s = sub.copy()
        for d in range(nlat):
            s[:, d] = sub[:, d] - (e_sub[e, d] - 1) * npp[d]
            off = N_axis[d] - (spec.array.N_elements_axis[d] - 1) * npp[d] + 1
            neg = s[:, d] <= 0
            s[neg, d] = off - s[neg, d]
        out[:, e] = np.ravel_multi_index(tuple((s - 1).T), N_axis, order="F") + 1
    return out


# ---------------------------------------------------------------------------
# operator tables (what operator_born holds and the MEX reads)
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class MixTables:
    idx_f_to_meas: np.ndarray      # 1-based into the measurement's unique freqs
    idx_active: np.ndarray         # 1-based element indices
    prefactors: np.ndarray         # (N_f_mix, N_active) complex


@dataclasses.dataclass
class MeasTables:
    idx_f_to_global: np.ndarray    # 1-based into global unique freqs
    p_incident: np.ndarray         # (N_f_meas, N_pts) complex
    mixes: List[MixTables]


@dataclasses.dataclass
class BornTables:
    h_ref: np.ndarray              # (N_f_unique, N_pts) complex
    shift: np.ndarray              # (N_pts, N_e) 1-based
    meas: List[MeasTables]
    N_points_axis: tuple
    N_elements_axis: tuple
    N_points_per_pitch_axis: tuple

    @property
    def N_points(self):
        return self.shift.shape[0]

    @property
    def N_elements(self):
        return self.shift.shape[1]

    def N_observations(self):
        return sum(len(mx.idx_f_to_meas) for ms in self.meas for mx in ms.mixes)


def normal_velocities(spec, m: int) -> np.ndarray:
    """v_d[f, a] = apod_a U(f) exp(-j 2 pi f tau_a) H_tx(f) with tau quantized
    to T_clk (setup.m:625, :635-646; tx.m:264-286); H_tx = 1 (delta response).
    Rows follow the measurement's unique frequencies."""
    wave = spec.waves[m]
    f = spec.frequencies[spec.measurement_frequencies(m)]
    tau_q = np.round(wave.delays / T_CLK) * T_CLK
    U = spec.excitation_spectrum(f)
    return U[:, None] * wave.apodization[None, :] * np.exp(-2j * math.pi * f[:, None] * tau_q[None, :])


def prefactors_unique(spec) -> np.ndarray:
    """-dV k(f)^2 over the global unique frequencies (setup.m:759)."""
    k = wavenumbers(spec.frequencies, spec.medium)
    return -spec.fov.cell_volume * k ** 2


def build_tables(spec, h_ref: Optional[np.ndarray] = None, h_tx: Optional[np.ndarray] = None,
                 with_p_inc: bool = True) -> BornTables:
    """Assemble what the operator object holds: h_ref with the rx filter
    (sequence.m:529-585), the shift map, p_inc per measurement from h_ref with
    the tx filter (sequence.m:768-861), prefactors times rx weights
    (sequence.m:590-634) and the 1-based frequency maps."""
    if h_ref is None:
        h_ref = transfer_function(spec, spec.frequencies, 0, spec.filter_rx)
    shift = shift_lateral(spec)
    if with_p_inc and h_tx is None:
        if (spec.filter_tx.kind, spec.filter_tx.param) == (spec.filter_rx.kind, spec.filter_rx.param):
            h_tx = h_ref
        else:
            h_tx = transfer_function(spec, spec.frequencies, 0, spec.filter_tx)
    pref_u = prefactors_unique(spec)
    meas = []
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        if with_p_inc:
            vd = normal_velocities(spec, m)
            p = incident_field(h_tx[fidx], shift, spec.waves[m].indices_active, vd)
        else:
            p = np.zeros((len(fidx), 0), dtype=np.complex128)
        mixes = []
        for mix in spec.mixes[m]:
            fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
            pref = pref_u[fidx[fk]][:, None] * np.asarray(mix.weights)[None, :]
            mixes.append(MixTables(fk + 1, np.asarray(mix.indices_active) + 1, pref.astype(np.complex128)))
        meas.append(MeasTables(fidx + 1, p, mixes))
    return BornTables(h_ref, shift, meas, tuple(spec.fov.N_points_axis),
                      tuple(spec.array.N_elements_axis), tuple(spec.N_points_per_pitch_axis))


def incident_field(h_tx_meas: np.ndarray, shift: np.ndarray, indices_active: np.ndarray,
                   v_d: np.ndarray) -> np.ndarray:
    """p_in[f, p] = sum_a h_tx[f, shift[p, e_a]] v_d[f, a] (sequence.m:825-840).
    `indices_active` 0-based."""
    p = np.zeros((h_tx_meas.shape[0], shift.shape[0]), dtype=np.complex128)
    for a, e in enumerate(indices_active):
        p += h_tx_meas[:, shift[:, e] - 1] * v_d[:, a][:, None]
    return p


# ---------------------------------------------------------------------------
# the operator pair
# ---------------------------------------------------------------------------
def occupied_points(gamma: np.ndarray, rule: str = "mex") -> np.ndarray:
    """Occupied grid points: 'mex' = any |Re| or |Im| above DBL_EPSILON over
    the columns (cu:466-488); 'cpu' = sum(|gamma|, 2) > eps
    (operator_born.m:361)."""
    g = np.asarray(gamma)
    if g.ndim == 1:
        g = g[:, None]
    if rule == "mex":
        occ = np.any((np.abs(g.real) > DBL_EPS) | (np.abs(g.imag) > DBL_EPS), axis=1)
    else:
        occ = np.sum(np.abs(g), axis=1) > DBL_EPS
    return np.nonzero(occ)[0]


def forward(T: BornTables, gamma: np.ndarray, rule: str = "mex",
            rows: Optional[np.ndarray] = None) -> np.ndarray:
    """u_M = A gamma, the loop nest measurement -> mix -> active element of
    forward_quick_cpu (operator_born.m:353-459): Phi = h(fmap, shift) .*
    p_inc(fmap, occ) .* pref(:, a); u += Phi * gamma(occ).  Frequencies always
    go through the maps, as in the MEX (cu:329-333, :1114-1120).  `rows`
    optionally restricts the output to those observation indices (0-based).
    Returns (N_obs, N_rhs) (or (len(rows), N_rhs))."""
    g = np.asarray(gamma, dtype=np.complex128)
    squeeze = g.ndim == 1
    if squeeze:
        g = g[:, None]
    occ = occupied_points(g, rule)
    N_obs = T.N_observations()
    out = np.zeros((N_obs, g.shape[1]), dtype=np.complex128)
    want = None if rows is None else np.zeros(N_obs, dtype=bool)
    if want is not None:
        want[np.asarray(rows)] = True
    obs = 0
    for ms in T.meas:
        fg_meas = ms.idx_f_to_global - 1
        p_occ = ms.p_incident[:, occ]
        for mx in ms.mixes:
            fm = mx.idx_f_to_meas - 1
            nf = len(fm)
            sel = np.arange(nf) if want is None else np.nonzero(want[obs:obs + nf])[0]
            if len(sel) and len(occ):
                acc = np.zeros((len(sel), g.shape[1]), dtype=np.complex128)
                for a, e1 in enumerate(mx.idx_active):
                    sh = T.shift[occ, e1 - 1] - 1
                    h = T.h_ref[np.ix_(fg_meas[fm[sel]], sh)]
                    Phi = h * p_occ[fm[sel]] * mx.prefactors[sel, a][:, None]
                    acc += Phi @ g[occ]
                out[obs + sel] = acc
            obs += nf
    if rows is not None:
        out = out[np.asarray(rows)]
    return out[:, 0] if squeeze else out


def energy_rx(T: BornTables, m: int = 0, kernels: Optional[Sequence[np.ndarray]] = None) -> np.ndarray:
    """Received energies of measurement m, canonical basis
    (energy_rx_canonical, operator_born.m:718-859): per mix, Phi_M = sum_a
    h_ref(fmap, shift(:, e_a)) .* p_inc(fm, :) .* pref(:, a) (:786, :830-831),
    with a diagonal TGC Phi_M = forward_transform(LT_tgc.transforms{k}, Phi_M)
    (:836-843; `kernels`: one per mix of measurement m, applied by tgc_apply),
    then E += vecnorm(Phi_M, 2, 1).^2 (:848).  Returns (N_pts,)."""
    ms = T.meas[m]
    fg_meas = ms.idx_f_to_global - 1
    E = np.zeros(T.N_points)
    for k, mx in enumerate(ms.mixes):
        fm = mx.idx_f_to_meas - 1
        Phi = np.zeros((len(fm), T.N_points), dtype=np.complex128)
        for a, e1 in enumerate(mx.idx_active):
            sh = T.shift[:, e1 - 1] - 1
            Phi += T.h_ref[np.ix_(fg_meas[fm], sh)] * ms.p_incident[fm] * mx.prefactors[:, a][:, None]
        if kernels is not None:
            Phi = tgc_apply([kernels[k]], [len(fm)], Phi, 1)
        E += np.sum(np.abs(Phi) ** 2, axis=0)
    return E


def time_signals(T: BornTables, frequencies: np.ndarray, u: np.ndarray, delta_f: float,
                 N_t: int) -> np.ndarray:
    """Time-domain mixed voltage signals (format_voltages, operator.m:900-1007,
    then signal_matrix.signal, signal_matrix.m:293-397, lbs_q = 0): for each
    mix of each measurement, its coefficients u(obs) placed at the integer
    frequency indices q = f / delta_f of a length floor(N_t/2)+1 spectrum,
    then N_t * irfft(., N_t) (numpy's real inverse FFT reads the spectrum as
    conjugate symmetric, as MATLAB's ifft(., 'symmetric') does).
    u: (N_obs,) or (N_obs, N_rhs).  Returns (N_rhs, N_mix_signals, N_t)."""
    uu = np.asarray(u)
    if uu.ndim == 1:
        uu = uu[:, None]
    q_fg = np.rint(np.asarray(frequencies) / delta_f).astype(np.int64)
    sig = []
    obs = 0
    for ms in T.meas:
        fg_meas = ms.idx_f_to_global - 1
        for mx in ms.mixes:
            fm = mx.idx_f_to_meas - 1
            X = np.zeros((N_t // 2 + 1, uu.shape[1]), dtype=np.complex128)
            X[q_fg[fg_meas[fm]]] = uu[obs:obs + len(fm)]
            sig.append(N_t * np.fft.irfft(X, N_t, axis=0))
            obs += len(fm)
    return np.stack(sig, axis=0).transpose(2, 0, 1)


def tgc_apply(kernels: Sequence[np.ndarray], mix_n_obs: Sequence[int], u: np.ndarray, mode: int) -> np.ndarray:
    """Time gain compensation as the reference computes it: per mix, the
    DFT-based discrete convolution linear_transforms.convolutions.fft with
    cut_off (fft.m:85-126; kernel truncation convolution.m:62-98).  Forward:
    y = ifft(fft(k, N_dft) .* fft(x, N_dft)) cut to rows M+1..end-M; adjoint:
    x padded by M zeros on both sides, y = ifft(conj(fft(k)) .* fft(x)) cut
    to its first N rows.  N_dft = len(k) + N - 1.  mode 1 forward, else
    adjoint.  u: (N_obs,) or (N_obs, N_rhs), mixes in observation order."""
    uu = np.asarray(u, dtype=np.complex128)
    squeeze = uu.ndim == 1
    if squeeze:
        uu = uu[:, None]
    out = np.empty_like(uu)
    obs = 0
    for k, N in zip(kernels, mix_n_obs):
        k = np.asarray(k, dtype=np.complex128).reshape(-1)
        M = (len(k) - 1) // 2
        if M + 1 > N:
            cut = M - N + 1
            k = k[cut:len(k) - cut]
            M = N - 1
        x = uu[obs:obs + N]
        n_dft = len(k) + N - 1
        kd = np.fft.fft(k, n_dft)
        if mode == 1:
            y = np.fft.ifft(kd[:, None] * np.fft.fft(x, n_dft, axis=0), axis=0)[M:n_dft - M]
        else:
            xp = np.concatenate([np.zeros((M, x.shape[1])), x, np.zeros((M, x.shape[1]))])
            y = np.fft.ifft(np.conj(kd)[:, None] * np.fft.fft(xp, n_dft, axis=0), axis=0)[:N]
        out[obs:obs + N] = y
        obs += N
    return out[:, 0] if squeeze else out


def adjoint(T: BornTables, u: np.ndarray, points: Optional[np.ndarray] = None) -> np.ndarray:
    """gamma_hat = A^H u_M over all grid points (or the 0-based subset
    `points`), the loop nest of adjoint_quick_cpu (operator_born.m:511-593):
    gamma_hat += Phi' * u_mk."""
    uu = np.asarray(u, dtype=np.complex128)
    squeeze = uu.ndim == 1
    if squeeze:
        uu = uu[:, None]
    pts = np.arange(T.N_points) if points is None else np.asarray(points)
    out = np.zeros((len(pts), uu.shape[1]), dtype=np.complex128)
    obs = 0
    for ms in T.meas:
        fg_meas = ms.idx_f_to_global - 1
        p_pts = ms.p_incident[:, pts]
        for mx in ms.mixes:
            fm = mx.idx_f_to_meas - 1
            nf = len(fm)
            u_mk = uu[obs:obs + nf]
            for a, e1 in enumerate(mx.idx_active):
                sh = T.shift[pts, e1 - 1] - 1
                h = T.h_ref[np.ix_(fg_meas[fm], sh)]
                Phi = h * p_pts[fm] * mx.prefactors[:, a][:, None]
                out += Phi.conj().T @ u_mk
            obs += nf
    return out[:, 0] if squeeze else out


def assemble_matrix(T: BornTables) -> np.ndarray:
    """Dense A (N_obs x N_pts) entry by entry from its definition; for tiny
    tables only."""
    A = np.zeros((T.N_observations(), T.N_points), dtype=np.complex128)
    obs = 0
    for ms in T.meas:
        for mx in ms.mixes:
            for fi, fm1 in enumerate(mx.idx_f_to_meas):
                fg = ms.idx_f_to_global[fm1 - 1] - 1
                for p in range(T.N_points):
                    s = 0j
                    for a, e1 in enumerate(mx.idx_active):
                        s += (mx.prefactors[fi, a] * T.h_ref[fg, T.shift[p, e1 - 1] - 1]
                              * ms.p_incident[fm1 - 1, p])
                    A[obs, p] = s
                obs += 1
    return A


def rel_l2(x: np.ndarray, ref: np.ndarray) -> float:
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))


def restricted_tables(spec, points: np.ndarray) -> BornTables:
    """Exact tables of the operator restricted to the grid points `points`
    (0-based): the forward of a gamma supported on `points`, and the adjoint
    at `points`, equal those of the full operator.  h_ref is evaluated only
    where the shift map sends `points` (transfer_function at those
    positions), p_inc only at `points` (sequence.m:825-840).  Makes
    full-size checks affordable on the CPU."""
    points = np.asarray(points)
    if mirror_map_defined(spec):
        shift, grid_pos = shift_lateral(spec), spec.fov.positions()
    else:
        shift, grid_pos = extended_shift(spec)
    need = np.unique(shift[points, :] - 1)
    pos = grid_pos[need]
    h_sub = transfer_function(spec, spec.frequencies, 0, spec.filter_rx, positions=pos)
    same = (spec.filter_tx.kind, spec.filter_tx.param) == (spec.filter_rx.kind, spec.filter_rx.param)
    htx_sub = h_sub if same else transfer_function(spec, spec.frequencies, 0, spec.filter_tx, positions=pos)
    remap = np.full(grid_pos.shape[0], -1, dtype=np.int64)
    remap[need] = np.arange(len(need))
    shift_sub = remap[shift[points, :] - 1] + 1
    pref_u = prefactors_unique(spec)
    meas = []
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        vd = normal_velocities(spec, m)
        p = incident_field(htx_sub[fidx], shift_sub, spec.waves[m].indices_active, vd)
        mixes = []
        for mix in spec.mixes[m]:
            fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
            pref = pref_u[fidx[fk]][:, None] * np.asarray(mix.weights)[None, :]
            mixes.append(MixTables(fk + 1, np.asarray(mix.indices_active) + 1, pref.astype(np.complex128)))
        meas.append(MeasTables(fidx + 1, p, mixes))
    return BornTables(h_sub, shift_sub, meas, (len(points), 1), tuple(spec.array.N_elements_axis),
                      tuple(spec.N_points_per_pitch_axis))


def direct_tables(spec, points: Optional[np.ndarray] = None) -> BornTables:
    """Tables of the operator restricted to the grid points `points` (0-based;
    all by default) built from every element's own transfer function at each
    point (setup.m:1216-1281 evaluated at the element's centre, with its own
    taper, setup.m:902), i.e. without any lateral shift or mirror map: column
    p*N_e + e of h_ref is element e at point p.  p_inc is the direct sum over
    the transmitting elements (sequence.m:825-840).  Valid for any FOV,
    including one narrower than the aperture."""
    N_e = spec.array.N_elements
    if points is None:
        points = np.arange(spec.fov.N_points)
    points = np.asarray(points)
    pos = spec.fov.positions()[points]
    f = spec.frequencies
    n = len(points)
    h_rx = np.empty((len(f), n * N_e), dtype=np.complex128)
    same = (spec.filter_tx.kind, spec.filter_tx.param) == (spec.filter_rx.kind, spec.filter_rx.param)
    h_tx = h_rx if same else np.empty_like(h_rx)
    for e in range(N_e):
        h_rx[:, e::N_e] = transfer_function(spec, f, e, spec.filter_rx, positions=pos)
        if not same:
            h_tx[:, e::N_e] = transfer_function(spec, f, e, spec.filter_tx, positions=pos)
    shift = (np.arange(n)[:, None] * N_e + np.arange(N_e)[None, :] + 1).astype(np.int64)
    pref_u = prefactors_unique(spec)
    meas = []
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        vd = normal_velocities(spec, m)
        p = incident_field(h_tx[fidx], shift, spec.waves[m].indices_active, vd)
        mixes = []
        for mix in spec.mixes[m]:
            fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
            pref = pref_u[fidx[fk]][:, None] * np.asarray(mix.weights)[None, :]
            mixes.append(MixTables(fk + 1, np.asarray(mix.indices_active) + 1, pref.astype(np.complex128)))
        meas.append(MeasTables(fidx + 1, p, mixes))
    return BornTables(h_rx, shift, meas, (n, 1), tuple(spec.array.N_elements_axis),
                      tuple(spec.N_points_per_pitch_axis))


def _taper_all(spec, f: float, positions: np.ndarray, filt) -> np.ndarray:
    """(n_pts, N_e) anti-aliasing taper of every element at one frequency:
    filter_samples(filter_flags(...)) of each element, vectorized over the
    elements (setup.m:902, :980-1017)."""
    n = positions.shape[0]
    N_e = spec.array.N_elements
    if filt is None or filt.kind == "off":
        return np.ones((n, N_e))
    nd = spec.fov.N_dimensions
    kr = float(np.real(wavenumbers(np.array([f]), spec.medium))[0])
    rel = [d for d in range(nd - 1) if spec.array.N_elements_axis[d] > 1]
    pitch = np.array([spec.array.pitch_axis[d] for d in rel])
    ctr = spec.array.positions_ctr()
    out = np.empty((n, N_e))
    for e0 in range(0, N_e, 64):
        c = ctr[e0:e0 + 64]
        c3 = np.concatenate([c, np.zeros((c.shape[0], 1))], axis=1)
        vec = positions[:, None, :] - c3[None, :, :]
        dist = np.linalg.norm(vec, axis=2)
        e = np.abs(vec[:, :, : nd - 1] / dist[..., None])[..., rel]
        out[:, e0:e0 + c.shape[0]] = filter_samples(kr * pitch * e, filt)
    return out


def sampled_tables(spec, points: np.ndarray, f_sel: np.ndarray):
    """The operator restricted to the grid points `points` and the global
    frequencies `f_sel` (both 0-based), built like direct_tables -- every
    element's own transfer function at each point, no shift or mirror map --
    for a 3-D sequence, with the face sum of setup.m:1216-1281 evaluated by
    the C routine ref_face_sum3 (pinned against transfer_function by the
    tests).  Returns (T, rows): T's observations are the full operator's
    observations at the sampled frequencies, in the full operator's order,
    and rows holds their 0-based indices in the full observation vector."""
    from . import ref_loop
    if spec.fov.N_dimensions != 3:
        raise ValueError("sampled_tables: 3-D sequences only")
    if face_weights(spec, spec.frequencies[:1]) is not None:
        raise ValueError("sampled_tables: uniform faces without a lens only")
    N_e = spec.array.N_elements
    points = np.asarray(points)
    pos = spec.fov.positions()[points]
    f_sel = np.asarray(sorted(set(int(i) for i in f_sel)))
    f = spec.frequencies[f_sel]
    k = wavenumbers(f, spec.medium)
    offs, dA = spec.array.face_offsets()
    ctr = spec.array.positions_ctr()
    n = len(points)
    same = (spec.filter_tx.kind, spec.filter_tx.param) == (spec.filter_rx.kind, spec.filter_rx.param)
    h_rx = np.empty((len(f), n * N_e), dtype=np.complex128)
    h_tx = h_rx if same else np.empty_like(h_rx)
    for i in range(len(f)):
        raw = ref_loop.face_sum3(pos, ctr, offs, dA, k[i])
        h_rx[i] = (raw * _taper_all(spec, f[i], pos, spec.filter_rx)).reshape(-1)
        if not same:
            h_tx[i] = (raw * _taper_all(spec, f[i], pos, spec.filter_tx)).reshape(-1)
    shift = (np.arange(n)[:, None] * N_e + np.arange(N_e)[None, :] + 1).astype(np.int64)
    pos_of = {int(g): i for i, g in enumerate(f_sel)}      # global frequency -> sampled row
    pref_u = prefactors_unique(spec)
    meas, rows = [], []
    obs = 0
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        keep = np.array([j for j, g in enumerate(fidx) if int(g) in pos_of], dtype=np.int64)  # measurement rows kept
        sub_of = {int(j): i for i, j in enumerate(keep)}
        vd = normal_velocities(spec, m)[keep]
        act = np.asarray(spec.waves[m].indices_active)
        hsel = h_tx[[pos_of[int(fidx[j])] for j in keep]].reshape(len(keep), n, N_e)
        p = np.einsum("fpa,fa->fp", hsel[:, :, act], vd)
        mixes = []
        for mix in spec.mixes[m]:
            fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
            jj = [j for j in range(len(fk)) if int(fk[j]) in sub_of]
            rows.extend(obs + j for j in jj)
            fk_kept = np.array([fk[j] for j in jj], dtype=np.int64)
            pref = pref_u[fidx[fk_kept]][:, None] * np.asarray(mix.weights)[None, :]
            mixes.append(MixTables(np.array([sub_of[int(x)] for x in fk_kept], dtype=np.int64) + 1,
                                   np.asarray(mix.indices_active) + 1, pref.astype(np.complex128)))
            obs += len(fk)
        meas.append(MeasTables(np.array([pos_of[int(fidx[j])] for j in keep], dtype=np.int64) + 1, p, mixes))
    T = BornTables(h_rx, shift, meas, (n, 1), tuple(spec.array.N_elements_axis),
                   tuple(spec.N_points_per_pitch_axis))
    return T, np.asarray(rows, dtype=np.int64)



# ---------------------------------------------------------------------------
# dense inputs over the whole FOV at a few frequencies
# ---------------------------------------------------------------------------
def frequency_tables(spec, f_sel: np.ndarray):
    """The operator restricted to the global frequencies `f_sel` (0-based),
    over every grid point, through the reference's own table layout: h_ref on
    the FOV with the shift/mirror map (setup_grid_symmetric.m:95-241) and
    p_inc from h_ref with the tx filter (sequence.m:825-840).  Every factor
    of a row of A -- transfer function, taper, prefactor, normal velocity --
    depends on that row's frequency alone, so these rows equal the full
    operator's.  p_inc is formed per frequency as the product of the normal
    velocities with the gathered element transfer functions.  Returns
    (T, rows): rows holds the 0-based indices of T's observations in the full
    observation vector."""
    f_sel = np.asarray(sorted(set(int(i) for i in f_sel)), dtype=np.int64)
    pos_of = {int(g): i for i, g in enumerate(f_sel)}
    f = spec.frequencies[f_sel]
    if mirror_map_defined(spec):
        shift = shift_lateral(spec)
        pos = None
    else:
        shift, pos = extended_shift(spec)
    h_rx = transfer_function(spec, f, 0, spec.filter_rx, positions=pos)
    same = (spec.filter_tx.kind, spec.filter_tx.param) == (spec.filter_rx.kind, spec.filter_rx.param)
    h_tx = h_rx if same else transfer_function(spec, f, 0, spec.filter_tx, positions=pos)
    pref_u = prefactors_unique(spec)
    keeps = []
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        keeps.append(np.array([j for j, g in enumerate(fidx) if int(g) in pos_of], dtype=np.int64))
    p_rows = [np.zeros((len(kp), spec.fov.N_points), dtype=np.complex128) for kp in keeps]
    vds = [normal_velocities(spec, m) for m in range(spec.N_measurements)]
    for i, g in enumerate(f_sel):
        He = h_tx[i][shift.T - 1]                      # (N_e, N_pts): element e's transfer function
        for m in range(spec.N_measurements):
            fidx = spec.measurement_frequencies(m)
            for r, j in enumerate(keeps[m]):
                if int(fidx[j]) != int(g):
                    continue
                v = np.zeros(spec.array.N_elements, dtype=np.complex128)
                v[np.asarray(spec.waves[m].indices_active)] = vds[m][j]
                p_rows[m][r] = v @ He
    meas, rows = [], []
    obs = 0
    for m in range(spec.N_measurements):
        fidx = spec.measurement_frequencies(m)
        sub_of = {int(j): r for r, j in enumerate(keeps[m])}
        mixes = []
        for mix in spec.mixes[m]:
            fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
            jj = [j for j in range(len(fk)) if int(fk[j]) in sub_of]
            rows.extend(obs + j for j in jj)
            fk_kept = np.array([fk[j] for j in jj], dtype=np.int64)
            pref = pref_u[fidx[fk_kept]][:, None] * np.asarray(mix.weights)[None, :]
            mixes.append(MixTables(np.array([sub_of[int(x)] for x in fk_kept], dtype=np.int64) + 1,
                                   np.asarray(mix.indices_active) + 1, pref.astype(np.complex128)))
            obs += len(fk)
        meas.append(MeasTables(np.array([pos_of[int(fidx[j])] for j in keeps[m]], dtype=np.int64) + 1,
                               p_rows[m], mixes))
    T = BornTables(h_rx, shift, meas, tuple(spec.fov.N_points_axis), tuple(spec.array.N_elements_axis),
                   tuple(spec.N_points_per_pitch_axis))
    return T, np.asarray(rows, dtype=np.int64)


def mirror_map_defined(spec) -> bool:
    """The mirror of setup_grid_symmetric.m:199-228 stays inside the FOV only
    when every lateral axis holds at least (N_elements - 1) pitches of
    points."""
    npp = spec.N_points_per_pitch_axis
    return all(spec.fov.N_points_axis[d] >= (spec.array.N_elements_axis[d] - 1) * npp[d]
               for d in range(spec.fov.N_dimensions - 1))


def extended_shift(spec):
    """Shift map into the FOV grid extended laterally by (N_elements_axis - 1)
    pitches before it, for FOVs narrower than the aperture: by translation
    invariance of the face integral and the taper, element e sees point
    (i_x, [i_y,] z) where element 1 sees (i_x - (e_x - 1) N_p, ...), which
    lies on the extended grid.  Returns the 1-based N_pts x N_e map and the
    positions of the extended grid (MATLAB order, x fastest)."""
    npp = spec.N_points_per_pitch_axis
    N_axis = list(spec.fov.N_points_axis)
    nd = len(N_axis)
    Ne_axis = list(spec.array.N_elements_axis)
    ext_axis = [N_axis[d] + (Ne_axis[d] - 1) * npp[d] for d in range(nd - 1)] + [N_axis[-1]]
    axes = []
    for d in range(nd):
        if d < nd - 1:
            j = np.arange(ext_axis[d]) - (Ne_axis[d] - 1) * npp[d]
            axes.append((j - (N_axis[d] - 1) / 2.0) * spec.fov.delta_axis[d])
        else:
            axes.append(spec.fov.z0 + np.arange(N_axis[d]) * spec.fov.delta_axis[d])
    mesh = np.meshgrid(*axes, indexing="ij")
    pos = np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)
    sub = np.stack(np.unravel_index(np.arange(spec.fov.N_points), N_axis, order="F"), axis=1)
    e_sub = np.stack(np.unravel_index(np.arange(spec.array.N_elements), Ne_axis, order="F"), axis=1)
    shift = np.empty((spec.fov.N_points, spec.array.N_elements), dtype=np.int64)
    for e in range(spec.array.N_elements):
        s = sub.copy()
        for d in range(nd - 1):
            s[:, d] = sub[:, d] - e_sub[e, d] * npp[d] + (Ne_axis[d] - 1) * npp[d]
        shift[:, e] = np.ravel_multi_index(tuple(s.T), ext_axis, order="F") + 1
    return shift, pos


def _rows_by_frequency(T: BornTables):
    """For each measurement m and each of its unique frequencies j: the list
    of (observation index, mix k, row i of the mix) reading it."""
    out = []
    obs = 0
    for ms in T.meas:
        per = [[] for _ in range(len(ms.idx_f_to_global))]
        for k, mx in enumerate(ms.mixes):
            for i, j1 in enumerate(mx.idx_f_to_meas):
                per[j1 - 1].append((obs + i, k, i))
            obs += len(mx.idx_f_to_meas)
        out.append(per)
    return out


def forward_by_frequency(T: BornTables, gamma: np.ndarray) -> np.ndarray:
    """u = A gamma over every grid point (no occupancy compaction), the loop
    nest of forward_quick_cpu (operator_born.m:353-459) reordered frequency
    first: for each global frequency, the element transfer functions H[e, p]
    = h_ref[f, shift[p, e]] are gathered once and V_m = H (p_inc_m .* gamma)
    holds every element's sum for measurement m; u(obs) = sum_a pref(i, a)
    V_m[e_a].  For dense inputs at full FOV size.  (N_obs, N_rhs)."""
    g = np.asarray(gamma, dtype=np.complex128)
    squeeze = g.ndim == 1
    if squeeze:
        g = g[:, None]
    out = np.zeros((T.N_observations(), g.shape[1]), dtype=np.complex128)
    by_f = _rows_by_frequency(T)
    for fg in range(T.h_ref.shape[0]):
        users = [(m, j) for m, ms in enumerate(T.meas) for j in range(len(ms.idx_f_to_global))
                 if ms.idx_f_to_global[j] - 1 == fg and by_f[m][j]]
        if not users:
            continue
        H = T.h_ref[fg][T.shift.T - 1]
        n_r = g.shape[1]
        B = np.concatenate([T.meas[m].p_incident[j][:, None] * g for m, j in users], axis=1)
        V_all = H @ B
        for c, (m, j) in enumerate(users):
            ms = T.meas[m]
            V = V_all[:, c * n_r:(c + 1) * n_r]
            for o, k, i in by_f[m][j]:
                mx = ms.mixes[k]
                out[o] = mx.prefactors[i] @ V[mx.idx_active - 1]
    return out[:, 0] if squeeze else out


def adjoint_by_frequency(T: BornTables, u: np.ndarray) -> np.ndarray:
    """gamma_hat = A^H u over every grid point, the loop nest of
    adjoint_quick_cpu (operator_born.m:511-593) reordered frequency first:
    W_m[e] = sum over the rows reading (m, f) of conj(pref(i, a)) u(obs) for
    e = e_a, then gamma_hat += conj(p_inc_m) .* (H^H W_m).  (N_pts, N_rhs)."""
    uu = np.asarray(u, dtype=np.complex128)
    squeeze = uu.ndim == 1
    if squeeze:
        uu = uu[:, None]
    out = np.zeros((T.N_points, uu.shape[1]), dtype=np.complex128)
    by_f = _rows_by_frequency(T)
    for fg in range(T.h_ref.shape[0]):
        users = [(m, j) for m, ms in enumerate(T.meas) for j in range(len(ms.idx_f_to_global))
                 if ms.idx_f_to_global[j] - 1 == fg and by_f[m][j]]
        if not users:
            continue
        H = T.h_ref[fg][T.shift.T - 1]
        n_r = uu.shape[1]
        W = np.zeros((T.N_elements, len(users) * n_r), dtype=np.complex128)
        for c, (m, j) in enumerate(users):
            ms = T.meas[m]
            for o, k, i in by_f[m][j]:
                mx = ms.mixes[k]
                np.add.at(W[:, c * n_r:(c + 1) * n_r], mx.idx_active - 1,
                          np.conj(mx.prefactors[i])[:, None] * uu[o][None, :])
        Z = (W.conj().T @ H).conj()                    # (users * N_rhs, N_pts) = (H^H W)^T
        for c, (m, j) in enumerate(users):
            out += np.conj(T.meas[m].p_incident[j])[:, None] * Z[c * n_r:(c + 1) * n_r].T
    return out[:, 0] if squeeze else out


def max_abs(x: np.ndarray, ref: np.ndarray) -> float:
    """Largest absolute deviation |x - ref| over all entries."""
    return float(np.max(np.abs(np.asarray(x) - np.asarray(ref)), initial=0.0))
