"""Measurement-sequence descriptions for the Born operator pair.

A `SequenceSpec` carries what `scattering.sequences.sequence` plus
`scattering.options` fix for the operator path: the transducer array
(+scattering/+sequences/+setups/+transducers/array_planar_regular.m:30-85),
the lossy homogeneous fluid (+materials/homogeneous_fluid.m:16-19,
+absorption_models/time_causal.m:58-84), the discretized field of view
(+geometry/orthotope.m:41-167), the global frequency list, the incident waves
(+syntheses/+deterministic/qpw.m:67-99, qsw.m:68-108, +random/rad.m:72-117),
the receive mixes (+settings/+controls/rx.m) and the anti-aliasing filters
(+scattering/+options/anti_aliasing.m:16-17).  All values are SI doubles.

The named workloads (`config0` ... `config4`) are the five synthetic
configurations this package is measured on.  Choices the reference leaves
open are fixed here once, and both the CPU restatement and the device table
build read them from this module:

* face integration: 4 points across the element width in 2-D, 4x4 in 3-D
  (grid_numbers.m:15 defaults to [4;53]);
* the first axial grid point sits at the stated depth, lateral axes are
  centred (orthotope.m:142-143 centring);
* excitation: Gaussian-windowed tone burst with a closed-form spectrum, unit
  transmit and receive impulse responses (delta_matrix.m:18-89);
* every delay is quantized to T_clk = 1/80 MHz (setup.m:22, :625);
* the time-causal model uses f_ref = 4 MHz, c_avg = c_ref, no constant term.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

T_CLK = 1.0 / 80e6


@dataclasses.dataclass
class LensSpec:
    """Acoustic lens on every element face (transducers/lens.m): lateral foci
    at the axial distances `axial_focus_axis`, one per face axis (inf = no
    focusing along that axis; +thickness/focus_axial.m), in a lossless lens
    material of sound speed `c_lens` (absorption_models/none.m:83).  The
    defaults are those of the L14-5/38 (parameters_L14_5_38.m:32-34)."""
    axial_focus_axis: Tuple[float, ...] = (math.inf, 16e-3)
    c_lens: float = 1540.0


@dataclasses.dataclass
class ArraySpec:
    """Regular planar orthogonal array; element centres are centred on the
    origin (array_planar_regular.m:63-66).  Every element carries the same
    face: `apodization` maps the face points' normalized positions (an
    (N_face, N_axes) array with entries in [-1, 1]) to real weights (None =
    uniform, +apodization/uniform.m), `lens` is an optional acoustic lens;
    together they give the complex face weights of face.m:359-376."""
    N_elements_axis: Tuple[int, ...]        # (N_x,) in 2-D, (N_x, N_y) in 3-D
    pitch_axis: Tuple[float, ...]
    width_axis: Tuple[float, ...]
    N_face_points_axis: Tuple[int, ...]     # integration points per face axis
    lens: Optional[LensSpec] = None
    apodization: Optional[Callable[[np.ndarray], np.ndarray]] = None

    @property
    def N_dimensions(self) -> int:
        return len(self.N_elements_axis)

    @property
    def N_elements(self) -> int:
        return int(np.prod(self.N_elements_axis))

    def positions_ctr(self) -> np.ndarray:
        """Element centres, MATLAB linear order (first axis fastest), shape
        (N_elements, N_dimensions)."""
        axes = [(np.arange(n) - (n - 1) / 2.0) * p
                for n, p in zip(self.N_elements_axis, self.pitch_axis)]
        mesh = np.meshgrid(*axes, indexing="ij")
        return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)

    def face_offsets(self) -> Tuple[np.ndarray, float]:
        """Face integration points relative to the element centre and the
        face cell measure (orthotope.m:104, :142-143; face.m:184-229)."""
        deltas = [w / n for w, n in zip(self.width_axis, self.N_face_points_axis)]
        axes = [(np.arange(n) - (n - 1) / 2.0) * d
                for n, d in zip(self.N_face_points_axis, deltas)]
        mesh = np.meshgrid(*axes, indexing="ij")
        pts = np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)
        return pts, float(np.prod(deltas))


@dataclasses.dataclass
class MediumSpec:
    """Homogeneous lossy fluid with the time-causal power-law model
    (time_causal.m:58-155)."""
    c_ref: float = 1500.0
    f_ref: float = 4e6
    absorption_constant: float = 0.0       # dB/cm
    absorption: float = 0.5                # dB/(MHz^y cm)
    exponent: float = 1.0
    dispersion: bool = True
    c_avg: float = 1500.0


@dataclasses.dataclass
class FovSpec:
    """Orthogonal regular grid: lateral axes centred, axial axis starting at
    `z0` (first grid point)."""
    N_points_axis: Tuple[int, ...]          # (N_x, N_z) or (N_x, N_y, N_z)
    delta_axis: Tuple[float, ...]
    z0: float

    @property
    def N_dimensions(self) -> int:
        return len(self.N_points_axis)

    @property
    def N_points(self) -> int:
        return int(np.prod(self.N_points_axis))

    @property
    def cell_volume(self) -> float:
        return float(np.prod(self.delta_axis))

    def axes(self) -> List[np.ndarray]:
        out = []
        for d, (n, dl) in enumerate(zip(self.N_points_axis, self.delta_axis)):
            if d < self.N_dimensions - 1:
                out.append((np.arange(n) - (n - 1) / 2.0) * dl)
            else:
                out.append(self.z0 + np.arange(n) * dl)
        return out

    def positions(self) -> np.ndarray:
        """Grid points in MATLAB linear order (grid_regular.m:191, :250)."""
        mesh = np.meshgrid(*self.axes(), indexing="ij")
        return np.stack([m.reshape(-1, order="F") for m in mesh], axis=1)


@dataclasses.dataclass
class FilterSpec:
    """Spatial anti-aliasing filter (+scattering/+anti_aliasing_filters)."""
    kind: str = "raised_cosine"            # off | raised_cosine | logistic | boxcar
    param: float = 0.5                     # roll-off factor or growth rate

    KINDS = ("off", "raised_cosine", "logistic", "boxcar")

    @property
    def code(self) -> int:
        return self.KINDS.index(self.kind)


@dataclasses.dataclass
class WaveSpec:
    """One incident wave: per-element apodization and (unquantized) delays
    of the active transmit elements."""
    indices_active: np.ndarray             # 0-based element indices
    apodization: np.ndarray
    delays: np.ndarray                     # seconds, before T_clk quantization
    label: str = ""


@dataclasses.dataclass
class MixSpec:
    """One mixed receive signal: active elements and their weights
    (the rx impulse responses are weighted deltas)."""
    indices_active: np.ndarray             # 0-based element indices
    weights: np.ndarray                    # complex or real, per active element
    indices_f: Optional[np.ndarray] = None  # 0-based into measurement freqs (None = all)


@dataclasses.dataclass
class SequenceSpec:
    name: str
    array: ArraySpec
    medium: MediumSpec
    fov: FovSpec
    frequencies: np.ndarray                # global unique frequency list (Hz)
    waves: List[WaveSpec]
    mixes: List[List[MixSpec]]             # per measurement
    meas_f_indices: Optional[List[np.ndarray]] = None  # per measurement, 0-based into frequencies
    pulse_fc: float = 4e6
    pulse_cycles: float = 2.0
    filter_tx: FilterSpec = dataclasses.field(default_factory=FilterSpec)
    filter_rx: FilterSpec = dataclasses.field(default_factory=FilterSpec)
    seed: int = 0

    @property
    def N_measurements(self) -> int:
        return len(self.waves)

    def measurement_frequencies(self, m: int) -> np.ndarray:
        if self.meas_f_indices is None:
            return np.arange(len(self.frequencies))
        return np.asarray(self.meas_f_indices[m])

    @property
    def N_points_per_pitch_axis(self) -> Tuple[int, ...]:
        """issymmetric: lateral spacing must divide the pitch (setup.m:1155-1160)."""
        out = []
        for d in range(self.array.N_dimensions):
            r = self.array.pitch_axis[d] / self.fov.delta_axis[d]
            n = int(round(r))
            if abs(r - n) > 1e-9 * max(1, n):
                raise ValueError("lateral spacing is not an integer fraction of the pitch")
            out.append(n)
        return tuple(out)

    def excitation_spectrum(self, f: np.ndarray) -> np.ndarray:
        """Closed-form continuous Fourier transform (e^{-j2 pi f t}) of a
        zero-phase Gaussian-windowed cosine burst whose envelope FWHM spans
        `pulse_cycles` periods of `pulse_fc`."""
        f = np.asarray(f, dtype=np.float64)
        fwhm = self.pulse_cycles / self.pulse_fc
        sigma = fwhm / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        a = sigma * math.sqrt(2.0 * math.pi) / 2.0
        g = lambda x: np.exp(-2.0 * (math.pi * sigma * x) ** 2)
        return (a * (g(f - self.pulse_fc) + g(f + self.pulse_fc))).astype(np.complex128)

    def N_observations(self) -> int:
        n = 0
        for m in range(self.N_measurements):
            nf = len(self.measurement_frequencies(m))
            for mix in self.mixes[m]:
                n += nf if mix.indices_f is None else len(mix.indices_f)
        return n


# ---------------------------------------------------------------------------
# wave syntheses (delays before quantization)
# ---------------------------------------------------------------------------
def qpw_wave(array: ArraySpec, theta: float, c_avg: float, phi: float = 0.0) -> WaveSpec:
    """Quasi-plane wave steered by theta from the axial axis (qpw.m:67-99):
    tau_e = <x_e, e_theta_lateral>/c_avg, shifted to start at zero."""
    pos = array.positions_ctr()
    if array.N_dimensions == 1:
        e_lat = np.array([math.sin(theta)])
    else:
        e_lat = np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi)])
    tau = pos @ e_lat / c_avg
    tau = tau - tau.min()
    n = array.N_elements
    return WaveSpec(np.arange(n), np.ones(n), tau, label=f"qpw({math.degrees(theta):.1f})")


def qsw_wave(array: ArraySpec, position_src: np.ndarray, c_avg: float,
             angles: Optional[np.ndarray] = None) -> WaveSpec:
    """Quasi-spherical wave from a virtual source (qsw.m:68-108): active
    elements lie inside the opening angles; tau = distance/c_avg, shifted."""
    pos = array.positions_ctr()
    pos3 = np.concatenate([pos, np.zeros((pos.shape[0], 1))], axis=1)
    vec = pos3 - np.asarray(position_src)[None, :]
    dist = np.linalg.norm(vec, axis=1)
    if angles is None:
        angles = np.full(array.N_dimensions, math.pi)
    ok = np.ones(len(dist), dtype=bool)
    far = dist >= np.finfo(float).eps
    ang = np.arcsin(np.abs(vec[far, :array.N_dimensions] / dist[far, None]))
    ok[far] = np.all(ang <= np.asarray(angles) / 2, axis=1)
    idx = np.nonzero(ok)[0]
    tau = dist[idx] / c_avg
    tau = tau - tau.min()
    return WaveSpec(idx, np.ones(len(idx)), tau, label="qsw")


def random_wave(array: ArraySpec, rng: np.random.Generator, t_max: float) -> WaveSpec:
    """The config-2 waves: random apodization +-1 (Bernoulli 0.5, as ra.m and
    rad.m draw it) and delays drawn U[0, t_max].  The toolbox's own rad
    synthesis permutes a linear delay ramp instead (rad_wave)."""
    n = array.N_elements
    apod = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    tau = rng.uniform(0.0, t_max, n)
    return WaveSpec(np.arange(n), apod, tau, label="random")


def _random_apodization(n: int, rng: np.random.Generator) -> np.ndarray:
    # rand >= 0.5 -> +1, else -1 (ra.m, rad.m compute_delays_scalar)
    return np.where(rng.random(n) >= 0.5, 1.0, -1.0)


def _permuted_ramp(array: ArraySpec, rng: np.random.Generator, e_theta_lateral, c_avg: float) -> np.ndarray:
    # t_shift_max = sum over lateral axes with more than one element of
    # (N_axis - 1) * pitch_axis * |e_theta component| / c_avg; the delays are a
    # random permutation of k * t_shift_max / (N_e - 1), k = 0..N_e-1
    # (rd.m, rad.m compute_delays_scalar)
    e = np.abs(np.atleast_1d(np.asarray(e_theta_lateral, dtype=np.float64)))
    t_max = sum((na - 1) * pa * e[d] for d, (na, pa) in enumerate(zip(array.N_elements_axis, array.pitch_axis))
                if na > 1) / c_avg
    n = array.N_elements
    t_inc = t_max / (n - 1) if n > 1 else 0.0
    return rng.permutation(n).astype(np.float64) * t_inc


def ra_wave(array: ArraySpec, rng: np.random.Generator) -> WaveSpec:
    """Random apodization, no delays (syntheses/random/ra.m)."""
    n = array.N_elements
    return WaveSpec(np.arange(n), _random_apodization(n, rng), np.zeros(n), label="ra")


def rd_wave(array: ArraySpec, rng: np.random.Generator, e_theta_lateral, c_avg: float) -> WaveSpec:
    """Unit apodization, randomly permuted delay ramp (syntheses/random/rd.m);
    e_theta_lateral: the lateral components of the steering direction."""
    n = array.N_elements
    return WaveSpec(np.arange(n), np.ones(n), _permuted_ramp(array, rng, e_theta_lateral, c_avg), label="rd")


def rad_wave(array: ArraySpec, rng_apo: np.random.Generator, rng_del: np.random.Generator, e_theta_lateral,
             c_avg: float) -> WaveSpec:
    """Random apodization and randomly permuted delay ramp, from two
    generators (syntheses/random/rad.m)."""
    n = array.N_elements
    return WaveSpec(np.arange(n), _random_apodization(n, rng_apo),
                    _permuted_ramp(array, rng_del, e_theta_lateral, c_avg), label="rad")


def identity_mixes(N_elements: int) -> List[MixSpec]:
    return [MixSpec(np.array([e]), np.array([1.0])) for e in range(N_elements)]


# ---------------------------------------------------------------------------
# the five workloads
# ---------------------------------------------------------------------------
PITCH_2D = 0.3048e-3
WIDTH_2D = 0.27e-3


def config0() -> SequenceSpec:
    """64-element array, 256x256 FOV at pitch/4 from 2 mm, 65 bins 2-6 MHz,
    one 0-degree QPW, identity receive (4,160 observations)."""
    arr = ArraySpec((64,), (PITCH_2D,), (WIDTH_2D,), (4,))
    med = MediumSpec()
    fov = FovSpec((256, 256), (PITCH_2D / 4, PITCH_2D / 4), 2e-3)
    f = 2.0e6 + 62.5e3 * np.arange(65)
    waves = [qpw_wave(arr, 0.0, med.c_avg)]
    return SequenceSpec("config0", arr, med, fov, f, waves, [identity_mixes(64)], seed=1)


def config1() -> SequenceSpec:
    """128-element array, 512x512 FOV at pitch/4 from 2 mm, 257 bins 2-10 MHz,
    11 QPWs at -10:2:10 degrees, identity receive (361,856 observations)."""
    arr = ArraySpec((128,), (PITCH_2D,), (WIDTH_2D,), (4,))
    med = MediumSpec()
    fov = FovSpec((512, 512), (PITCH_2D / 4, PITCH_2D / 4), 2e-3)
    f = 2.0e6 + 31.25e3 * np.arange(257)
    waves = [qpw_wave(arr, math.radians(a), med.c_avg) for a in range(-10, 11, 2)]
    mixes = [identity_mixes(128) for _ in waves]
    return SequenceSpec("config1", arr, med, fov, f, waves, mixes, seed=2)


def config2() -> SequenceSpec:
    """Config-1 geometry, 32 random waves (+-1 apodization, delays U[0,5 us]),
    16 receive mixes of all 128 elements with random +-1 weights."""
    base = config1()
    rng = np.random.Generator(np.random.PCG64(3))
    waves = [random_wave(base.array, rng, 5e-6) for _ in range(32)]
    mixes = []
    for _ in waves:
        mm = []
        for _k in range(16):
            w = np.where(rng.random(128) < 0.5, -1.0, 1.0)
            mm.append(MixSpec(np.arange(128), w))
        mixes.append(mm)
    return SequenceSpec("config2", base.array, base.medium, base.fov, base.frequencies,
                        waves, mixes, seed=3)


def config3_instances(n: int = 24) -> List[SequenceSpec]:
    """24 uneven 2-D instances: FOV side in {128,256,384,512} at pitch/4,
    33-257 bins, 1-11 plane waves, the config-1 array."""
    rng = np.random.Generator(np.random.PCG64(4))
    arr = ArraySpec((128,), (PITCH_2D,), (WIDTH_2D,), (4,))
    med = MediumSpec()
    out = []
    for i in range(n):
        side = int(rng.choice([128, 256, 384, 512]))
        nf = int(rng.integers(33, 258))
        npw = int(rng.integers(1, 12))
        fov = FovSpec((side, side), (PITCH_2D / 4, PITCH_2D / 4), 2e-3)
        f = 2.0e6 + (8.0e6 / max(nf - 1, 1)) * np.arange(nf)
        angles = np.linspace(-10, 10, npw) if npw > 1 else np.array([0.0])
        waves = [qpw_wave(arr, math.radians(a), med.c_avg) for a in angles]
        out.append(SequenceSpec(f"config3_{i}", arr, med, fov, f, waves,
                                [identity_mixes(128) for _ in waves], seed=4))
    return out


def config4() -> SequenceSpec:
    """32x32 matrix array (pitch 0.3 mm), 128x128x192 FOV at pitch/2 from
    3 mm, 65 bins 2.5-5.5 MHz, 5 QSWs from 10 mm behind the array."""
    arr = ArraySpec((32, 32), (0.3e-3, 0.3e-3), (0.27e-3, 0.27e-3), (4, 4))
    med = MediumSpec()
    fov = FovSpec((128, 128, 192), (0.15e-3, 0.15e-3, 0.15e-3), 3e-3)
    f = np.linspace(2.5e6, 5.5e6, 65)
    waves = [qsw_wave(arr, np.array([x, 0.0, -10e-3]), med.c_avg)
             for x in (0.0, -3e-3, 3e-3, -6e-3, 6e-3)]
    mixes = [identity_mixes(arr.N_elements) for _ in waves]
    return SequenceSpec("config4", arr, med, fov, f, waves, mixes, seed=5)


CONFIGS = {"config0": config0, "config1": config1, "config2": config2, "config4": config4}


# ---------------------------------------------------------------------------
# synthetic phantoms
# ---------------------------------------------------------------------------
def phantom_points(spec: SequenceSpec, n_points: int = 30, seed: Optional[int] = None) -> np.ndarray:
    """Zero except `n_points` distinct uniformly-drawn grid points with N(0,1)
    real amplitudes."""
    rng = np.random.Generator(np.random.PCG64(spec.seed if seed is None else seed))
    g = np.zeros(spec.fov.N_points)
    idx = rng.choice(spec.fov.N_points, size=n_points, replace=False)
    g[idx] = rng.standard_normal(n_points)
    return g


def phantom_speckle(spec: SequenceSpec, n_disks: int = 6, radius: float = 2e-3,
                    n_bright: int = 20, amp_bright: float = 10.0,
                    seed: Optional[int] = None) -> np.ndarray:
    """i.i.d. N(0,1) speckle scaled by 0.05 inside `n_disks` anechoic disks
    (spheres in 3-D) at random positions, plus `n_bright` points of amplitude
    `amp_bright`."""
    rng = np.random.Generator(np.random.PCG64(spec.seed if seed is None else seed))
    pos = spec.fov.positions()
    g = rng.standard_normal(spec.fov.N_points)
    lo = pos.min(axis=0)
    hi = pos.max(axis=0)
    for _ in range(n_disks):
        c = lo + rng.random(pos.shape[1]) * (hi - lo)
        inside = np.sum((pos - c) ** 2, axis=1) <= radius ** 2
        g[inside] *= 0.05
    if n_bright:
        idx = rng.choice(spec.fov.N_points, size=n_bright, replace=False)
        g[idx] = amp_bright
    return g


def phantom(spec: SequenceSpec) -> np.ndarray:
    """The phantom each named workload is stated with: config 0, 30 point
    scatterers; configs 1 and 2 (and config-3 instances), speckle with 6
    anechoic disks of radius 2 mm and 20 bright points of amplitude 10;
    config 4, speckle with 4 anechoic spheres of radius 2 mm and no bright
    points."""
    if spec.name == "config0":
        return phantom_points(spec, 30)
    if spec.name == "config4":
        return phantom_speckle(spec, n_disks=4, n_bright=0)
    return phantom_speckle(spec)
