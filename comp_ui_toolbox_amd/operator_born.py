"""Host-side mirror of scattering.operator_born's operator path.

`OperatorBorn` owns a libborn plan (device-resident tables).  It is built
either from the tables an operator_born object holds -- exactly what the MEX
reads (combined_quick_gpu.cu:201-391, :817-823) -- or from a `SequenceSpec`,
in which case the transfer function and the incident fields are built on the
device.  `combined_quick_gpu` keeps the MEX's calling convention and error
identifiers (combined_quick_gpu.m:11-23, cu:201-441); `op_A` is the solver
contract of operator.combined_scalar (operator.m:1019-1072).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import BornError, check, dptr
from . import physics


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128_colmajor(mat) -> np.ndarray:
    """Column-major memory of a 2-D complex matrix as an interleaved float64
    buffer (MATLAB's layout)."""
    m = np.asarray(mat, dtype=np.complex128)
    if m.ndim == 1:
        m = m[:, None]
    return np.ascontiguousarray(m.T).view(np.float64)


class _DescKeeper:
    """Keeps the numpy buffers a descriptor points at alive for one call."""

    def __init__(self):
        self.refs = []

    def keep(self, a):
        self.refs.append(a)
        return a


def _measurement_descs(keeper, meas_list, with_p_inc: bool):
    """Build born_measurement_desc[] from objects with the attributes of the
    operator's tables: idx_f_to_global (1-based), p_incident (N_f_meas x
    N_pts), mixes[k].idx_f_to_meas / idx_active (1-based) / prefactors."""
    M = (_lib.MeasDesc * len(meas_list))()
    for m, ms in enumerate(meas_list):
        mixes = (_lib.MixDesc * len(ms.mixes))()
        for k, mx in enumerate(ms.mixes):
            fi = keeper.keep(_f64(mx.idx_f_to_meas))
            ai = keeper.keep(_f64(mx.idx_active))
            pf = keeper.keep(_c128_colmajor(mx.prefactors))
            mixes[k] = _lib.MixDesc(len(fi), dptr(fi), len(ai), dptr(ai), dptr(pf))
        keeper.keep(mixes)
        fg = keeper.keep(_f64(ms.idx_f_to_global))
        p = keeper.keep(_c128_colmajor(ms.p_incident)) if with_p_inc else None
        M[m] = _lib.MeasDesc(len(fg), dptr(fg), dptr(p), len(ms.mixes), mixes)
    keeper.keep(M)
    return M


class _Mix:
    def __init__(self, idx_f_to_meas, idx_active, prefactors):
        self.idx_f_to_meas = idx_f_to_meas
        self.idx_active = idx_active
        self.prefactors = prefactors


class _Meas:
    def __init__(self, idx_f_to_global, mixes):
        self.idx_f_to_global = idx_f_to_global
        self.p_incident = None
        self.mixes = mixes


def _min_gap(f: np.ndarray) -> float:
    g = np.diff(np.sort(f))
    g = g[g > 0]
    return float(g.min()) if len(g) else float(f.min())


class OperatorBorn:
    """Born scattering operator A (N_obs x N_pts) on one MI355X."""

    def __init__(self, plan: int, n_obs: int, n_pts: int, device: int, meta: Optional[dict] = None):
        self._plan = C.c_void_p(plan)
        self.N_observations = n_obs
        self.N_points = n_pts
        self.device = device
        self.meta = meta or {}

    # ---------------------------------------------------------- creation
    @classmethod
    def from_tables(cls, T, device: int = 0, precision: str = "fp32") -> "OperatorBorn":
        """Plan from an operator's tables: T.h_ref (N_f_unique x N_pts),
        T.shift (N_pts x N_e, 1-based), T.meas[m] (idx_f_to_global,
        p_incident, mixes), T.N_points_axis, T.N_elements_axis."""
        L = _lib.lib()
        keeper = _DescKeeper()
        h = keeper.keep(_c128_colmajor(T.h_ref))
        sh = keeper.keep(np.ascontiguousarray(np.asarray(T.shift, dtype=np.float64).T))
        M = _measurement_descs(keeper, T.meas, True)
        npa = list(T.N_points_axis) + [0] * (3 - len(T.N_points_axis))
        nea = list(T.N_elements_axis) + [1] * (2 - len(T.N_elements_axis))
        n_pts, n_e = np.asarray(T.shift).shape
        d = _lib.TablesDesc(len(T.N_points_axis), (C.c_int32 * 3)(*npa), (C.c_int32 * 2)(*nea),
                            n_pts, n_e, np.asarray(T.h_ref).shape[0], dptr(h), dptr(sh.reshape(-1)),
                            len(T.meas), M, device, _lib.PRECISIONS[precision])
        plan = C.c_void_p()
        check(L.born_plan_create_tables(C.byref(d), C.byref(plan)))
        return cls._wrap(plan, device, {"rows_h": np.asarray(T.h_ref).shape[0],
                                        "rows_p": [len(ms.idx_f_to_global) for ms in T.meas],
                                        "n_mix_signals": sum(len(ms.mixes) for ms in T.meas),
                                        "meas_n_mix": [len(ms.mixes) for ms in T.meas],
                                        "mix_n_obs": [len(mx.idx_f_to_meas) for ms in T.meas for mx in ms.mixes]})

    @classmethod
    def from_sequence(cls, spec, device: int = 0, precision: str = "fp32") -> "OperatorBorn":
        """Plan from a SequenceSpec: h_ref and every p_incident are built on
        the device (FP64 evaluation, FP32 storage)."""
        L = _lib.lib()
        keeper = _DescKeeper()
        meas = []
        for m in range(spec.N_measurements):
            fidx = spec.measurement_frequencies(m)
            mixes = []
            for k, mix in enumerate(spec.mixes[m]):
                fk = np.arange(len(fidx)) if mix.indices_f is None else np.asarray(mix.indices_f)
                mixes.append(_Mix(fk + 1, np.asarray(mix.indices_active) + 1, physics.prefactors_mix(spec, m, k)))
            meas.append(_Meas(fidx + 1, mixes))
        M = _measurement_descs(keeper, meas, False)
        f = keeper.keep(_f64(spec.frequencies))
        k = keeper.keep(np.ascontiguousarray(physics.wavenumbers(spec.frequencies, spec.medium)).view(np.float64))
        nd = spec.fov.N_dimensions
        npa = list(spec.fov.N_points_axis) + [0] * (3 - nd)
        dla = list(spec.fov.delta_axis) + [0.0] * (3 - nd)
        nea = list(spec.array.N_elements_axis) + [1] * (2 - spec.array.N_dimensions)
        pitch = list(spec.array.pitch_axis) + [0.0] * (2 - spec.array.N_dimensions)
        nfp = list(spec.array.N_face_points_axis) + [1] * (2 - spec.array.N_dimensions)
        fdl = [w / n for w, n in zip(spec.array.width_axis, spec.array.N_face_points_axis)] + [0.0] * (2 - spec.array.N_dimensions)
        fw = physics.face_weights(spec, spec.frequencies)
        fw_ptr = None if fw is None else dptr(keeper.keep(_c128_colmajor(fw)))
        ntx = (C.c_int32 * spec.N_measurements)()
        txi = (C.POINTER(C.c_double) * spec.N_measurements)()
        vd = (C.POINTER(C.c_double) * spec.N_measurements)()
        for m in range(spec.N_measurements):
            idx = keeper.keep(_f64(np.asarray(spec.waves[m].indices_active) + 1))
            v = keeper.keep(_c128_colmajor(physics.normal_velocities(spec, m)))
            ntx[m] = len(idx)
            txi[m] = dptr(idx)
            vd[m] = dptr(v)
        d = _lib.GeometryDesc(
            nd, (C.c_int32 * 3)(*npa), (C.c_double * 3)(*dla), spec.fov.z0,
            (C.c_int32 * 2)(*nea), (C.c_double * 2)(*pitch), (C.c_int32 * 2)(*nfp), (C.c_double * 2)(*fdl),
            fw_ptr, len(f), dptr(f), dptr(k),
            _lib.FILTER_CODES[spec.filter_rx.kind], _lib.FILTER_CODES[spec.filter_tx.kind],
            spec.filter_rx.param, spec.filter_tx.param,
            spec.N_measurements, M, ntx, txi, vd, device, _lib.PRECISIONS[precision])
        plan = C.c_void_p()
        check(L.born_plan_create_geometry(C.byref(d), C.byref(plan)))
        return cls._wrap(plan, device, {"spec": spec.name, "rows_h": len(spec.frequencies),
                                        "rows_p": [len(spec.measurement_frequencies(m))
                                                   for m in range(spec.N_measurements)],
                                        "n_mix_signals": sum(len(mm) for mm in spec.mixes),
                                        "meas_n_mix": [len(mm) for mm in spec.mixes],
                                        "mix_n_obs": [len(mx.idx_f_to_meas) for ms in meas for mx in ms.mixes]})

    @classmethod
    def _wrap(cls, plan, device, meta=None):
        L = _lib.lib()
        n_obs, n_pts = C.c_int64(), C.c_int64()
        check(L.born_plan_size(plan, C.byref(n_obs), C.byref(n_pts)))
        return cls(plan.value, n_obs.value, n_pts.value, device, meta)

    def close(self):
        if self._plan is not None and self._plan.value:
            check(_lib.lib().born_plan_destroy(self._plan))
        self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- queries
    @property
    def size(self):
        """[N_observations, N_points] (op_A mode 0, operator.m:1035-1046)."""
        return (self.N_observations, self.N_points)

    @property
    def path(self) -> str:
        return {0: "structured", 1: "general"}.get(_lib.lib().born_plan_path(self._plan), "invalid")

    def set_sparse_forward(self, max_points: int = -1):
        """Host forward with few occupied points as gathers (the MEX's
        occupancy compaction): -1 by estimated cost, 0 never, k > 0 whenever
        at most k points are occupied."""
        check(_lib.lib().born_plan_set_sparse_forward(self._plan, int(max_points)))

    def force_general(self, on: bool = True):
        check(_lib.lib().born_plan_force_general(self._plan, int(on)))

    def set_timing(self, on: bool = True):
        """Record HIP events around every contraction launch (on its stream)."""
        check(_lib.lib().born_plan_set_timing(self._plan, int(on)))

    def get_timing(self):
        """(ms, launches) summed per contraction kernel since the last call:
        {'forward': (ms, n), 'adjoint': (ms, n)}."""
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        check(_lib.lib().born_plan_get_timing(self._plan, ms, n))
        return {"forward": (ms[0], n[0]), "adjoint": (ms[1], n[1])}

    def set_precision(self, precision: str):
        check(_lib.lib().born_plan_set_precision(self._plan, _lib.PRECISIONS[precision]))

    def table(self, which: str, index: int = 0) -> np.ndarray:
        """Device copy (complex64) of h_ref ('h_ref', N_f_unique x N_h) or of
        p_incident of measurement `index` ('p_incident', N_f_meas x N_pts).
        N_h = N_pts, or the laterally extended table grid when the FOV is
        narrower than the aperture."""
        w = {"h_ref": 0, "p_incident": 1}[which]
        rows, cols = C.c_int64(), C.c_int64()
        check(_lib.lib().born_plan_table_shape(self._plan, w, index, C.byref(rows), C.byref(cols)))
        cap = rows.value * cols.value
        buf = np.empty(cap, dtype=np.complex64)
        check(_lib.lib().born_plan_get_table(self._plan, w, index, buf.ctypes.data_as(C.c_void_p), cap))
        return buf.reshape(cols.value, rows.value).T.copy()

    def energy_rx(self, measurement: int = 0, tgc: bool = False) -> np.ndarray:
        """Received energies of one measurement, canonical basis
        (energy_rx_canonical, operator_born.m:718-859): per grid point the
        squared norm of its column of the measurement's rows of A, or of
        TGC applied to them when tgc=True (kernels of set_tgc, :836-843).
        `measurement` is 0-based among the plan's measurements.  (N_pts,)
        float64."""
        out = np.empty(self.N_points, dtype=np.float64)
        fn = _lib.lib().born_energy_rx_tgc if tgc else _lib.lib().born_energy_rx
        check(fn(self._plan, int(measurement), dptr(out)))
        return out

    def energy_rx_dev(self, measurement: int = 0, out=None, stream=None, tgc: bool = False):
        """Device variant: float64 tensor of N_pts energies on the plan's device."""
        import torch
        if out is None:
            out = torch.empty(self.N_points, dtype=torch.float64, device=f"cuda:{self.device}")
        if out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous() or out.numel() != self.N_points:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix", "expected a contiguous float64 device tensor")
        s = torch.cuda.current_stream(out.device).cuda_stream if stream is None else stream
        fn = _lib.lib().born_energy_rx_tgc_dev if tgc else _lib.lib().born_energy_rx_dev
        check(fn(self._plan, int(measurement), C.c_void_p(out.data_ptr()), C.c_void_p(s)))
        return out

    def measurement_rows(self, measurement: int) -> range:
        """Observation indices of one measurement (0-based among the plan's
        measurements): observations are measurement-major (cu:773, :978)."""
        n_mix = self.meta["meas_n_mix"]
        if not 0 <= measurement < len(n_mix):
            raise BornError(10, "born:InvalidArgument", "measurement index out of range")
        first = sum(n_mix[:measurement])
        obs = self.meta["mix_n_obs"]
        start = sum(obs[:first])
        return range(start, start + sum(obs[first:first + n_mix[measurement]]))

    def energy_rx_arbitrary(self, dictionary, measurements=None, tgc: bool = False, batch: int = 64):
        """Received energies of the coefficients of a dictionary
        (energy_rx_arbitrary, operator_born.m:606-713): for every coefficient
        j, E[j] = ||TGC(A Psi^T e_j)||^2 over the rows of the measurements,
        summed over the measurements as energy_rx_scalar does (:136-151; all
        of the plan's by default).  As in the reference, the unit coefficient
        vectors go through the forward in batches (of `batch` right-hand sides
        here; N_objects = 1024 there, :609): synthesis by
        dictionary.adjoint_transform, the device forward with one right-hand
        side per atom, the plan's TGC when `tgc` (operator_born.m:222-236),
        and the squared norms of the measurements' rows, accumulated in FP64
        on the device.  `dictionary` needs adjoint_transform on device
        tensors of shape (N_pts, n) (e.g. wavelet.D4Dictionary) and
        N_coefficients = N_points.  (N_points,) float64 numpy array."""
        import torch
        if measurements is None:
            measurements = range(len(self.meta["meas_n_mix"]))
        rows = torch.tensor(np.concatenate([np.arange(r.start, r.stop) for r in map(self.measurement_rows, measurements)]),
                            dtype=torch.int64, device=f"cuda:{self.device}")
        n = self.N_points
        E = torch.empty(n, dtype=torch.float64, device=f"cuda:{self.device}")
        for j0 in range(0, n, batch):
            nb = min(batch, n - j0)
            theta = torch.zeros((n, nb), dtype=torch.float64, device=f"cuda:{self.device}")
            theta[torch.arange(j0, j0 + nb, device=theta.device), torch.arange(nb, device=theta.device)] = 1.0
            atoms = dictionary.adjoint_transform(theta)                 # (N_pts, nb)
            g = atoms.T.to(torch.complex64).contiguous()
            u = self.forward_dev(g)
            if tgc:
                u = self.tgc_dev(u, mode=1)
            E[j0:j0 + nb] = (u.index_select(1, rows).to(torch.complex128).abs() ** 2).sum(dim=1)
        return E.cpu().numpy()

    def time_signals(self, u, frequencies, delta_f: float = 0.0, delta_t: float = 0.0, stream=None):
        """Time-domain mixed voltage signals (format_voltages + signal,
        operator.m:900-1007, signal_matrix.m:293-397): u (N_rhs, N_obs)
        complex64 device tensor -> float32 tensor (N_rhs, N_mix_signals, N_t)
        and N_t.  `frequencies`: the plan's global unique frequencies (Hz)."""
        import torch
        uu = u.reshape(1, -1) if u.dim() == 1 else u
        self._check_dev(uu, self.N_observations)
        f = np.ascontiguousarray(np.asarray(frequencies, dtype=np.float64))
        n_t = C.c_int32()
        # size query: N_t is at most 2 * max(q) for the default, or 1 / (delta_f delta_t)
        df = delta_f if delta_f > 0 else _min_gap(f)
        nt_max = int(round(1.0 / (df * delta_t))) if delta_t > 0 else 2 * int(round(f.max() / df))
        n_sig = self._n_mix_signals()
        out = torch.empty((uu.shape[0], n_sig, max(nt_max, 1)), dtype=torch.float32, device=uu.device)
        s = torch.cuda.current_stream(uu.device).cuda_stream if stream is None else stream
        check(_lib.lib().born_time_signals(self._plan, dptr(f), float(delta_f), float(delta_t),
                                           C.c_void_p(uu.data_ptr()), uu.shape[0], C.c_void_p(out.data_ptr()),
                                           out.numel(), C.byref(n_t), C.c_void_p(s)))
        return out[:, :, : n_t.value], n_t.value

    def _n_mix_signals(self) -> int:
        return int(self.meta["n_mix_signals"])

    # ---------------------------------------------------------- columns
    def columns(self, points) -> np.ndarray:
        """A[:, points] (0-based grid points), complex128 (N_obs, n): the
        forward of the unit vectors at those points, gathered directly
        (born_columns)."""
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1) + 1.0)
        out = np.empty((len(pts), self.N_observations), dtype=np.complex128)
        check(_lib.lib().born_columns(self._plan, dptr(pts), len(pts), dptr(out.view(np.float64))))
        return np.ascontiguousarray(out.T)

    def columns_dev(self, points, out=None, stream=None):
        """Device variant: complex64 tensor (n, N_obs), row j = A[:, points[j]]."""
        import torch
        pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1) + 1.0)
        if out is None:
            out = torch.empty((len(pts), self.N_observations), dtype=torch.complex64, device=f"cuda:{self.device}")
        self._check_dev(out, self.N_observations)
        if out.numel() != len(pts) * self.N_observations:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix", "output must hold n x N_obs entries")
        s = torch.cuda.current_stream(out.device).cuda_stream if stream is None else stream
        check(_lib.lib().born_columns_dev(self._plan, dptr(pts), len(pts), C.c_void_p(out.data_ptr()), C.c_void_p(s)))
        return out

    # ---------------------------------------------------------- TGC
    def set_tgc(self, kernels=None):
        """Time gain compensation applied by op_A(..., tgc=True): one odd-length
        complex kernel per mix signal (measurement-major), e.g. from
        tgc.exponential_kernels; None clears it (regularization.tgc.off)."""
        L = _lib.lib()
        if kernels is None:
            check(L.born_tgc_set(self._plan, None, None))
            return
        if len(kernels) != self._n_mix_signals():
            raise BornError(10, "born:InvalidArgument",
                            f"expected {self._n_mix_signals()} TGC kernels, one per mix signal")
        half = (C.c_int32 * len(kernels))()
        for i, k in enumerate(kernels):
            n = np.asarray(k).size
            if n % 2 != 1:
                raise BornError(10, "born:InvalidArgument", "TGC kernels must have odd lengths 2 M + 1")
            half[i] = (n - 1) // 2
        taps = np.ascontiguousarray(np.concatenate([np.asarray(k, dtype=np.complex128).reshape(-1)
                                                    for k in kernels]))
        check(L.born_tgc_set(self._plan, half, dptr(taps.view(np.float64))))

    def _tgc_host(self, u, mode: int) -> np.ndarray:
        uu = np.asarray(u)
        squeeze = uu.ndim == 1
        u2 = uu[:, None] if squeeze else uu
        if u2.shape[0] != self.N_observations:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix",
                            f"input must be a double matrix with {self.N_observations} rows!")
        uin = _c128_colmajor(u2)
        out = np.empty((u2.shape[1], self.N_observations), dtype=np.complex128)
        check(_lib.lib().born_tgc_apply(self._plan, int(mode), dptr(uin), u2.shape[1], dptr(out.view(np.float64))))
        res = out.T
        return res[:, 0].copy() if squeeze else np.ascontiguousarray(res)

    def tgc_forward(self, u) -> np.ndarray:
        """TGC of u (N_obs,) or (N_obs, N_rhs), FP64 (forward_transform)."""
        return self._tgc_host(u, 1)

    def tgc_adjoint(self, u) -> np.ndarray:
        """Adjoint TGC (adjoint_transform), FP64."""
        return self._tgc_host(u, 2)

    def tgc_dev(self, u, mode: int = 1, out=None, stream=None):
        """TGC on the device: u (N_obs,) or (N_rhs, N_obs) complex64; mode 1
        forward, 2 adjoint.  Returns a new tensor unless `out` is given."""
        import torch
        squeeze = u.dim() == 1
        uu = u.reshape(1, -1) if squeeze else u
        self._check_dev(uu, self.N_observations)
        if out is None:
            out = torch.empty_like(uu)
        s = torch.cuda.current_stream(uu.device).cuda_stream if stream is None else stream
        check(_lib.lib().born_tgc_apply_dev(self._plan, int(mode), C.c_void_p(uu.data_ptr()),
                                            C.c_void_p(out.data_ptr()), uu.shape[0], C.c_void_p(s)))
        return out.reshape(-1) if squeeze else out

    # ---------------------------------------------------------- host FP64 path
    def forward(self, gamma) -> np.ndarray:
        """u_M = A gamma with MEX semantics (occupancy rule, zero output when
        nothing is occupied).  gamma: (N_pts,) or (N_pts, N_rhs)."""
        g = np.asarray(gamma)
        squeeze = g.ndim == 1
        g2 = g[:, None] if squeeze else g
        if g2.shape[0] != self.N_points:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix",
                            f"input must be a double matrix with {self.N_points} rows!")
        nrhs = g2.shape[1]
        gin = _c128_colmajor(g2)
        out = np.empty((nrhs, self.N_observations), dtype=np.complex128)
        check(_lib.lib().born_forward(self._plan, dptr(gin), nrhs, dptr(out.view(np.float64))))
        res = out.T
        return res[:, 0].copy() if squeeze else np.ascontiguousarray(res)

    def adjoint(self, u) -> np.ndarray:
        """gamma_hat = A^H u_M over all N_pts points."""
        uu = np.asarray(u)
        squeeze = uu.ndim == 1
        u2 = uu[:, None] if squeeze else uu
        if u2.shape[0] != self.N_observations:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix",
                            f"input must be a double matrix with {self.N_observations} rows!")
        nrhs = u2.shape[1]
        uin = _c128_colmajor(u2)
        out = np.empty((nrhs, self.N_points), dtype=np.complex128)
        check(_lib.lib().born_adjoint(self._plan, dptr(uin), nrhs, dptr(out.view(np.float64))))
        res = out.T
        return res[:, 0].copy() if squeeze else np.ascontiguousarray(res)

    def __matmul__(self, gamma):
        return self.forward(gamma)

    mtimes = forward

    # ---------------------------------------------------------- device path
    def forward_dev(self, gamma, out=None, stream=None):
        """Device forward on torch complex64 tensors: gamma (N_pts,) or
        (N_rhs, N_pts) contiguous -> u (N_obs,) or (N_rhs, N_obs)."""
        import torch
        squeeze = gamma.dim() == 1
        g = gamma.reshape(1, -1) if squeeze else gamma
        self._check_dev(g, self.N_points)
        nrhs = g.shape[0]
        if out is None:
            out = torch.empty((nrhs, self.N_observations), dtype=torch.complex64, device=g.device)
        s = torch.cuda.current_stream(g.device).cuda_stream if stream is None else stream
        check(_lib.lib().born_forward_dev(self._plan, C.c_void_p(g.data_ptr()), nrhs,
                                          C.c_void_p(out.data_ptr()), C.c_void_p(s)))
        return out.reshape(-1) if squeeze else out

    def adjoint_dev(self, u, out=None, stream=None):
        """Device adjoint: u (N_obs,) or (N_rhs, N_obs) complex64 -> gamma_hat."""
        import torch
        squeeze = u.dim() == 1
        uu = u.reshape(1, -1) if squeeze else u
        self._check_dev(uu, self.N_observations)
        nrhs = uu.shape[0]
        if out is None:
            out = torch.empty((nrhs, self.N_points), dtype=torch.complex64, device=uu.device)
        s = torch.cuda.current_stream(uu.device).cuda_stream if stream is None else stream
        check(_lib.lib().born_adjoint_dev(self._plan, C.c_void_p(uu.data_ptr()), nrhs,
                                          C.c_void_p(out.data_ptr()), C.c_void_p(s)))
        return out.reshape(-1) if squeeze else out

    def autotune(self, stream=None):
        """Choose the forward work-item count, persistent grid size and
        adjoint segment width of this plan by timing (structured path).
        Returns the best (forward, adjoint) contraction times in ms."""
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        f, a = C.c_double(), C.c_double()
        check(_lib.lib().born_plan_autotune(self._plan, C.c_void_p(s), C.byref(f), C.byref(a)))
        return f.value, a.value

    @staticmethod
    def _check_dev(t, rows):
        import torch
        if t.dtype != torch.complex64 or not t.is_cuda or not t.is_contiguous() or t.shape[-1] != rows:
            raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix",
                            f"expected a contiguous complex64 device tensor with last dimension {rows}")

    # ---------------------------------------------------------- solver contract
    def op_A(self, x, mode: int, tgc: bool = False, dictionary=None):
        """operator.combined_scalar: mode 0 size, 1 forward, 2 adjoint
        (operator.m:1019-1072).  tgc=True applies the kernels of set_tgc
        after the forward and their adjoint before the adjoint; a dictionary
        (forward_transform / adjoint_transform, e.g. wavelet.D4Dictionary)
        synthesizes gamma = Psi^T x before the forward and analyses the
        adjoint's output (forward_scalar / adjoint_scalar,
        operator_born.m:190-236, :263-306).  Mode 0 returns the operator's
        own size, as the reference does (operator.m:1044-1045)."""
        if mode == 0:
            return list(self.size)
        if mode == 1:
            g = dictionary.adjoint_transform(x) if dictionary is not None else x
            u = self.forward(g)
            return self.tgc_forward(u) if tgc else u
        if mode == 2:
            g = self.adjoint(self.tgc_adjoint(x) if tgc else x)
            return dictionary.forward_transform(g) if dictionary is not None else g
        raise BornError(6, "combined_scalar:InvalidMode", "mode must be 0, 1 or 2")


def combined_quick_gpu(operator_born, mode, input_matrix):
    """output = combined_quick_gpu(operator_born, mode, input_matrix)
    (combined_quick_gpu.m:11-23): mode 1 forward, any other value adjoint;
    the input must be a nonempty 2-D double matrix with N_points (forward)
    or N_observations (adjoint) rows; real input is promoted to complex."""
    if not isinstance(operator_born, OperatorBorn):
        raise BornError(5, "combined_quick_gpu:NoOperatorBorn", "operator_born must be a single scattering.operator_born!")
    m = np.asarray(mode)
    if m.size != 1 or not np.issubdtype(m.dtype, np.number) or np.iscomplexobj(m):
        raise BornError(6, "combined_quick_gpu:NoNumericScalarMode", "mode must be a real-valued numeric scalar!")
    mode_i = int(m.reshape(()).item())
    x = np.asarray(input_matrix)
    rows = operator_born.N_points if mode_i == 1 else operator_born.N_observations
    if x.dtype not in (np.float64, np.complex128) or x.ndim != 2 or x.shape[0] != rows or x.size == 0:
        raise BornError(7, "combined_quick_gpu:NoValidDoubleMatrix", f"input must be a double matrix with {rows} rows!")
    return operator_born.forward(x) if mode_i == 1 else operator_born.adjoint(x)


class OperatorBatch:
    """Several operator instances applied together, one right-hand side each:
    a few persistent contraction launches per direction cover all of them,
    one per group of instances that take the same kernel form (born_batch_*).
    The plans must stay alive while the batch exists."""

    def __init__(self, ops, graph: bool = False):
        self.ops = list(ops)
        n = len(self.ops)
        arr = (C.c_void_p * n)(*[op._plan for op in self.ops])
        b = C.c_void_p()
        check(_lib.lib().born_batch_create(arr, n, C.byref(b)))
        self._batch = b
        if graph:
            self.set_graph(True)

    def close(self):
        if getattr(self, "_batch", None):
            _lib.lib().born_batch_destroy(self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool = True):
        check(_lib.lib().born_batch_set_graph(self._batch, int(bool(on))))

    def set_timing(self, on: bool = True):
        check(_lib.lib().born_batch_set_timing(self._batch, int(bool(on))))

    def get_timing(self):
        ms = (C.c_double * 2)()
        n = (C.c_int64 * 2)()
        check(_lib.lib().born_batch_get_timing(self._batch, ms, n))
        return {"forward": (ms[0], n[0]), "adjoint": (ms[1], n[1])}

    def autotune(self, stream=None):
        import torch
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        f, a = C.c_double(), C.c_double()
        check(_lib.lib().born_batch_autotune(self._batch, C.c_void_p(s), C.byref(f), C.byref(a)))
        return f.value, a.value

    def _apply(self, fn, xs, rows_in, rows_out, outs, stream):
        import torch
        n = len(self.ops)
        if len(xs) != n:
            raise ValueError("one input per instance is required")
        for x, r in zip(xs, rows_in):
            OperatorBorn._check_dev(x.reshape(1, -1), r)
        if outs is None:
            outs = [torch.empty(r, dtype=torch.complex64, device=x.device) for x, r in zip(xs, rows_out)]
        s = torch.cuda.current_stream().cuda_stream if stream is None else stream
        pin = (C.c_void_p * n)(*[x.data_ptr() for x in xs])
        pout = (C.c_void_p * n)(*[o.data_ptr() for o in outs])
        check(fn(self._batch, pin, pout, C.c_void_p(s)))
        return outs

    def forward_dev(self, gammas, outs=None, stream=None):
        """gammas[i]: complex64 (N_points(i),) -> list of u_i (N_observations(i),)."""
        return self._apply(_lib.lib().born_batch_forward_dev, gammas, [op.N_points for op in self.ops],
                           [op.N_observations for op in self.ops], outs, stream)

    def adjoint_dev(self, us, outs=None, stream=None):
        """us[i]: complex64 (N_observations(i),) -> list of gamma_hat_i (N_points(i),)."""
        return self._apply(_lib.lib().born_batch_adjoint_dev, us, [op.N_observations for op in self.ops],
                           [op.N_points for op in self.ops], outs, stream)

