"""Exponential time gain compensation (TGC) as convolution kernels.

regularization.tgc.exponential compensates the decay of the echoes by the
gain curve g(t) = exp(exponent (t - lb)) over each mix's recording interval
[lb, ub].  In the frequency domain the product g(t) s(t) of periodic signals
(period T_ref = 1 / delta_f of the mix) is the convolution of their Fourier
coefficients, so the transform is one discrete convolution per mix with the
Fourier coefficients of g, truncated where their magnitude has fallen by
`decay_dB` (+regularization/+tgc/exponential.m:118-213,
+regularization/+tgc/+curves/exponential.m:150-214).

The kernels go to the device with `OperatorBorn.set_tgc`; the convolution
itself runs there (born_tgc_apply / born_tgc_apply_dev).
"""
from __future__ import annotations

import math

import numpy as np

__all__ = ["exponential_coefficients", "exponential_kernels", "kernel_from_coefficients"]


def exponential_coefficients(exponent: float, interval_t, T_ref: float, decay_dB: float = -40.0) -> np.ndarray:
    """Fourier coefficients c_k, k = 0..K, of g(t) = exp(exponent (t - lb))
    on [lb, ub] with period T_ref >= ub - lb:
        c_k = (1 / T_ref) int_lb^ub g(t) exp(-2j pi k t / T_ref) dt
            = exp(-2j pi f_k lb) (exp((a - 2j pi f_k) T) - 1) / ((a - 2j pi f_k) T_ref),
    f_k = k / T_ref, T = ub - lb.  K = ceil(ub_f T_ref) with ub_f the
    frequency beyond which |c_k| / |c_0| stays below 10^(decay_dB / 20)
    (fourier_coefficients, curves/exponential.m:150-214).  complex128 (K+1,)."""
    lb, ub = float(interval_t[0]), float(interval_t[1])
    T = ub - lb
    a = float(exponent)
    if not (a > 0.0):
        raise ValueError("exponent must be positive")
    if not (T > 0.0):
        raise ValueError("the recording interval must have positive length")
    if T_ref < T:
        raise ValueError("T_ref must be greater than or equal to T!")
    if not (decay_dB < 0.0):
        raise ValueError("decay_dB must be negative")
    # exact-period case (curves/exponential.m:188-194; eps(0) * T tolerance)
    if abs(T_ref - T) <= np.finfo(np.float64).smallest_subnormal * T:
        ub_f = a * math.sqrt(10.0 ** (-decay_dB / 10.0) - 1.0) / (2.0 * math.pi)
    else:
        e = math.exp(-a * T)
        ub_f = a * math.sqrt(((1.0 + e) * 10.0 ** (-decay_dB / 20.0) / (1.0 - e)) ** 2 - 1.0) / (2.0 * math.pi)
    K = int(math.ceil(ub_f * T_ref))
    f = np.arange(K + 1, dtype=np.float64) / T_ref
    den = a - 2j * np.pi * f
    return np.exp(-2j * np.pi * f * lb) * (np.exp(den * T) - 1.0) / (den * T_ref)


def kernel_from_coefficients(c: np.ndarray) -> np.ndarray:
    """Hermitian kernel [conj(c_K) .. conj(c_1), c_0 .. c_K] of length 2K+1
    (tgc/exponential.m:177-183): taps k[d] = c_d, k[-d] = conj(c_d)."""
    c = np.asarray(c, dtype=np.complex128)
    return np.concatenate([np.conj(c[:0:-1]), c])


def exponential_kernels(exponents, intervals_t, deltas_f, decays_dB=-60.0):
    """One kernel per mix signal (measurement-major, the order of the
    observations) for regularization.tgc.exponential(exponents, decays_dB)
    (get_LT_scalar, tgc/exponential.m:118-213).  `intervals_t`: (N_mix, 2)
    recording intervals [lb, ub] of the mixes' receivers in s; `deltas_f`:
    the spacing of each mix's frequency axis in Hz (T_ref = 1 / delta_f);
    `exponents` (1/s) and `decays_dB` are scalars or one per mix."""
    iv = np.asarray(intervals_t, dtype=np.float64).reshape(-1, 2)
    n = iv.shape[0]
    ex = np.broadcast_to(np.asarray(exponents, dtype=np.float64), (n,))
    dd = np.broadcast_to(np.asarray(decays_dB, dtype=np.float64), (n,))
    df = np.broadcast_to(np.asarray(deltas_f, dtype=np.float64), (n,))
    return [kernel_from_coefficients(exponential_coefficients(ex[i], iv[i], 1.0 / df[i], dd[i])) for i in range(n)]
