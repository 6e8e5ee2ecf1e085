"""Orthogonal periodized Daubechies-4 wavelet transform of 2-D images (torch).

The regularizer of the config-3 loss, lambda ||Psi gamma||_1, uses Psi = the
separable 2-D discrete wavelet transform with the four-tap Daubechies filter
(Daubechies, "Ten Lectures on Wavelets", 1992, table 6.1, N = 2)

    h = [1 + sqrt3, 3 + sqrt3, 3 - sqrt3, 1 - sqrt3] / (4 sqrt2),
    g[j] = (-1)^j h[3 - j],

with periodic extension (Mallat's algorithm on the circle), so Psi is
orthogonal: Psi^T = Psi^-1.  A level halves both axes of the current
approximation block; the depth is limited by how often both sides stay even
(a 384-point side allows 7 levels).  The reference takes its wavelets from
WaveLab, which is not part of it; this module is this package's own.

The operations are gathers and index-adds on torch tensors, so autograd
differentiates through them on either device.
"""
from __future__ import annotations

import math

import torch

_S3 = math.sqrt(3.0)
H = [(1 + _S3) / (4 * math.sqrt(2.0)), (3 + _S3) / (4 * math.sqrt(2.0)),
     (3 - _S3) / (4 * math.sqrt(2.0)), (1 - _S3) / (4 * math.sqrt(2.0))]
G = [((-1) ** j) * H[3 - j] for j in range(4)]


def max_levels(shape) -> int:
    """Largest depth for which every level's block has even sides >= 4."""
    n = min(shape)
    m = list(shape)
    lv = 0
    while all(s % 2 == 0 and s >= 4 for s in m):
        m = [s // 2 for s in m]
        lv += 1
    return lv if n >= 4 else 0


def _analysis_last(x: torch.Tensor) -> torch.Tensor:
    """One level along the last axis: [approximation | detail]."""
    n = x.shape[-1]
    k = torch.arange(n // 2, device=x.device)
    a = torch.zeros(x.shape[:-1] + (n // 2,), dtype=x.dtype, device=x.device)
    d = torch.zeros_like(a)
    for j in range(4):
        xj = x.index_select(-1, (2 * k + j) % n)
        a = a + H[j] * xj
        d = d + G[j] * xj
    return torch.cat([a, d], dim=-1)


def _synthesis_last(y: torch.Tensor) -> torch.Tensor:
    """Inverse (= transpose) of _analysis_last."""
    n = y.shape[-1]
    k = torch.arange(n // 2, device=y.device)
    a, d = y[..., : n // 2], y[..., n // 2:]
    x = torch.zeros_like(y)
    for j in range(4):
        x = x.index_add(-1, (2 * k + j) % n, H[j] * a + G[j] * d)
    return x


def dwt2(img: torch.Tensor, levels: int) -> torch.Tensor:
    """Psi img for an (..., N0, N1) real tensor; coefficients in the usual
    nested layout (approximation in the top-left block)."""
    if levels > max_levels(img.shape[-2:]):
        raise ValueError("too many levels for this image size")
    out = img
    n0, n1 = img.shape[-2:]
    for _ in range(levels):
        blk = out[..., :n0, :n1]
        blk = _analysis_last(blk)
        blk = _analysis_last(blk.transpose(-1, -2)).transpose(-1, -2)
        out = torch.cat([torch.cat([blk, out[..., :n0, n1:]], dim=-1), out[..., n0:, :]], dim=-2)
        n0, n1 = n0 // 2, n1 // 2
    return out


def idwt2(coef: torch.Tensor, levels: int) -> torch.Tensor:
    """Psi^T coef = Psi^-1 coef."""
    if levels > max_levels(coef.shape[-2:]):
        raise ValueError("too many levels for this image size")
    N0, N1 = coef.shape[-2:]
    out = coef
    for lv in reversed(range(levels)):
        n0, n1 = N0 >> lv, N1 >> lv
        blk = out[..., :n0, :n1]
        blk = _synthesis_last(blk.transpose(-1, -2)).transpose(-1, -2)
        blk = _synthesis_last(blk)
        out = torch.cat([torch.cat([blk, out[..., :n0, n1:]], dim=-1), out[..., n0:, :]], dim=-2)
    return out


class D4Dictionary:
    """The orthogonal D4 wavelet as the dictionary of op_A, with the
    linear_transforms interface forward_scalar / adjoint_scalar use
    (operator_born.m:190-200, :300-306): forward_transform = analysis Psi,
    adjoint_transform = synthesis Psi^T.  It acts on FOV vectors of a 2-D grid
    (N_x, N_z), points x fastest, as (N_pts,) or (N_pts, N_rhs) numpy arrays
    (real or complex; real and imaginary parts are transformed separately)
    or torch tensors, on `device`."""

    def __init__(self, grid_shape, levels=None, device="cuda"):
        if len(grid_shape) != 2:
            raise ValueError("the D4 dictionary acts on 2-D grids")
        self.nx, self.nz = int(grid_shape[0]), int(grid_shape[1])
        self.levels = max_levels((self.nz, self.nx)) if levels is None else int(levels)
        if self.levels > max_levels((self.nz, self.nx)):
            raise ValueError("too many levels for this grid")
        self.device = device

    def _apply(self, x, fn):
        import numpy as np
        is_np = not isinstance(x, torch.Tensor)
        t = torch.as_tensor(np.asarray(x) if is_np else x)
        squeeze = t.dim() == 1
        cols = t.reshape(-1, 1) if squeeze else t
        if cols.shape[0] != self.nx * self.nz:
            raise ValueError(f"expected {self.nx * self.nz} rows")
        dev = cols.device if not is_np else torch.device(self.device)
        imgs = cols.T.reshape(-1, self.nz, self.nx).to(dev)
        if torch.is_complex(imgs):
            out = torch.complex(fn(imgs.real.double(), self.levels), fn(imgs.imag.double(), self.levels))
        else:
            out = fn(imgs.double(), self.levels)
        out = out.reshape(-1, self.nx * self.nz).T
        out = out.reshape(-1) if squeeze else out
        return out.cpu().numpy() if is_np else out

    def forward_transform(self, x):
        """Psi x (analysis)."""
        return self._apply(x, dwt2)

    def adjoint_transform(self, x):
        """Psi^T x (synthesis, = Psi^-1 x)."""
        return self._apply(x, idwt2)
