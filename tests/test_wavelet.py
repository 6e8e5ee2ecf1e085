"""The orthogonal periodized Daubechies-4 transform: filter identities,
orthogonality on dyadic and non-dyadic sides, vanishing moments."""
import math

import numpy as np
import pytest
import torch

from comp_ui_toolbox_amd import wavelet as W


def test_filter_identities():
    h, g = np.array(W.H), np.array(W.G)
    assert abs(h.sum() - math.sqrt(2)) < 1e-15
    assert abs((h * h).sum() - 1) < 1e-15
    assert abs(h[0] * h[2] + h[1] * h[3]) < 1e-15          # double-shift orthogonality
    assert abs(g.sum()) < 1e-15 and abs((np.arange(4) * g).sum()) < 1e-14   # two vanishing moments


@pytest.mark.parametrize("shape,levels", [((16, 16), 2), ((384, 384), 7), ((64, 32), 3)])
def test_orthogonal(shape, levels):
    rng = np.random.Generator(np.random.PCG64(5))
    x = torch.tensor(rng.standard_normal(shape))
    c = W.dwt2(x, levels)
    assert abs(float(c.norm() / x.norm()) - 1.0) < 1e-13
    assert float((W.idwt2(c, levels) - x).norm() / x.norm()) < 1e-13
    y = torch.tensor(rng.standard_normal(shape))
    # <Psi x, y> = <x, Psi^T y>
    assert abs(float((c * y).sum() - (x * W.idwt2(y, levels)).sum())) < 1e-10


def test_linear_ramp_has_no_interior_details():
    """Two vanishing moments: the finest details of a linear ramp vanish
    except where the periodic extension wraps."""
    n = 32
    x = torch.arange(n, dtype=torch.float64).repeat(n, 1)       # ramp along the last axis
    y = W._analysis_last(x)
    d = y[:, n // 2:]
    assert float(d[:, :-1].abs().max()) < 1e-12 and float(d[:, -1].abs().max()) > 1.0


def test_max_levels_non_dyadic():
    assert W.max_levels((384, 384)) == 7
    assert W.max_levels((128, 128)) == 6
    assert W.max_levels((6, 6)) == 1


def test_d4_dictionary_on_fov_vectors():
    """The dictionary acts on FOV vectors (x fastest), complex columns part
    by part, and is orthogonal: Psi^T Psi = I and <Psi x, y> = <x, Psi^T y>."""
    import numpy as np
    from comp_ui_toolbox_amd.wavelet import D4Dictionary
    D = D4Dictionary((16, 8), device="cpu")
    rng = np.random.default_rng(3)
    x = rng.standard_normal((128, 2)) + 1j * rng.standard_normal((128, 2))
    y = rng.standard_normal((128, 2)) + 1j * rng.standard_normal((128, 2))
    assert np.allclose(D.adjoint_transform(D.forward_transform(x)), x, atol=1e-12)
    assert abs(np.vdot(D.forward_transform(x), y) - np.vdot(x, D.adjoint_transform(y))) < 1e-10
    img = x[:, 0].reshape(8, 16)                      # rows z, columns x
    from comp_ui_toolbox_amd.wavelet import dwt2
    ref = dwt2(torch.tensor(img.real), D.levels).numpy().reshape(-1)
    assert np.allclose(D.forward_transform(x[:, 0].real), ref, atol=1e-14)
