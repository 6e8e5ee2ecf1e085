"""Wave syntheses of the configs: the random ones follow ra.m / rd.m / rad.m
(apodization +-1, delays a permuted linear ramp), the deterministic ones
their closed forms."""
import math

import numpy as np

from comp_ui_toolbox_amd import configs as C


def test_random_syntheses():
    arr = C.ArraySpec((16,), (3e-4,), (2.7e-4,), (4,))
    c = 1540.0
    e = [math.sin(0.3)]
    ra = C.ra_wave(arr, np.random.default_rng(1))
    assert set(np.unique(ra.apodization)) <= {-1.0, 1.0} and np.all(ra.delays == 0)
    rd = C.rd_wave(arr, np.random.default_rng(2), e, c)
    t_inc = 15 * 3e-4 * abs(e[0]) / c / 15
    assert np.all(rd.apodization == 1.0)
    assert np.allclose(np.sort(rd.delays), np.arange(16) * t_inc, rtol=0, atol=1e-18)
    assert not np.all(np.diff(rd.delays) > 0)          # permuted, not the plain ramp
    rad = C.rad_wave(arr, np.random.default_rng(1), np.random.default_rng(2), e, c)
    assert np.array_equal(rad.apodization, ra.apodization) and np.array_equal(rad.delays, rd.delays)


def test_random_syntheses_2d_array():
    arr = C.ArraySpec((8, 4), (3e-4, 2e-4), (2.7e-4, 1.8e-4), (2, 2))
    e = [0.2, -0.1]
    rd = C.rd_wave(arr, np.random.default_rng(3), e, 1500.0)
    t_max = (7 * 3e-4 * 0.2 + 3 * 2e-4 * 0.1) / 1500.0
    assert math.isclose(rd.delays.max(), t_max, rel_tol=1e-12)


def test_qpw_delays_closed_form():
    arr = C.ArraySpec((8,), (3e-4,), (2.7e-4,), (4,))
    w = C.qpw_wave(arr, 0.2, 1500.0)
    x = arr.positions_ctr()[:, 0]
    assert np.allclose(w.delays, (x - x.min()) * math.sin(0.2) / 1500.0, rtol=0, atol=1e-18)


def test_config4_phantom_has_four_spheres_and_no_bright_points():
    """The config-4 phantom: N(0,1) speckle scaled by 0.05 inside 4 spheres of
    radius 2 mm, no bright points; configs 1 and 2 keep their 20 points of 10."""
    import numpy as np
    from comp_ui_toolbox_amd import configs as C
    s4 = C.config4()
    g = C.phantom(s4)
    assert g.shape == (s4.fov.N_points,) and np.max(np.abs(g)) < 8.0
    assert np.array_equal(g, C.phantom_speckle(s4, n_disks=4, radius=2e-3, n_bright=0))
    # N(0,1) puts 1.6 % of the points below 0.02; the damped spheres add about 0.4 %
    assert 0.017 < np.mean(np.abs(g) < 0.02) < 0.025
    g1 = C.phantom(C.config1())
    assert np.sum(g1 == 10.0) == 20
