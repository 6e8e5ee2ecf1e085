"""Writes the committed golden vectors under tests/golden/.

* hankel_mpmath.npz -- H0^(2)(z) at 40 significant digits (mpmath) over the
  argument range of the workloads (|z| from 0.05 to 3000, Im z <= 0).  It is
  independent of scipy and of the oracle.
* config0.npz -- the config-0 phantom (30 points), its forward data u
  (4,160 observations) and the adjoint image on 512 fixed grid points, all
  from the FP64 oracle.  This is a regression vector for the oracle and the
  parity target of the device path at config-0 size.

Run from the repository root:  python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import mpmath
import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from comp_ui_toolbox_amd import configs  # noqa: E402
from oracle import born_oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def hankel_vectors():
    rng = np.random.Generator(np.random.PCG64(11))
    mags = np.concatenate([np.geomspace(0.05, 3000.0, 160), [11.99, 12.0, 12.01, 16.76, 2370.0]])
    # the workloads have Im z = Im(k) D >= -3.5; keep Im z in [-min(5, 0.05|z|), 0]
    z = mags - 1j * rng.uniform(0.0, 1.0, len(mags)) * np.minimum(5.0, 0.05 * mags)
    mpmath.mp.dps = 40
    h = np.array([complex(mpmath.hankel2(0, mpmath.mpc(v.real, v.imag))) for v in z])
    np.savez_compressed(OUT / "hankel_mpmath.npz", z=z, h=h)


def config0_vectors():
    spec = configs.config0()
    T = O.build_tables(spec)
    g = configs.phantom_points(spec, 30)
    u = O.forward(T, g)
    pts = np.sort(np.random.Generator(np.random.PCG64(12)).choice(spec.fov.N_points, 512, replace=False))
    gh = O.adjoint(T, u, points=pts)
    np.savez_compressed(OUT / "config0.npz", gamma=g, u=u, points=pts, gamma_hat=gh)


if __name__ == "__main__":
    hankel_vectors()
    config0_vectors()
    print("written to", OUT)
