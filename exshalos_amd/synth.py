"""Synthetic inputs for the field -> LPT -> halos -> P(k) path.

Two generators, both declared so that any run can be reproduced:

* ``bbks_spectrum``: a linear P(k) = A k^ns T_BBKS(k / (Om0 h))^2 on
  log-spaced k, with A fixed so that the top-hat variance at R = 8 Mpc/h,
  integrated in float64, equals sigma8^2.
* ``white_noise``: unit-variance complex modes on the half-complex lattice
  [nd][nd][nd//2+1] drawn by ``numpy.random.default_rng(seed)``
  (real and imaginary parts ~ N(0, 1)).  The reference draws its modes from
  GSL's taus generator inside its own loop (density_grid.c:86-185); the
  modes here are an explicit input so that the GPU path and the CPU oracle
  consume identical numbers.
"""

from __future__ import annotations

import numpy as np


def bbks_transfer(q: np.ndarray) -> np.ndarray:
    """BBKS (Bardeen et al. 1986) transfer function of q = k / (Om0 h)."""
    q = np.asarray(q, dtype=np.float64)
    x = 2.34 * q
    poly = 1.0 + 3.89 * q + (16.1 * q) ** 2 + (5.46 * q) ** 3 + (6.71 * q) ** 4
    return np.log1p(x) / x * poly ** (-0.25)


def tophat_window(x: np.ndarray) -> np.ndarray:
    """Fourier transform of a normalised top-hat sphere, 3 (sin x - x cos x)/x^3."""
    x = np.asarray(x, dtype=np.float64)
    return 3.0 * (np.sin(x) - x * np.cos(x)) / x**3


def sigma_R(k: np.ndarray, P: np.ndarray, R: float) -> float:
    """Top-hat rms at radius R by the trapezoid rule in ln k (float64)."""
    lnk = np.log(k)
    f = k**3 * P * tophat_window(k * R) ** 2 / (2.0 * np.pi**2)
    return float(np.sqrt(np.trapezoid(f, lnk)))


def bbks_spectrum(
    nk: int = 500,
    kmin: float = 1e-4,
    kmax: float = 10.0,
    Om0: float = 0.31,
    h: float = 0.677,
    ns: float = 0.965,
    sigma8: float = 0.81,
) -> tuple[np.ndarray, np.ndarray]:
    """Return (k, P) in float64 on ``nk`` log-spaced wavenumbers [h/Mpc].

    The amplitude is fixed on a dense grid (20001 points over [1e-5, 1e2])
    so that it does not depend on ``nk``.
    """
    kd = np.logspace(-5, 2, 20001)
    shape_d = kd**ns * bbks_transfer(kd / (Om0 * h)) ** 2
    A = (sigma8 / sigma_R(kd, shape_d, 8.0)) ** 2
    k = np.logspace(np.log10(kmin), np.log10(kmax), nk)
    P = A * k**ns * bbks_transfer(k / (Om0 * h)) ** 2
    return k, P


def white_noise(nd: int, seed: int, dtype=np.float32, ndx: int | None = None,
                ndy: int | None = None) -> np.ndarray:
    """Unit complex normal modes, shape (ndx, ndy, nd//2+1, 2), C order."""
    ndx = nd if ndx is None else ndx
    ndy = nd if ndy is None else ndy
    rng = np.random.default_rng(seed)
    g = rng.standard_normal((ndx, ndy, nd // 2 + 1, 2))
    return np.ascontiguousarray(g.astype(dtype, copy=False))
