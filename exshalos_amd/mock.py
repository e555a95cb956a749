"""Halo mocks: the hot-path part of pyexshalos.mock (mock.py:11-353).

Same signatures and defaults as the reference; the arguments are cast by
``check_precision()`` as the reference's wrappers do (mock.py:113-142) and
passed positionally to ``lib.exshalos``.  Extra keyword-only arguments:
``noise`` (explicit white-noise modes), ``device_out`` (torch device
tensors instead of numpy arrays) and ``sigma_radii`` (> 0: the halo masses
use an FP64 sigma(R) table over that many radii instead of the reference's
one radius per input k in the grid's precision).
"""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np


def _cast(prec, *vals):
    R = np.float32 if prec == 4 else np.float64
    return [R(v) for v in vals]


def Generate_Halos_Box_from_Pk(
    k: np.ndarray,
    P: np.ndarray,
    R_max: float = 100000.0,
    nd: int = 256,
    ndx: int = 0,
    ndy: int = 0,
    ndz: int = 0,
    Lc: float = 2.0,
    Om0: float = 0.31,
    z: float = 0.0,
    k_smooth: float = 10000.0,
    delta_c: float = -1.0,
    Nmin: int = 1,
    a: float = 1.0,
    beta: float = 0.0,
    alpha: float = 0.0,
    seed: int = 12345,
    fixed: bool = False,
    phase: float = 0.0,
    OUT_DEN: bool = False,
    OUT_LPT: bool = False,
    OUT_VEL: bool = False,
    OUT_PROF: bool = False,
    DO_2LPT: bool = False,
    OUT_FLAG: bool = False,
    verbose: bool = False,
    nthreads: int = 1,
    *,
    noise=None,
    device_out: bool = False,
    sigma_radii: int = 0,
) -> Dict[str, np.ndarray]:
    """Halos in a box from a linear power spectrum (reference mock.py:11-184).

    Returns a dict with "posh", "Mh" and, on request, "velh", "pos", "vel",
    "flag", "grid".
    """
    from .lib.exshalos import check_precision, halos_box_from_pk

    precision = check_precision()
    R = np.float32 if precision == 4 else np.float64
    k = np.asarray(k).astype(R)
    P = np.asarray(P).astype(R)
    R_max, Lc, Om0, z, k_smooth, delta_c, a, beta, alpha, phase = _cast(
        precision, R_max, Lc, Om0, z, k_smooth, delta_c, a, beta, alpha, phase)
    if ndx <= 0:
        ndx = nd
    if ndy <= 0:
        ndy = nd
    if ndz <= 0:
        ndz = nd
    return halos_box_from_pk(
        k, P, R_max, np.int32(ndx), np.int32(ndy), np.int32(ndz), Lc, np.int32(seed), k_smooth,
        Om0, z, delta_c, np.int32(Nmin), a, beta, alpha, np.int32(fixed), phase,
        np.int32(OUT_DEN), np.int32(OUT_LPT), np.int32(OUT_VEL), np.int32(DO_2LPT),
        np.int32(OUT_FLAG), np.int32(OUT_PROF), np.int32(verbose), np.int32(nthreads),
        noise=noise, device_out=device_out, sigma_radii=sigma_radii)


def Generate_Halos_Box_from_Grid(
    grid: np.ndarray,
    k: np.ndarray,
    P: np.ndarray,
    S: Optional[np.ndarray] = None,
    V: Optional[np.ndarray] = None,
    Lc: float = 2.0,
    Om0: float = 0.31,
    z: float = 0.0,
    k_smooth: float = 10000.0,
    delta_c: float = -1.0,
    Nmin: int = 1,
    a: float = 1.0,
    beta: float = 0.0,
    alpha: float = 0.0,
    OUT_LPT: bool = False,
    OUT_VEL: bool = False,
    DO_2LPT: bool = False,
    OUT_FLAG: bool = False,
    OUT_PROF: bool = False,
    verbose: bool = False,
    nthreads: int = 1,
    *,
    device_out: bool = False,
    sigma_radii: int = 0,
) -> Dict[str, np.ndarray]:
    """Halos from a density grid (reference mock.py:188-353)."""
    In_disp = S is not None
    if In_disp:
        if S.shape[0] != grid.shape[0] * grid.shape[1] * grid.shape[2]:
            raise ValueError("The number of particles in S is different of the number of grid cells!")
        if S.shape[1] != 3:
            raise ValueError("The number of components of S is different from 3!")
        if V is not None:
            if V.shape[0] != grid.shape[0] * grid.shape[1] * grid.shape[2]:
                raise ValueError("The number of particles in V is different of the number of grid cells!")
            if V.shape[1] != 3:
                raise ValueError("The number of components of V is different from 3!")
    from .lib.exshalos import check_precision, halos_box_from_grid

    precision = check_precision()
    Lc, Om0, z, k_smooth, delta_c, a, beta, alpha = _cast(
        precision, Lc, Om0, z, k_smooth, delta_c, a, beta, alpha)
    return halos_box_from_grid(
        k, P, grid, S, V, Lc, k_smooth, Om0, z, delta_c, np.int32(Nmin), a, beta, alpha,
        np.int32(OUT_LPT), np.int32(OUT_VEL), np.int32(DO_2LPT), np.int32(OUT_FLAG),
        np.int32(In_disp), np.int32(OUT_PROF), np.int32(verbose), np.int32(nthreads),
        device_out=device_out, sigma_radii=sigma_radii)
