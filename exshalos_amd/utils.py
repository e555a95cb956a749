"""Field, halo-finder and LPT wrappers: the hot-path part of pyexshalos.utils
(utils.py:8-288), same signatures and defaults as the reference."""

from __future__ import annotations

from typing import Dict, Optional

import numpy as np


def Generate_Density_Grid(
    k: np.ndarray,
    P: np.ndarray,
    R_max: float = 100000.0,
    nd: int = 256,
    ndx: Optional[int] = None,
    ndy: Optional[int] = None,
    ndz: Optional[int] = None,
    Lc: float = 2.0,
    outk: bool = False,
    seed: int = 12345,
    fixed: bool = False,
    phase: float = 0.0,
    k_smooth: float = 100000.0,
    verbose: bool = False,
    nthreads: int = 1,
    *,
    noise=None,
    device_out: bool = False,
) -> np.ndarray:
    """Gaussian density grid from P(k) (reference utils.py:8-109)."""
    from .lib.exshalos import check_precision, density_grid_compute

    precision = check_precision()
    R = np.float32 if precision == 4 else np.float64
    k = np.asarray(k).astype(R)
    P = np.asarray(P).astype(R)
    ndx = nd if ndx is None else ndx
    ndy = nd if ndy is None else ndy
    ndz = nd if ndz is None else ndz
    return density_grid_compute(k, P, R(R_max), np.int32(ndx), np.int32(ndy), np.int32(ndz),
                                R(Lc), np.int32(outk), np.int32(seed), np.int32(fixed), R(phase),
                                R(k_smooth), np.int32(verbose), np.int32(nthreads), noise=noise,
                                device_out=device_out)


def Find_Halos_from_Grid(
    grid: np.ndarray,
    k: np.ndarray,
    P: np.ndarray,
    Lc: float = 2.0,
    Om0: float = 0.31,
    z: float = 0.0,
    delta_c: Optional[float] = None,
    Nmin: int = 10,
    a: float = 1.0,
    beta: float = 0.0,
    alpha: float = 0.0,
    OUT_FLAG: bool = False,
    verbose: bool = False,
    *,
    device_out: bool = False,
) -> Dict[str, np.ndarray]:
    """Lagrangian halo catalogue of a density grid (reference utils.py:113-209).

    As in the reference, ``delta_c=None`` becomes NaN, which yields no
    halos (utils.py:175); pass -1 for the default delta_c(z)."""
    from .lib.exshalos import check_precision, find_halos

    precision = check_precision()
    R = np.float32 if precision == 4 else np.float64
    k = np.asarray(k).astype(R)
    P = np.asarray(P).astype(R)
    dc = R(np.nan) if delta_c is None else R(delta_c)
    return find_halos(grid, k, P, R(Lc), Om0, R(z), dc, np.int32(Nmin), R(a), R(beta), R(alpha),
                      np.int32(OUT_FLAG), np.int32(False), verbose, device_out=device_out)


def Displace_LPT(
    grid: np.ndarray,
    Lc: float = 2.0,
    Om0: float = 0.31,
    z: float = 0.0,
    k_smooth: float = 10000.0,
    DO_2LPT: bool = False,
    OUT_VEL: bool = False,
    Input_k: bool = False,
    OUT_POS: bool = True,
    verbose: bool = False,
    *,
    device_out: bool = False,
) -> Dict[str, np.ndarray]:
    """LPT positions (or displacements) of the grid's particles
    (reference utils.py:213-288)."""
    from .lib.exshalos import check_precision, lpt_compute

    precision = check_precision()
    R = np.float32 if precision == 4 else np.float64
    return lpt_compute(grid, R(Lc), Om0, R(z), R(k_smooth), np.int32(DO_2LPT), np.int32(OUT_VEL),
                       np.int32(Input_k), np.int32(OUT_POS), np.int32(verbose),
                       device_out=device_out)
