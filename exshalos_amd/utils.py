"""Field, halo-finder and LPT wrappers and the barrier fit: the hot-path part
of pyexshalos.utils (utils.py:8-455), same signatures and defaults as the
reference."""

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
    sigma_radii: int = 0,
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
                      np.int32(OUT_FLAG), np.int32(False), verbose, device_out=device_out,
                      sigma_radii=sigma_radii)


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


def _Omz(z: float, Om0: float) -> float:
    """Matter density parameter at z (reference theory.py:40-52)."""
    a3 = np.power(1.0 + z, 3.0)
    return Om0 * a3 / (Om0 * a3 + (1.0 - Om0))


def _deltac(z: float, Om0: float) -> float:
    """delta_c(z) fit 1.686 Omz^0.0055 (reference theory.py:55-67)."""
    return 1.686 * np.power(_Omz(z, Om0), 0.0055)


def Fit_Barrier(
    k: np.ndarray,
    P: np.ndarray,
    M: np.ndarray,
    dndlnM: np.ndarray,
    dn_err: Optional[np.ndarray] = None,
    grid=None,
    R_max: float = 100000.0,
    Mmin: Optional[float] = None,
    Mmax: Optional[float] = None,
    Nm: int = 25,
    nd: int = 256,
    Lc: float = 2.0,
    Om0: float = 0.31,
    z: float = 0.0,
    delta_c: Optional[float] = None,
    Nmin: int = 10,
    seed: int = 12345,
    x0: Optional[np.ndarray] = None,
    verbose: bool = False,
    nthreads: int = 1,
    Max_iter: int = 100,
    tol: Optional[float] = None,
    *,
    noise=None,
) -> np.ndarray:
    """Best-fit ellipsoidal-barrier parameters (a, beta, alpha) for a target
    mass function (reference utils.py:291-455): Nelder-Mead over the chi^2 of
    the abundance of the halos found in one density grid.

    The grid stays on the device for every trial, and each trial's masses go
    straight from the halo finder to the device histogram (no catalogue copy
    to the host).  As in the reference, the errors of the target are
    interpolated at M rather than ln M (utils.py:401).  ``noise=`` (keyword)
    feeds explicit white-noise modes to the grid as Generate_Density_Grid
    does."""
    from scipy.interpolate import interp1d
    from scipy.optimize import minimize

    from .simulation import Compute_Abundance

    if grid is None:
        grid = Generate_Density_Grid(k, P, R_max, nd=nd, Lc=Lc, seed=seed, verbose=False,
                                     nthreads=nthreads, noise=noise, device_out=True)
    M = np.asarray(M)
    dndlnM = np.asarray(dndlnM)
    dn_err = np.zeros(len(M)) if dn_err is None else np.asarray(dn_err)
    Mmin = -1.0 if Mmin is None else Mmin
    Mmax = -1.0 if Mmax is None else Mmax
    fdn = interp1d(np.log(M[M > 0.0]), dndlnM[M > 0.0], bounds_error=False, fill_value=0.0)
    fdn_err = interp1d(np.log(M[M > 0.0]), dn_err[M > 0.0], bounds_error=False,
                       fill_value=0.0)
    if delta_c is None:
        delta_c = _deltac(z, Om0)

    def Chi2(theta):
        a, beta, alpha = theta
        Find_Halos_from_Grid(grid, k, P, Lc=Lc, Om0=Om0, z=z, delta_c=delta_c, Nmin=Nmin, a=a,
                             beta=beta, alpha=alpha, verbose=False, device_out=True)
        dnh = Compute_Abundance(None, Mmin=Mmin, Mmax=Mmax, Nm=Nm, Lc=Lc, nd=nd, verbose=False)
        mask = dnh["dn"] > 0.0
        chi2 = np.sum(np.power(dnh["dn"][mask] - fdn(np.log(dnh["Mh"][mask])), 2.0) /
                      (np.power(dnh["dn_err"][mask], 2.0) +
                       np.power(fdn_err(dnh["Mh"][mask]), 2.0))) / (Nm - 4)
        if verbose:
            print("Current try: (%f, %f, %f) with chi2 = %f" % (a, beta, alpha, chi2))
        return chi2

    if x0 is None:
        x0 = [0.55, 0.4, 0.7]
    bounds = [(0.1, 2.0), (0.0, 1.0), (0.0, 1.0)]
    x = minimize(Chi2, x0=x0, bounds=bounds, method="Nelder-Mead",
                 options={"maxiter": Max_iter}, tol=tol)
    return x.x
