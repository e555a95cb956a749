"""Density grids, power spectra, bispectra and halo abundances: the hot-path
part of pyexshalos.simulation (simulation.py:9-311, 467-588, 719-806), same
signatures and defaults.

Extra keyword-only arguments: ``device_out`` (torch device tensors) and,
for Compute_Power_Spectrum, ``kbin_precision`` (4: float bin index as the
reference's FP32 build; 8: double).  Window "TSC" (5) is an addition.
"""

from __future__ import annotations

from typing import Dict, Optional, Union

import numpy as np

_WINDOWS = {"NO": 0, "NGP": 1, "CIC": 2, "SPHERICAL": 3, "EXPONENTIAL": 4, "TSC": 5}


def _window_id(window, allow_none):
    if isinstance(window, str):
        w = window.upper()
        if w not in _WINDOWS:
            raise ValueError(f"unknown window {window!r}")
        wid = _WINDOWS[w]
    else:
        wid = int(window)
    if wid == 0 and not allow_none:
        return None
    return wid


def _direction_id(direction):
    if direction is None or direction == -1:
        return -1
    if direction in ("x", "X", 0):
        return 0
    if direction in ("y", "Y", 1):
        return 1
    if direction in ("z", "Z", 2):
        return 2
    raise ValueError("Direction must be None, x, y or z!")


def Compute_Density_Grid(
    pos: np.ndarray,
    vel: Optional[np.ndarray] = None,
    mass: Optional[np.ndarray] = None,
    types: Optional[np.ndarray] = None,
    nd: int = 256,
    L: float = 1000.0,
    Om0: float = 0.31,
    z: float = 0.0,
    direction: Optional[int] = None,
    window: Union[str, int] = "CIC",
    R: float = 4.0,
    R_times: float = 5.0,
    interlacing: bool = False,
    folds: int = 1,
    verbose: bool = False,
    nthreads: int = 1,
    *,
    device_out: bool = False,
) -> np.ndarray:
    """Density grid of a set of tracers (reference simulation.py:9-166)."""
    from .lib.spectrum import check_precision, grid_compute

    precision = check_precision()
    Rt = np.float32 if precision == 4 else np.float64
    nmass = 0 if mass is None else 1
    if types is not None:
        types = np.asarray(types)
        types = np.fabs(types)
        types = types - np.min(types)
        types = types.astype("int32")
        ntypes = len(np.unique(types))
    else:
        ntypes = 1
    direction = _direction_id(direction)
    wid = _window_id(window, allow_none=False)
    if wid is None:
        print("You need to choose some density assigment method to construct the density grid!")
        return None
    grid = grid_compute(pos, vel, mass, np.int32(nmass), types, np.int32(ntypes), np.int32(nd),
                        Rt(L), Rt(Om0), Rt(z), np.int32(direction), np.int32(wid), Rt(R),
                        Rt(R_times), np.int32(interlacing), np.int32(folds), np.int32(verbose),
                        np.int32(nthreads), device_out=device_out)
    if not interlacing and ntypes == 1:
        grid = grid.reshape([nd, nd, nd])
    elif not interlacing:
        grid = grid.reshape([ntypes, nd, nd, nd])
    elif ntypes == 1:
        grid = grid.reshape([2, nd, nd, nd])
    return grid


def Compute_Power_Spectrum(
    grid: np.ndarray,
    L: float = 1000.0,
    window: Union[str, int] = "CIC",
    R: float = 4.0,
    Nk: int = 25,
    k_min: Optional[float] = None,
    k_max: Optional[float] = None,
    l_max: int = 0,
    direction: Optional[Union[str, int]] = None,
    folds: int = 1,
    verbose: bool = False,
    nthreads: int = 1,
    ntypes: int = 1,
    *,
    kbin_precision: Optional[int] = None,
    device_out: bool = False,
) -> Dict[str, np.ndarray]:
    """Power spectra of density grids (reference simulation.py:170-311)."""
    shape = tuple(grid.shape)
    if len(shape) == 3:
        interlacing = 0
        ntypes = 1
    elif len(shape) == 5:
        interlacing = 1
        ntypes = shape[0]
    elif len(shape) == 4 and ntypes == 1:
        interlacing = 1
    elif len(shape) == 4 and ntypes > 1:
        interlacing = 0
    else:
        raise ValueError("grid must have 3, 4 or 5 dimensions")
    nd = shape[-1]
    L = L / folds
    if k_min is None:
        k_min = 2.0 * np.pi / L
    if k_max is None:
        k_max = np.pi / L * nd
    from .lib.spectrum import check_precision, power_compute

    precision = check_precision()
    Rt = np.float32 if precision == 4 else np.float64
    wid = _window_id(window, allow_none=True)
    direction = _direction_id(direction)
    x = power_compute(grid, np.int32(ntypes), np.int32(nd), Rt(L), np.int32(wid), Rt(R),
                      np.int32(interlacing), np.int32(Nk), Rt(k_min), Rt(k_max), np.int32(l_max),
                      np.int32(direction), np.int32(verbose), np.int32(nthreads),
                      kbin_precision=kbin_precision, device_out=device_out)
    if ntypes == 1 and l_max == 0:
        x["Pk"] = x["Pk"].reshape([Nk])
    elif ntypes == 1:
        x["Pk"] = x["Pk"].reshape([int(l_max / 2) + 1, Nk])
    elif l_max == 0:
        x["Pk"] = x["Pk"].reshape([int(ntypes * (ntypes + 1) / 2), Nk])
    return x


def Compute_Abundance(
    Mh,
    Mmin: Optional[float] = None,
    Mmax: Optional[float] = None,
    Nm: int = 25,
    Lc: float = 2.0,
    nd: int = 256,
    ndx: Optional[int] = None,
    ndy: Optional[int] = None,
    ndz: Optional[int] = None,
    verbose: bool = False,
) -> Dict[str, np.ndarray]:
    """Differential abundance of halo masses (reference simulation.py:719-806):
    dict Mh (mean mass per bin), dn, dn_err.  Mh may be a numpy array, a
    torch device tensor, or None for the last halo catalogue's masses (kept
    on the device)."""
    from .lib.spectrum import abundance_compute, check_precision

    Mmin = -1.0 if Mmin is None else Mmin
    Mmax = -1.0 if Mmax is None else Mmax
    Rt = np.float32 if check_precision() == 4 else np.float64
    ndx = nd if ndx is None else ndx
    ndy = nd if ndy is None else ndy
    ndz = nd if ndz is None else ndz
    return abundance_compute(Mh, Rt(Mmin), Rt(Mmax), np.int32(Nm), Rt(Lc), np.int32(ndx),
                             np.int32(ndy), np.int32(ndz), np.int32(verbose))


def Compute_BiSpectrum(
    grid,
    L: float = 1000.0,
    window: Union[str, int] = "CIC",
    R: float = 4.0,
    Nk: int = 25,
    k_min: Optional[float] = None,
    k_max: Optional[float] = None,
    folds: int = 1,
    verbose: bool = False,
    nthreads: int = 1,
    ntypes: int = 1,
) -> Dict[str, np.ndarray]:
    """Bispectra and shell power spectra of density grids (reference
    simulation.py:467-588): dict kP, Pk, Nk, kB, Bk, Ntri."""
    shape = tuple(grid.shape)
    if len(shape) == 3:
        interlacing, ntypes = 0, 1
    elif len(shape) == 5:
        interlacing, ntypes = 1, shape[0]
    elif len(shape) == 4 and ntypes == 1:
        interlacing = 1
    elif len(shape) == 4 and ntypes > 1:
        interlacing = 0
    else:
        raise ValueError("grid must have 3, 4 or 5 dimensions")
    nd = shape[-1]
    L = L / folds
    if k_min is None:
        k_min = 2.0 * np.pi / L
    if k_max is None:
        k_max = np.pi / L * nd
    from .lib.spectrum import bi_compute, check_precision

    Rt = np.float32 if check_precision() == 4 else np.float64
    wid = _window_id(window, allow_none=True)
    x = bi_compute(grid, np.int32(ntypes), np.int32(nd), Rt(L), np.int32(wid), Rt(R),
                   np.int32(interlacing), np.int32(Nk), Rt(k_min), Rt(k_max), np.int32(verbose),
                   np.int32(nthreads))
    if ntypes == 1:
        x["Pk"] = x["Pk"].reshape([Nk])
        x["Bk"] = x["Bk"].reshape([len(x["Ntri"])])
    return x
