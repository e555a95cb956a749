"""Mirror of the reference's ``pyexshalos.lib.spectrum`` extension module
(hot-path methods of src/spectrum/spectrum.c:28-38): check_precision,
grid_compute, power_compute, bi_compute, abundance_compute, with the same
positional arguments.

Additions (keyword-only): ``device_out`` (return torch device tensors) and
``kbin_precision`` for power_compute (4: the reference's float bin index,
the default in FP32 mode; 8: double).  Inputs may be numpy arrays or torch
device tensors.  Divergences (documented): window 5 is TSC (not in the
reference); positions are
not modified in place by the redshift-space shift (gridmodule.c:167-169
edits the caller's array); type indices must lie in [0, ntype).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._common import get_precision, np_real, to_device, to_numpy


def check_precision() -> int:
    """spectrum.c:41-43."""
    return get_precision()


def grid_compute(pos, vel, mass, nmass, type, ntype, nd, L, Om0, z, direction, window, R,
                 R_times, interlacing, folds, verbose, nthreads, *, device_out=False):
    """Density grids [ntype][1+interlacing][nd][nd][nd] (spectrum.c:46-108)."""
    import torch

    prec = get_precision()
    ctx = N.context()
    Rt = np_real(prec)
    p = to_device(pos, Rt)
    if p.dim() != 2 or p.shape[1] != 3:
        raise ValueError("pos must have shape (N, 3)")
    npart = int(p.shape[0])
    ntype = int(ntype)
    direction = int(direction)
    v = to_device(vel, Rt) if direction != -1 else None
    m = to_device(mass, Rt) if nmass else None
    t = to_device(type, np.int32) if ntype > 1 else None
    if t is not None:
        if int(t.min()) < 0 or int(t.max()) >= ntype:
            raise ValueError("tracer types must lie in [0, ntype)")
    il = int(bool(interlacing))
    grid = torch.empty((ntype, 1 + il, int(nd), int(nd), int(nd)),
                       dtype=torch.float32 if prec == 4 else torch.float64, device=p.device)
    N.check(N.load().exh_grid_compute(ctx, N.ptr(p), N.ptr(v), N.ptr(m), N.ptr(t), npart, ntype,
                                      int(nd), float(L), float(Om0), float(z), direction,
                                      int(window), float(R), float(R_times), il, int(folds), prec,
                                      N.ptr(grid)),
            "grid_compute")
    return grid if device_out else to_numpy(grid)


def power_compute(grid, ntype, nd, L, window, R, interlacing, Nk, k_min, k_max, l_max, direction,
                  verbose, nthreads, *, kbin_precision=None, device_out=False):
    """All auto and cross spectra and multipoles (spectrum.c:111-197)."""
    prec = get_precision()
    ctx = N.context()
    g = to_device(grid, np_real(prec))
    ntype, nd, Nk, l_max = int(ntype), int(nd), int(Nk), int(l_max)
    il = int(bool(interlacing))
    if g.numel() != ntype * (1 + il) * nd**3:
        raise ValueError("grid size does not match ntype, interlacing and nd")
    NPs = ntype * (ntype + 1) // 2
    ls = l_max // 2 + 1
    kout = np.zeros(Nk, dtype=np.float64)
    Pout = np.zeros((NPs, ls, Nk), dtype=np.float64)
    cnt = np.zeros(Nk, dtype=np.int64)
    kb = prec if kbin_precision is None else int(kbin_precision)
    N.check(N.load().exh_power_compute(ctx, N.ptr(g), ntype, nd, float(L), int(window), float(R),
                                       il, Nk, float(k_min), float(k_max), l_max, int(direction),
                                       prec, kb, N.ptr(kout), N.ptr(Pout), N.ptr(cnt)),
            "power_compute")
    Rt = np_real(prec)
    return {"k": kout.astype(Rt), "Pk": Pout.astype(Rt), "Nk": cnt}


def abundance_compute(Mh_array, Mmin, Mmax, Nm, Lc, ndx, ndy, ndz, verbose):
    """Differential halo abundance (spectrum.c:535-600): dict Mh (mean mass
    per bin), dn, dn_err.  Mh_array: numpy or torch device masses; None takes
    the masses of the last halo catalogue, which stay on the device."""
    prec = get_precision()
    ctx = N.context()
    Nm = int(Nm)
    if Mh_array is None:
        m, nh = None, 0
    else:
        m = to_device(Mh_array, np_real(prec)).reshape(-1)
        nh = int(m.numel())
    Mmean = np.zeros(Nm, dtype=np.float64)
    dn = np.zeros(Nm, dtype=np.float64)
    dn_err = np.zeros(Nm, dtype=np.float64)
    N.check(N.load().exh_abundance_compute(ctx, N.ptr(m), nh, float(Mmin), float(Mmax), Nm,
                                           float(Lc), int(ndx), int(ndy), int(ndz), prec,
                                           N.ptr(Mmean), N.ptr(dn), N.ptr(dn_err)),
            "abundance_compute")
    Rt = np_real(prec)
    return {"Mh": Mmean.astype(Rt), "dn": dn.astype(Rt), "dn_err": dn_err.astype(Rt)}


def bi_compute(grid, ntype, nd, L, window, R, interlacing, Nk, k_min, k_max, verbose, nthreads):
    """Bispectra of all tracer triples and the shell power spectra
    (spectrum.c:291-406): dict kP, Pk [NPs][Nk], Nk, kB [3][Ntri],
    Bk [ntype^3][Ntri], Ntri."""
    prec = get_precision()
    ctx = N.context()
    g = to_device(grid, np_real(prec))
    ntype, nd, Nk = int(ntype), int(nd), int(Nk)
    il = int(bool(interlacing))
    if g.numel() != ntype * (1 + il) * nd**3:
        raise ValueError("grid size does not match ntype, interlacing and nd")
    lib = N.load()
    ntri = lib.exh_bispectrum_size(Nk)
    N.check(min(ntri, 0), "bispectrum_size")
    NPs, NBs = ntype * (ntype + 1) // 2, ntype**3
    kP = np.zeros(Nk)
    Pk = np.zeros((NPs, Nk))
    cnt = np.zeros(Nk)
    kB = np.zeros((3, ntri))
    Bk = np.zeros((NBs, ntri))
    ctri = np.zeros(ntri)
    got = C.c_int()
    N.check(lib.exh_bispectrum_compute(ctx, N.ptr(g), ntype, nd, float(L), int(window), float(R),
                                       il, Nk, float(k_min), float(k_max), prec, ntri,
                                       N.ptr(kP), N.ptr(Pk), N.ptr(cnt), N.ptr(kB), N.ptr(Bk),
                                       N.ptr(ctri), C.byref(got)),
            "bi_compute")
    Rt = np_real(prec)
    return {"kP": kP.astype(Rt), "Pk": Pk.astype(Rt), "Nk": cnt.astype(Rt), "kB": kB.astype(Rt),
            "Bk": Bk.astype(Rt), "Ntri": ctri.astype(Rt)}
