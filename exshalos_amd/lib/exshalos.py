"""Mirror of the reference's ``pyexshalos.lib.exshalos`` extension module.

Same function names, positional argument order and output dictionaries as
the CPython methods in src/exshalos/exshalos.c:27-38; the work runs in the
gfx950 library (libexhip.so) on the current torch HIP device.

Additions (keyword-only, additive):
  noise=       explicit unit white-noise modes [nd][nd][nd//2+1][2] (numpy or
               device tensor).  Without it the modes are drawn from
               ``numpy.random.default_rng(seed)`` (exshalos_amd.synth); the
               reference draws them from GSL's taus + gsl_ran_gaussian,
               a stream this library does not reproduce.
  device_out=  return torch device tensors instead of numpy arrays.
Divergences (documented): every requested output is returned for any flag
combination (the reference returns {} for combinations it does not list,
exshalos.c:865-937); OUT_PROF raises NotImplementedError; the flag grid of
find_halos has shape [nd, nd, nd]
(exshalos.c:558 passes a pointer as the dimension).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._common import get_precision, host_f64, np_real, to_device, to_numpy


def check_precision() -> int:
    """exshalos.c:41-43: bytes per fft_real (4 float, 8 double)."""
    return get_precision()


def _noise_for(nd3, seed, prec, noise):
    from ..synth import white_noise

    if noise is None:
        noise = white_noise(nd3[2], int(seed), np_real(prec), ndx=nd3[0], ndy=nd3[1])
    t = to_device(noise, np_real(prec))
    want = (nd3[0], nd3[1], nd3[2] // 2 + 1, 2)
    if tuple(t.shape) != want:
        raise ValueError(f"noise must have shape {want}, got {tuple(t.shape)}")
    return t


def _modes_src(kh, Ph, noise_t, fixed, phase, R_max=100000.0):
    """White-noise modes with P(k); R_max < 1e5 cuts xi(r) beyond R_max
    (Inter_Power, density_grid.c:52-59)."""
    s = N.FieldSrc()
    s.kind = N.SRC_MODES
    s.data = noise_t.data_ptr()
    s.k = kh.ctypes.data
    s.P = Ph.ctypes.data
    s.Nk = len(kh)
    s.fixed = int(bool(fixed))
    s.phase = float(phase)
    s.R_max = float(R_max)
    return s


def _halo_opts(kh, Ph, Nmin, a, beta, alpha, seed, prec, sigma_radii=0):
    """sigma_radii > 0: the FP64 sigma(R) table over that many radii (an
    option of this library; 0 is the reference's table, one radius per k)."""
    o = N.HaloOpts()
    R = np_real(prec)
    o.Nmin = int(Nmin)
    o.a, o.beta, o.alpha = float(R(a)), float(R(beta)), float(R(alpha))
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o.k = kh.ctypes.data
    o.P = Ph.ctypes.data
    o.Nk = len(kh)
    o.sigma_radii = int(sigma_radii)
    return o


def _out(t, device_out):
    return t if device_out else to_numpy(t)


def density_grid_compute(k, P, R_max, Ndx, Ndy, Ndz, Lc, outk, seed, fixed, phase, k_smooth,
                         verbose, nthreads, *, noise=None, device_out=False):
    """Gaussian density grid (exshalos.c:46-121, density_grid.c:71-217)."""
    import torch

    prec = get_precision()
    ctx = N.context()
    nd3 = (int(Ndx), int(Ndy), int(Ndz))
    b = N.box(nd3, Lc, 0.31, 0.0, 1.686, prec)
    kh, Ph = host_f64(k), host_f64(P)
    noise_t = _noise_for(nd3, seed, prec, noise)
    src = _modes_src(kh, Ph, noise_t, fixed, phase, np_real(prec)(R_max))
    dt = torch.float32 if prec == 4 else torch.float64
    grid = torch.empty(nd3, dtype=dt, device=noise_t.device)
    gridk = torch.empty((nd3[0], nd3[1], nd3[2] // 2 + 1, 2), dtype=dt, device=noise_t.device) \
        if outk else None
    N.check(N.load().exh_density_grid(ctx, C.byref(b), C.byref(src), float(k_smooth),
                                      N.ptr(grid), N.ptr(gridk)), "density_grid_compute")
    out = {"grid": _out(grid, device_out)}
    if outk:
        out["gridk"] = _out(gridk, device_out)
    return out


def _catalogue(ctx, prec, want_posh=True, want_vel=False, device_out=False):
    import torch

    lib = N.load()
    nh = int(_last["nh"])
    dt = torch.float32 if prec == 4 else torch.float64
    dev = torch.device("cuda", torch.cuda.current_device())
    x = torch.empty((nh, 3), dtype=torch.int32, device=dev)
    cnt = torch.empty((nh,), dtype=torch.int32, device=dev)
    Mh = torch.empty((nh,), dtype=dt, device=dev)
    N.check(lib.exh_halos_get(ctx, N.ptr(x), N.ptr(cnt), None, N.ptr(Mh), 0), "halos_get")
    out = {"x": x, "count": cnt, "Mh": Mh}
    if want_posh:
        posh = torch.empty((nh, 3), dtype=dt, device=dev)
        velh = torch.empty((nh, 3), dtype=dt, device=dev) if want_vel else None
        N.check(lib.exh_halos_posh(ctx, N.ptr(posh), N.ptr(velh), 0), "halos_posh")
        out["posh"] = posh
        if want_vel:
            out["velh"] = velh
    return out


_last = {"nh": 0, "prec": 4}


def _flag(ctx, nd3, device_out):
    import torch

    f = torch.empty(nd3, dtype=torch.int64, device=torch.device("cuda", torch.cuda.current_device()))
    N.check(N.load().exh_halos_flag(ctx, N.ptr(f), 0), "halos_flag")
    return _out(f, device_out)


def find_halos(delta, k, P, Lc, Om0, redshift, dc, Nmin, a, beta, alpha, OUT_FLAG, OUT_PROF,
               verbose, *, device_out=False, sigma_radii=0):
    """Halo catalogue of a density grid (exshalos.c:509-610); Lagrangian
    positions Lc (x + 1/2), mass-draw seed 12345 (exshalos.c:552)."""
    if OUT_PROF:
        raise NotImplementedError("OUT_PROF (Lagrangian profiles) is not implemented")
    prec = get_precision()
    ctx = N.context()
    d = to_device(delta, np_real(prec))
    if d.dim() != 3:
        raise ValueError("delta must be a 3D grid")
    nd3 = tuple(int(v) for v in d.shape)
    b = N.box(nd3, Lc, Om0, redshift, dc, prec)
    kh, Ph = host_f64(k), host_f64(P)
    o = _halo_opts(kh, Ph, Nmin, a, beta, alpha, 12345, prec, sigma_radii)
    nh = C.c_longlong()
    N.check(N.load().exh_find_halos(ctx, C.byref(b), N.ptr(d), C.byref(o), C.byref(nh)),
            "find_halos")
    _last.update(nh=nh.value, prec=prec)
    cat = _catalogue(ctx, prec, True, False)
    out = {"posh": _out(cat["posh"], device_out), "Mh": _out(cat["Mh"], device_out)}
    if OUT_FLAG:
        out["flag"] = _flag(ctx, nd3, device_out)
    return out


def lpt_compute(delta, Lc, Om0, redshift, k_smooth, DO_2LPT, OUT_VEL, Ink, OUT_POS, verbose, *,
                device_out=False):
    """LPT displacements of the grid's particles (exshalos.c:613-707)."""
    import torch

    prec = get_precision()
    ctx = N.context()
    d = to_device(delta, np_real(prec))
    if Ink:
        if d.dim() != 4 or d.shape[-1] != 2:
            raise ValueError("Ink grids are [nd][nd][nd/2+1][2]")
        nd3 = (int(d.shape[0]), int(d.shape[1]), 2 * (int(d.shape[2]) - 1))
        kind = N.SRC_DELTAK
    else:
        nd3 = tuple(int(v) for v in d.shape)
        kind = N.SRC_DELTA
    b = N.box(nd3, Lc, Om0, redshift, -1.0, prec)
    s = N.FieldSrc()
    s.kind = kind
    s.data = d.data_ptr()
    ng = nd3[0] * nd3[1] * nd3[2]
    dt = torch.float32 if prec == 4 else torch.float64
    S = torch.empty((ng, 3), dtype=dt, device=d.device)
    V = torch.empty((ng, 3), dtype=dt, device=d.device) if OUT_VEL else None
    N.check(N.load().exh_lpt(ctx, C.byref(b), C.byref(s), float(k_smooth), int(bool(DO_2LPT)),
                             int(bool(OUT_POS)), N.ptr(S), N.ptr(V)), "lpt_compute")
    out = {"pos": _out(S, device_out)}
    if OUT_VEL:
        out["vel"] = _out(V, device_out)
    return out


def halos_box_from_pk(k, P, R_max, Ndx, Ndy, Ndz, Lc, seed, k_smooth, Om0, redshift, dc, Nmin,
                      a, beta, alpha, fixed, phase, OUT_DEN, OUT_LPT, OUT_VEL, DO_2LPT, OUT_FLAG,
                      OUT_PROF, verbose, nthreads, *, noise=None, device_out=False,
                      sigma_radii=0):
    """Halos (and particles) from a linear P(k) (exshalos.c:710-940)."""
    import torch

    if OUT_PROF:
        raise NotImplementedError("OUT_PROF (Lagrangian profiles) is not implemented")
    prec = get_precision()
    ctx = N.context()
    nd3 = (int(Ndx), int(Ndy), int(Ndz))
    b = N.box(nd3, Lc, Om0, redshift, dc, prec)
    kh, Ph = host_f64(k), host_f64(P)
    noise_t = _noise_for(nd3, seed, prec, noise)
    src = _modes_src(kh, Ph, noise_t, fixed, phase, np_real(prec)(R_max))
    o = _halo_opts(kh, Ph, Nmin, a, beta, alpha, seed, prec, sigma_radii)
    dt = torch.float32 if prec == 4 else torch.float64
    ng = nd3[0] * nd3[1] * nd3[2]
    dev = noise_t.device
    grid = torch.empty(nd3, dtype=dt, device=dev) if OUT_DEN else None
    S = torch.empty((ng, 3), dtype=dt, device=dev) if OUT_LPT else None
    V = torch.empty((ng, 3), dtype=dt, device=dev) if (OUT_LPT and OUT_VEL) else None
    nh = C.c_longlong()
    N.check(N.load().exh_halos_box_from_pk(ctx, C.byref(b), C.byref(src), C.byref(o),
                                           float(k_smooth), int(bool(DO_2LPT)),
                                           int(bool(OUT_VEL)), N.ptr(grid), N.ptr(S), N.ptr(V),
                                           C.byref(nh)), "halos_box_from_pk")
    _last.update(nh=nh.value, prec=prec)
    cat = _catalogue(ctx, prec, True, bool(OUT_VEL))
    out = {"posh": _out(cat["posh"], device_out), "Mh": _out(cat["Mh"], device_out)}
    if OUT_VEL:
        out["velh"] = _out(cat["velh"], device_out)
    if OUT_LPT:
        out["pos"] = _out(S, device_out)
        if OUT_VEL:
            out["vel"] = _out(V, device_out)
    if OUT_FLAG:
        out["flag"] = _flag(ctx, nd3, device_out)
    if OUT_DEN:
        out["grid"] = _out(grid, device_out)
    return out


def halos_box_from_grid(k, P, grid, S, V, Lc, k_smooth, Om0, redshift, dc, Nmin, a, beta, alpha,
                        OUT_LPT, OUT_VEL, DO_2LPT, OUT_FLAG, IN_disp, OUT_PROF, verbose, nthreads,
                        *, device_out=False, sigma_radii=0):
    """Halos from a density grid (exshalos.c:943-1103); mass-draw seed 1234."""
    import torch

    if OUT_PROF:
        raise NotImplementedError("OUT_PROF (Lagrangian profiles) is not implemented")
    prec = get_precision()
    ctx = N.context()
    R = np_real(prec)
    d = to_device(grid, R)
    nd3 = tuple(int(v) for v in d.shape)
    b = N.box(nd3, Lc, Om0, redshift, dc, prec)
    kh, Ph = host_f64(k), host_f64(P)
    o = _halo_opts(kh, Ph, Nmin, a, beta, alpha, 1234, prec, sigma_radii)
    dt = torch.float32 if prec == 4 else torch.float64
    ng = nd3[0] * nd3[1] * nd3[2]
    S_in = to_device(S, R) if IN_disp else None
    V_in = to_device(V, R) if (IN_disp and OUT_VEL and V is not None) else None
    S_out = torch.empty((ng, 3), dtype=dt, device=d.device) if (OUT_LPT and not IN_disp) else None
    V_out = torch.empty((ng, 3), dtype=dt, device=d.device) \
        if (OUT_LPT and OUT_VEL and not IN_disp) else None
    nh = C.c_longlong()
    N.check(N.load().exh_halos_box_from_grid(ctx, C.byref(b), N.ptr(d), C.byref(o),
                                             float(k_smooth), int(bool(DO_2LPT)),
                                             int(bool(OUT_VEL)), N.ptr(S_in), N.ptr(V_in),
                                             N.ptr(S_out), N.ptr(V_out), C.byref(nh)),
            "halos_box_from_grid")
    _last.update(nh=nh.value, prec=prec)
    want_vel = bool(OUT_VEL) and (not IN_disp or V_in is not None)
    cat = _catalogue(ctx, prec, True, want_vel)
    out = {"posh": _out(cat["posh"], device_out), "Mh": _out(cat["Mh"], device_out)}
    if want_vel:
        out["velh"] = _out(cat["velh"], device_out)
    if OUT_LPT and not IN_disp:
        out["pos"] = _out(S_out, device_out)
        if OUT_VEL:
            out["vel"] = _out(V_out, device_out)
    if OUT_FLAG:
        out["flag"] = _flag(ctx, nd3, device_out)
    return out


def halo_catalogue(device_out=False):
    """Centres (cell indices), cell counts and masses of the last halo run
    on this device (an addition; the reference keeps them internal)."""
    ctx = N.context()
    cat = _catalogue(ctx, _last["prec"], False)
    return {k: _out(v, device_out) for k, v in cat.items()}
