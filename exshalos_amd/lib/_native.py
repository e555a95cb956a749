"""ctypes binding of the HIP library ``libexhip.so`` (C-ABI in include/exhip.h).

The library is loaded after ``torch`` so that both share the HIP runtime
torch already loaded into the process; device buffers are torch tensors,
and every library call runs on torch's current HIP stream.  There is no CPU
fallback: without the built library or without a GPU the calls raise.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libexhip.so")

EXH_OK = 0
ERRORS = {
    -1: ValueError,
    -2: RuntimeError,
    -3: RuntimeError,
    -4: ValueError,
    -5: NotImplementedError,
    -6: RuntimeError,
}

SRC_MODES, SRC_DELTA, SRC_DELTAK = 0, 1, 2
TUNE_DEPOSIT_TILE, TUNE_DEPOSIT_ACC, TUNE_HALO_CLAIMS = 1, 2, 3


class Box(C.Structure):
    _fields_ = [
        ("nd", C.c_int * 3), ("nz2", C.c_int), ("nmin", C.c_int), ("precision", C.c_int),
        ("ng", C.c_longlong),
        ("Om0", C.c_double), ("redshift", C.c_double), ("dc", C.c_double), ("rhoc", C.c_double),
        ("Hz", C.c_double), ("Omz", C.c_double), ("rhomz", C.c_double), ("Dv", C.c_double),
        ("Lc", C.c_double), ("Mcell", C.c_double), ("L", C.c_double * 3), ("Mtot", C.c_double),
        ("kl", C.c_double * 3), ("Normx", C.c_double), ("Normk", C.c_double),
    ]


class FieldSrc(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("data", C.c_void_p), ("k", C.c_void_p), ("P", C.c_void_p),
        ("Nk", C.c_int), ("fixed", C.c_int), ("phase", C.c_double), ("R_max", C.c_double),
    ]


class HaloOpts(C.Structure):
    _fields_ = [
        ("Nmin", C.c_int), ("a", C.c_double), ("beta", C.c_double), ("alpha", C.c_double),
        ("seed", C.c_ulong), ("k", C.c_void_p), ("P", C.c_void_p), ("Nk", C.c_int),
        ("sigma_radii", C.c_int),
    ]


_P = C.c_void_p
_SIGS = {
    "exh_version": (C.c_char_p, []),
    "exh_last_error": (C.c_char_p, []),
    "exh_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "exh_ctx_destroy": (C.c_int, [_P]),
    "exh_ctx_set_stream": (C.c_int, [_P, _P]),
    "exh_ctx_synchronize": (C.c_int, [_P]),
    "exh_ctx_set_timing": (C.c_int, [_P, C.c_int]),
    "exh_ctx_timings": (C.c_int, [_P, C.c_char_p, C.c_int, _P, C.c_int]),
    "exh_ctx_halo_stats": (C.c_int, [_P, _P]),
    "exh_ctx_set_tunable": (C.c_int, [_P, C.c_int, C.c_int]),
    "exh_box_init": (C.c_int, [C.POINTER(Box), C.c_int, C.c_int, C.c_int, C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_int]),
    "exh_density_grid": (C.c_int, [_P, C.POINTER(Box), C.POINTER(FieldSrc), C.c_double, _P, _P]),
    "exh_halo_tables": (C.c_int, [C.POINTER(Box), C.POINTER(HaloOpts), _P, _P, C.c_int, _P, _P,
                                  C.POINTER(C.c_int)]),
    "exh_find_halos": (C.c_int, [_P, C.POINTER(Box), _P, C.POINTER(HaloOpts),
                                 C.POINTER(C.c_longlong)]),
    "exh_halos_get": (C.c_int, [_P, _P, _P, _P, _P, C.c_int]),
    "exh_halos_posh": (C.c_int, [_P, _P, _P, C.c_int]),
    "exh_halos_flag": (C.c_int, [_P, _P, C.c_int]),
    "exh_lpt": (C.c_int, [_P, C.POINTER(Box), C.POINTER(FieldSrc), C.c_double, C.c_int, C.c_int,
                          _P, _P]),
    "exh_pk_adjoint": (C.c_int, [_P, C.POINTER(Box), C.POINTER(FieldSrc), C.c_double, C.c_int,
                                 C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P, _P,
                                 _P]),
    "exh_halos_box_from_pk": (C.c_int, [_P, C.POINTER(Box), C.POINTER(FieldSrc),
                                        C.POINTER(HaloOpts), C.c_double, C.c_int, C.c_int, _P,
                                        _P, _P, C.POINTER(C.c_longlong)]),
    "exh_halos_box_power": (C.c_int, [_P, C.POINTER(Box), C.POINTER(FieldSrc),
                                      C.POINTER(HaloOpts), C.c_double, C.c_int, _P, C.c_int,
                                      C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _P, _P,
                                      _P, _P, C.POINTER(C.c_longlong)]),
    "exh_halos_box_from_grid": (C.c_int, [_P, C.POINTER(Box), _P, C.POINTER(HaloOpts),
                                          C.c_double, C.c_int, C.c_int, _P, _P, _P, _P,
                                          C.POINTER(C.c_longlong)]),
    "exh_grid_compute": (C.c_int, [_P, _P, _P, _P, _P, C.c_longlong, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _P]),
    "exh_grid_compute_noalloc": (C.c_int, [_P, _P, _P, _P, _P, C.c_longlong, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_double, C.c_int,
                                           C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_int, _P]),
    "exh_power_compute": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                    C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                    C.c_int, C.c_int, _P, _P, _P]),
    "exh_abundance_compute": (C.c_int, [_P, _P, C.c_longlong, C.c_double, C.c_double, C.c_int,
                                        C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P,
                                        _P]),
    "exh_bispectrum_size": (C.c_int, [C.c_int]),
    "exh_bispectrum_compute": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_double, C.c_int,
                                         C.c_double, C.c_int, C.c_int, C.c_double, C.c_double,
                                         C.c_int, C.c_longlong, _P, _P, _P, _P, _P, _P,
                                         C.POINTER(C.c_int)]),
}

_lib = None
_lock = threading.Lock()
_ctx = {}


def exported_symbols():
    return list(_SIGS)


def load():
    """Load libexhip.so (after torch) and declare every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        import torch  # noqa: F401  (loads the HIP runtime the library binds to)

        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m exshalos_amd.build` "
                "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().exh_last_error().decode()


def check(rc: int, what: str = ""):
    if rc != EXH_OK:
        exc = ERRORS.get(rc, RuntimeError)
        raise exc(f"{what}: {last_error()}" if what else last_error())


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("exshalos_amd needs a HIP device (MI355X); none is visible")


def context(device: int | None = None):
    """The library context of `device`, bound to torch's current stream."""
    import torch

    require_gpu()
    lib = load()
    dev = torch.cuda.current_device() if device is None else int(device)
    ctx = _ctx.get(dev)
    if ctx is None:
        h = C.c_void_p()
        check(lib.exh_ctx_create(dev, C.byref(h)), "exh_ctx_create")
        ctx = h
        _ctx[dev] = ctx
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib.exh_ctx_set_stream(ctx, C.c_void_p(stream)), "exh_ctx_set_stream")
    return ctx


def box(nd, Lc, Om0=0.31, z=0.0, dc=-1.0, precision=4) -> Box:
    nd3 = (nd, nd, nd) if np.isscalar(nd) else tuple(int(v) for v in nd)
    b = Box()
    check(load().exh_box_init(C.byref(b), nd3[0], nd3[1], nd3[2], float(Lc), float(Om0),
                              float(z), float(dc), int(precision)), "exh_box_init")
    return b


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array, or None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return C.c_void_p(t.ctypes.data)
    return C.c_void_p(t.data_ptr())


def timings(ctx):
    lib = load()
    names = C.create_string_buffer(1 << 16)
    ms = (C.c_double * 4096)()
    n = lib.exh_ctx_timings(ctx, names, len(names), ms, 4096)
    if n < 0:
        check(n, "exh_ctx_timings")
    labels = names.value.decode().split("\n")[:n]
    return list(zip(labels, [ms[i] for i in range(n)]))


def halo_stats(ctx):
    out = (C.c_longlong * 5)()
    check(load().exh_ctx_halo_stats(ctx, out), "exh_ctx_halo_stats")
    return dict(peaks=out[0], candidates=out[1], rounds=out[2], sort_levels=out[3],
                sort_doubling=out[4])
