"""CPU oracle driver (TEST INFRASTRUCTURE ONLY).

Runs the reference's algorithms for the field -> LPT -> halos -> grid ->
P(k) path on the CPU, stage by stage, through the C restatement in
``oracle/exh_oracle.c`` (built into ``oracle/liboracle.so`` by
``oracle/Makefile``), with the 3D FFTs done by ``numpy.fft`` in float64.
Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module; the product never does.

``precision=4`` follows the reference's single-precision build (float
fields with its double/long-double islands, file:line citations in the C
source); ``precision=8`` is the FP64 oracle.

Entry points mirror the reference's C functions:

* ``pk_truncate``        -- Inter_Power's FFTLog truncation (density_grid.c:52-59,
                            fftlog.c:49-157)
* ``field_from_modes``   -- Compute_Den (density_grid.c:71-217)
* ``halo_tables``        -- Compute_Sig / Compute_nh / Barrier (find_halos.c:4-93, 200-212)
* ``find_halos``         -- Find_Halos (find_halos.c:479-563)
* ``lpt``                -- Compute_1LPT / Compute_2LPT / Compute_Pos (lpt.c)
* ``halos_box_from_pk``  -- Generate_Halos_Box_from_Pk (box.c:4-68)
* ``grid_compute``       -- Tracer_Grid (gridmodule.c:146-375)
* ``power_compute``      -- Compute_gridk + Power_Spectrum (spectrum_h.c:83-167,
                            powermodule.c:4-108)
* ``abundance``          -- abundance_compute / Measure_Abundance (spectrum.c:535-600,
                            abundance.c:4-36)
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_OMP_PATH = os.path.join(_HERE, "liboracle_omp.so")
_lib = None
_lib_omp = None
_use_omp = False

M_MAX = 1e16


def build() -> str:
    """Compile the oracle library with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib, _lib_omp
    if _use_omp:
        if _lib_omp is None:
            if not os.path.exists(_OMP_PATH):
                build()
            _lib_omp = C.CDLL(_OMP_PATH)
        return _lib_omp
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def use_openmp(on: bool) -> None:
    """Route the oracle's calls to the OpenMP build (the CPU baseline: the
    deposit, deconvolution, binning and normalisation loops threaded as the
    reference threads them, OMP_NUM_THREADS threads) or to the sequential
    parity build."""
    global _use_omp
    _use_omp = bool(on)


def _real(precision: int):
    return np.float32 if precision == 4 else np.float64


def _sfx(precision: int) -> str:
    return "_f32" if precision == 4 else "_f64"


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def _params_struct(precision: int):
    R = C.c_float if precision == 4 else C.c_double

    class Params(C.Structure):
        _fields_ = [
            ("Om0", R), ("redshift", R), ("dc", R), ("rhoc", R), ("Hz", R),
            ("Omz", R), ("rhomz", R), ("Dv", R),
            ("Lc", R), ("Mcell", R), ("L", R * 3), ("Mtot", R), ("kl", R * 3),
            ("Normx", R), ("Normk", R),
            ("nd", C.c_int * 3), ("nz2", C.c_int), ("nmin", C.c_int),
            ("ng", C.c_longlong),
        ]

    return Params


_PARAMS = {4: _params_struct(4), 8: _params_struct(8)}


def params(nd, Lc, Om0=0.31, z=0.0, dc=-1.0, precision=8):
    """Derived cosmology/box parameters (exshalos_h.c:43-106)."""
    nd3 = (nd, nd, nd) if np.isscalar(nd) else tuple(nd)
    p = _PARAMS[precision]()
    R = _real(precision)
    fn = getattr(lib(), "ora_set_params" + _sfx(precision))
    RC = C.c_float if precision == 4 else C.c_double
    rc = fn(C.byref(p), C.c_int(nd3[0]), C.c_int(nd3[1]), C.c_int(nd3[2]),
            RC(R(Lc)), RC(R(Om0)), RC(R(z)), RC(R(dc)))
    if rc != 0:
        raise ValueError("Lc must be a cell size < 100 Mpc/h or a cell mass > 1e5 Msun/h")
    return p


def _rc(precision):
    return C.c_float if precision == 4 else C.c_double


def spline_eval(x, y, xq):
    """Natural cubic spline through (x, y) evaluated at xq (the oracle's
    restatement of GSL's cspline, density_grid.c:62)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    xq = np.ascontiguousarray(xq, dtype=np.float64)
    out = np.zeros(len(xq), dtype=np.float64)
    rc = lib().ora_spline_eval_many_f64(_ptr(x), _ptr(y), C.c_int(len(x)), _ptr(xq),
                                        _ptr(out), C.c_int(len(xq)))
    if rc < 0:
        raise RuntimeError("oracle spline failed")
    if rc == 1:
        raise ValueError("spline evaluated outside its knots")
    return out


# --------------------------------------------------------------------------
# P(k) truncation in configuration space (FFTLog)
# --------------------------------------------------------------------------
def _fht(x, a, mu):
    """Hamilton's (2000) FFTLog fast Hankel transform of order mu, bias 0, on
    the log-spaced grid x (fftlog.c:93-131): low-ringing k r (fftlog.c:49-61),
    coefficients u_m = (k0 r0)^(-2 pi i m / L) U_mu(2 pi i m / L) with
    U_mu(z) = 2^z Gamma((mu+1+z)/2) / Gamma((mu+1-z)/2), the convolution by
    FFTs, the reversal, and the output grid (k0 r0 / x_0) e^(n L / N).
    Returns (y, b)."""
    from scipy.special import loggamma

    n = len(x)
    L = np.log(x[-1] / x[0]) * n / (n - 1.0)
    h = (mu + 1.0) / 2.0
    # low-ringing kr: ln(2/kr) N/L + 2 arg Gamma(h + i pi N/(2L)) / pi an integer
    kr = 1.0
    ga = np.imag(loggamma(h + 1j * np.pi * n / (2.0 * L)))
    arg = np.log(2.0 / kr) * n / L + 2.0 * ga / np.pi
    kr *= np.exp((arg - np.round(arg)) * L / n)
    k0r0 = kr * np.exp(-L)
    m = np.arange(n // 2 + 1)
    y = np.pi / L
    u = np.exp(1j * (-2.0 * m * y * np.log(k0r0 / 2.0) + 2.0 * np.imag(loggamma(h + 1j * m * y))))
    full = np.concatenate([u, np.conj(u[1:n - n // 2][::-1])])
    if n % 2 == 0:
        full[n // 2] = full[n // 2].real
    b = np.fft.ifft(np.fft.fft(a) * full)[::-1]
    out = k0r0 / x[0] * np.exp(np.arange(n) * L / n)
    return out, b


def pk_truncate(k, P, R_max):
    """Inter_Power's truncation (density_grid.c:52-59, fftlog.c:133-157):
    xi(r) of P(k) by FFTLog (l = 0, k^(3/2) weights), xi = 0 for r > R_max,
    then P(k) again on the transform's k grid.  Returns (k', P') in double."""
    k = np.asarray(k, dtype=np.float64)
    P = np.asarray(P, dtype=np.float64)
    r, b = _fht(k, k**1.5 * P, 0.5)
    xi = np.real((2.0 * np.pi * r) ** -1.5 * b)
    xi[r > R_max] = 0.0
    k2, b2 = _fht(r, r**1.5 * xi, 0.5)
    P2 = np.real((2.0 * np.pi * k2) ** -1.5 * b2) * (2.0 * np.pi) ** 3
    return k2, P2, r, xi


# --------------------------------------------------------------------------
# field
# --------------------------------------------------------------------------
def field_modes(p, k, P, noise, precision=8, k_smooth=1e5, fixed=False, phase=0.0,
                R_max=1e5):
    """Compute_Den's modes; R_max < 1e5 truncates xi(r) first (Inter_Power),
    splining the double-precision result as the reference does."""
    R = _real(precision)
    nd = tuple(p.nd)
    k64 = np.ascontiguousarray(np.asarray(k, dtype=R).astype(np.float64))
    P64 = np.ascontiguousarray(np.asarray(P, dtype=R).astype(np.float64))
    if R(R_max) < 100000.0:
        k64, P64, _, _ = pk_truncate(k64, P64, float(R(R_max)))
        k64, P64 = np.ascontiguousarray(k64), np.ascontiguousarray(P64)
    noise = np.ascontiguousarray(noise, dtype=R)
    assert noise.shape == (nd[0], nd[1], nd[2] // 2 + 1, 2)
    deltak = np.zeros_like(noise)
    rc = getattr(lib(), "ora_field_modes" + _sfx(precision))(
        C.byref(p), _ptr(k64), _ptr(P64), C.c_int(len(k64)), _ptr(noise),
        _rc(precision)(k_smooth), C.c_int(int(fixed)), _rc(precision)(phase),
        _rc(precision)(R(R_max)), _ptr(deltak))
    if rc < 0:
        raise RuntimeError("oracle field_modes failed")
    if rc == 1:
        raise ValueError("|k| of the grid lies outside the tabulated P(k)")
    return deltak


def _fft_workers() -> int:
    """Threads of the oracle's FFTs (pocketfft through scipy.fft): at most 16,
    the CPU share of one GPU box."""
    return max(1, min(16, os.cpu_count() or 1))


def c2r(ck: np.ndarray, nd) -> np.ndarray:
    """Unnormalised complex-to-real 3D transform (FFTW c2r convention), in
    float64."""
    import scipy.fft as sf

    z = ck[..., 0].astype(np.float64) + 1j * ck[..., 1].astype(np.float64)
    n = int(np.prod(nd))
    out = sf.irfftn(z, s=tuple(nd), axes=(0, 1, 2), workers=_fft_workers())
    del z
    out *= n
    return out


def r2c(x: np.ndarray) -> np.ndarray:
    """Unnormalised real-to-complex 3D transform in float64, returned as
    [..., 2]."""
    import scipy.fft as sf

    z = sf.rfftn(x.astype(np.float64), axes=(0, 1, 2), workers=_fft_workers())
    out = np.empty(z.shape + (2,), dtype=np.float64)
    out[..., 0] = z.real
    out[..., 1] = z.imag
    return out


def field_from_modes(k, P, nd, Lc, noise, Om0=0.31, z=0.0, precision=8, k_smooth=1e5,
                     fixed=False, phase=0.0, R_max=1e5):
    """Compute_Den: returns (delta [nd]^3, deltak [nd][nd][nd/2+1][2], params)."""
    R = _real(precision)
    p = params(nd, Lc, Om0, z, -1.0, precision)
    dk = field_modes(p, k, P, noise, precision, k_smooth, fixed, phase, R_max)
    delta = (c2r(dk, tuple(p.nd)).astype(R) * R(p.Normx)).astype(R)
    return delta, dk, p


# --------------------------------------------------------------------------
# halo finder
# --------------------------------------------------------------------------
def cells_in_spheres(smax=10000):
    cap = 20000
    out = np.zeros(cap, dtype=np.int32)
    m = lib().ora_cells_in_spheres(_ptr(out), C.c_int(cap), C.c_int(smax))
    return out[:m].copy()


def halo_tables(p, k, P, a=1.0, beta=0.0, alpha=0.0, precision=8, nB=None, sigma_radii=0):
    """sigma^2 per cell count, barrier table, and the mass-function knots.

    sigma_radii > 0 selects the FP64 table: that many radii, sigma^2 and the
    Tinker integral in double (the knots then number sigma_radii)."""
    R = _real(precision)
    k = np.ascontiguousarray(k, dtype=R)
    P = np.ascontiguousarray(P, dtype=R)
    a, beta, alpha = R(a), R(beta), R(alpha)
    do_eb = 0 if (a == 1.0 and beta == 0.0 and alpha == 0.0) else 1
    Ncells = int(np.floor(M_MAX / float(p.Mcell)))
    if nB is None:
        nB = max(Ncells, 1) + 8192
    sig = np.zeros(max(Ncells, 1), dtype=R)
    Bt = np.zeros(nB, dtype=R)
    Nr = sigma_radii if sigma_radii > 0 else len(k)
    Mk = np.zeros(Nr, dtype=np.float64)
    Int = np.zeros(Nr, dtype=np.float64)
    RC = _rc(precision)
    rc = getattr(lib(), "ora_halo_tables" + _sfx(precision))(
        C.byref(p), _ptr(k), _ptr(P), C.c_int(len(k)), C.c_int(do_eb), RC(a), RC(beta),
        RC(alpha), _ptr(sig), _ptr(Bt), C.c_int(nB), _ptr(Mk), _ptr(Int), C.c_int(sigma_radii))
    if rc < 0:
        raise ValueError("sigma(M) spline evaluated outside its knots")
    return dict(Ncells=rc, Sig_grid=sig, Bt=Bt, Mknots=Mk, Int=Int, do_eb=do_eb)


def find_peaks(delta, precision=8):
    R = _real(precision)
    delta = np.ascontiguousarray(delta, dtype=R)
    nd = (C.c_int * 3)(*delta.shape)
    fn = getattr(lib(), "ora_find_peaks" + _sfx(precision))
    fn.restype = C.c_longlong
    n = fn(nd, _ptr(delta), None, None, C.c_longlong(0))
    px = np.zeros((max(n, 1), 3), dtype=np.int32)
    pd = np.zeros(max(n, 1), dtype=R)
    fn(nd, _ptr(delta), _ptr(px), _ptr(pd), C.c_longlong(n))
    return px[:n], pd[:n]


def lomuto_sort(keys, precision=8):
    """Return (sorted keys, perm) with perm[new] = old, as quickSort_peaks."""
    R = _real(precision)
    key = np.ascontiguousarray(keys, dtype=R).copy()
    perm = np.arange(len(key), dtype=np.int32)
    getattr(lib(), "ora_lomuto_sort" + _sfx(precision))(_ptr(key), _ptr(perm),
                                                          C.c_longlong(len(key)))
    return key, perm


def taus_uniforms(seed, n):
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib().ora_taus_uniforms(C.c_ulong(seed), _ptr(out), C.c_longlong(n))
    return out[:n]


def find_halos(delta, k, P, Lc, Om0=0.31, z=0.0, dc=-1.0, Nmin=1, a=1.0, beta=0.0,
               alpha=0.0, seed=12345, precision=8, p=None, sigma_radii=0):
    """Find_Halos: peaks, Lomuto sort, growth, masses.

    Returns a dict with the halo centres ``x`` [nh,3] (cell indices),
    ``count``, ``r2`` (cells with r^2 < r2 belong to the halo), ``Mh``, the
    int64 ``flag`` grid, and the sorted peak list (``peak_x``, ``peak_den``).
    """
    R = _real(precision)
    delta = np.ascontiguousarray(delta, dtype=R)
    nd = delta.shape
    if p is None:
        p = params(nd, Lc, Om0, z, dc, precision)
    tabs = halo_tables(p, k, P, a, beta, alpha, precision, sigma_radii=sigma_radii)
    px, pd = find_peaks(delta, precision)
    skey, perm = lomuto_sort(pd, precision)
    spx = np.ascontiguousarray(px[perm])
    ng = int(np.prod(nd))
    flag = np.full(ng, ng, dtype=np.int64)
    cap = max(len(pd), 1)
    hx = np.zeros((cap, 3), dtype=np.int32)
    hc = np.zeros(cap, dtype=np.int32)
    hr = np.zeros(cap, dtype=np.int32)
    fn = getattr(lib(), "ora_grow_halos" + _sfx(precision))
    fn.restype = C.c_longlong
    nh = fn(C.byref(p), C.c_longlong(len(pd)), _ptr(spx), _ptr(skey), _ptr(delta),
            _ptr(tabs["Bt"]), C.c_int(len(tabs["Bt"])), C.c_int(tabs["Ncells"]),
            C.c_int(Nmin), _ptr(flag), _ptr(hx), _ptr(hc), _ptr(hr))
    if nh < 0:
        raise RuntimeError(f"oracle grow_halos failed ({nh})")
    hx, hc, hr = hx[:nh].copy(), hc[:nh].copy(), hr[:nh].copy()
    sph = cells_in_spheres()
    u = taus_uniforms(seed, nh)
    Mh = np.zeros(max(nh, 1), dtype=R)
    fn2 = getattr(lib(), "ora_halo_masses" + _sfx(precision))
    err = fn2(C.byref(p), _ptr(tabs["Mknots"]), _ptr(tabs["Int"]), C.c_int(len(tabs["Int"])),
              _ptr(sph), C.c_int(len(sph)), C.c_int(tabs["Ncells"]), _ptr(hc),
              C.c_longlong(nh), _ptr(u), _ptr(Mh))
    return dict(x=hx, count=hc, r2=hr, Mh=Mh[:nh], flag=flag.reshape(nd), peak_x=spx,
                peak_den=skey, mass_err=err, tables=tabs, params=p)


# --------------------------------------------------------------------------
# LPT
# --------------------------------------------------------------------------
def lpt(deltak, p, k_smooth=1e4, do_2lpt=False, precision=8):
    """Compute_1LPT (+ Compute_2LPT): returns (S1, S2 or None) as [ng,3]."""
    R = _real(precision)
    nd = tuple(p.nd)
    ck = np.ascontiguousarray(deltak, dtype=R).copy()
    RC = _rc(precision)
    sfx = _sfx(precision)
    getattr(lib(), "ora_poisson_k" + sfx)(C.byref(p), _ptr(ck), RC(k_smooth), C.c_int(0),
                                          RC(1.0))
    phi = np.ascontiguousarray(c2r(ck, nd).astype(R))
    S1 = np.zeros((int(np.prod(nd)), 3), dtype=R)
    getattr(lib(), "ora_lpt1_disp" + sfx)(C.byref(p), _ptr(phi), _ptr(S1))
    if not do_2lpt:
        return S1, None
    src = np.zeros(nd, dtype=R)
    getattr(lib(), "ora_lpt2_source" + sfx)(C.byref(p), _ptr(S1), _ptr(src))
    ck2 = np.ascontiguousarray(r2c(src).astype(R))
    getattr(lib(), "ora_poisson_k" + sfx)(C.byref(p), _ptr(ck2), RC(k_smooth), C.c_int(1),
                                          RC(p.Normk))
    phi2 = np.ascontiguousarray(c2r(ck2, nd).astype(R))
    S2 = np.zeros_like(S1)
    getattr(lib(), "ora_lpt2_disp" + sfx)(C.byref(p), _ptr(phi2), _ptr(S2))
    return S1, S2


def positions(S, p, precision=8):
    """Compute_Pos: q + S wrapped once into the box."""
    R = _real(precision)
    out = np.ascontiguousarray(S, dtype=R).copy()
    getattr(lib(), "ora_lpt_positions" + _sfx(precision))(C.byref(p), _ptr(out))
    return out


def halo_positions(p, flag, S1, S2, halos, precision=8, with_vel=False):
    R = _real(precision)
    nh = len(halos["count"])
    posh = np.zeros((max(nh, 1), 3), dtype=R)
    velh = np.zeros((max(nh, 1), 3), dtype=R) if with_vel else None
    getattr(lib(), "ora_halo_disp" + _sfx(precision))(
        C.byref(p), _ptr(np.ascontiguousarray(flag.reshape(-1), dtype=np.int64)),
        _ptr(np.ascontiguousarray(S1, dtype=R)),
        _ptr(None if S2 is None else np.ascontiguousarray(S2, dtype=R)), C.c_longlong(nh),
        _ptr(np.ascontiguousarray(halos["x"], dtype=np.int32)),
        _ptr(np.ascontiguousarray(halos["count"], dtype=np.int32)), _ptr(posh), _ptr(velh))
    return posh[:nh], (velh[:nh] if with_vel else None)


def particle_velocities(S1, S2, p, precision=8):
    """Particle velocities of Compute_1LPT / Compute_2LPT (lpt.c:82-86,
    210-216): V = pow(Omz, 0.5454) Hz / (1 + z) S1 evaluated in double and
    rounded to the float type, then V += 2 pow(Omz, 6/11) Hz / (1 + z) S2."""
    R = _real(precision)
    f64 = np.float64
    Omz, Hz, z = f64(R(p.Omz)), f64(R(p.Hz)), f64(R(p.redshift))
    V = (Omz**0.5454 * Hz / (1.0 + z) * np.asarray(S1, dtype=R).astype(f64)).astype(R)
    if S2 is not None:
        V = (V.astype(f64) + 2.0 * Omz ** (6.0 / 11.0) * Hz / (1.0 + z) *
             np.asarray(S2, dtype=R).astype(f64)).astype(R)
    return V


def halo_positions_given(p, flag, S, V, halos, precision=8):
    """Halo positions (and velocities when V is given) from given
    displacements: the IN_disp branch of Generate_Halos_Box_from_Grid
    (box.c:96-110) and Compute_Posh (lpt.c:245-255)."""
    R = _real(precision)
    nh = len(halos["count"])
    posh = np.zeros((max(nh, 1), 3), dtype=R)
    velh = np.zeros((max(nh, 1), 3), dtype=R) if V is not None else None
    getattr(lib(), "ora_halo_disp_given" + _sfx(precision))(
        C.byref(p), _ptr(np.ascontiguousarray(flag.reshape(-1), dtype=np.int64)),
        _ptr(np.ascontiguousarray(S, dtype=R)),
        _ptr(None if V is None else np.ascontiguousarray(V, dtype=R)), C.c_longlong(nh),
        _ptr(np.ascontiguousarray(halos["x"], dtype=np.int32)),
        _ptr(np.ascontiguousarray(halos["count"], dtype=np.int32)), _ptr(posh), _ptr(velh))
    return posh[:nh], (velh[:nh] if V is not None else None)


def halos_box_from_grid(delta, k, P, Lc, S=None, V=None, Om0=0.31, z=0.0, k_smooth=1e4,
                        dc=-1.0, Nmin=1, a=1.0, beta=0.0, alpha=0.0, do_2lpt=False,
                        precision=8, sigma_radii=0):
    """Generate_Halos_Box_from_Grid (box.c:71-150) with the mass-draw seed
    1234 (exshalos.c:987).  With S given (IN_disp) the halos take the means
    of S (and V) over their cells; otherwise the LPT of the grid
    (Compute_Denk, then Compute_1LPT/2LPT) gives the displacements.
    Returns posh, velh (V given, or the LPT velocities), the halo dict, and
    S1/S2 when the LPT ran."""
    R = _real(precision)
    delta = np.ascontiguousarray(delta, dtype=R)
    nd = delta.shape
    p = params(nd, Lc, Om0, z, dc, precision)
    hal = find_halos(delta, k, P, Lc, Om0, z, dc, Nmin, a, beta, alpha, 1234, precision, p,
                     sigma_radii)
    out = dict(halos=hal, params=p, Mh=hal["Mh"])
    if S is not None:
        out["posh"], out["velh"] = halo_positions_given(p, hal["flag"], S, V, hal, precision)
        return out
    dk = (r2c(delta) * R(p.Normk)).astype(R)
    S1, S2 = lpt(dk, p, k_smooth, do_2lpt, precision)
    out["posh"], out["velh"] = halo_positions(p, hal["flag"], S1, S2, hal, precision, True)
    out["S1"], out["S2"] = S1, S2
    return out


def halos_box_from_pk(k, P, nd, Lc, noise, Om0=0.31, z=0.0, k_smooth=1e4, dc=-1.0,
                      Nmin=1, a=1.0, beta=0.0, alpha=0.0, seed=12345, do_2lpt=False,
                      out_lpt=False, precision=8, R_max=1e5, sigma_radii=0):
    """Generate_Halos_Box_from_Pk (box.c:4-68) with explicit noise modes.
    R_max < 1e5 truncates the field's P(k) only; the halo tables use the
    P(k) as given (box.c:18-21)."""
    R = _real(precision)
    p = params(nd, Lc, Om0, z, dc, precision)
    dk = field_modes(p, k, P, noise, precision, 1e5, R_max=R_max)
    delta = (c2r(dk, tuple(p.nd)).astype(R) * R(p.Normx)).astype(R)
    hal = find_halos(delta, k, P, Lc, Om0, z, dc, Nmin, a, beta, alpha, seed, precision, p,
                     sigma_radii)
    S1, S2 = lpt(dk, p, k_smooth, do_2lpt, precision)
    posh, _ = halo_positions(p, hal["flag"], S1, S2, hal, precision)
    out = dict(posh=posh, Mh=hal["Mh"], halos=hal, delta=delta, deltak=dk, params=p)
    if out_lpt:
        S = S1 if S2 is None else (S1 + S2).astype(R)
        out["pos"] = positions(S, p, precision)
    return out


# --------------------------------------------------------------------------
# deposit and power spectrum
# --------------------------------------------------------------------------
def grid_compute(pos, nd, L, window=2, interlacing=False, mass=None, types=None,
                 ntype=1, direction=-1, vel=None, Om0=0.31, z=0.0, folds=1, precision=8,
                 R=4.0, R_times=5.0):
    """Tracer_Grid: returns grid [ntype][1+interlacing][nd][nd][nd].  window 3
    (sphere) and 4 (exponential) use R and R_times (gridmodule.c:57-119)."""
    dt = _real(precision)
    pos = np.ascontiguousarray(pos, dtype=dt).copy()
    npart = pos.shape[0]
    grid = np.zeros((ntype, 1 + int(interlacing), nd, nd, nd), dtype=dt)
    mass = None if mass is None else np.ascontiguousarray(mass, dtype=dt)
    types = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
    vel = None if vel is None else np.ascontiguousarray(vel, dtype=dt)
    RC = _rc(precision)
    rc = getattr(lib(), "ora_tracer_grid" + _sfx(precision))(
        _ptr(grid), C.c_int(nd), RC(L), C.c_int(direction), _ptr(pos), _ptr(vel),
        C.c_longlong(npart), _ptr(mass), _ptr(types), C.c_int(ntype), C.c_int(window), RC(R),
        RC(R_times), C.c_int(int(interlacing)), RC(Om0), RC(z), C.c_int(folds))
    if rc != 0:
        raise ValueError(f"oracle tracer_grid failed ({rc})")
    return grid


def power_compute(grid, ntype, nd, L, window=2, interlacing=False, Nk=25, k_min=None,
                  k_max=None, l_max=0, direction=-1, precision=8, R=4.0):
    """Compute_gridk + Power_Spectrum.  grid: [ntype][1+il][nd][nd][nd].  R: the
    smoothing length of windows 3 and 4 (spectrum_h.c:51-70)."""
    dt = _real(precision)
    grid = np.asarray(grid, dtype=dt).reshape(ntype, 1 + int(interlacing), nd, nd, nd)
    if k_min is None:
        k_min = 2.0 * np.pi / L
    if k_max is None:
        k_max = np.pi / L * nd
    RC = _rc(precision)
    sfx = _sfx(precision)
    gks = []
    for t in range(ntype):
        f1 = np.ascontiguousarray(r2c(grid[t, 0]).astype(dt))
        f2 = np.ascontiguousarray(r2c(grid[t, 1]).astype(dt)) if interlacing else None
        out = np.zeros_like(f1)
        getattr(lib(), "ora_gridk" + sfx)(C.c_int(nd), RC(L), C.c_int(window), RC(R),
                                          _ptr(f1), _ptr(f2), _ptr(out))
        gks.append(out)
    arr = (C.c_void_p * ntype)(*[g.ctypes.data for g in gks])
    NPs = ntype * (ntype + 1) // 2
    ls = l_max // 2 + 1
    Pk = np.zeros((NPs, ls, Nk), dtype=np.float64)
    Km = np.zeros(Nk, dtype=np.float64)
    cnt = np.zeros(Nk, dtype=np.int64)
    getattr(lib(), "ora_power_bins" + sfx)(
        arr, C.c_int(ntype), C.c_int(nd), RC(L), C.c_int(Nk), RC(dt(k_min)), RC(dt(k_max)),
        C.c_int(l_max), C.c_int(direction), _ptr(Pk), _ptr(Km), _ptr(cnt))
    return dict(k=Km.astype(dt), Pk=Pk.astype(dt), Nk=cnt, Pk64=Pk, k64=Km)


def abundance(Mh, Mmin=-1.0, Mmax=-1.0, Nm=25, Lc=2.0, ndx=256, ndy=256, ndz=256, precision=8):
    """Differential halo abundance: dict Mh (mean mass per bin), dn, dn_err."""
    dt = _real(precision)
    RC = _rc(precision)
    m = np.ascontiguousarray(Mh, dtype=dt)
    Mmean = np.zeros(Nm, dtype=dt)
    dn = np.zeros(Nm, dtype=dt)
    dn_err = np.zeros(Nm, dtype=dt)
    getattr(lib(), "ora_abundance" + _sfx(precision))(
        _ptr(m), C.c_longlong(len(m)), RC(Mmin), RC(Mmax), C.c_int(Nm), RC(Lc), C.c_int(ndx),
        C.c_int(ndy), C.c_int(ndz), _ptr(Mmean), _ptr(dn), _ptr(dn_err))
    return {"Mh": Mmean, "dn": dn, "dn_err": dn_err}


# --------------------------------------------------------------------------
# bispectrum
# --------------------------------------------------------------------------
def bispectrum_triangles(Nk):
    """The reference's configuration order (bimodule.c:49-317): for each bin
    a, every b < a gives (a, a, b) when b < floor(a/2), else (a, b, c) for
    max(a-b-1, 0) <= c < b, then (a, b, b) and (a, a, b); each a ends with
    (a, a, a).  Returns a list of bin triples."""
    tri = []
    for a in range(Nk):
        for b in range(a):
            if b < a // 2:
                tri.append((a, a, b))
                continue
            for c in range(max(a - b - 1, 0), b):
                tri.append((a, b, c))
            tri.append((a, b, b))
            tri.append((a, a, b))
        tri.append((a, a, a))
    return tri


def shell_index(nd, L, Nk, k_min, k_max, precision=8):
    """Bin of every half-complex mode, Indice(|k|, k_min, dk) in the grid's
    precision (bimodule.c:72-75, spectrum_h.c:39-48), and |k|."""
    dt = _real(precision)
    kn = dt(2 * np.pi / dt(L))
    dk = dt((dt(k_max) - dt(k_min)) / dt(Nk))
    nz = nd // 2 + 1
    f = np.arange(nd)
    f = np.where(2 * f < nd, f, f - nd).astype(dt) * kn
    fz = np.arange(nz)
    fz = np.where(2 * fz < nd, fz, fz - nd).astype(dt) * kn
    kx, ky, kz = np.meshgrid(f, f, fz, indexing="ij")
    kmod = np.sqrt(kx * kx + ky * ky + kz * kz).astype(dt)
    ind = np.floor(((kmod - dt(k_min)) / dk).astype(dt)).astype(np.int64)
    return ind, kmod


def bispectrum(grid, ntype, nd, L, window=2, R=4.0, interlacing=False, Nk=25, k_min=None,
               k_max=None, precision=8):
    """bi_compute / Bi_Spectrum (spectrum.c:291-406, bimodule.c:4-373): every
    shell's field by c2r of its modes, sums of triple (and pair) products over
    the cells in double, B = sum / sum(I I I) L^(3/2).  Returns the
    reference's keys kP, Pk [NPs][Nk], Nk, kB [3][Ntri], Bk [NBs][Ntri],
    Ntri."""
    dt = _real(precision)
    RC = _rc(precision)
    sfx = _sfx(precision)
    il = bool(interlacing)
    grid = np.asarray(grid, dtype=dt).reshape(ntype, 1 + int(il), nd, nd, nd)
    L = dt(L)
    k_min = 2.0 * np.pi / L if k_min is None else k_min
    k_max = np.pi / L * nd if k_max is None else k_max
    gks = []
    for t in range(ntype):
        f1 = np.ascontiguousarray(r2c(grid[t, 0]).astype(dt))
        f2 = np.ascontiguousarray(r2c(grid[t, 1]).astype(dt)) if il else None
        out = np.zeros_like(f1)
        getattr(lib(), "ora_gridk" + sfx)(C.c_int(nd), RC(L), C.c_int(window), RC(R),
                                          _ptr(f1), _ptr(f2), _ptr(out))
        gks.append(out[..., 0].astype(np.float64) + 1j * out[..., 1].astype(np.float64))
    ind, kmod = shell_index(nd, L, Nk, k_min, k_max, precision)
    F, I, km = [], [], []
    for a in range(Nk):
        m = ind == a
        F.append([(np.fft.irfftn(np.where(m, g, 0), s=(nd, nd, nd), axes=(0, 1, 2)) * nd**3)
                  .astype(dt).astype(np.float64) for g in gks])
        I.append((np.fft.irfftn(m.astype(np.complex128), s=(nd, nd, nd), axes=(0, 1, 2)) *
                  nd**3).astype(dt).astype(np.float64))
        cnt = int(m.sum())
        km.append(np.float64(kmod[m].astype(np.float64).sum()) / cnt if cnt else np.nan)
    tri = bispectrum_triangles(Nk)
    NBs, NPs, ng = ntype**3, ntype * (ntype + 1) // 2, nd**3
    B = np.zeros((NBs, len(tri)))
    IB = np.zeros(len(tri))
    for n, (i1, i2, i3) in enumerate(tri):
        for f1 in range(ntype):
            for f2 in range(ntype):
                for f3 in range(ntype):
                    B[(f1 * ntype + f2) * ntype + f3, n] = np.sum(F[i1][f1] * F[i2][f2] *
                                                               F[i3][f3])
        IB[n] = np.sum(I[i1] * I[i2] * I[i3])
    P = np.zeros((NPs, Nk))
    IP = np.zeros(Nk)
    for a in range(Nk):
        q = 0
        for f1 in range(ntype):
            for f2 in range(f1 + 1):
                P[q, a] = np.sum(F[a][f1] * F[a][f2])
                q += 1
        IP[a] = np.sum(I[a] * I[a])
    kB = np.array([[km[t[j]] for t in tri] for j in range(3)])
    return {"kP": np.array(km).astype(dt), "Pk": (P / IP).astype(dt), "Nk": (IP / ng).astype(dt),
            "kB": kB.astype(dt), "Bk": (B / IB * np.float64(L) ** 1.5).astype(dt),
            "Ntri": (IB / ng).astype(dt), "B64": B, "IB64": IB}
