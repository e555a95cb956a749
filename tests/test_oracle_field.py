"""Pins of the oracle's field stage (density_grid.c:33-217) on the CPU.

Each check compares the oracle with something it does not share code with:
scipy's natural cubic spline, closed forms of single Fourier modes, and the
Hermitian symmetry a real field requires.
"""

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import pipeline as O


def test_spline_matches_scipy_natural():
    """Inter_Power's GSL cspline (density_grid.c:62) is the natural cubic
    spline: second derivative zero at both ends."""
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(0.0, 10.0, 60))
    x = np.unique(x)
    y = np.sin(x) + 0.1 * x**2
    xq = np.linspace(x[0], x[-1], 997)
    ref = CubicSpline(x, y, bc_type="natural")(xq)
    np.testing.assert_allclose(O.spline_eval(x, y, xq), ref, rtol=0, atol=1e-12)
    # the BBKS knots of the benchmark (log-spaced, 500 points)
    from exshalos_amd.synth import bbks_spectrum

    k, P = bbks_spectrum()
    kq = np.geomspace(k[0], k[-1], 2001)
    ref = CubicSpline(k, P, bc_type="natural")(kq)
    np.testing.assert_allclose(O.spline_eval(k, P, kq), ref, rtol=1e-12, atol=0)


def test_spline_reproduces_linear_data_and_rejects_extrapolation():
    x = np.linspace(0.0, 3.0, 7)
    np.testing.assert_allclose(O.spline_eval(x, 2.0 + 5.0 * x, [0.1, 1.7, 2.99]),
                               2.0 + 5.0 * np.array([0.1, 1.7, 2.99]), atol=1e-13)
    with pytest.raises(ValueError):
        O.spline_eval(x, x, [3.5])


def _linear_pk():
    # a natural spline through linear data is that line, so A(k) is exact
    k = np.linspace(1e-3, 10.0, 500)
    return k, 100.0 + 50.0 * k


def _one_mode(nd, idx, g):
    noise = np.zeros((nd, nd, nd // 2 + 1, 2))
    noise[idx] = g
    return noise


@pytest.mark.parametrize("idx", [(3, 14, 5), (11, 2, 7), (0, 5, 3)])
def test_single_generic_mode_is_a_cosine(idx):
    """One generated mode X = A(|k|) (g1 + i g2), A = sqrt(P/2), gives
    delta(n) = Normx * 2 A (g1 cos(theta) - g2 sin(theta)) with
    theta = 2 pi (i n0 + j n1 + k n2) / nd (density_grid.c:155-216):
    pins the k-vector map (i >= nd/2 -> i - nd), the amplitude, Normx and
    the sign of the transform."""
    nd, Lc = 16, 4.0
    k, P = _linear_pk()
    g = (0.7, -1.3)
    delta, dk, p = O.field_from_modes(k, P, nd, Lc, _one_mode(nd, idx, g))
    L = nd * Lc
    kf = 2 * np.pi / L
    kv = np.array([i if i < nd // 2 else i - nd for i in idx[:2]] + [idx[2]]) * kf
    A = np.sqrt((100.0 + 50.0 * np.linalg.norm(kv)) / 2.0)
    n = np.indices((nd, nd, nd))
    theta = 2 * np.pi * (idx[0] * n[0] + idx[1] * n[1] + idx[2] * n[2]) / nd
    expect = 2.0 * A * (g[0] * np.cos(theta) - g[1] * np.sin(theta)) / np.sqrt(L**3)
    np.testing.assert_allclose(delta, expect, rtol=0, atol=1e-12 * np.abs(expect).max())
    assert np.isclose(p.Normx, 1.0 / np.sqrt(L**3), rtol=1e-15)


def test_kz0_plane_mode_and_its_hermitian_copy():
    """In the kz = 0 plane the mode (i, j) with i > nd/2 copies the
    conjugate of (nd - i, nd - j) (density_grid.c:109-150); the field is the
    same cosine as a generic mode."""
    nd, Lc = 16, 4.0
    k, P = _linear_pk()
    g = (1.1, 0.4)
    idx = (3, 5, 0)
    delta, dk, _ = O.field_from_modes(k, P, nd, Lc, _one_mode(nd, idx, g))
    part = dk[nd - 3, nd - 5, 0]
    np.testing.assert_array_equal(part, [dk[3, 5, 0, 0], -dk[3, 5, 0, 1]])
    L = nd * Lc
    A = np.sqrt((100.0 + 50.0 * 2 * np.pi / L * np.sqrt(9 + 25)) / 2.0)
    n = np.indices((nd, nd, nd))
    theta = 2 * np.pi * (3 * n[0] + 5 * n[1]) / nd
    expect = 2.0 * A * (g[0] * np.cos(theta) - g[1] * np.sin(theta)) / np.sqrt(L**3)
    np.testing.assert_allclose(delta, expect, atol=1e-12 * np.abs(expect).max())


def test_self_conjugate_mode_is_real_with_sqrt2():
    """Modes with every index in {0, nd/2} are their own conjugates: the
    reference sets sqrt(2) A g1 and a zero imaginary part
    (density_grid.c:169-175)."""
    nd, Lc = 16, 4.0
    k, P = _linear_pk()
    delta, dk, _ = O.field_from_modes(k, P, nd, Lc, _one_mode(nd, (8, 0, 0), (0.9, 5.0)))
    L = nd * Lc
    A = np.sqrt((100.0 + 50.0 * 8 * 2 * np.pi / L) / 2.0)
    assert dk[8, 0, 0, 1] == 0.0
    n0 = np.indices((nd, nd, nd))[0]
    expect = np.sqrt(2.0) * A * 0.9 * (-1.0) ** n0 / np.sqrt(L**3)
    np.testing.assert_allclose(delta, expect, atol=1e-13)
    # without the R_max truncation the DC mode is zero
    d2, dk2, _ = O.field_from_modes(k, P, nd, Lc, _one_mode(nd, (0, 0, 0), (3.0, 0.0)))
    assert np.all(dk2 == 0.0) and np.all(d2 == 0.0)


@pytest.mark.parametrize("prec", [4, 8])
def test_dc_mode_with_truncated_spectrum_is_uniform(prec):
    """With R_max < 1e5 the DC mode is not zero: |k| is replaced by
    k0 = (kl0 kl1 kl2)^(1/3) / 4, A = sqrt(P_trunc(k0) / 2), and the
    self-conjugate rule writes sqrt(2) A g (density_grid.c:158-161,
    169-175).  A single DC mode is then a uniform field
    delta = sqrt(2) A g Normx.  P_trunc is the FFTLog-truncated spectrum
    (pinned against scipy.fft.fht below), splined here by scipy's natural
    cubic spline."""
    from exshalos_amd.synth import bbks_spectrum

    nd, Lc, R_max, g = 16, 4.0, 60.0, 3.0
    k, P = bbks_spectrum()
    Rt = np.float32 if prec == 4 else np.float64
    d, dk, p = O.field_from_modes(k, P, nd, Lc, _one_mode(nd, (0, 0, 0), (g, 0.7)),
                                  precision=prec, R_max=R_max)
    k2, P2, _, _ = O.pk_truncate(np.asarray(k, Rt).astype(np.float64),
                                 np.asarray(P, Rt).astype(np.float64), R_max)
    kf = 2 * np.pi / (nd * Lc)
    k0 = kf / 4.0
    A = np.sqrt(CubicSpline(k2, P2, bc_type="natural")(k0) / 2.0)
    expect = np.sqrt(2.0) * A * g / np.sqrt((nd * Lc) ** 3)
    tol = 2e-6 if prec == 4 else 1e-12
    assert dk[0, 0, 0, 1] == 0.0 and np.count_nonzero(dk) == 1
    np.testing.assert_allclose(d, np.full(d.shape, expect), rtol=tol, atol=0)
    # the DC amplitude follows the truncated spectrum (not the given one)
    A_full = np.sqrt(CubicSpline(k, P, bc_type="natural")(k0) / 2.0)
    assert abs(A_full / A - 1.0) > 1e-3


@pytest.mark.parametrize("nd", [8, 16])
def test_random_modes_are_hermitian_in_the_kz_planes(nd):
    """With random noise everywhere the kz in {0, nd/2} planes are exactly
    Hermitian, so the half-complex grid describes a real field: the full
    complex inverse FFT of the completed spectrum has no imaginary part."""
    from exshalos_amd.synth import bbks_spectrum

    k, P = bbks_spectrum()
    noise = np.random.default_rng(3).standard_normal((nd, nd, nd // 2 + 1, 2))
    delta, dk, p = O.field_from_modes(k, P, nd, 4.0, noise)
    z = dk[..., 0] + 1j * dk[..., 1]
    for kz in (0, nd // 2):
        plane = z[:, :, kz]
        mirror = np.conj(plane[(-np.arange(nd)) % nd][:, (-np.arange(nd)) % nd])
        np.testing.assert_array_equal(plane, mirror)
    full = np.zeros((nd, nd, nd), dtype=complex)
    full[:, :, : nd // 2 + 1] = z
    for kz in range(nd // 2 + 1, nd):
        full[:, :, kz] = np.conj(z[(-np.arange(nd)) % nd][:, (-np.arange(nd)) % nd, nd - kz])
    x = np.fft.ifftn(full) * nd**3 * p.Normx
    assert np.abs(x.imag).max() < 1e-12 * np.abs(x.real).max()
    np.testing.assert_allclose(x.real, delta, atol=1e-12 * np.abs(delta).max())


def test_field_variance_matches_the_spectrum():
    """<delta^2> = sum over modes of |X|^2 Normx^2 (Parseval); with unit
    Gaussian noise its expectation is sum P(k)/V over the grid modes."""
    from exshalos_amd.synth import bbks_spectrum

    nd, Lc = 32, 4.0
    k, P = bbks_spectrum()
    noise = np.random.default_rng(11).standard_normal((nd, nd, nd // 2 + 1, 2))
    delta, dk, p = O.field_from_modes(k, P, nd, Lc, noise)
    # Parseval for the unnormalised c2r: sum x^2 = N sum_full |X|^2
    z = dk[..., 0] + 1j * dk[..., 1]
    w = np.full(nd // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    lhs = np.sum(delta.astype(np.float64) ** 2)
    rhs = nd**3 * np.sum(w * np.abs(z) ** 2) * p.Normx**2
    assert abs(lhs / rhs - 1.0) < 1e-10


def test_fftlog_matches_scipy_fht():
    """The oracle's FFTLog (fftlog.c:49-131) equals scipy.fft.fht, an
    independent implementation of Hamilton's algorithm, with the same
    low-ringing offset (scipy.fft.fhtoffset), value for value."""
    from scipy.fft import fht, fhtoffset

    n = 500
    k = np.logspace(-4, 2, n)
    a = k**1.5 * np.exp(-0.5 * (3.0 * k) ** 2) * k**0.3
    dln = np.log(k[1] / k[0])
    r, b = O._fht(k, a, 0.5)
    A = fht(a, dln, mu=0.5, offset=fhtoffset(dln, 0.5, initial=0.0))
    assert np.max(np.abs(b.real - A)) < 1e-11 * np.max(np.abs(A))
    assert np.max(np.abs(b.imag)) < 1e-12 * np.max(np.abs(A))


def test_fftlog_gaussian_pair_and_grid_label():
    """P(k) = exp(-k^2 s^2 / 2) has xi(r) = (2 pi s^2)^(-3/2) exp(-r^2 / 2 s^2).
    The values of the reference's transform sit one log step above the grid
    it labels them with (fht's k[0] = k0 r0 / r[0] with L = N dln,
    fftlog.c:93-124); the round trip P -> xi -> P returns the input grid."""
    n, s = 500, 2.0
    k = np.logspace(-4, 2, n)
    P = np.exp(-0.5 * (k * s) ** 2)
    r, b = O._fht(k, k**1.5 * P, 0.5)
    rt = r * (k[1] / k[0])
    xi = np.real((2 * np.pi * rt) ** -1.5 * b)
    exact = (2 * np.pi * s * s) ** -1.5 * np.exp(-0.5 * (rt / s) ** 2)
    sel = (rt > 0.1) & (rt < 5 * s)
    assert np.max(np.abs(xi - exact)[sel]) < 1e-4 * exact.max()
    k2, P2, _, _ = O.pk_truncate(k, P, 1e30)
    np.testing.assert_allclose(k2, k, rtol=1e-13)
    sel = (k > 1e-3) & (k < 1.0)
    np.testing.assert_allclose(P2[sel], P[sel], rtol=1e-9, atol=1e-12)


def test_pk_truncation_is_a_projection():
    """Truncating xi at R_max twice equals once (both transforms share the
    grid), and the truncated spectrum's xi vanishes beyond R_max."""
    from exshalos_amd import synth

    k, P = synth.bbks_spectrum()
    k, P = np.asarray(k, np.float64), np.asarray(P, np.float64)
    k1, P1, r, xi1 = O.pk_truncate(k, P, 60.0)
    k2, P2, r2, xi2 = O.pk_truncate(k1, P1, 60.0)
    np.testing.assert_allclose(k2, k1, rtol=1e-13)
    np.testing.assert_allclose(P2, P1, rtol=0, atol=1e-9 * np.max(np.abs(P1)))
    r0, b = O._fht(k1, k1**1.5 * P1, 0.5)
    xi = np.real((2 * np.pi * r0) ** -1.5 * b)
    assert np.max(np.abs(xi[r0 > 60.0])) < 1e-12 * np.max(np.abs(xi))
    # truncation removes power: P1 differs from P where xi had support
    assert np.max(np.abs(P1 - P)) > 1e-3 * np.max(P)
