"""Pins of the oracle's LPT stage (lpt.c:4-255) against plane-wave closed
forms on the CPU.

For delta = A cos(k x) the 1LPT displacement of the reference's
sixth-order finite-difference gradient is S_x = -A D6(k) sin(k x) / k^2
with D6(k) = (90 sin kh - 18 sin 2kh + 2 sin 3kh) / (60 h).  A single plane
wave has no second-order source; two orthogonal waves have the separable
source Sxx Syy and a closed-form 2LPT displacement.
"""

import numpy as np
import pytest

from oracle import pipeline as O

LC = 4.0


def _d6(k, h):
    return (90 * np.sin(k * h) - 18 * np.sin(2 * k * h) + 2 * np.sin(3 * k * h)) / (60 * h)


def _deltak(delta, p):
    """Compute_Denk: r2c times Normk (density_grid.c:220-235)."""
    return O.r2c(delta) * p.Normk


@pytest.mark.parametrize("axis,m", [(0, 1), (1, 3), (2, 5)])
def test_1lpt_plane_wave(axis, m):
    nd = 32
    p = O.params(nd, LC)
    L = nd * LC
    kk = 2 * np.pi * m / L
    x = np.indices((nd, nd, nd))[axis] * LC
    A = 0.37
    delta = A * np.cos(kk * x)
    S1, S2 = O.lpt(_deltak(delta, p), p, k_smooth=1e4, do_2lpt=True)
    S1 = S1.reshape(nd, nd, nd, 3)
    expect = -A * _d6(kk, LC) * np.sin(kk * x) / kk**2
    for a in range(3):
        ref = expect if a == axis else np.zeros_like(expect)
        np.testing.assert_allclose(S1[..., a], ref, atol=1e-12 * np.abs(expect).max())
    # a single plane wave has no second-order source
    assert np.abs(S2).max() < 1e-12 * np.abs(expect).max()


def test_2lpt_two_orthogonal_waves():
    """delta = A cos(k1 x) + B cos(k2 y): Sxx = -a1 D6(k1) cos(k1 x),
    Syy = -b1 D6(k2) cos(k2 y) with a1 = A D6(k1)/k1^2, b1 = B D6(k2)/k2^2;
    the source is C cos(k1 x) cos(k2 y), C = a1 b1 D6(k1) D6(k2), and
    S2_x = -(3/7) Om^(-1/143) C D6(k1) sin(k1 x) cos(k2 y) / (k1^2 + k2^2)
    (lpt.c:116-134, 141-173, 182-223)."""
    nd = 32
    p = O.params(nd, LC)
    L = nd * LC
    k1, k2 = 2 * np.pi * 2 / L, 2 * np.pi * 3 / L
    A, B = 0.5, 0.8
    ix = np.indices((nd, nd, nd))
    x, y = ix[0] * LC, ix[1] * LC
    delta = A * np.cos(k1 * x) + B * np.cos(k2 * y)
    S1, S2 = O.lpt(_deltak(delta, p), p, k_smooth=1e4, do_2lpt=True)
    S2 = S2.reshape(nd, nd, nd, 3)
    d1, d2 = _d6(k1, LC), _d6(k2, LC)
    a1, b1 = A * d1 / k1**2, B * d2 / k2**2
    Cc = a1 * b1 * d1 * d2
    f = -3.0 / 7.0 * p.Omz ** (-1.0 / 143)
    ex = f * Cc * d1 * np.sin(k1 * x) * np.cos(k2 * y) / (k1**2 + k2**2)
    ey = f * Cc * d2 * np.cos(k1 * x) * np.sin(k2 * y) / (k1**2 + k2**2)
    scale = max(np.abs(ex).max(), np.abs(ey).max())
    np.testing.assert_allclose(S2[..., 0], ex, atol=1e-10 * scale)
    np.testing.assert_allclose(S2[..., 1], ey, atol=1e-10 * scale)
    assert np.abs(S2[..., 2]).max() < 1e-10 * scale


def test_k_smooth_removes_short_waves():
    """Modes with |k| > k_smooth are zeroed (lpt.c:22-26)."""
    nd = 32
    p = O.params(nd, LC)
    L = nd * LC
    x = np.indices((nd, nd, nd))[0] * LC
    delta = np.cos(2 * np.pi * 6 * x / L)
    S1, _ = O.lpt(_deltak(delta, p), p, k_smooth=2 * np.pi * 5 / L)
    full, _ = O.lpt(_deltak(delta, p), p, k_smooth=1e4)
    assert np.abs(S1).max() < 1e-12 * np.abs(full).max()


def test_positions_are_lattice_plus_displacement_wrapped():
    """Compute_Pos (lpt.c:229-242): q = (i + 1/2) Lc, pos = q + S wrapped
    once into [0, L)."""
    nd = 8
    p = O.params(nd, LC)
    L = nd * LC
    rng = np.random.default_rng(2)
    S = rng.uniform(-0.9 * L, 0.9 * L, (nd**3, 3))
    pos = O.positions(S, p)
    q = (np.indices((nd, nd, nd)).reshape(3, -1).T + 0.5) * LC
    np.testing.assert_allclose(pos, np.mod(q + S, L), atol=1e-12)
    assert (pos >= 0).all() and (pos < L).all()
