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


def _numpy_halo_disp(p, flag, S1, S2, x, count, R):
    """Compute_1LPT/Compute_2LPT halo sums and Compute_Posh restated with
    numpy (lpt.c:56-80, 182-210, 245-255): each cell whose |flag| < ng adds
    its displacement to halo |flag| in cell order, in the grid's float type
    (posh += kx); velocities add f * kx in double, rounded to the float type
    each time (velh += pow(Omz, 0.5454) Hz / (1 + z) * kx, and
    2 pow(Omz, 6/11) Hz / (1 + z) * kx for the second order); the 1LPT pass
    over all cells precedes the 2LPT pass.  Then posh = wrap(x Lc + Lc/2 +
    posh / count) with the centre rounded to the float type, velh /= count.
    The order of additions per halo is kept by walking the members by their
    rank within their halo."""
    ng = flag.size
    h = np.abs(flag.reshape(-1))
    cells = np.nonzero(h < ng)[0]
    hh = h[cells]
    order = np.argsort(hh, kind="stable")  # members of each halo in cell order
    first = np.searchsorted(hh[order], hh[order])
    rank = np.empty(len(cells), dtype=np.int64)
    rank[order] = np.arange(len(cells)) - first
    nh = len(count)
    posh = np.zeros((nh, 3), dtype=R)
    velh = np.zeros((nh, 3), dtype=R)
    f64 = np.float64
    Omz, Hz, z = f64(R(p.Omz)), f64(R(p.Hz)), f64(R(p.redshift))
    facs = (Omz**0.5454 * Hz / (1.0 + z), 2.0 * Omz ** (6.0 / 11.0) * Hz / (1.0 + z))
    for S, fac in ((S1, facs[0]), (S2, facs[1])):
        if S is None:
            continue
        S = np.asarray(S, dtype=R).reshape(-1, 3)
        for r in range(int(rank.max()) + 1 if len(rank) else 0):
            sel = rank == r
            c, t = cells[sel], hh[sel]
            posh[t] = posh[t] + S[c]
            velh[t] = (velh[t].astype(f64) + fac * S[c].astype(f64)).astype(R)
    Lc, L = R(p.Lc), np.array(p.L, dtype=R)
    base = (x.astype(R) * Lc).astype(f64) + f64(Lc) / 2.0
    base = base.astype(R)
    out = base + posh / count.astype(R)[:, None]
    out = np.where(out >= L, out - L, out)
    out = np.where(out < 0, out + L, out).astype(R)
    return out, (velh / count.astype(R)[:, None]).astype(R)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("do2", [False, True])
def test_halo_positions_and_velocities_numpy_restatement(bbks, prec, do2):
    """The oracle's per-halo displacement means (ora_halo_disp) equal the
    numpy restatement bit for bit, on a random field's catalogue (halos of
    many sizes, centres flagged -h, rejected cells ng + 1 with Nmin = 3)
    with 1LPT and 2LPT displacements.  Dropping the centre cells (flag
    -h) or the second-order velocity term breaks the equality."""
    from exshalos_amd import synth

    k, P = bbks
    R = np.float32 if prec == 4 else np.float64
    nd = 32
    noise = synth.white_noise(nd, 8, R)
    delta, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    res = O.find_halos(delta, k, P, LC, Nmin=3, seed=5, precision=prec, p=p)
    assert len(res["count"]) > 50 and (res["flag"] == nd**3 + 1).any()
    S1, S2 = O.lpt(dk, p, 1e4, do2, prec)
    posh, velh = O.halo_positions(p, res["flag"], S1, S2, res, precision=prec, with_vel=True)
    ep, ev = _numpy_halo_disp(p, res["flag"], S1, S2, res["x"], res["count"], R)
    np.testing.assert_array_equal(posh, ep)
    np.testing.assert_array_equal(velh, ev)
    assert (posh >= 0).all() and (posh < nd * LC).all()
    # plausible mistakes are caught
    wrong, _ = _numpy_halo_disp(p, np.where(res["flag"] < 0, nd**3, res["flag"]), S1, S2,
                                res["x"], res["count"], R)
    assert not np.array_equal(wrong, posh)  # centre cells dropped
    if do2:
        _, wv = _numpy_halo_disp(p, res["flag"], S1, None, res["x"], res["count"], R)
        assert not np.array_equal(wv, velh)
