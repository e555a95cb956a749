"""GPU parity of every stage against the CPU oracle on identical inputs.

Bars: integer outputs (peaks, halo centres, counts, radii, flags, mode
counts) bit-exact; halo masses bit-exact (same double arithmetic); fields,
displacements, grids and P(k) within the float tolerance written in each
test (FP64: the FFT implementations differ at ~1e-15; FP32: float
rounding).
"""

import numpy as np
import pytest

from oracle import pipeline as O
from exshalos_amd import synth
from exshalos_amd.lib import exshalos as X
from exshalos_amd.lib import spectrum as SP
from exshalos_amd.lib import set_precision

pytestmark = pytest.mark.gpu

LC = 4.0


def _noise(nd, seed, prec):
    return synth.white_noise(nd, seed, np.float32 if prec == 4 else np.float64)


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("nd", [32, 64])
def test_density_grid(bbks, prec, nd):
    k, P = bbks
    set_precision(prec)
    noise = _noise(nd, 3, prec)
    g = X.density_grid_compute(k, P, 1e5, nd, nd, nd, LC, 1, 3, 0, 0.0, 1e5, 0, 1, noise=noise)
    delta, dk, _ = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    tol = 2e-6 if prec == 4 else 1e-12
    assert _rel(g["grid"], delta) < tol
    assert _rel(g["gridk"], dk) < (1e-6 if prec == 4 else 1e-14)


@pytest.fixture
def halo_claims():
    """Sets the context's claim width (EXH_TUNE_HALO_CLAIMS) for one test and
    restores the automatic choice afterwards."""
    from exshalos_amd.lib import _native as N

    def set_(v):
        N.check(N.load().exh_ctx_set_tunable(N.context(), N.TUNE_HALO_CLAIMS, v),
                "exh_ctx_set_tunable")

    yield set_
    set_(0)


# 0: 32-bit claims (fewer than 2^24 candidates); 64: the wide claims
@pytest.mark.parametrize("claims", [0, 64])
@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("nd,seed", [(32, 1), (64, 1), (64, 7)])
def test_find_halos_bit_exact(bbks, prec, nd, seed, claims, halo_claims):
    k, P = bbks
    set_precision(prec)
    halo_claims(claims)
    delta, _, _ = O.field_from_modes(k, P, nd, LC, _noise(nd, seed, prec), precision=prec)
    ref = O.find_halos(delta, k, P, LC, seed=12345, precision=prec)
    out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    cat = X.halo_catalogue()
    assert len(cat["count"]) == len(ref["count"]) > 0
    np.testing.assert_array_equal(cat["x"], ref["x"])
    np.testing.assert_array_equal(cat["count"], ref["count"])
    np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))
    np.testing.assert_array_equal(out["Mh"], ref["Mh"])
    R = np.float32 if prec == 4 else np.float64
    lag = (R(LC) * (ref["x"] + 0.5)).astype(R)
    np.testing.assert_array_equal(out["posh"], lag)


def test_halo_claims_width_switch_and_epoch_wrap(bbks, halo_claims):
    """The claim array keeps its contents between calls: switching the claim
    width clears it, and 32-bit claims clear it again before their 8-bit
    round tag wraps (every 255 rounds).  Calls on one box, with the width
    switched on every fifth of the first twenty and then more than 300
    rounds of 32-bit claims, must each give the oracle's catalogue."""
    from exshalos_amd.lib import _native as N

    k, P = bbks
    set_precision(4)
    delta, _, _ = O.field_from_modes(k, P, 32, LC, _noise(32, 3, 4), precision=4)
    ref = O.find_halos(delta, k, P, LC, seed=12345, precision=4)
    rounds32, call = 0, 0
    # every round takes a tag, so more than 255 rounds of 32-bit claims
    # after the last width switch wrap the tag at least once
    while rounds32 <= 300 and call < 400:
        wide = call % 5 == 4 and call < 20
        halo_claims(64 if wide else 0)
        out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
        cat = X.halo_catalogue()
        np.testing.assert_array_equal(cat["x"], ref["x"])
        np.testing.assert_array_equal(cat["count"], ref["count"])
        np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))
        if call >= 20:
            rounds32 += N.halo_stats(N.context())["rounds"]
        call += 1
    assert rounds32 > 300


def test_find_halos_repeated_calls_mixed_sizes(bbks):
    """The growth claims persist across calls and are not cleared between
    them (round tags keep counting): a large box, a small one in the same
    buffers, then the large one again must each match the oracle."""
    k, P = bbks
    set_precision(4)
    cases = []
    for nd, seed in [(64, 2), (32, 4), (64, 2)]:
        delta, _, _ = O.field_from_modes(k, P, nd, LC, _noise(nd, seed, 4), precision=4)
        cases.append((delta, O.find_halos(delta, k, P, LC, seed=12345, precision=4)))
    for delta, ref in cases:
        out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
        cat = X.halo_catalogue()
        np.testing.assert_array_equal(cat["x"], ref["x"])
        np.testing.assert_array_equal(cat["count"], ref["count"])
        np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))


@pytest.mark.parametrize("prec", [4, 8])
def test_find_halos_ties_and_barrier(bbks, prec):
    """Quantised density values force many ties (Lomuto tie order) and an
    ellipsoidal barrier with Nmin > 1 exercises shrink and rejection."""
    k, P = bbks
    set_precision(prec)
    nd = 64
    delta, _, _ = O.field_from_modes(k, P, nd, LC, _noise(nd, 11, prec), precision=prec)
    R = np.float32 if prec == 4 else np.float64
    delta = (np.round(delta * 64) / 64).astype(R)
    for (a, beta, alpha, Nmin) in [(1.0, 0.0, 0.0, 1), (0.80, 0.29, 0.53, 3)]:
        ref = O.find_halos(delta, k, P, LC, Nmin=Nmin, a=a, beta=beta, alpha=alpha,
                           seed=12345, precision=prec)
        _, cnts = np.unique(ref["peak_den"], return_counts=True)
        assert (cnts > 1).sum() > 100  # ties present
        out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, Nmin, a, beta, alpha, 1, 0, 0)
        cat = X.halo_catalogue()
        np.testing.assert_array_equal(cat["x"], ref["x"])
        np.testing.assert_array_equal(cat["count"], ref["count"])
        np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))
        np.testing.assert_array_equal(out["Mh"], ref["Mh"])


@pytest.mark.parametrize("prec", [4, 8])
def test_find_halos_fp64_sigma_table(bbks, prec):
    """sigma_radii = 64 (the FP64 sigma(R) table) under an ellipsoidal
    barrier: catalogue, flags and masses bit-exact against the oracle's
    table of the same option; switching the option off and on again
    rebuilds the cached tables."""
    k, P = bbks
    set_precision(prec)
    nd = 64
    delta, _, _ = O.field_from_modes(k, P, nd, LC, _noise(nd, 9, prec), precision=prec)
    eb = (0.80, 0.29, 0.53)
    refs = {r: O.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, *eb, seed=12345,
                            precision=prec, sigma_radii=r) for r in (0, 64)}
    for r in (64, 0, 64):
        out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, *eb, 1, 0, 0, sigma_radii=r)
        cat = X.halo_catalogue()
        assert len(cat["count"]) == len(refs[r]["count"]) > 0
        np.testing.assert_array_equal(cat["count"], refs[r]["count"])
        np.testing.assert_array_equal(out["flag"].reshape(-1), refs[r]["flag"].reshape(-1))
        np.testing.assert_array_equal(out["Mh"], refs[r]["Mh"])


@pytest.mark.parametrize("prec", [4, 8])
def test_cached_tables_follow_pk(bbks, prec):
    """The P(k) spline and the halo tables are kept on the device between
    calls; changing P(k), the box and back must rebuild them each time."""
    k, P = bbks
    set_precision(prec)
    nd = 32
    noise = _noise(nd, 5, prec)
    for (Pi, Lc) in [(P, LC), (2.0 * P, LC), (2.0 * P, 5.0), (P, LC)]:
        g = X.density_grid_compute(k, Pi, 1e5, nd, nd, nd, Lc, 1, 3, 0, 0.0, 1e5, 0, 1,
                                   noise=noise)
        delta, _, _ = O.field_from_modes(k, Pi, nd, Lc, noise, precision=prec)
        assert _rel(g["grid"], delta) < (2e-6 if prec == 4 else 1e-12)
        ref = O.find_halos(delta, k, Pi, Lc, seed=12345, precision=prec)
        out = X.find_halos(delta, k, Pi, Lc, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
        np.testing.assert_array_equal(X.halo_catalogue()["count"], ref["count"])
        np.testing.assert_array_equal(out["Mh"], ref["Mh"])


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("do2", [0, 1])
@pytest.mark.parametrize("nd", [32, 64])
def test_lpt(bbks, prec, do2, nd):
    """nd = 32 has partial (16 x 64) stencil tiles along k; nd = 64 only
    full ones, where the FP32 2LPT source keeps its own cells in registers
    and stages only the halos."""
    k, P = bbks
    set_precision(prec)
    noise = _noise(nd, 5, prec)
    delta, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    S1, S2 = O.lpt(dk, p, 1e4, bool(do2), prec)
    S = S1 if S2 is None else S1.astype(np.float64) + S2
    out = X.lpt_compute(dk, LC, 0.31, 0.0, 1e4, do2, 0, 1, 0, 0)
    tol = 5e-5 if prec == 4 else 1e-10
    assert _rel(out["pos"], S) < tol
    # from the real-space grid (Compute_Denk path) the displacements agree too
    out2 = X.lpt_compute(delta, LC, 0.31, 0.0, 1e4, do2, 0, 0, 0, 0)
    assert _rel(out2["pos"], S) < (5e-4 if prec == 4 else 1e-9)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("nd", [48, 96, 192])
def test_fft_lengths_three_times_power_of_two(bbks, prec, nd):
    """Sides 3 * 2^a run the FFT passes with a leading radix-3 stage (c2r of
    the field, the Poisson-fused c2r and the r2c of the 2LPT source, the r2c
    of the power spectrum): field, 2LPT displacements and P(k) against the
    oracle (numpy FFTs) at the bars of the power-of-two sides."""
    k, P = bbks
    set_precision(prec)
    noise = _noise(nd, 4, prec)
    g = X.density_grid_compute(k, P, 1e5, nd, nd, nd, LC, 1, 4, 0, 0.0, 1e5, 0, 1, noise=noise)
    delta, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    assert _rel(g["grid"], delta) < (2e-6 if prec == 4 else 1e-12)
    assert _rel(g["gridk"], dk) < (1e-6 if prec == 4 else 1e-14)
    if nd <= 96:
        S1, S2 = O.lpt(dk, p, 1e4, True, prec)
        out = X.lpt_compute(dk, LC, 0.31, 0.0, 1e4, 1, 0, 1, 0, 0)
        assert _rel(out["pos"], S1.astype(np.float64) + S2) < (5e-5 if prec == 4 else 1e-10)
    R = np.float32 if prec == 4 else np.float64
    grid = np.random.default_rng(nd).normal(size=(1, 2, nd, nd, nd)).astype(R)
    L = nd * LC
    ref = O.power_compute(grid, 1, nd, L, window=2, interlacing=True, Nk=16, precision=prec)
    out = SP.power_compute(grid, 1, nd, L, 2, 4.0, 1, 16, 2 * np.pi / L, np.pi / L * nd, 0, -1,
                           0, 1)
    np.testing.assert_array_equal(out["Nk"], ref["Nk"])
    assert _rel(out["Pk"], ref["Pk64"]) < (1e-5 if prec == 4 else 1e-10)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("window", [1, 2, 5])
@pytest.mark.parametrize("il", [0, 1])
@pytest.mark.parametrize("nd", [32, 24, 20])
def test_grid_compute(prec, window, il, nd):
    """Particles in random order (the direct atomic scatter) on grids of
    several sizes."""
    set_precision(prec)
    rng = np.random.default_rng(2)
    L = 128.0
    R = np.float32 if prec == 4 else np.float64
    pos = (rng.random((20000, 3)) * L).astype(R)
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=bool(il), precision=prec)
    g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, window, 4.0, 5.0, il,
                        1, 0, 1)
    tol = 2e-5 if prec == 4 else 1e-12
    assert g.shape == ref.shape
    assert _rel(g, ref) < tol


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("window", [1, 2, 5])
@pytest.mark.parametrize("il", [0, 1])
@pytest.mark.parametrize("nL,nd", [(32, 32), (32, 64), (64, 64), (64, 128), (64, 32)])
def test_grid_compute_lattice_tiles(prec, window, il, nL, nd):
    """Particles in Lagrangian lattice order (np = nL^3, as LPT writes
    them) take the LDS-tiled deposits.  (32, 32) and (64, 64): fixed
    windows of 24 cells at each tile's mean displacement; the antipodal
    outlier in the first tile (also the tile's sampled particle) leaves its
    window and deposits directly.  (64, 32): 8-cell tiles in 16-cell
    windows.  (32, 64) and (64, 128): 16-particle tiles span 32 cells a
    side, the window exceeds the LDS and extent-sized boxes take
    over (boxes that do not fit deposit directly).  Weighted and
    redshift-space variants run through the same tiles."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    L = 128.0
    rng = np.random.default_rng(13 + window + 2 * il + nd + nL)
    q = (np.indices((nL, nL, nL)).reshape(3, -1).T + 0.5) * (L / nL)
    pos = q + rng.normal(0, 1.5 * L / nd, q.shape)
    pos[0] += L / 2  # antipodal outlier in the first tile
    pos[5 * nL * nL + 3 * nL + 7] -= (L / 3, 0.0, L / 4)
    pos = np.mod(pos, L).astype(R)
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=bool(il), precision=prec)
    g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, window, 4.0, 5.0, il,
                        1, 0, 1)
    tol = 2e-5 if prec == 4 else 1e-12
    assert _rel(g, ref) < tol
    vel = (rng.normal(size=pos.shape) * 300).astype(R)
    mass = rng.random(len(pos)).astype(R)
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=bool(il), mass=mass,
                         direction=1, vel=vel, precision=prec)
    g = SP.grid_compute(pos, vel, mass, 1, None, 1, nd, L, 0.31, 0.0, 1, window, 4.0, 5.0, il,
                        1, 0, 1)
    assert _rel(g, ref) < tol


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("folds", [2, 3])
@pytest.mark.parametrize("window", [2, 5])
def test_grid_compute_folds(prec, folds, window):
    """folds > 1 (gridmodule.c:152): positions wrapped into the folded box
    L / folds, grid cells of L / (folds nd).  Lattice-ordered particles (the
    tiled path: a tile's particles then spread over folds times its cells,
    so most stencils leave the window and deposit directly) and a random
    set (the direct scatter), interlaced, against the oracle."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    L, nL, nd = 128.0, 32, 32
    rng = np.random.default_rng(41 + folds + window)
    q = (np.indices((nL, nL, nL)).reshape(3, -1).T + 0.5) * (L / nL)
    lattice = np.mod(q + rng.normal(0, 1.0, q.shape), L).astype(R)
    scattered = (rng.random((20000, 3)) * L).astype(R)
    for pos in (lattice, scattered):
        ref = O.grid_compute(pos, nd, L, window=window, interlacing=True, folds=folds,
                             precision=prec)
        g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, window, 4.0, 5.0,
                            1, folds, 0, 1)
        assert _rel(g, ref) < (2e-5 if prec == 4 else 1e-12)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("il", [0, 1])
def test_grid_compute_lattice_bulk_flows(prec, il):
    """Lattice-ordered particles (nL = nd = 64, 4^3 tiles of 16 cells) moved
    by a smooth flow of up to 15 cells, with shell crossing: the tiles'
    windows sit up to 14 cells from their tiles and many stencils leave
    their window and deposit directly.  Together they equal the oracle's
    grid."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    nd = nL = 64
    L = 256.0
    cell = L / nd
    rng = np.random.default_rng(31 + il)
    q = (np.indices((nL, nL, nL)).reshape(3, -1).T + 0.5) * (L / nL)
    w = 2 * np.pi / L
    flow = 15.0 * cell * np.stack([np.sin(w * q[:, 1] + np.pi / 4),
                                   np.sin(w * q[:, 2] + np.pi / 4),
                                   np.cos(w * q[:, 0] + np.pi / 4)], axis=1)
    pos = np.mod(q + flow + rng.normal(0, 0.7 * cell, q.shape), L).astype(R)
    ref = O.grid_compute(pos, nd, L, window=2, interlacing=bool(il), precision=prec)
    g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, 2, 4.0, 5.0, il,
                        1, 0, 1)
    tol = 2e-5 if prec == 4 else 1e-12
    assert _rel(g, ref) < tol


@pytest.mark.parametrize("prec", [4, 8])
def test_grid_compute_weighted_types_rsd(prec):
    set_precision(prec)
    rng = np.random.default_rng(4)
    nd, L = 32, 128.0
    R = np.float32 if prec == 4 else np.float64
    n = 30000
    pos = (rng.random((n, 3)) * L).astype(R)
    vel = (rng.normal(size=(n, 3)) * 300).astype(R)
    mass = rng.random(n).astype(R)
    types = rng.integers(0, 3, n).astype(np.int32)
    ref = O.grid_compute(pos, nd, L, window=2, interlacing=True, mass=mass, types=types,
                         ntype=3, direction=2, vel=vel, precision=prec)
    g = SP.grid_compute(pos, vel, mass, 1, types, 3, nd, L, 0.31, 0.0, 2, 2, 4.0, 5.0, 1, 1, 0, 1)
    assert _rel(g, ref) < (2e-5 if prec == 4 else 1e-12)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("window,R,R_times", [(3, 5.0, 1.0), (3, 9.0, 1.0), (4, 3.0, 2.0)])
@pytest.mark.parametrize("il", [0, 1])
def test_grid_compute_sphere_exponential(prec, window, R, R_times, il):
    """Windows 3 and 4 (gridmodule.c:57-119): every cell centre within R (or
    R R_times, weighted by exp(-r^2 / 2R^2)); unweighted grids normalised by
    the deposited weight.  Then masses, three tracer types and the
    redshift-space shift on the same kernel."""
    set_precision(prec)
    rng = np.random.default_rng(21 + window + il)
    nd, L, n = 24, 96.0, 3000
    RT = np.float32 if prec == 4 else np.float64
    pos = (rng.random((n, 3)) * L).astype(RT)
    pos[0] = (0.001, L - 0.001, 0.5 * L / nd)  # near the edges and on a cell centre
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=bool(il), precision=prec, R=R,
                         R_times=R_times)
    g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, window, R, R_times, il,
                        1, 0, 1)
    tol = 2e-5 if prec == 4 else 1e-12
    assert g.shape == ref.shape
    assert _rel(g, ref) < tol
    vel = (rng.normal(size=(n, 3)) * 300).astype(RT)
    mass = rng.random(n).astype(RT)
    types = rng.integers(0, 3, n).astype(np.int32)
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=bool(il), mass=mass, types=types,
                         ntype=3, direction=2, vel=vel, precision=prec, R=R, R_times=R_times)
    g = SP.grid_compute(pos, vel, mass, 1, types, 3, nd, L, 0.31, 0.0, 2, window, R, R_times, il,
                        1, 0, 1)
    assert _rel(g, ref) < tol


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("window,il,ntype,lmax,direction",
                         [(2, 0, 1, 0, -1), (2, 1, 1, 0, -1), (5, 1, 2, 4, 2), (1, 0, 3, 2, 0),
                          (3, 0, 1, 2, 1), (4, 1, 2, 0, -1)])
def test_power_compute(prec, window, il, ntype, lmax, direction):
    set_precision(prec)
    rng = np.random.default_rng(9)
    nd, L, Nk = 32, 256.0, 16
    R = np.float32 if prec == 4 else np.float64
    grid = rng.normal(size=(ntype, 1 + il, nd, nd, nd)).astype(R)
    ref = O.power_compute(grid, ntype, nd, L, window=window, interlacing=bool(il), Nk=Nk,
                          l_max=lmax, direction=direction, precision=prec)
    kmin, kmax = 2 * np.pi / L, np.pi / L * nd
    out = SP.power_compute(grid, ntype, nd, L, window, 4.0, il, Nk, kmin, kmax, lmax,
                           direction, 0, 1)
    np.testing.assert_array_equal(out["Nk"], ref["Nk"])
    tol = 1e-5 if prec == 4 else 1e-10
    assert _rel(out["Pk"], ref["Pk64"]) < tol
    assert _rel(out["k"], ref["k64"]) < (1e-6 if prec == 4 else 1e-13)


@pytest.mark.parametrize("prec", [4, 8])
def test_halos_box_from_pk_config1(bbks, prec):
    """Configuration 1 shape (nd=64, Lc=4, 1LPT) end to end."""
    k, P = bbks
    set_precision(prec)
    nd = 64
    noise = _noise(nd, 1, prec)
    out = X.halos_box_from_pk(k, P, 1e5, nd, nd, nd, LC, 1, 1e4, 0.31, 0.0, -1.0, 1, 1.0, 0.0,
                              0.0, 0, 0.0, 1, 1, 0, 0, 1, 0, 0, 1, noise=noise)
    delta = out["grid"]
    # halos bit-exact on the GPU's own density grid
    ref = O.find_halos(delta, k, P, LC, seed=1, precision=prec)
    cat = X.halo_catalogue()
    np.testing.assert_array_equal(cat["x"], ref["x"])
    np.testing.assert_array_equal(cat["count"], ref["count"])
    np.testing.assert_array_equal(out["Mh"], ref["Mh"])
    # halo positions from the oracle's displacements on the same grid
    p = ref["params"]
    _, dk, _ = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    S1, _ = O.lpt(dk, p, 1e4, False, prec)
    posh, _ = O.halo_positions(p, ref["flag"], S1, None, ref, precision=prec)
    d = np.abs(out["posh"].astype(np.float64) - posh)
    d = np.minimum(d, 256.0 - d)
    assert d.max() < (2e-3 if prec == 4 else 1e-8)
    pos = O.positions(S1, p, prec)
    d = np.abs(out["pos"].astype(np.float64) - pos)
    d = np.minimum(d, 256.0 - d)
    assert d.max() < (2e-3 if prec == 4 else 1e-8)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("do2", [0, 1])
def test_halos_box_power_matches_three_calls(bbks, prec, do2):
    """exh_halos_box_power (LPT, grid and P(k) on a second stream beside the
    halo finder) gives what halos_box_from_pk + grid_compute +
    power_compute give: positions and halo count bit for bit, mode counts
    exactly, grids and P(k) to the rounding of the deposit's float atomics."""
    from exshalos_amd.pipeline import BoxPipeline

    k, P = bbks
    nd = 64
    pipe = BoxPipeline(k, P, nd=nd, Lc=LC, precision=prec, do_2lpt=bool(do2), Nk=24)
    pipe.set_modes(_noise(nd, 9, prec))
    k1, P1, c1, nh1 = pipe.run(fused=False)
    pos1 = pipe.pos.cpu().numpy().copy()
    g1 = pipe.grid.cpu().numpy().copy()
    k2, P2, c2, nh2 = pipe.run(fused=True)
    pos2 = pipe.pos.cpu().numpy()
    g2 = pipe.grid.cpu().numpy()
    assert nh1 == nh2 > 0
    np.testing.assert_array_equal(pos1, pos2)
    np.testing.assert_array_equal(c1, c2)
    tol = 1e-5 if prec == 4 else 1e-12
    assert _rel(g2, g1) < tol
    assert _rel(P2, P1) < tol
    assert _rel(k2, k1) < tol


@pytest.mark.parametrize("prec", [4, 8])
def test_halos_box_power_against_the_oracle(bbks, prec):
    """The benchmarked call itself (exh_halos_box_power, 2LPT, interlaced
    CIC, FP64 bin index) against the oracle at nd = 64: the catalogue and
    masses bit-exact on the GPU's density grid (the same field kernels as
    the call's), particle positions against the oracle's 2LPT, and P(k)
    against the FP64 oracle's pipeline with exact mode counts."""
    from exshalos_amd.pipeline import BoxPipeline

    k, P = bbks
    set_precision(prec)
    nd, Nk = 64, 24
    L = nd * LC
    noise = _noise(nd, 9, prec)
    pipe = BoxPipeline(k, P, nd=nd, Lc=LC, precision=prec, do_2lpt=True, Nk=Nk)
    pipe.set_modes(noise)
    kk, Pk, cnt, nh = pipe.run(fused=True)
    x, count, Mh = pipe.catalogue()
    pos = pipe.pos.cpu().numpy().astype(np.float64)
    g = X.density_grid_compute(k, P, 1e5, nd, nd, nd, LC, 0, 1, 0, 0.0, 1e5, 0, 1, noise=noise)
    ref = O.find_halos(g["grid"], k, P, LC, seed=1, precision=prec)
    assert nh == len(ref["count"]) > 0
    np.testing.assert_array_equal(x, ref["x"])
    np.testing.assert_array_equal(count, ref["count"])
    np.testing.assert_array_equal(Mh, ref["Mh"])
    p8 = O.params(nd, LC, precision=8)
    dk = O.field_modes(p8, k, P, noise.astype(np.float64), 8)
    S1, S2 = O.lpt(dk, p8, 1e4, True, 8)
    opos = O.positions((S1 + S2), p8, 8)
    d = np.abs(pos - opos)
    assert np.minimum(d, L - d).max() < (2e-3 if prec == 4 else 1e-8)
    gr = O.grid_compute(opos, nd, L, window=2, interlacing=True, precision=8)
    pr = O.power_compute(gr, 1, nd, L, window=2, interlacing=True, Nk=Nk, precision=8)
    np.testing.assert_array_equal(cnt, pr["Nk"])
    assert _rel(Pk, pr["Pk64"][0, 0]) < (1e-4 if prec == 4 else 1e-5)
    assert _rel(kk, pr["k64"]) < (1e-6 if prec == 4 else 1e-12)


@pytest.mark.parametrize("prec", [4, 8])
def test_sort_extreme_pivots_and_ties(bbks, prec):
    """Isolated one-cell spikes on every other cell, heights nearly sorted in
    lexicographic order and quantised: Lomuto's last-element pivots are then
    extreme (long queue rotations, the pointer-doubling path) and ties are
    everywhere.  Every spike is its own one-cell halo, so the catalogue order
    is the sort order: it must equal the oracle's exactly."""
    from exshalos_amd.lib import _native as N

    k, P = bbks
    set_precision(prec)
    nd = 64
    R = np.float32 if prec == 4 else np.float64
    rng = np.random.default_rng(5)
    n = (nd // 2) ** 3
    v = 1.7 + 1.3 * np.arange(n) / n + rng.normal(0, 0.02, n)
    v = np.round(v * 4096) / 4096
    delta = np.zeros((nd, nd, nd), dtype=R)
    delta[::2, ::2, ::2] = v.reshape(nd // 2, nd // 2, nd // 2)
    ref = O.find_halos(delta, k, P, LC, seed=12345, precision=prec)
    out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    cat = X.halo_catalogue()
    stats = N.halo_stats(N.context())
    assert stats["sort_doubling"] > 0  # the extreme-pivot path ran
    assert len(cat["count"]) == len(ref["count"]) > n // 2
    np.testing.assert_array_equal(cat["x"], ref["x"])
    np.testing.assert_array_equal(out["Mh"], ref["Mh"])
    np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))


@pytest.mark.parametrize("prec", [4, 8])
def test_sort_with_peaks_below_barrier(bbks, prec):
    """Spikes on every other cell, alternately above the barrier (nearly
    sorted, quantised: long rotations and ties) and below it (random,
    quantised).  The sort leaves the order of keys below the smallest
    barrier open; the catalogue, masses and flags must still equal the
    oracle's exactly."""
    from exshalos_amd.lib import _native as N

    k, P = bbks
    set_precision(prec)
    nd = 64
    R = np.float32 if prec == 4 else np.float64
    rng = np.random.default_rng(9)
    n = (nd // 2) ** 3
    hi = 1.7 + 1.3 * np.arange(n) / n + rng.normal(0, 0.02, n)
    lo = rng.uniform(0.2, 1.2, n)
    v = np.where(np.arange(n) % 2 == 0, hi, lo)
    v[-1] = 0.5  # the first pivot lies below the barrier
    v = np.round(v * 1024) / 1024
    delta = np.zeros((nd, nd, nd), dtype=R)
    delta[::2, ::2, ::2] = v.reshape(nd // 2, nd // 2, nd // 2)
    ref = O.find_halos(delta, k, P, LC, seed=12345, precision=prec)
    out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    cat = X.halo_catalogue()
    stats = N.halo_stats(N.context())
    assert stats["peaks"] == n
    assert n // 4 < stats["candidates"] < 3 * n // 4  # about half below the barrier
    assert len(cat["count"]) == len(ref["count"])
    np.testing.assert_array_equal(cat["x"], ref["x"])
    np.testing.assert_array_equal(cat["count"], ref["count"])
    np.testing.assert_array_equal(out["Mh"], ref["Mh"])
    np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("given_range", [False, True])
def test_abundance_compute(prec, given_range):
    """abundance_compute (spectrum.c:535-600): identical bin counts (dn and
    dn_err are functions of the count in the reference's arithmetic), mean
    masses to the summation order's rounding."""
    set_precision(prec)
    RT = np.float32 if prec == 4 else np.float64
    rng = np.random.default_rng(31)
    M = np.exp(rng.uniform(np.log(2e12), np.log(3e15), 20000)).astype(RT)
    Mmin, Mmax = (RT(5e12), RT(1e15)) if given_range else (-1.0, -1.0)
    ref = O.abundance(M, Mmin, Mmax, 20, 4.0, 64, 64, 64, precision=prec)
    out = SP.abundance_compute(M, Mmin, Mmax, 20, 4.0, 64, 64, 64, 0)
    np.testing.assert_array_equal(out["dn"], ref["dn"])
    np.testing.assert_array_equal(out["dn_err"], ref["dn_err"])
    np.testing.assert_allclose(out["Mh"], ref["Mh"], rtol=1e-4 if prec == 4 else 1e-12)


@pytest.mark.parametrize("prec", [4, 8])
def test_abundance_of_the_last_catalogue_and_barrier_chi2(bbks, prec):
    """Masses left on the device by the halo finder give the abundance of the
    catalogue; the barrier fit's chi^2 at a fixed (a, beta, alpha) equals the
    one built from the oracle's halo finder and abundance (utils.py:374-404),
    and a short Nelder-Mead run returns its best evaluated point, no worse
    than the start."""
    from exshalos_amd import simulation as SIM
    from exshalos_amd import utils as U
    from scipy.interpolate import interp1d

    k, P = bbks
    set_precision(prec)
    nd, Nm, Nmin = 64, 8, 1
    delta, _, _ = O.field_from_modes(k, P, nd, LC, _noise(nd, 3, prec), precision=prec)
    X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, Nmin, 0.8, 0.2, 0.5, 0, 0, 0)
    ref_h = O.find_halos(delta, k, P, LC, dc=-1.0, Nmin=Nmin, a=0.8, beta=0.2, alpha=0.5,
                         seed=12345, precision=prec)
    out = SIM.Compute_Abundance(None, Nm=Nm, Lc=LC, nd=nd)
    ref = O.abundance(ref_h["Mh"], -1.0, -1.0, Nm, LC, nd, nd, nd, precision=prec)
    np.testing.assert_array_equal(out["dn"], ref["dn"])
    np.testing.assert_allclose(out["Mh"], ref["Mh"], rtol=1e-5 if prec == 4 else 1e-12)
    # target: the oracle's abundance at (0.8, 0.2, 0.5) with 10% errors
    Mt, dnt = ref["Mh"][ref["dn"] > 0], ref["dn"][ref["dn"] > 0]
    err = 0.1 * dnt
    dc = U._deltac(0.0, 0.31)
    theta = (0.6, 0.35, 0.65)
    h = O.find_halos(delta, k, P, LC, dc=dc, Nmin=Nmin, a=theta[0], beta=theta[1],
                     alpha=theta[2], seed=12345, precision=prec)
    dnh = O.abundance(h["Mh"], -1.0, -1.0, Nm, LC, nd, nd, nd, precision=prec)
    fdn = interp1d(np.log(Mt), dnt, bounds_error=False, fill_value=0.0)
    fdn_err = interp1d(np.log(Mt), err, bounds_error=False, fill_value=0.0)
    mask = dnh["dn"] > 0
    chi2_ref = np.sum((dnh["dn"][mask] - fdn(np.log(dnh["Mh"][mask]))) ** 2 /
                      (dnh["dn_err"][mask] ** 2 + fdn_err(dnh["Mh"][mask]) ** 2)) / (Nm - 4)
    import scipy.optimize as so

    real_minimize = so.minimize
    evals = []

    def recording(fun, x0, **kw):  # theta first, then the fit; every chi^2 kept
        def f(x):
            v = fun(x)
            evals.append((np.array(x), v))
            return v

        f(np.array(theta))
        return real_minimize(f, x0, **kw)

    x0 = np.array([0.55, 0.4, 0.7])
    so.minimize = recording
    try:
        xb = U.Fit_Barrier(k, P, Mt, dnt, err, grid=delta, Nm=Nm, nd=nd, Lc=LC, Nmin=Nmin,
                           x0=x0, Max_iter=12)
    finally:
        so.minimize = real_minimize
    np.testing.assert_allclose(evals[0][1], chi2_ref, rtol=1e-4 if prec == 4 else 1e-9)
    f0 = [v for x, v in evals[1:] if np.allclose(x, x0)][0]
    fb = min(v for x, v in evals[1:] if np.allclose(x, xb))
    assert fb <= f0 and fb == min(v for _, v in evals[1:])
    assert np.all(xb >= [0.1, 0.0, 0.0]) and np.all(xb <= [2.0, 1.0, 1.0])


@pytest.mark.parametrize("prec", [4, 8])
def test_pk_truncation_field_and_halos(bbks, prec):
    """R_max < 1e5 (Inter_Power, density_grid.c:52-59): the field's P(k) is
    cut in configuration space by FFTLog before the spline; the halo
    tables keep the P(k) as given (box.c:18-21).  Field against the
    oracle's FFTLog (pinned against scipy.fft.fht), halos bit-exact."""
    k, P = bbks
    set_precision(prec)
    nd, R_max = 64, 60.0
    noise = _noise(nd, 4, prec)
    g = X.density_grid_compute(k, P, R_max, nd, nd, nd, LC, 1, 3, 0, 0.0, 1e5, 0, 1,
                               noise=noise)
    delta, dk, _ = O.field_from_modes(k, P, nd, LC, noise, precision=prec, R_max=R_max)
    full, _, _ = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    tol = 2e-6 if prec == 4 else 1e-11
    assert _rel(g["grid"], delta) < tol
    assert _rel(g["gridk"], dk) < (1e-6 if prec == 4 else 1e-13)
    assert _rel(delta, full) > 1e-3  # the cut changes the field
    out = X.halos_box_from_pk(k, P, R_max, nd, nd, nd, LC, 1, 1e4, 0.31, 0.0, -1.0, 1, 1.0, 0.0,
                              0.0, 0, 0.0, 1, 0, 0, 0, 1, 0, 0, 1, noise=noise)
    ref = O.halos_box_from_pk(k, P, nd, LC, noise, seed=1, precision=prec, R_max=R_max)
    assert _rel(out["grid"], ref["delta"]) < tol
    ref_h = O.find_halos(out["grid"], k, P, LC, seed=1, precision=prec)
    np.testing.assert_array_equal(out["Mh"], ref_h["Mh"])


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("ntype,il,window", [(1, 0, 2), (1, 1, 2), (2, 1, 1), (1, 0, 3)])
def test_bispectrum(prec, ntype, il, window):
    """bi_compute (spectrum.c:291-406, bimodule.c): configurations in the
    reference's order, shell power spectra, bispectra of every tracer triple
    and the triangle counts against the oracle (whose real-space sums are
    pinned by Fourier-space triangle sums)."""
    from exshalos_amd.lib import _native as N

    set_precision(prec)
    RT = np.float32 if prec == 4 else np.float64
    rng = np.random.default_rng(40 + ntype + il + window)
    nd, L, Nk = 24, 120.0, 7
    grid = rng.normal(size=(ntype, 1 + il, nd, nd, nd)).astype(RT)
    grid[:, :, :4] += 1.5  # some phase coherence, so that B is not noise only
    assert N.load().exh_bispectrum_size(Nk) == len(O.bispectrum_triangles(Nk))
    ref = O.bispectrum(grid, ntype, nd, L, window=window, R=6.0, interlacing=bool(il), Nk=Nk,
                       precision=prec)
    kmin, kmax = 2 * np.pi / L, np.pi / L * nd
    out = SP.bi_compute(grid, ntype, nd, L, window, 6.0, il, Nk, kmin, kmax, 0, 1)
    tol = 1e-4 if prec == 4 else 1e-10
    for key in ("kP", "Nk", "kB", "Ntri"):
        assert out[key].shape == ref[key].shape
        np.testing.assert_allclose(out[key], ref[key], rtol=1e-6 if prec == 4 else 1e-12)
    assert _rel(out["Pk"], ref["Pk"]) < tol
    assert out["Bk"].shape == ref["Bk"].shape
    assert _rel(out["Bk"], ref["Bk"]) < tol


@pytest.mark.parametrize("window", [2, 5])
@pytest.mark.parametrize("nL,nd", [(32, 32), (64, 64)])
def test_grid_compute_lattice_dense_clump(window, nL, nd):
    """Lattice-ordered particles whose first tile collapses onto one point:
    thousands of unit weights land in one window cell, more than the
    window's 32-bit fixed-point sums hold (1023 particles), so that tile
    deposits straight to the grid; the grid still equals the oracle's."""
    set_precision(4)
    L = 128.0
    rng = np.random.default_rng(3 + nd + window)
    q = (np.indices((nL, nL, nL)).reshape(3, -1).T + 0.5) * (L / nL)
    pos = q + rng.normal(0, 0.5 * L / nd, q.shape)
    lat = np.indices((nL, nL, nL)).reshape(3, -1).T
    first = np.all(lat < 16, axis=1)  # the 16^3 corner of the lattice
    pos[first] = (10.3, 11.7, 9.2) + rng.normal(0, 0.01, (first.sum(), 3))
    pos = np.mod(pos, L).astype(np.float32)
    ref = O.grid_compute(pos, nd, L, window=window, interlacing=True, precision=4)
    g = SP.grid_compute(pos, None, None, 0, None, 1, nd, L, 0.31, 0.0, -1, window, 4.0, 5.0, 1,
                        1, 0, 1)
    assert _rel(g, ref) < 2e-5
