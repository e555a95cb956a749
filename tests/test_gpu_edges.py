"""Edge cases of the path against the oracle: non-cubic boxes (with a
3 * 2^a side) and the smallest FFT side, a field without halos, an empty
particle set, a binning range without modes, and one bin holding every
mode."""

import numpy as np
import pytest

from exshalos_amd import synth
from exshalos_amd.lib import exshalos as X
from exshalos_amd.lib import set_precision
from exshalos_amd.lib import spectrum as SP
from oracle import pipeline as O

pytestmark = pytest.mark.gpu

LC = 4.0


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("shape", [(48, 32, 64), (32, 64, 48), (16, 16, 16)])
def test_non_cubic_field_and_halos(bbks, prec, shape):
    """Density grid of an (n0, n1, n2) box (density_grid.c:71-217 takes
    three sides) against the oracle, and the halo catalogue, flags and
    masses bit-exact on the GPU's grid."""
    k, P = bbks
    set_precision(prec)
    n0, n1, n2 = shape
    R = np.float32 if prec == 4 else np.float64
    noise = synth.white_noise(n2, 7, R, ndx=n0, ndy=n1)
    g = X.density_grid_compute(k, P, 1e5, n0, n1, n2, LC, 1, 7, 0, 0.0, 1e5, 0, 1, noise=noise)
    delta, dk, _ = O.field_from_modes(k, P, shape, LC, noise, precision=prec)
    assert g["grid"].shape == shape
    assert _rel(g["grid"], delta) < (2e-6 if prec == 4 else 1e-12)
    assert _rel(g["gridk"], dk) < (1e-6 if prec == 4 else 1e-14)
    ref = O.find_halos(g["grid"], k, P, LC, seed=12345, precision=prec)
    out = X.find_halos(g["grid"], k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    cat = X.halo_catalogue()
    assert len(cat["count"]) == len(ref["count"]) > 0
    np.testing.assert_array_equal(cat["x"], ref["x"])
    np.testing.assert_array_equal(cat["count"], ref["count"])
    np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))
    np.testing.assert_array_equal(out["Mh"], ref["Mh"])


@pytest.mark.parametrize("prec", [4, 8])
def test_field_without_halos(bbks, prec):
    """A field below the barrier everywhere: no halo, every cell free
    (flag ng), an empty mass array."""
    k, P = bbks
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    nd = 24
    delta = (np.random.default_rng(2).standard_normal((nd, nd, nd)) * 0.05).astype(R)
    ref = O.find_halos(delta, k, P, LC, precision=prec)
    assert len(ref["count"]) == 0
    out = X.find_halos(delta, k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    assert len(out["Mh"]) == 0 and len(X.halo_catalogue()["count"]) == 0
    np.testing.assert_array_equal(out["flag"].reshape(-1), ref["flag"].reshape(-1))


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("window", [1, 2, 5])
def test_empty_particle_set(prec, window):
    """No particles: the normalisation divides zero by a zero mean as the
    reference does (NaN in every cell), in both sub-grids."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    pos = np.zeros((0, 3), dtype=R)
    ref = O.grid_compute(pos, 16, 64.0, window=window, interlacing=True, precision=prec)
    g = SP.grid_compute(pos, None, None, 0, None, 1, 16, 64.0, 0.31, 0.0, -1, window, 4.0, 5.0, 1,
                        1, 0, 1)
    assert g.shape == ref.shape
    np.testing.assert_array_equal(np.isnan(g), np.isnan(ref))
    assert np.isnan(g).all()


@pytest.mark.parametrize("prec", [4, 8])
def test_binning_range_without_modes(prec):
    """k_min beyond every mode: counts, |k| and P(k) all zero, as the
    oracle gives."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    nd, L = 16, 64.0
    grid = np.random.default_rng(0).normal(size=(1, 1, nd, nd, nd)).astype(R)
    ref = O.power_compute(grid, 1, nd, L, window=2, Nk=4, k_min=10.0, k_max=20.0,
                          precision=prec)
    out = SP.power_compute(grid, 1, nd, L, 2, 4.0, 0, 4, 10.0, 20.0, 0, -1, 0, 1)
    np.testing.assert_array_equal(out["Nk"], ref["Nk"])
    assert not np.any(out["Nk"])
    np.testing.assert_array_equal(np.asarray(out["Pk"]).reshape(-1), 0.0)
    np.testing.assert_array_equal(np.asarray(out["k"]).reshape(-1), 0.0)


@pytest.mark.parametrize("prec", [4, 8])
def test_one_bin_holds_every_mode(prec):
    """Nk = 1 over the whole range: the single bin's count is the number of
    unmasked modes below k_max, its P(k) the oracle's."""
    set_precision(prec)
    R = np.float32 if prec == 4 else np.float64
    nd, L = 32, 128.0
    grid = np.random.default_rng(4).normal(size=(1, 2, nd, nd, nd)).astype(R)
    kmin, kmax = 0.0, np.pi / L * nd * np.sqrt(3.0)
    ref = O.power_compute(grid, 1, nd, L, window=2, interlacing=True, Nk=1, k_min=kmin,
                          k_max=kmax, precision=prec)
    out = SP.power_compute(grid, 1, nd, L, 2, 4.0, 1, 1, kmin, kmax, 0, -1, 0, 1)
    np.testing.assert_array_equal(out["Nk"], ref["Nk"])
    assert _rel(out["Pk"], ref["Pk64"]) < (1e-5 if prec == 4 else 1e-10)
