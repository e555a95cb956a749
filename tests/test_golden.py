"""The committed golden fixtures (tests/golden, written by make_golden.py
from the oracle) against the oracle on the CPU and against the HIP library
on the GPU.  The CPU half catches regressions of the oracle itself; the GPU
half anchors parity to fixed data."""

import os

import numpy as np
import pytest

from oracle import pipeline as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LC = 4.0


def _load(name):
    return np.load(os.path.join(GOLD, name))


def _rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# --------------------------------------------------------------------- CPU
@pytest.mark.parametrize("prec", [4, 8])
def test_oracle_reproduces_halo_fixture(prec):
    f = _load(f"halos_nd32_p{prec}.npz")
    delta, _, _ = O.field_from_modes(f["k"], f["P"], 32, LC, f["noise"], precision=prec)
    assert _rel(delta, f["delta"]) < (1e-6 if prec == 4 else 1e-13)
    hal = O.find_halos(f["delta"], f["k"], f["P"], LC, seed=12345, precision=prec)
    for key in ("x", "count", "r2", "Mh"):
        np.testing.assert_array_equal(hal[key], f[key])
    np.testing.assert_array_equal(hal["flag"].astype(np.int32), f["flag"])


@pytest.mark.parametrize("prec", [4, 8])
def test_oracle_reproduces_lpt_fixture(prec):
    from exshalos_amd.synth import bbks_spectrum

    f = _load(f"lpt_nd16_p{prec}.npz")
    k, P = bbks_spectrum()
    _, dk, p = O.field_from_modes(k, P, 16, LC, f["noise"], precision=prec)
    S1, S2 = O.lpt(dk, p, 1e4, True, prec)
    assert _rel(S1, f["S1"]) < (1e-5 if prec == 4 else 1e-12)
    assert _rel(S2, f["S2"]) < (1e-5 if prec == 4 else 1e-12)


@pytest.mark.parametrize("prec", [4, 8])
def test_oracle_reproduces_grid_power_fixture(prec):
    f = _load(f"grid_power_nd16_p{prec}.npz")
    for win in (1, 2, 5):
        g = O.grid_compute(f["pos"], 16, 64.0, window=win, interlacing=True, precision=prec)
        assert _rel(g, f[f"grid_w{win}"]) < (1e-6 if prec == 4 else 1e-13)
        pk = O.power_compute(f[f"grid_w{win}"], 1, 16, 64.0, window=win, interlacing=True, Nk=8,
                             l_max=2, direction=2, precision=prec)
        np.testing.assert_array_equal(pk["Nk"], f[f"Nk_w{win}"])
        assert _rel(pk["Pk64"], f[f"Pk_w{win}"]) < (1e-6 if prec == 4 else 1e-12)


# --------------------------------------------------------------------- GPU
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
def test_hip_matches_halo_fixture(prec):
    from exshalos_amd.lib import exshalos as X
    from exshalos_amd.lib import set_precision

    set_precision(prec)
    f = _load(f"halos_nd32_p{prec}.npz")
    k, P = f["k"], f["P"]
    g = X.density_grid_compute(k, P, 1e5, 32, 32, 32, LC, 1, 3, 0, 0.0, 1e5, 0, 1,
                               noise=f["noise"])
    assert _rel(g["grid"], f["delta"]) < (2e-6 if prec == 4 else 1e-12)
    out = X.find_halos(f["delta"], k, P, LC, 0.31, 0.0, -1.0, 1, 1.0, 0.0, 0.0, 1, 0, 0)
    cat = X.halo_catalogue()
    np.testing.assert_array_equal(cat["x"], f["x"])
    np.testing.assert_array_equal(cat["count"], f["count"])
    np.testing.assert_array_equal(out["Mh"], f["Mh"])
    np.testing.assert_array_equal(out["flag"].reshape(-1).astype(np.int64),
                                  f["flag"].reshape(-1).astype(np.int64))


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
def test_hip_matches_lpt_fixture(prec):
    from exshalos_amd.lib import exshalos as X
    from exshalos_amd.lib import set_precision

    set_precision(prec)
    f = _load(f"lpt_nd16_p{prec}.npz")
    out = X.lpt_compute(f["delta"], LC, 0.31, 0.0, 1e4, 1, 0, 0, 0, 0)
    S = f["S1"].astype(np.float64) + f["S2"]
    assert _rel(out["pos"], S) < (5e-4 if prec == 4 else 1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [4, 8])
def test_hip_matches_grid_power_fixture(prec):
    from exshalos_amd.lib import set_precision
    from exshalos_amd.lib import spectrum as SP

    set_precision(prec)
    f = _load(f"grid_power_nd16_p{prec}.npz")
    L = 64.0
    for win in (1, 2, 5):
        g = SP.grid_compute(f["pos"], None, None, 0, None, 1, 16, L, 0.31, 0.0, -1, win, 4.0,
                            5.0, 1, 1, 0, 1)
        assert _rel(g, f[f"grid_w{win}"]) < (2e-5 if prec == 4 else 1e-12)
        out = SP.power_compute(f[f"grid_w{win}"], 1, 16, L, win, 4.0, 1, 8, 2 * np.pi / L,
                               np.pi / L * 16, 2, 2, 0, 1)
        np.testing.assert_array_equal(out["Nk"], f[f"Nk_w{win}"])
        assert _rel(out["Pk"], f[f"Pk_w{win}"]) < (1e-5 if prec == 4 else 1e-10)
