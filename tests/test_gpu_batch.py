"""Concurrent batches of realisations (exshalos_amd.batch.BatchRunner).

Each box: field -> halos -> 2LPT particles -> CIC of the particles and of
the halo positions -> matter/halo auto and cross spectra (ntype = 2).
Checks: a batch run on several contexts at once gives every box's result
of a one-context run (halo counts exactly, spectra to the deposit's float
rounding), uneven sizes including 2^a x 3 sides; an FP64 box equals the
oracle's chain (halos on the oracle's own grid, which the GPU's matches to
FFT rounding) within 1e-8.
"""

import numpy as np
import pytest
import torch

from exshalos_amd import synth
from exshalos_amd.batch import BatchRunner
from oracle import pipeline as O

pytestmark = pytest.mark.gpu

LC = 4.0


def _noises(boxes, prec):
    dt = np.float32 if prec == 4 else np.float64
    return [torch.as_tensor(synth.white_noise(nd, seed, dt), device="cuda") for nd, seed in boxes]


def test_batch_concurrent_equals_one_context(bbks):
    k, P = bbks
    boxes = [(48, 1), (32, 2), (64, 3), (32, 4), (96, 5), (48, 6)]
    noises = _noises(boxes, 4)
    seq = BatchRunner(k, P, Nk=16, workers=1, table={})
    r1, _ = seq.run(boxes, noises)
    seq.close()
    par = BatchRunner(k, P, Nk=16, workers=4,
                      table={"sizes": {"48": {"deposit_tile": 16}, "96": {"deposit_tile": 32}}})
    r4, _ = par.run(boxes, noises)
    par.close()
    for a, b in zip(r1, r4):
        assert (a.nd, a.seed) == (b.nd, b.seed)
        assert a.nh == b.nh > 0
        np.testing.assert_array_equal(a.counts, b.counts)
        np.testing.assert_allclose(b.Pk, a.Pk, rtol=1e-5)
        assert np.all(np.isfinite(a.Pk[:, a.counts > 0]))


def test_batch_fp64_box_matches_the_oracle(bbks):
    k, P = bbks
    nd, seed, Nk = 32, 7, 12
    L = nd * LC
    r = BatchRunner(k, P, Nk=Nk, precision=8, workers=2, table={})
    res, _ = r.run([(nd, seed), (nd, seed + 1)], _noises([(nd, seed), (nd, seed + 1)], 8))
    r.close()
    got = res[0]
    noise = synth.white_noise(nd, seed, np.float64)
    ref = O.halos_box_from_pk(k, P, nd, LC, noise, k_smooth=1e4, seed=seed, do_2lpt=True,
                              out_lpt=True, precision=8)
    assert got.nh == len(ref["Mh"])
    g = np.zeros((2, 1, nd, nd, nd))
    g[0] = O.grid_compute(ref["pos"], nd, L, window=2, precision=8)[0]
    g[1] = O.grid_compute(ref["posh"], nd, L, window=2, precision=8)[0]
    pr = O.power_compute(g, 2, nd, L, window=2, Nk=Nk, precision=8)
    np.testing.assert_array_equal(got.counts, pr["Nk"])
    np.testing.assert_allclose(got.Pk, pr["Pk64"][:, 0, :], rtol=1e-8,
                               atol=1e-12 * np.abs(pr["Pk64"]).max())
