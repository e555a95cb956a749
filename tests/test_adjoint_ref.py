"""Pins of the FP64 torch reference of the differentiable chain
(oracle/adjoint_ref.py) on the CPU.

Its forward (modes -> 1LPT/2LPT -> positions -> CIC/TSC -> monopole P(k))
equals the C oracle's pipeline (oracle/pipeline.py, itself pinned by the
closed forms of tests/test_oracle_*.py) on the same modes, and its autograd
gradient equals central differences of that forward.
"""

import numpy as np
import pytest
import torch

from exshalos_amd import synth
from oracle import adjoint_ref as AR
from oracle import pipeline as O

LC = 4.0


@pytest.mark.parametrize("do2,window", [(False, 2), (True, 2), (True, 5)])
def test_forward_matches_the_c_oracle(bbks, do2, window):
    k, P = bbks
    nd, Nk = 32, 12
    L = nd * LC
    noise = synth.white_noise(nd, 4, np.float64)
    c = AR.ChainConsts(k, P, nd, LC, Nk=Nk, window=window)
    # the modes equal Compute_Den's
    _, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=8)
    z = AR.delta_k(torch.as_tensor(noise), c).numpy()
    np.testing.assert_allclose(z.real, dk[..., 0], rtol=0, atol=1e-14 * np.abs(dk).max())
    np.testing.assert_allclose(z.imag, dk[..., 1], rtol=0, atol=1e-14 * np.abs(dk).max())
    # displacements, positions, grid and P(k)
    S1, S2 = O.lpt(dk, p, 1e4, do2, 8)
    S = S1 if S2 is None else S1 + S2
    St = AR.displacements(torch.as_tensor(noise), c, do2)
    got = np.stack([s.numpy().reshape(-1) for s in St], axis=1)
    np.testing.assert_allclose(got, S, rtol=0, atol=1e-11 * np.abs(S).max())
    pos = O.positions(S, p, 8)
    g = O.grid_compute(pos, nd, L, window=window, precision=8)[0, 0]
    gt = AR.deposit(AR.positions(St, c), c).numpy()
    np.testing.assert_allclose(gt, g, rtol=0, atol=1e-10 * np.abs(g).max())
    pr = O.power_compute(g, 1, nd, L, window=window, Nk=Nk, precision=8)
    np.testing.assert_array_equal(c.count, pr["Nk"])
    Pt = AR.power(torch.as_tensor(gt), c).numpy()
    np.testing.assert_allclose(Pt, pr["Pk64"][0, 0], rtol=1e-9)
    np.testing.assert_allclose(c.kmean, pr["k64"], rtol=1e-12)


@pytest.mark.parametrize("do2,window", [(True, 2), (True, 5), (False, 2)])
def test_gradient_matches_central_differences(bbks, do2, window):
    """Autograd through the chain equals central differences of chi^2 along
    random directions, in float64 with eps = 1e-6: small enough that no
    particle crosses a kink of the CIC weights (a cell centre; on this box
    one particle lies 1.3e-6 cells from one, so eps = 1e-4 does cross)."""
    k, P = bbks
    nd, Nk = 16, 8
    noise = torch.as_tensor(synth.white_noise(nd, 2, np.float64))
    c = AR.ChainConsts(k, P, nd, LC, Nk=Nk, window=window)
    target = 1.05 * np.interp(c.kmean, k, P)

    def loss(m):
        return AR.chi2(AR.pk_of_modes(m, c, do2), target, c.count)

    m = noise.clone().requires_grad_(True)
    loss(m).backward()
    g = m.grad.numpy()
    rng = np.random.default_rng(7)
    eps = 1e-6
    fd, gu = [], []
    for _ in range(6):
        u = rng.standard_normal(noise.shape)
        u /= np.linalg.norm(u)
        ut = torch.as_tensor(u)
        with torch.no_grad():
            fd.append(float(loss(noise + eps * ut) - loss(noise - eps * ut)) / (2 * eps))
        gu.append(float(np.sum(g * u)))
    fd, gu = np.array(fd), np.array(gu)
    assert np.linalg.norm(fd - gu) <= 1e-6 * np.linalg.norm(fd)
    # the copied modes of the kz = 0, nd/2 planes do not enter the field
    assert np.all(g[c.copied.numpy()] == 0.0)
