"""The hand-written adjoint of modes -> 1LPT/2LPT -> CIC/TSC -> P(k) on the
GPU (exh_pk_adjoint through exshalos_amd.adjoint.PowerFromModes).

Checks: against torch autograd of the FP64 reference chain
(oracle/adjoint_ref.py, whose forward is pinned against the C oracle and
whose gradient is pinned by central differences on the CPU), and on the
differentiation workload (nd=128, L=512, FP32 forward, chi^2 of the CIC
P(k) in 64 bins against 1.05 x the input spectrum) against central
differences of the library's FP64 forward along 64 random unit directions
with eps = 1e-3.  Bars: 1e-8 (FP64 adjoint vs FP64 reference), 1e-3 (FP32
adjoint vs FP64 reference and vs the differences), written at each
comparison.
"""

import sys

import numpy as np
import pytest
import torch

from exshalos_amd import adjoint as ADJ
from exshalos_amd import synth
from oracle import adjoint_ref as AR

pytestmark = pytest.mark.gpu

LC = 4.0


def _say(msg):
    print(msg, file=sys.stderr, flush=True)


def _relL2(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _spectrum_for(prec, k, P):
    R = np.float32 if prec == 4 else np.float64
    return np.asarray(k).astype(R).astype(np.float64), np.asarray(P).astype(R).astype(np.float64)


def _library_grad(noise64, cfg, target):
    """chi^2 and its gradient through PowerFromModes + pk_chi2."""
    dt = torch.float32 if cfg.precision == 4 else torch.float64
    m = torch.as_tensor(noise64, dtype=dt, device="cuda").requires_grad_(True)
    Pk = ADJ.power_from_modes(m, cfg)
    _, kk, cnt = ADJ.power_of_modes(m.detach(), cfg)
    loss = ADJ.pk_chi2(Pk, target, cnt)
    loss.backward()
    return float(loss.detach()), m.grad.detach().double().cpu().numpy(), \
        Pk.detach().cpu().numpy(), cnt


def _reference_grad(noise64, kP, cfg, target, do2):
    k, P = kP
    c = AR.ChainConsts(k, P, cfg.nd, cfg.Lc, k_smooth=cfg.k_smooth, Nk=cfg.Nk, window=cfg.window,
                       device="cuda")
    m = torch.as_tensor(noise64, dtype=torch.float64, device="cuda").requires_grad_(True)
    Pk = AR.pk_of_modes(m, c, do2)
    loss = AR.chi2(Pk, target, c.count)
    loss.backward()
    return float(loss.detach()), m.grad.detach().cpu().numpy(), Pk.detach().cpu().numpy(), c.count


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("do2,window", [(0, 2), (1, 2), (1, 5)])
def test_adjoint_matches_fp64_reference(bbks, prec, do2, window):
    k, P = bbks
    nd, Nk = 32, 12
    cfg = ADJ.ChainConfig(k, P, nd=nd, Lc=LC, do_2lpt=bool(do2), window=window, Nk=Nk,
                          precision=prec)
    noise = synth.white_noise(nd, 3, np.float64)
    _, kk, cnt = ADJ.power_of_modes(
        torch.as_tensor(noise, dtype=torch.float32 if prec == 4 else torch.float64,
                        device="cuda"), cfg)
    target = 1.05 * np.interp(kk, k, P)
    L1, g1, P1, c1 = _library_grad(noise, cfg, target)
    L0, g0, P0, c0 = _reference_grad(noise, _spectrum_for(prec, k, P), cfg, target, do2)
    np.testing.assert_array_equal(c1, c0)
    assert _relL2(P1, P0) < (1e-10 if prec == 8 else 1e-4)
    e = _relL2(g1, g0)
    _say(f"adjoint nd={nd} prec={prec} 2lpt={do2} window={window}: grad rel L2 {e:.2e}")
    assert e < (1e-8 if prec == 8 else 1e-3)
    # the copied modes of the kz = 0, nd/2 planes carry no gradient
    c = AR.ChainConsts(k, P, nd, LC, Nk=Nk, window=window)
    assert np.all(g1[c.copied.numpy()] == 0.0)


@pytest.mark.timeout(900)
def test_adjoint_differentiation_workload(bbks):
    """nd=128, L=512, 2LPT, CIC, 64 bins, chi^2 against 1.05 P(k): the FP32
    adjoint gradient against central differences (eps = 1e-3) of the FP64
    forward along 64 random unit directions, and against the FP64 torch
    reference gradient; the FP64 adjoint against the same reference."""
    k, P = bbks
    nd, Nk = 128, 64
    cfg32 = ADJ.ChainConfig(k, P, nd=nd, Lc=LC, do_2lpt=True, window=2, Nk=Nk, precision=4)
    cfg64 = ADJ.ChainConfig(k, P, nd=nd, Lc=LC, do_2lpt=True, window=2, Nk=Nk, precision=8)
    noise = synth.white_noise(nd, 1, np.float32).astype(np.float64)
    m64 = torch.as_tensor(noise, dtype=torch.float64, device="cuda")
    _, kk, cnt = ADJ.power_of_modes(m64, cfg64)
    target = 1.05 * np.interp(kk, k, P)
    _say("adjoint nd=128: FP32 and FP64 library gradients")
    L32, g32, _, _ = _library_grad(noise, cfg32, target)
    L64, g64, _, _ = _library_grad(noise, cfg64, target)
    assert abs(L32 / L64 - 1) < 1e-4

    def loss64(m):
        Pk, _, c = ADJ.power_of_modes(m, cfg64)
        w = np.where(c > 0, c / 2.0, 0.0)
        return float(np.sum(w * (Pk / np.where(c > 0, target, 1.0) - 1.0) ** 2))

    _say("adjoint nd=128: central differences along 64 directions")
    rng = np.random.default_rng(11)
    eps = 1e-3
    fd, gu32, gu64 = [], [], []
    for _ in range(64):
        u = rng.standard_normal(noise.shape)
        u /= np.linalg.norm(u)
        ut = torch.as_tensor(u, device="cuda")
        fd.append((loss64(m64 + eps * ut) - loss64(m64 - eps * ut)) / (2 * eps))
        gu32.append(float(np.sum(g32 * u)))
        gu64.append(float(np.sum(g64 * u)))
    fd = np.array(fd)
    e32, e64 = _relL2(gu32, fd), _relL2(gu64, fd)
    _say(f"adjoint nd=128: FD agreement FP32 {e32:.2e}, FP64 {e64:.2e}")
    assert e32 < 1e-3
    assert e64 < 1e-3
    _say("adjoint nd=128: FP64 torch reference gradient")
    _, gref, _, _ = _reference_grad(noise, _spectrum_for(8, k, P), cfg64, target, True)
    r32, r64 = _relL2(g32, gref), _relL2(g64, gref)
    _say(f"adjoint nd=128: vs FP64 reference: FP32 {r32:.2e}, FP64 {r64:.2e}")
    assert r32 < 1e-3
    assert r64 < 1e-8
