"""FP64 reference of the differentiable chain (TEST INFRASTRUCTURE ONLY).

White-noise modes -> delta_k (Compute_Den's amplitudes and Hermitian copy
rule, density_grid.c:94-198) -> 1LPT / 2LPT displacements
(lpt.c:4-226, sixth-order differences) -> positions (lpt.c:229-242) ->
unweighted, normalised CIC or TSC mass assignment on the box's grid
(gridmodule.c:21-53, 146-270) -> monopole P(k) (spectrum_h.c:83-167,
powermodule.c:4-108), written with torch operations in float64 so that
torch's autograd gives the reference gradient the HIP adjoint is checked
against.  Only the tests import this module.

The forward values are pinned against the C oracle (oracle/pipeline.py) on
the CPU; the gradient against central differences of this forward.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import pipeline as O


def _d6(f, axis):
    """The reference's difference f[+3] - 9 f[+2] + 45 f[+1] - 45 f[-1] +
    9 f[-2] - f[-3] along `axis` (periodic); torch.roll(f, -s)[i] = f[i+s]."""
    r = lambda s: torch.roll(f, shifts=-s, dims=axis)  # noqa: E731
    return r(3) - 9.0 * r(2) + 45.0 * r(1) - 45.0 * r(-1) + 9.0 * r(-2) - r(-3)


class ChainConsts:
    """Mode tables of one box: amplitudes, copy rule, Poisson factors, bins."""

    def __init__(self, k, P, nd, Lc, Om0=0.31, z=0.0, k_smooth=1e4, Nk=64, k_min=None,
                 k_max=None, window=2, device="cpu"):
        p = O.params(nd, Lc, Om0, z, -1.0, precision=8)
        self.p = p
        n = nd
        nz = n // 2 + 1
        self.n, self.nz = n, nz
        self.L = float(p.L[0])
        self.Lc = float(p.Lc)
        kl = float(p.kl[0])
        i = np.arange(n)[:, None, None]
        j = np.arange(n)[None, :, None]
        kk = np.arange(nz)[None, None, :]
        kx = np.where(i < n // 2, i, i - n) * kl
        ky = np.where(j < n // 2, j, j - n) * kl
        kz = np.where(kk == n // 2, -n / 2.0 * kl, kk * kl)
        kx, ky, kz = np.broadcast_arrays(kx, ky, kz)
        kmod = np.sqrt(kx * kx + ky * ky + kz * kz)
        # amplitudes of the field (its own k_smooth is 1e5, box.c:18)
        A = np.zeros_like(kmod)
        pos = (kmod > 0) & (kmod <= 1e5)
        A[pos] = np.sqrt(O.spline_eval(np.asarray(k, np.float64), np.asarray(P, np.float64),
                                       kmod[pos]) / 2.0)
        I, J, K = np.broadcast_arrays(i, j, kk)
        plane = (K == 0) | (K == n // 2)
        h = n // 2
        copied = plane & (((I > h) & (J > h)) | (((I == 0) | (I == h)) & (J > h)) |
                          (((J == 0) | (J == h)) & (I > h)) | ((I > h) & (J < h)))
        selfc = ((I == 0) | (I == h)) & ((J == 0) | (J == h)) & plane
        mi, mj = (-I) % n, (-J) % n
        mirror = (mi * n + mj) * nz + K
        # Poisson factor Normx / (-k^2) of the LPT (0 at k = 0 and beyond k_smooth)
        fact = np.zeros_like(kmod)
        ok = (kmod > 0) & (kmod <= k_smooth)
        fact[ok] = float(p.Normx) / (-(kmod[ok] ** 2))
        t = lambda a: torch.as_tensor(a, device=device)  # noqa: E731
        self.A = t(A)
        self.copied = t(copied)
        self.selfc = t(selfc)
        self.mirror = t(mirror.reshape(-1).astype(np.int64))
        self.fact = t(fact)
        self.Normk = float(p.Normk)
        self.Normx = float(p.Normx)
        self.inv60 = 1.0 / (60.0 * self.Lc)
        self.coef = -3.0 / 7.0 * float(p.Omz) ** (-1.0 / 143)
        # bins of the estimator: double bin index (powermodule.c:47-55 mask)
        L = self.L
        kn = 2 * np.pi / L
        k_min = 2 * np.pi / L if k_min is None else k_min
        k_max = np.pi / L * n if k_max is None else k_max
        f = np.where(2 * np.arange(n) < n, np.arange(n), np.arange(n) - n)
        fz = np.where(2 * np.arange(nz) < n, np.arange(nz), np.arange(nz) - n)
        FX, FY, FZ = np.meshgrid(f, f, fz, indexing="ij")
        km = np.sqrt((FX * kn) ** 2 + (FY * kn) ** 2 + (FZ * kn) ** 2)
        dk = (k_max - k_min) / Nk
        with np.errstate(invalid="ignore"):
            b = np.floor((km - k_min) / dk).astype(np.int64)
        masked = plane & (((I > h) & (J > h)) | ((I == 0) | (I == h)) | ((J == 0) | (J == h)) |
                          ((I > h) & (J < h)))
        valid = (~masked) & (b >= 0) & (b < Nk)
        self.Nk = Nk
        self.bin = t(np.where(valid, b, Nk).reshape(-1))
        self.count = np.bincount(b[valid], minlength=Nk)
        self.kmean = np.bincount(b[valid], weights=km[valid], minlength=Nk) / np.maximum(
            self.count, 1)
        Lb = L / n
        pw = {1: 1, 2: 2, 5: 3}[window]
        sx = np.ones(n)
        ax = f * kn * Lb / 2
        sx[ax != 0] = np.sin(ax[ax != 0]) / ax[ax != 0]
        sz = np.ones(nz)
        az = fz * kn * Lb / 2
        sz[az != 0] = np.sin(az[az != 0]) / az[az != 0]
        W = (sx[:, None, None] * sx[None, :, None] * sz[None, None, :]) ** pw
        normk = (L / (n * n)) ** 1.5
        self.wk2 = t(((normk / W) ** 2).reshape(-1))
        self.window = window
        self.device = device


def delta_k(noise, c: ChainConsts):
    """Compute_Den's modes from the noise (complex [n][n][nz])."""
    g0, g1 = noise[..., 0], noise[..., 1]
    re = torch.where(c.selfc, math.sqrt(2.0) * c.A * g0, c.A * g0)
    im = torch.where(c.selfc, torch.zeros_like(g1), c.A * g1)
    re_m = re.reshape(-1)[c.mirror].reshape(re.shape)
    im_m = im.reshape(-1)[c.mirror].reshape(im.shape)
    re = torch.where(c.copied, re_m, re)
    im = torch.where(c.copied, -im_m, im)
    return torch.complex(re, im)


def displacements(noise, c: ChainConsts, do_2lpt=True):
    n = c.n
    dk = delta_k(noise, c)
    phi = torch.fft.irfftn(dk * c.fact, s=(n, n, n), norm="forward")
    S1 = [-_d6(phi, a) * c.inv60 for a in range(3)]
    if not do_2lpt:
        return S1
    sxx = _d6(S1[0], 0) * c.inv60
    sxy = _d6(S1[0], 1) * c.inv60
    sxz = _d6(S1[0], 2) * c.inv60
    syy = _d6(S1[1], 1) * c.inv60
    syz = _d6(S1[1], 2) * c.inv60
    szz = _d6(S1[2], 2) * c.inv60
    src = sxx * syy + sxx * szz + syy * szz - sxy**2 - sxz**2 - syz**2
    phi2 = torch.fft.irfftn(torch.fft.rfftn(src) * (c.fact * c.Normk), s=(n, n, n),
                            norm="forward")
    return [S1[a] + c.coef * _d6(phi2, a) * c.inv60 for a in range(3)]


def positions(S, c: ChainConsts):
    n = c.n
    idx = torch.arange(n, device=S[0].device, dtype=torch.float64)
    q = [(idx.view(-1, 1, 1), idx.view(1, -1, 1), idx.view(1, 1, -1))[a] * c.Lc + c.Lc / 2.0
         for a in range(3)]
    return [torch.remainder(q[a] + S[a], c.L).reshape(-1) for a in range(3)]


def deposit(pos, c: ChainConsts):
    """delta = rho ng / np - 1 with CIC (2) or TSC (5) weights."""
    n = c.n
    Ld = c.L / n
    ng = n**3
    grid = torch.zeros(ng, dtype=torch.float64, device=pos[0].device)
    if c.window == 2:
        cells, wts = [], []
        for a in range(3):
            u = pos[a] / Ld
            xt = torch.floor(u).detach().long()
            xt = torch.where(xt == n, xt - 1, xt)
            dx = u - (xt.double() + 0.5)
            s = torch.where(dx >= 0, 1, -1)
            o = (xt + s) % n
            cells.append((xt, o))
            wts.append((1.0 - dx.abs(), dx.abs()))
        for sel in range(8):
            b = [(sel >> a) & 1 for a in range(3)]
            id3 = (cells[0][b[0]] * n + cells[1][b[1]]) * n + cells[2][b[2]]
            w = wts[0][b[0]] * wts[1][b[1]] * wts[2][b[2]]
            grid = grid.index_add(0, id3, w)
    else:
        cells, wts = [], []
        for a in range(3):
            u = pos[a] / Ld - 0.5
            cf = torch.floor(u + 0.5).detach()
            d = u - cf
            ci = cf.long()
            cells.append([(ci + m - 1) % n for m in range(3)])
            wts.append([0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2])
        for ia in range(3):
            for ib in range(3):
                for ie in range(3):
                    id3 = (cells[0][ia] * n + cells[1][ib]) * n + cells[2][ie]
                    grid = grid.index_add(0, id3, wts[0][ia] * wts[1][ib] * wts[2][ie])
    npart = pos[0].numel()
    return (grid * (ng / npart) - 1.0).reshape(n, n, n)


def power(delta, c: ChainConsts):
    F = torch.fft.rfftn(delta).reshape(-1)
    v = (F.real**2 + F.imag**2) * c.wk2
    s = torch.zeros(c.Nk + 1, dtype=torch.float64, device=delta.device).index_add(0, c.bin, v)
    cnt = torch.as_tensor(np.maximum(c.count, 1), dtype=torch.float64, device=delta.device)
    return s[: c.Nk] / cnt


def pk_of_modes(noise, c: ChainConsts, do_2lpt=True):
    """Monopole P(k) of the LPT particles of `noise` (float64 [n][n][nz][2])."""
    return power(deposit(positions(displacements(noise, c, do_2lpt), c), c), c)


def chi2(Pk, target, count):
    """sum_b N_b / 2 (P_b / T_b - 1)^2 over bins with modes."""
    m = torch.as_tensor(count > 0, device=Pk.device)
    w = torch.as_tensor(count / 2.0, dtype=torch.float64, device=Pk.device)
    t = torch.as_tensor(np.where(count > 0, target, 1.0), dtype=torch.float64, device=Pk.device)
    return torch.sum(torch.where(m, w * (Pk / t - 1.0) ** 2, torch.zeros_like(Pk)))
