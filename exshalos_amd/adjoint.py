"""Differentiable power spectrum of LPT particles with respect to the
white-noise modes (an addition: the reference has no adjoint).

``PowerFromModes`` is a ``torch.autograd.Function`` over unit white-noise
modes [nd][nd][nd//2+1][2] on the HIP device.  Its forward is the
library's own chain -- ``exh_lpt`` (modes -> 1LPT/2LPT positions,
lpt.c:4-242), ``exh_grid_compute`` (unweighted CIC/TSC mass assignment,
gridmodule.c:146-270) and ``exh_power_compute`` (monopole P(k),
powermodule.c:4-108) -- and its backward calls ``exh_pk_adjoint``, the
hand-written adjoint kernels, which recompute the displaced positions
instead of storing them.  The discrete halo finder is not on this chain.

``pk_chi2`` is the loss of the differentiation workload:
chi^2 = sum_b N_b / 2 (P_b / T_b - 1)^2, the Gaussian chi^2 of a measured
P(k) against a target T with the cosmic variance 2 T^2 / N_b of N_b modes.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .lib import _native as N
from .lib._common import host_f64


@dataclass
class ChainConfig:
    """Box, spectrum and estimator of the differentiable chain."""

    k: np.ndarray
    P: np.ndarray
    nd: int = 128
    Lc: float = 4.0
    Om0: float = 0.31
    z: float = 0.0
    k_smooth: float = 1e4
    do_2lpt: bool = True
    window: int = 2  # CIC (2) or TSC (5)
    Nk: int = 64
    k_min: float | None = None
    k_max: float | None = None
    kbin_precision: int = 8
    precision: int = 4
    _host: tuple = field(default=None, repr=False)

    def __post_init__(self):
        L = self.nd * self.Lc
        if self.k_min is None:
            self.k_min = 2.0 * np.pi / L
        if self.k_max is None:
            self.k_max = np.pi / L * self.nd
        R = np.float32 if self.precision == 4 else np.float64
        self._host = (host_f64(np.asarray(self.k).astype(R)),
                      host_f64(np.asarray(self.P).astype(R)))

    @property
    def L(self) -> float:
        return self.nd * self.Lc


def _box_src(cfg: ChainConfig, noise):
    b = N.box(cfg.nd, cfg.Lc, cfg.Om0, cfg.z, -1.0, cfg.precision)
    kh, Ph = cfg._host
    s = N.FieldSrc()
    s.kind = N.SRC_MODES
    s.data = noise.data_ptr()
    s.k = kh.ctypes.data
    s.P = Ph.ctypes.data
    s.Nk = len(kh)
    s.fixed = 0
    s.phase = 0.0
    s.R_max = 1e5
    return b, s


def _check_noise(noise, cfg: ChainConfig):
    import torch

    dt = torch.float32 if cfg.precision == 4 else torch.float64
    want = (cfg.nd, cfg.nd, cfg.nd // 2 + 1, 2)
    if tuple(noise.shape) != want or noise.dtype != dt or not noise.is_cuda:
        raise ValueError(f"noise must be a {dt} device tensor of shape {want}")
    return noise.contiguous()


def power_of_modes(noise, cfg: ChainConfig):
    """Forward: (P [Nk] float64, k [Nk], counts [Nk]) of the particles of
    `noise` through exh_lpt, exh_grid_compute and exh_power_compute."""
    import torch

    noise = _check_noise(noise, cfg)
    lib = N.load()
    ctx = N.context()
    b, s = _box_src(cfg, noise)
    ng = cfg.nd**3
    pos = torch.empty((ng, 3), dtype=noise.dtype, device=noise.device)
    N.check(lib.exh_lpt(ctx, C.byref(b), C.byref(s), float(cfg.k_smooth), int(cfg.do_2lpt), 1,
                        N.ptr(pos), None), "exh_lpt")
    grid = torch.empty((cfg.nd,) * 3, dtype=noise.dtype, device=noise.device)
    N.check(lib.exh_grid_compute(ctx, N.ptr(pos), None, None, None, ng, 1, cfg.nd,
                                 float(cfg.L), float(cfg.Om0), float(cfg.z), -1,
                                 int(cfg.window), 0.0, 0.0, 0, 1, int(cfg.precision),
                                 N.ptr(grid)), "exh_grid_compute")
    del pos
    kk = np.zeros(cfg.Nk)
    Pk = np.zeros(cfg.Nk)
    cnt = np.zeros(cfg.Nk, dtype=np.int64)
    N.check(lib.exh_power_compute(ctx, N.ptr(grid), 1, cfg.nd, float(cfg.L), int(cfg.window),
                                  0.0, 0, int(cfg.Nk), float(cfg.k_min), float(cfg.k_max), 0,
                                  -1, int(cfg.precision), int(cfg.kbin_precision), N.ptr(kk),
                                  N.ptr(Pk), N.ptr(cnt)), "exh_power_compute")
    return Pk, kk, cnt


def power_adjoint(noise, cfg: ChainConfig, gP):
    """dL/dnoise (a device tensor like `noise`) for dL/dP_b = gP [Nk]
    through exh_pk_adjoint; also returns the recomputed P and counts."""
    import torch

    noise = _check_noise(noise, cfg)
    lib = N.load()
    ctx = N.context()
    b, s = _box_src(cfg, noise)
    g = torch.empty_like(noise)
    gph = np.ascontiguousarray(np.asarray(gP, dtype=np.float64).reshape(-1))
    if len(gph) != cfg.Nk:
        raise ValueError(f"gP must have {cfg.Nk} values")
    Pk = np.zeros(cfg.Nk)
    cnt = np.zeros(cfg.Nk, dtype=np.int64)
    N.check(lib.exh_pk_adjoint(ctx, C.byref(b), C.byref(s), float(cfg.k_smooth),
                               int(cfg.do_2lpt), int(cfg.window), int(cfg.Nk),
                               float(cfg.k_min), float(cfg.k_max), int(cfg.kbin_precision),
                               N.ptr(gph), N.ptr(g), N.ptr(Pk), N.ptr(cnt)), "exh_pk_adjoint")
    return g, Pk, cnt


def _function():
    import torch

    class PowerFromModes(torch.autograd.Function):
        """P(k) monopole [Nk] (float64, on the noise's device) of the LPT
        particles of the white-noise modes; differentiable in the modes."""

        @staticmethod
        def forward(ctx, noise, cfg):
            Pk, _, cnt = power_of_modes(noise, cfg)
            ctx.save_for_backward(noise)
            ctx.cfg = cfg
            return torch.as_tensor(Pk, dtype=torch.float64, device=noise.device)

        @staticmethod
        def backward(ctx, gP):
            (noise,) = ctx.saved_tensors
            g, _, _ = power_adjoint(noise, ctx.cfg, gP.detach().double().cpu().numpy())
            return g, None

    return PowerFromModes


_PF = None


def power_from_modes(noise, cfg: ChainConfig):
    """Differentiable P(k) of `noise` (see PowerFromModes)."""
    global _PF
    if _PF is None:
        _PF = _function()
    return _PF.apply(noise, cfg)


def pk_chi2(Pk, target, counts):
    """chi^2 = sum_b N_b / 2 (P_b / T_b - 1)^2 over the bins with modes
    (torch, differentiable in Pk)."""
    import torch

    counts = np.asarray(counts)
    m = counts > 0
    w = torch.as_tensor(np.where(m, counts / 2.0, 0.0), dtype=torch.float64, device=Pk.device)
    t = torch.as_tensor(np.where(m, target, 1.0), dtype=torch.float64, device=Pk.device)
    return torch.sum(w * (Pk / t - 1.0) ** 2)
