"""Helpers shared by the extension-module mirrors (precision, transfers)."""

from __future__ import annotations

import os

import numpy as np

_PRECISION = int(os.environ.get("EXH_PRECISION", "4"))


def get_precision() -> int:
    """Bytes per real of the field path: 4 (float, the reference's default
    build, include/exshalos/exshalos_h.h:17) or 8 (double)."""
    return _PRECISION


def set_precision(p: int) -> None:
    global _PRECISION
    if p not in (4, 8):
        raise ValueError("precision must be 4 or 8")
    _PRECISION = int(p)


def np_real(prec: int):
    return np.float32 if prec == 4 else np.float64


def torch_real(prec: int):
    import torch

    return torch.float32 if prec == 4 else torch.float64


def to_device(x, dtype_np, device=None):
    """numpy array or torch tensor -> contiguous device tensor of dtype."""
    import torch

    tdt = {np.float32: torch.float32, np.float64: torch.float64, np.int32: torch.int32,
           np.int64: torch.int64}[dtype_np]
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=tdt).contiguous()
    a = np.ascontiguousarray(np.asarray(x), dtype=dtype_np)
    return torch.from_numpy(a).to(dev)


def host_f64(x) -> np.ndarray:
    """1-D float64 host copy (P(k) tables)."""
    import torch

    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))


def as_real(x, prec):
    """Round a scalar the way the reference's wrappers do (np.float32/64)."""
    return float(np_real(prec)(x))


def to_numpy(t):
    return t.detach().cpu().numpy()
