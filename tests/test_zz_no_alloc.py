"""exh_grid_compute_noalloc, the mass assignment that allocates nothing,
against exh_grid_compute on the same particles, through the C-ABI.

The allocation-free call keeps the deposit's counters in a static device
array and runs on the caller's stream without a context; a context-free
call that would need a library buffer raises instead of allocating.  Two
calls in a row, the second on a smaller input (the array keeps the first
call's contents and must be cleared), then a third on the spherical
window's weight sums; each must equal exh_grid_compute's grid to the float
rounding of the deposit's atomics (2e-5 FP32, 1e-12 FP64).
"""

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from exshalos_amd.lib import _native as N

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.double().cpu().numpy()
    b = b.double().cpu().numpy()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _args(pos, types, ntype, nd, L, window, il, prec):
    npart = pos.shape[0]
    return (N.ptr(pos), None, None, N.ptr(types), npart, ntype, nd, float(L), 0.31, 0.0, -1,
            window, 5.0, 1.0, il, 1, prec)


def _grid(ntype, il, nd, prec):
    dt = torch.float32 if prec == 4 else torch.float64
    return torch.full((ntype, 1 + il, nd, nd, nd), 7.0, dtype=dt, device="cuda")


def _noalloc(pos, types, ntype, nd, L, window, il, prec):
    lib = N.load()
    g = _grid(ntype, il, nd, prec)
    stream = torch.cuda.current_stream().cuda_stream
    N.check(lib.exh_grid_compute_noalloc(C.c_void_p(stream),
                                         *_args(pos, types, ntype, nd, L, window, il, prec),
                                         N.ptr(g)), "exh_grid_compute_noalloc")
    torch.cuda.synchronize()
    return g


def _existing(pos, types, ntype, nd, L, window, il, prec):
    lib = N.load()
    ctx = N.context()
    g = _grid(ntype, il, nd, prec)
    N.check(lib.exh_grid_compute(ctx, *_args(pos, types, ntype, nd, L, window, il, prec),
                                 N.ptr(g)), "exh_grid_compute")
    torch.cuda.synchronize()
    return g


def _lattice(nL, L, rng, dt):
    q = (np.indices((nL, nL, nL)).reshape(3, -1).T + 0.5) * (L / nL)
    return torch.as_tensor(np.mod(q + rng.normal(0, 1.2 * L / nL, q.shape), L), dtype=dt,
                           device="cuda")


@pytest.mark.parametrize("prec", [4, 8])
def test_noalloc_deposit_equals_grid_compute(prec):
    dt = torch.float32 if prec == 4 else torch.float64
    tol = 2e-5 if prec == 4 else 1e-12
    rng = np.random.default_rng(17)
    # 1: 64^3 particles in lattice order onto 64^3, CIC, interlaced (the
    #    LDS window deposit; counts of one type)
    pos1 = _lattice(64, 256.0, rng, dt)
    # 2 (smaller): 20000 random particles of 3 types onto 24^3, TSC (the
    #    atomic scatter; counts of three types)
    pos2 = torch.as_tensor(rng.random((20000, 3)) * 96.0, dtype=dt, device="cuda")
    t2 = torch.as_tensor(rng.integers(0, 3, 20000), dtype=torch.int32, device="cuda")
    # 3: 3000 particles onto 16^3 with the spherical window (weight sums)
    pos3 = torch.as_tensor(rng.random((3000, 3)) * 64.0, dtype=dt, device="cuda")
    cases = [(pos1, None, 1, 64, 256.0, 2, 1), (pos2, t2, 3, 24, 96.0, 5, 0),
             (pos3, None, 1, 16, 64.0, 3, 0)]
    free0 = torch.cuda.mem_get_info()[0]
    outs = []
    for (pos, types, ntype, nd, L, w, il) in cases:
        print(f"noalloc: np={pos.shape[0]} nd={nd} window={w} prec={prec}", file=sys.stderr,
              flush=True)
        outs.append(_noalloc(pos, types, ntype, nd, L, w, il, prec))
    free1 = torch.cuda.mem_get_info()[0]
    print(f"noalloc: device free memory {free0} -> {free1}", file=sys.stderr, flush=True)
    for (pos, types, ntype, nd, L, w, il), got in zip(cases, outs):
        ref = _existing(pos, types, ntype, nd, L, w, il, prec)
        assert _rel(got, ref) < tol
        assert abs(float(got[0, 0].double().mean())) < 1e-5  # normalised: mean delta 0


def test_noalloc_rejects_sizes_beyond_its_scratch():
    lib = N.load()
    pos = torch.zeros((10, 3), dtype=torch.float32, device="cuda")
    types = torch.zeros(10, dtype=torch.int32, device="cuda")
    g = torch.zeros((65, 1, 8, 8, 8), dtype=torch.float32, device="cuda")
    rc = lib.exh_grid_compute_noalloc(None, N.ptr(pos), None, None, N.ptr(types), 10, 65, 8,
                                      32.0, 0.31, 0.0, -1, 2, 4.0, 5.0, 0, 1, 4, N.ptr(g))
    assert rc == -1 and "64 tracer types" in N.last_error()
