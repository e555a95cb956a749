"""Many realisations of uneven size, run concurrently on one GPU (an
addition: the reference runs one box per call).

Each realisation is the full path of ``Generate_Halos_Box_from_Pk``
(box.c:4-68: field, excursion-set halos, 2LPT particles) followed by the
mass assignment of two tracers -- the particles (matter) and the halo
positions -- and their auto- and cross power spectra in one
``power_compute`` call (ntype = 2, powermodule.c:62-74 pairs).

The halo finder is latency-bound (a sort level or a growth round is a few
short kernels), so one box leaves most of the GPU idle.  ``BatchRunner``
runs ``workers`` library contexts, each with its own HIP stream and
buffers, from host threads (the C calls release the GIL): the boxes'
kernels interleave on the GPU.  Boxes are dealt largest first.  The worker
count and each grid size's deposit tile come from the autotuner's table
(``exshalos_amd.autotune``, ``exshalos_amd/tuning/mi355x.json``).

HIP maps a process's streams onto ``GPU_MAX_HW_QUEUES`` hardware queues (4
by default), and streams that share a queue serialise; every context uses
up to three streams.  Set ``GPU_MAX_HW_QUEUES=16`` (``HW_QUEUES``) before
the HIP runtime starts: the 64-box workload ran 1.2x faster than with 4
queues, and 24 or 32 queues were slower than 16.
"""

from __future__ import annotations

import ctypes as C
import json
import os
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from .lib import _native as N
from .lib._common import host_f64

TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "mi355x.json")
HW_QUEUES = 16  # GPU_MAX_HW_QUEUES for batched runs


@dataclass
class BoxResult:
    """One realisation: halo count, bin k, the three spectra [mm, hm, hh] x
    Nk (pairs (0,0), (1,0), (1,1) in the reference's order), mode counts."""

    nd: int
    seed: int
    nh: int
    k: np.ndarray
    Pk: np.ndarray
    counts: np.ndarray


def load_table(path: str = TUNING_FILE) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


@dataclass
class _Worker:
    ctx: C.c_void_p
    bufs: dict = field(default_factory=dict)
    cap: int = 0


class BatchRunner:
    """Run boxes [(nd, seed), ...] with Lc, P(k) and Nk fixed.

    precision 4: FP32 fields with FP64 bin accumulators and an FP64 bin
    index; 8: FP64 throughout."""

    def __init__(self, k, P, Lc=4.0, Nk=64, precision=4, workers=None, table=None, device=0,
                 Om0=0.31, z=0.0, k_smooth=1e4):
        import torch

        N.require_gpu()
        if os.environ.get("GPU_MAX_HW_QUEUES", "4") == "4":
            import warnings

            warnings.warn(f"GPU_MAX_HW_QUEUES is 4 (HIP's default): the contexts' "
                          f"streams share queues; set it to {HW_QUEUES} before the HIP "
                          f"runtime starts", RuntimeWarning, stacklevel=2)
        self.lib = N.load()
        self.device = device
        self.prec = precision
        self.dt = torch.float32 if precision == 4 else torch.float64
        self.Lc, self.Nk, self.Om0, self.z, self.k_smooth = Lc, Nk, Om0, z, k_smooth
        R = np.float32 if precision == 4 else np.float64
        self.kh = host_f64(np.asarray(k).astype(R))
        self.Ph = host_f64(np.asarray(P).astype(R))
        self.table = load_table() if table is None else table
        self.nworkers = int(workers or self.table.get("workers", 4))
        self.workers = []
        for _ in range(self.nworkers):
            h = C.c_void_p()
            N.check(self.lib.exh_ctx_create(device, C.byref(h)), "exh_ctx_create")
            self.workers.append(_Worker(ctx=h))

    def close(self):
        for w in self.workers:
            self.lib.exh_ctx_destroy(w.ctx)
        self.workers = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _tile_for(self, nd):
        per = self.table.get("sizes", {}).get(str(nd), {})
        return int(per.get("deposit_tile", 0))

    def _buffers(self, w: _Worker, nd: int):
        import torch

        ng = nd**3
        if w.cap < ng:
            dev = torch.device("cuda", self.device)
            w.bufs = {
                "pos": torch.empty(3 * ng, dtype=self.dt, device=dev),
                "posh": torch.empty(3 * ng, dtype=self.dt, device=dev),
                "grid": torch.empty(2 * ng, dtype=self.dt, device=dev),
            }
            w.cap = ng
        # views of this box's size at the start of the worker's buffers (the
        # two tracer grids contiguous, as power_compute reads them)
        return {"pos": w.bufs["pos"][:3 * ng], "posh": w.bufs["posh"][:3 * ng],
                "grid": w.bufs["grid"][:2 * ng].view(2, ng)}

    def _run_box(self, w: _Worker, nd: int, seed: int, noise) -> BoxResult:
        lib, prec = self.lib, self.prec
        b = N.box(nd, self.Lc, self.Om0, self.z, -1.0, prec)
        s = N.FieldSrc()
        s.kind = N.SRC_MODES
        s.data = noise.data_ptr()
        s.k = self.kh.ctypes.data
        s.P = self.Ph.ctypes.data
        s.Nk = len(self.kh)
        s.R_max = 1e5
        o = N.HaloOpts()
        o.Nmin = 1
        o.a, o.beta, o.alpha = 1.0, 0.0, 0.0
        o.seed = int(seed)
        o.k = self.kh.ctypes.data
        o.P = self.Ph.ctypes.data
        o.Nk = len(self.kh)
        buf = self._buffers(w, nd)
        ng = nd**3
        L = float(nd * self.Lc)
        N.check(lib.exh_ctx_set_tunable(w.ctx, N.TUNE_DEPOSIT_TILE, self._tile_for(nd)),
                "exh_ctx_set_tunable")
        nh = C.c_longlong()
        N.check(lib.exh_halos_box_from_pk(w.ctx, C.byref(b), C.byref(s), C.byref(o),
                                          float(self.k_smooth), 1, 0, None, N.ptr(buf["pos"]),
                                          None, C.byref(nh)), "exh_halos_box_from_pk")
        N.check(lib.exh_halos_posh(w.ctx, N.ptr(buf["posh"]), None, 0), "exh_halos_posh")
        g0 = buf["grid"][0]
        g1 = buf["grid"][1]
        for pos, npart, g in ((buf["pos"], ng, g0), (buf["posh"], nh.value, g1)):
            N.check(lib.exh_grid_compute(w.ctx, N.ptr(pos), None, None, None, int(npart), 1, nd,
                                         L, float(self.Om0), float(self.z), -1, 2, 4.0, 5.0, 0,
                                         1, prec, N.ptr(g)), "exh_grid_compute")
        kk = np.zeros(self.Nk)
        Pk = np.zeros(3 * self.Nk)
        cnt = np.zeros(self.Nk, dtype=np.int64)
        N.check(lib.exh_power_compute(w.ctx, N.ptr(buf["grid"]), 2, nd, L, 2, 4.0, 0, self.Nk,
                                      2 * np.pi / L, np.pi / L * nd, 0, -1, prec, 8, N.ptr(kk),
                                      N.ptr(Pk), N.ptr(cnt)), "exh_power_compute")
        return BoxResult(nd=nd, seed=seed, nh=int(nh.value), k=kk, Pk=Pk.reshape(3, self.Nk),
                         counts=cnt)

    def run(self, boxes, noises):
        """boxes: [(nd, seed)]; noises: device tensors [nd][nd][nd//2+1][2]
        of the same order.  Returns (results in the input order, seconds)."""
        import torch

        torch.cuda.synchronize(self.device)  # the noise uploads ran on torch's stream
        order = sorted(range(len(boxes)), key=lambda i: -boxes[i][0])
        results = [None] * len(boxes)
        errors = []
        nxt = [0]
        lock = threading.Lock()

        def work(w: _Worker):
            while True:
                with lock:
                    if nxt[0] >= len(order) or errors:
                        return
                    i = order[nxt[0]]
                    nxt[0] += 1
                try:
                    nd, seed = boxes[i]
                    results[i] = self._run_box(w, nd, seed, noises[i])
                except Exception as e:  # surfaced after the join
                    errors.append(e)
                    return

        t0 = time.perf_counter()
        threads = [threading.Thread(target=work, args=(w,)) for w in self.workers]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        dt = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return results, dt
