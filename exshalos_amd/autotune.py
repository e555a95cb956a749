"""Autotuner of the batched runs (exshalos_amd.batch).

Measures, on the GPU it runs on, (1) each grid size's window-deposit tile
(EXH_TUNE_DEPOSIT_TILE: 16^3, 32 x 16 x 16 or 32 x 32 x 16 particles per
workgroup) as the time per box of a few boxes of that size run together,
and (2) the number of concurrent contexts (HIP streams) for a mix of all
the sizes (16 boxes of each by default, the batching workload), as cells
per second, best of three runs.  The fastest settings, with every
measurement, go to a JSON table that BatchRunner reads:

    python -m exshalos_amd.autotune [--out exshalos_amd/tuning/mi355x.json]

It raises GPU_MAX_HW_QUEUES from HIP's default of 4 to batch.HW_QUEUES
(another value in the environment is kept) and records it in the table.
"""

from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

from . import synth
from .batch import HW_QUEUES, TUNING_FILE, BatchRunner

SIZES = (64, 96, 128, 192)
TILES = (16, 3216, 32)
WORKERS = (1, 2, 4, 8, 12, 16)


def _noises(boxes, precision):
    import torch

    dt = np.float32 if precision == 4 else np.float64
    return [torch.as_tensor(synth.white_noise(nd, seed, dt), device="cuda") for nd, seed in boxes]


def _timed(runner, boxes, noises, reps=2):
    runner.run(boxes, noises)  # warm-up: buffers, plans, tables
    best = None
    for _ in range(reps):
        _, dt = runner.run(boxes, noises)
        best = dt if best is None else min(best, dt)
    return best


def tune(k, P, sizes=SIZES, tiles=TILES, workers=WORKERS, per_size=4, mix_per_size=16,
         precision=4, log=print):
    table = {"device": "MI355X (gfx950)", "precision": precision, "sizes": {},
             "measured": time.strftime("%Y-%m-%d"),
             "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))}
    for nd in sizes:
        boxes = [(nd, s + 1) for s in range(per_size)]
        noises = _noises(boxes, precision)
        res = {}
        for t in tiles:
            r = BatchRunner(k, P, precision=precision, workers=per_size,
                            table={"sizes": {str(nd): {"deposit_tile": t}}})
            res[str(t)] = round(_timed(r, boxes, noises) / per_size * 1e3, 3)
            r.close()
            log(f"[autotune] nd={nd} tile={t}: {res[str(t)]} ms per box")
        best = min(res, key=res.get)
        table["sizes"][str(nd)] = {"deposit_tile": int(best), "ms_per_box": res}
    mix = [(nd, s + 1) for nd in sizes for s in range(mix_per_size)]
    noises = _noises(mix, precision)
    cells = sum(nd**3 for nd, _ in mix)
    wres = {}
    for w in workers:
        r = BatchRunner(k, P, precision=precision, workers=w, table=table)
        wres[str(w)] = round(cells / _timed(r, mix, noises, reps=3) / 1e6, 1)
        r.close()
        log(f"[autotune] workers={w}: {wres[str(w)]} Mcells/s")
    table["workers"] = int(max(wres, key=wres.get))
    table["mcells_per_s_by_workers"] = wres
    return table


def main():
    if os.environ.get("GPU_MAX_HW_QUEUES", "4") == "4":  # HIP's default; before the runtime starts
        os.environ["GPU_MAX_HW_QUEUES"] = str(HW_QUEUES)
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=TUNING_FILE)
    ap.add_argument("--per-size", type=int, default=4,
                    help="boxes of each size in a tile measurement")
    ap.add_argument("--mix-per-size", type=int, default=16,
                    help="boxes of each size in the worker-count measurement")
    args = ap.parse_args()
    k, P = synth.bbks_spectrum()
    table = tune(k, P, per_size=args.per_size, mix_per_size=args.mix_per_size,
                 log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps(table))


if __name__ == "__main__":
    main()
