"""Write the golden fixtures of this directory from the CPU oracle.

    python tests/golden/make_golden.py

Each fixture stores its inputs (noise modes, P(k) knots, particles) with
the oracle's outputs, so tests can check both the oracle (regressions) and
the HIP library (parity) against fixed data without recomputing the
oracle.  Sizes are kept small (nd = 16 and 32).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pipeline as O  # noqa: E402
from exshalos_amd.synth import bbks_spectrum  # noqa: E402

LC = 4.0


def field_and_halos(prec):
    R = np.float32 if prec == 4 else np.float64
    k, P = bbks_spectrum()
    nd = 32
    noise = np.random.default_rng(21).standard_normal((nd, nd, nd // 2 + 1, 2)).astype(R)
    delta, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    hal = O.find_halos(delta, k, P, LC, seed=12345, precision=prec)
    np.savez_compressed(
        os.path.join(HERE, f"halos_nd32_p{prec}.npz"), k=k, P=P, noise=noise, delta=delta,
        x=hal["x"], count=hal["count"], r2=hal["r2"], Mh=hal["Mh"],
        flag=hal["flag"].astype(np.int32))
    # LPT at nd = 16: 1LPT + 2LPT displacements from the same kind of modes
    nd = 16
    noise = np.random.default_rng(23).standard_normal((nd, nd, nd // 2 + 1, 2)).astype(R)
    delta, dk, p = O.field_from_modes(k, P, nd, LC, noise, precision=prec)
    S1, S2 = O.lpt(dk, p, 1e4, True, prec)
    np.savez_compressed(os.path.join(HERE, f"lpt_nd16_p{prec}.npz"), noise=noise, delta=delta,
                        S1=S1, S2=S2)


def grid_and_power(prec):
    R = np.float32 if prec == 4 else np.float64
    nd, L = 16, 64.0
    rng = np.random.default_rng(22)
    pos = rng.uniform(0, L, (3000, 3)).astype(R)
    out = {"pos": pos}
    for win in (1, 2, 5):
        g = O.grid_compute(pos, nd, L, window=win, interlacing=True, precision=prec)
        pk = O.power_compute(g, 1, nd, L, window=win, interlacing=True, Nk=8, l_max=2,
                             direction=2, precision=prec)
        out[f"grid_w{win}"] = g
        out[f"Pk_w{win}"] = pk["Pk64"]
        out[f"k_w{win}"] = pk["k64"]
        out[f"Nk_w{win}"] = pk["Nk"]
    np.savez_compressed(os.path.join(HERE, f"grid_power_nd16_p{prec}.npz"), **out)


def main():
    for prec in (4, 8):
        field_and_halos(prec)
        grid_and_power(prec)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
