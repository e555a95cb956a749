"""exshalos_amd: MI355X-native (gfx950) implementation of the ExSHalos
field -> 1LPT/2LPT -> excursion-set halos -> mass assignment -> P(k) path.

API mirror of pyexshalos (mock, simulation, utils) over a C-ABI HIP library
(include/exhip.h, built into exshalos_amd/lib/libexhip.so)."""

from . import mock, simulation, synth, utils  # noqa: F401
from .lib import get_precision, set_precision  # noqa: F401

__version__ = "0.1.0"
