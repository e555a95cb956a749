"""Native layer: the same module names as the reference's compiled
extensions (pyexshalos/lib/__init__.py:1-4) for the hot path, backed by the
gfx950 library libexhip.so through ctypes."""

from . import exshalos, spectrum  # noqa: F401
from ._common import get_precision, set_precision  # noqa: F401
