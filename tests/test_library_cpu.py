"""Checks of the HIP library that need no GPU: it loads, exports every
function include/exhip.h declares, the ctypes prototypes cover exactly that
set, and the host-side entry points (box parameters, halo tables) agree
bit for bit with the oracle.  Error paths return status codes with a
message and never exit."""

import ctypes as C
import os
import re

import numpy as np
import pytest

from exshalos_amd.lib import _native as N
from oracle import pipeline as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "exhip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(exh_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libexhip.so is not built")
    return N.load()


def test_header_declares_the_boundary():
    names = _declared()
    assert len(names) >= 20
    for must in ("exh_find_halos", "exh_halos_box_from_pk", "exh_grid_compute",
                 "exh_power_compute", "exh_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    for name in _declared():
        assert hasattr(lib, name), name


def test_ctypes_prototypes_match_header():
    assert sorted(N.exported_symbols()) == _declared()


def test_version_and_error_reporting(lib):
    assert "gfx950" in lib.exh_version().decode()
    b = N.Box()
    rc = lib.exh_box_init(C.byref(b), 64, 64, 64, 500.0, 0.31, 0.0, -1.0, 4)
    assert rc == N.EXH_OK - 1
    assert "Lc/Mc" in N.last_error()
    assert lib.exh_box_init(C.byref(b), 64, 64, 64, 4.0, 0.31, 0.0, -1.0, 5) == -1
    assert "precision" in N.last_error()
    with pytest.raises(ValueError):
        N.box(64, 500.0)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("Lc", [4.0, 2.0, 3e12])
def test_box_init_matches_oracle(lib, prec, Lc):
    b = N.box(64, Lc, precision=prec)
    p = O.params(64, Lc, precision=prec)
    R = np.float32 if prec == 4 else np.float64
    for f in ("Om0", "dc", "rhoc", "Hz", "Omz", "rhomz", "Dv", "Lc", "Mcell", "Mtot", "Normx",
              "Normk"):
        assert R(getattr(b, f)) == R(getattr(p, f)), f
    assert list(b.nd) == list(p.nd) and b.nz2 == p.nz2 and b.ng == p.ng
    for a in range(3):
        assert R(b.L[a]) == R(p.L[a]) and R(b.kl[a]) == R(p.kl[a])


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("eb", [(1.0, 0.0, 0.0), (0.8, 0.29, 0.53)])
@pytest.mark.parametrize("radii", [0, 64])
def test_halo_tables_bit_identical_to_oracle(lib, bbks, prec, eb, radii):
    """radii 0: the reference's table (one radius per k, the grid's float
    type); 64: the FP64 table over 64 radii."""
    k, P = bbks
    R = np.float32 if prec == 4 else np.float64
    b = N.box(64, 4.0, precision=prec)
    p = O.params(64, 4.0, precision=prec)
    ref = O.halo_tables(p, k, P, *eb, precision=prec, sigma_radii=radii)
    Ncells = ref["Ncells"]
    nB = len(ref["Bt"])
    kk = np.ascontiguousarray(np.asarray(k, dtype=R).astype(np.float64))
    PP = np.ascontiguousarray(np.asarray(P, dtype=R).astype(np.float64))
    o = N.HaloOpts(Nmin=1, a=eb[0], beta=eb[1], alpha=eb[2], seed=1, k=kk.ctypes.data,
                   P=PP.ctypes.data, Nk=len(kk), sigma_radii=radii)
    sig = np.zeros(Ncells, dtype=R)
    Bt = np.zeros(nB, dtype=R)
    Mk = np.zeros(len(ref["Mknots"]))
    Int = np.zeros(len(ref["Int"]))
    nc = C.c_int()
    N.check(lib.exh_halo_tables(C.byref(b), C.byref(o), N.ptr(sig), N.ptr(Bt), nB, N.ptr(Mk),
                                N.ptr(Int), C.byref(nc)))
    assert nc.value == Ncells
    np.testing.assert_array_equal(sig, ref["Sig_grid"])
    np.testing.assert_array_equal(Bt, ref["Bt"])
    np.testing.assert_array_equal(Mk, ref["Mknots"])
    np.testing.assert_array_equal(Int, ref["Int"])


def test_context_needs_a_gpu_and_fails_loudly():
    """Without a HIP device the product path raises; it never falls back to
    a CPU implementation."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        N.context()
