"""C-ABI checks that need no GPU: the library loads and exports exactly the
entry points include/mdforce.h declares, and the ctypes binding covers them."""
import ctypes as ct
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mdforce.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mdf_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("mdf_p3m_create", "mdf_p3m_force", "mdf_pp_force", "mdf_pm_force", "mdf_md_run",
              "mdf_p3m_build_green", "mdf_density_modes"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from mdforce import _lib
    if not os.path.isfile(_lib.LIB_PATH):
        pytest.fail("libmdforce.so is not built (run __graft_entry__.build())")
    so = ct.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_binding_covers_header_and_abi_version():
    from mdforce import _lib
    L = _lib.lib()
    assert L.mdf_abi_version() == _lib.ABI_VERSION
    bound = set(_lib.SIGNATURES) | {"mdf_abi_version", "mdf_last_error"}
    assert set(declared()) == bound


def test_params_struct_layout():
    from mdforce._lib import P3MParams
    # int n_max; int mesh[3]; double box[3]; double kappa, alpha, rc; int order, precision
    assert P3MParams.box.offset == 16
    assert P3MParams.kappa.offset == 40
    assert P3MParams.order.offset == 64
    assert ct.sizeof(P3MParams) == 72


def test_create_rejects_bad_geometry_without_gpu_work():
    # argument validation happens before any device call
    from mdforce._lib import P3MParams, lib
    p = P3MParams()
    p.n_max = 10
    for d in range(3):
        p.mesh[d] = 48   # not a power of two
        p.box[d] = 10.0
    p.kappa, p.alpha, p.rc, p.order, p.precision = 1.0, 0.5, 3.0, 6, 0
    ctx = ct.c_void_p()
    assert lib().mdf_p3m_create(ct.byref(p), ct.byref(ctx)) == -1
    assert b"power of two" in lib().mdf_last_error()
    for d in range(3):
        p.mesh[d] = 32
    p.rc = 4.0   # L/rc = 2.5 < 3
    assert lib().mdf_p3m_create(ct.byref(p), ct.byref(ctx)) == -1
    assert b"L/rc >= 3" in lib().mdf_last_error()
    p.rc, p.order = 3.0, 4
    assert lib().mdf_p3m_create(ct.byref(p), ct.byref(ctx)) == -1


def test_create_rejects_too_many_cells_before_allocating():
    # a large box at a small cutoff: 8 floor(L/rc)^3 sub-cells exceed the cell
    # table, refused before any allocation or device call
    from mdforce._lib import P3MParams, lib
    p = P3MParams()
    p.n_max = 1000
    for d in range(3):
        p.mesh[d] = 64
        p.box[d] = 330.0
    p.kappa, p.alpha, p.rc, p.order, p.precision = 0.5, 0.8, 3.0, 6, 1
    ctx = ct.c_void_p()
    assert lib().mdf_p3m_create(ct.byref(p), ct.byref(ctx)) == -1
    msg = lib().mdf_last_error()
    assert b"too many cells" in msg and b"increase rc" in msg
    assert b"sub-cells" in msg
    assert not ctx.value


def test_sqw_rejects_calls_beyond_the_grid_limit():
    # 2^31 series samples or more are refused before any device call (the
    # pointers are placeholders that are never dereferenced)
    from mdforce._lib import lib
    L = lib()
    fake = ct.c_void_p(0x1000)
    big = ct.c_size_t(1 << 62)
    # 3 * 40000 series of 20000 samples = 2.4e9 > 2^31
    rc = L.mdf_sqw(fake, 1, 20000, 40000, 1, 40000, 1.0, 0.1, 1.0, 1, fake, fake, big, None)
    assert rc == -1 and b"exceed" in L.mdf_last_error()
    # the workspace is two halves of 16 B per series sample; a null workspace is refused
    nbytes = ct.c_size_t()
    assert L.mdf_sqw_workspace(1, 20000, 40000, ct.byref(nbytes)) == 0
    assert nbytes.value == 2 * 3 * 40000 * 20000 * 16
    assert L.mdf_sqw(fake, 1, 20000, 40000, 1, 40000, 1.0, 0.1, 1.0, 1, fake, None, big, None) == -1
    assert b"null workspace" in L.mdf_last_error()
