"""CPU checks of the drop-in boundary: the HIP library builds for gfx950,
loads, and exports every entry point include/cylpic.h declares.  No compute
call is made (there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "cylpic.h")).read()
    return sorted(set(re.findall(r"\b(cylpic_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def libpath():
    from cylindrical_epoch_amd import build
    if not os.path.exists(build.LIB):
        build.build()
    return build.LIB


def test_every_declared_symbol_is_exported(libpath):
    L = C.CDLL(libpath)
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header(libpath):
    from cylindrical_epoch_amd import _lib
    names = set(_declared()) - {"cylpic_last_error"}
    assert names == set(_lib.SIGNATURES), names ^ set(_lib.SIGNATURES)


def test_init_rejects_bad_arguments_without_gpu(libpath):
    from cylindrical_epoch_amd._lib import lib
    L = lib()
    h = C.c_void_p()
    bc = (C.c_int * 4)(1, 1, 2, 10)
    # ng must be 5: rejected before any device call
    rc = L.cylpic_init(C.byref(h), 32, 16, 2, 4, 1e-7, 1e-7, 1e-17, 0.0, 3.2e-6, 0.0, 1.6e-6,
                       C.cast(bc, C.c_void_p), C.cast(bc, C.c_void_p), 0, 0)
    assert rc == 1
    assert b"ng must be 5" in L.cylpic_last_error()
    assert not h.value


def test_init_rejects_cpml_without_gpu(libpath):
    """CPML field boundaries would grow the grid and change the particle
    edges (mpi_routines.F90:280-290); they are refused, not mapped."""
    from cylindrical_epoch_amd._lib import lib
    L = lib()
    h = C.c_void_p()
    bp = (C.c_int * 4)(5, 5, 2, 5)
    for code in (12, 13):
        bf = (C.c_int * 4)(code, 4, 2, 4)
        rc = L.cylpic_init(C.byref(h), 32, 16, 2, 5, 1e-7, 1e-7, 1e-17, 0.0, 3.2e-6, 0.0, 1.6e-6,
                           C.cast(bf, C.c_void_p), C.cast(bp, C.c_void_p), 0, 0)
        assert rc == 1
        assert b"CPML" in L.cylpic_last_error()
        assert not h.value


def test_null_handle_calls_fail_cleanly(libpath):
    from cylindrical_epoch_amd._lib import lib
    L = lib()
    assert L.cylpic_push_particles(None) == 1
    assert L.cylpic_step(None, None, None) == 1


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "cylindrical_epoch_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in txt.lower(), f
