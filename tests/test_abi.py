"""CPU-side checks of the C ABI: the library builds, loads, exports every
function include/glu_gfix.h declares, and refuses to run without a HIP
device instead of falling back to the CPU."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT

import glu_amd as gl


def declared_functions():
    text = open(os.path.join(ROOT, "include", "glu_gfix.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "sizeof", "defined"}
    glu_lower = {"cuts_struct"}  # GLU's own lower-case entry points
    return sorted({n for n in names if n not in skip and (n.startswith("glu_gfix") or n[0].isupper() or n in glu_lower)})


def test_header_lists_the_glu_entry_points():
    names = declared_functions()
    for n in ("Landau", "Coulomb", "FACG", "FASD", "Coulomb_FA", "cuts_struct"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(gl.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(gl.EXPORTS)


def test_no_cpu_fallback_without_device():
    h = C.c_void_p()
    dims = (C.c_size_t * 4)(4, 4, 4, 4)
    rc = gl.lib().glu_gfix_create(C.byref(h), dims, 0, 0)
    if rc == 0:  # a device is present: nothing to check here
        gl.lib().glu_gfix_destroy(h)
        pytest.skip("HIP device present")
    assert rc == -3
    assert b"no HIP device" in gl.lib().glu_gfix_last_error()
    with pytest.raises(gl.GfixError):
        gl.Gfix((4, 4, 4, 4))


def test_argument_errors_are_reported():
    h = C.c_void_p()
    assert gl.lib().glu_gfix_create(C.byref(h), (C.c_size_t * 4)(4, 4, 4, 4), 7, 0) == -1
    assert b"gftype" in gl.lib().glu_gfix_last_error()
    assert gl.footprint((32, 32, 32, 64)) > 32**3 * 64 * 576


@pytest.mark.parametrize("dims,gftype,word", [
    ((2560, 4, 4, 4), 0, "2048"),      # beyond the LDS FFT's pencil length
    ((4, 4, 4, 4096), 0, "2048"),      # t is transformed in Landau ...
    ((4, 4, 4, 5000), 1, "range"),     # ... not in Coulomb, but still capped
    ((11, 8, 8, 8), 0, "factor"),      # no radix 11 kernel
    ((8, 8, 8, 13), 1, None),          # Coulomb does not transform t: accepted
    ((2048, 2048, 512, 2), 0, "volume"),
    ((1, 8, 8, 8), 0, "range"),
])
def test_extent_refusals_need_no_device(dims, gftype, word):
    """Extents the kernels cannot handle are refused before any device
    work, with a message naming the limit."""
    h = C.c_void_p()
    rc = gl.lib().glu_gfix_create(C.byref(h), (C.c_size_t * 4)(*dims), gftype, 0)
    if word is None:
        if rc == 0:
            gl.lib().glu_gfix_destroy(h)
        assert rc in (0, -3)  # accepted; -3 only for the missing device
        return
    assert rc == -1 and not h.value
    assert word in gl.lib().glu_gfix_last_error().decode()


def test_oracle_is_not_linked_into_the_product():
    out = os.popen(f"nm -D --undefined-only {gl.LIB_PATH}").read()
    names = [l.split()[-1] for l in out.splitlines() if l.strip()]
    # the oracle's C symbols all start with or_
    assert not [n for n in names if n.startswith("or_")]


class _CutInfo(C.Structure):
    """struct cut_info, Headers/GLU_types.h:94-103"""
    _fields_ = [("dir", C.c_int), ("type", C.c_int), ("max_mom", C.c_size_t), ("max_t", C.c_size_t),
                ("where", C.c_char * 256), ("definition", C.c_int), ("angle", C.c_size_t),
                ("cyl_width", C.c_double)]


def test_cuts_struct_refusals_need_no_device():
    """GLU's cuts_struct refuses a missing lattice, the modes this library
    does not provide and the LOG definition with GLU_FAILURE before it
    touches the device (GLUON_PROPS = 6, NONEXCEPTIONAL = 1, LOG_DEF = 1)."""
    lib = C.CDLL(gl.LIB_PATH)
    lib.cuts_struct.restype = C.c_int
    lib.cuts_struct.argtypes = [C.c_void_p, _CutInfo]
    site = (C.c_char * 72)()
    assert lib.cuts_struct(None, _CutInfo(dir=6, type=1, max_mom=3)) == -1
    assert lib.cuts_struct(C.addressof(site), _CutInfo(dir=1, type=1, max_mom=3)) == -1
    assert lib.cuts_struct(C.addressof(site), _CutInfo(dir=6, type=1, max_mom=3, definition=1)) == -1
    assert lib.cuts_struct(C.addressof(site), _CutInfo(dir=6, type=9, max_mom=3)) == -1


@pytest.mark.parametrize("dims,nlat,word", [
    ((8, 8, 8, 32), 1, "extent"),      # a pencil longer than the direct DFT holds
    ((1, 8, 8, 8), 1, "extent"),
    ((8, 8, 8, 8), 70000, "65535"),    # more lattices than one launch's grid
])
def test_facg_batch_refusals_need_no_device(dims, nlat, word):
    """glu_gfix_facg_batch refuses shapes and batch sizes its one-workgroup
    kernel cannot hold before any device work; an empty batch of a valid
    shape is a no-op."""
    lib = gl.lib()
    it = (C.c_size_t * 1)()
    th = (C.c_double * 1)()
    d = (C.c_size_t * 4)(*dims)
    rc = lib.glu_gfix_facg_batch(d, nlat, C.c_void_p(8), 0.08, 1e-14, 100, it, th, None, 0)
    assert rc == -1 and word in lib.glu_gfix_last_error().decode()
    ok = (C.c_size_t * 4)(8, 8, 8, 8)
    assert lib.glu_gfix_facg_batch(ok, 0, C.c_void_p(8), 0.08, 1e-14, 100, it, th, None, 0) == 0
