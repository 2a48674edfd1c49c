"""GLU's own entry points, called as GLU calls them: `struct site` arrays of
link pointers and the global `Latt`, through the library's C ABI.

FACG() and Landau() (Gfix/FACG.c:341-446, Gfix/Landau.c:261-374) must end
below the accuracy with the oracle's iteration count (within 2, or 5%) and
functional (1e-10 relative), the plaquette preserved to 1e-14; Coulomb()
(Gfix/Coulomb.c:33-100) must fix every slice and agree with the oracle's
spatial link trace to 1e-10 relative.  cuts_struct() (Cuts/cuts.c:114-184)
must write the file GLU names (Cuts/cut_output.c:115-258) with the oracle's
momentum list byte for byte and its values to 1e-12 of the largest, and
refuse the modes and the LOG definition it does not provide."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import heatbath_config, rel

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

SZ = C.c_size_t


class LattInfo(C.Structure):
    """struct latt_info, Headers/GLU_types.h:229-243"""
    _fields_ = [("dims", SZ * 4), ("Lsq", SZ), ("Lcu", SZ), ("Volume", SZ), ("flow", SZ),
                ("gf_alpha", C.c_double), ("sm_alpha", C.c_double * 3), ("head", C.c_int),
                ("Seed", C.c_uint32 * 1), ("twiddles", C.c_double * 4), ("Nthreads", C.c_uint32),
                ("su2_data", C.c_void_p), ("argc", C.c_int)]


class CutInfo(C.Structure):
    """struct cut_info, Headers/GLU_types.h:94-103"""
    _fields_ = [("dir", C.c_int), ("type", C.c_int), ("max_mom", SZ), ("max_t", SZ),
                ("where", C.c_char * 256), ("definition", C.c_int), ("angle", SZ), ("cyl_width", C.c_double)]


EXCEPTIONAL, NONEXCEPTIONAL, GLUON_PROPS = 0, 1, 6
LINEAR_DEF, LOG_DEF = 0, 1
CUT_TYPE = {"hypercubic": 0, "psq": 1, "cylinder": 2, "conical": 3}


@pytest.fixture(scope="module")
def glu():
    lib = C.CDLL(gl.LIB_PATH)
    lib.FACG.restype = SZ
    lib.FACG.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_double, SZ]
    lib.Landau.restype = SZ
    lib.Landau.argtypes = [C.c_void_p, C.c_double, SZ, C.c_char_p]
    lib.Coulomb.restype = SZ
    lib.Coulomb.argtypes = [C.c_void_p, C.c_double, SZ]
    lib.cuts_struct.restype = C.c_int
    lib.cuts_struct.argtypes = [C.c_void_p, CutInfo]
    return lib


def set_latt(lib, dims, gf_alpha, flow=0):
    L = LattInfo.in_dll(lib, "Latt")
    for mu in range(4):
        L.dims[mu] = dims[mu]
    L.Lsq = dims[0] * dims[1]
    L.Lcu = L.Lsq * dims[2]
    L.Volume = L.Lcu * dims[3]
    L.flow = flow
    L.gf_alpha = gf_alpha
    L.Seed[0] = 1
    L.Nthreads = 1


class Sites:
    """GLU's struct site array over the links U[V, 4, 18] (in place): each
    site's O points at 4 pointers, one per direction, to 9 complex numbers"""

    def __init__(self, U):
        assert U.flags.c_contiguous and U.dtype == np.float64
        self.U = U
        V = U.shape[0]
        self.ptr = (U.ctypes.data + 144 * np.arange(V * 4, dtype=np.uint64)).reshape(V, 4)
        dt = np.dtype([("O", "u8"), ("neighbor", "u8", (4,)), ("back", "u8", (4,))])
        assert dt.itemsize == 72
        self.site = np.zeros(V, dt)
        self.site["O"] = self.ptr.ctypes.data + 32 * np.arange(V, dtype=np.uint64)

    @property
    def p(self):
        return self.site.ctypes.data


@pytest.fixture(scope="module")
def landau_fixed(glu):
    dims = (8, 8, 8, 16)
    U, plaq = heatbath_config(dims, sweeps=20)
    L = O.Lattice(dims, nth=4)
    W = U.copy()
    r = L.facg(W, 0.08, 1e-14, 1000)
    set_latt(glu, dims, 0.08)
    lat = Sites(U.copy())
    th = C.c_double()
    print(f"FACG() {dims}", flush=True)
    it = glu.FACG(lat.p, None, C.byref(th), 1e-14, 1000)
    return dims, U, plaq, L, W, r, lat.U, it, th.value


def test_facg_entry_point_matches_the_oracle(landau_fixed):
    dims, U, plaq, L, W, r, Ug, it, th = landau_fixed
    assert th < 1e-14
    assert abs(it - r["iters"]) <= max(2, r["iters"] // 20), (it, r["iters"])
    assert rel(L.functional(Ug), L.functional(W)) < 1e-10
    assert not np.array_equal(Ug, U)  # the links were fixed in place


def test_landau_entry_point(glu, landau_fixed):
    """Landau() fixes, reunitarises and reports; from the same input it takes
    FACG()'s iteration count"""
    dims, U, plaq, L, W, r, _, it, _ = landau_fixed
    set_latt(glu, dims, 0.08)
    lat = Sites(U.copy())
    print(f"Landau() {dims}", flush=True)
    n = glu.Landau(lat.p, 1e-14, 1000, b"")
    assert n == it
    assert L.theta_lin(lat.U) < 1e-14
    assert rel(L.functional(lat.U), L.functional(W)) < 1e-10
    assert abs(L.plaquette(lat.U) - plaq) < 1e-14


def test_coulomb_entry_point_matches_the_oracle(glu):
    dims = (6, 6, 6, 8)
    U, plaq = heatbath_config(dims, sweeps=20)
    L = O.Lattice(dims, nth=4)
    W = U.copy()
    r = L.coulomb(W, 0.09, 1e-14, 650, facg=True)
    set_latt(glu, dims, 0.09)
    lat = Sites(U.copy())
    print(f"Coulomb() {dims}", flush=True)
    n = glu.Coulomb(lat.p, 1e-14, 650)
    assert abs(n - r["iters"]) <= 2 * dims[3], (n, r["iters"])
    tr_o = (W[..., 0] + W[..., 8] + W[..., 16])[:, :3].mean()
    tr_g = (lat.U[..., 0] + lat.U[..., 8] + lat.U[..., 16])[:, :3].mean()
    assert rel(tr_g, tr_o) < 1e-10
    assert abs(L.plaquette(lat.U) - plaq) < 1e-13


def glu_file_name(prefix, dims, max_mom, tag, flow):
    """Cuts/cut_output.c:115-258 with SIN_MOM and the LINEAR definition"""
    d = "x".join(str(x) for x in dims)
    return f"{prefix}_SU3_{d}_{max_mom}_{tag}_SINMOM_LIN.{flow}.bin"


@pytest.mark.parametrize("mode,cut,max_mom,tag", [(GLUON_PROPS, "psq", 6, "cutPSQ"),
                                                  (GLUON_PROPS, "cylinder", 12, "cutCYL_w1"),
                                                  (EXCEPTIONAL, "conical", 12, "cutCON_20"),
                                                  (EXCEPTIONAL, "hypercubic", 3, "cutHYP")])
def test_cuts_struct_writes_glu_file(glu, landau_fixed, tmp_path, mode, cut, max_mom, tag):
    dims, _, _, L, _, _, Ug, _, _ = landau_fixed
    set_latt(glu, dims, 0.08, flow=1200)
    ci = CutInfo(dir=mode, type=CUT_TYPE[cut], max_mom=max_mom, max_t=0, where=(str(tmp_path) + "/").encode(),
                 definition=LINEAR_DEF, angle=20, cyl_width=1.0)
    lat = Sites(Ug.copy())
    print(f"cuts_struct mode={mode} cut={cut} max_mom={max_mom}", flush=True)
    assert glu.cuts_struct(lat.p, ci) == 0
    assert np.array_equal(lat.U, Ug)  # this library leaves the links alone
    name = glu_file_name("G2F2" if mode == GLUON_PROPS else "MOMgg", dims, max_mom, tag, 1200)
    assert os.listdir(tmp_path) == [name]
    f = O.gluon_prop if mode == GLUON_PROPS else O.gluon_exceptional
    mo, A, B = f(L, Ug, cut, max_mom, 1.0, 20.0)
    q = str(tmp_path / "oracle.bin")
    O.write_glueprop(q, mo, A, B)
    a, b = open(str(tmp_path / name), "rb").read(), open(q, "rb").read()
    n = len(mo)
    head = 4 + 20 * n
    assert len(a) == len(b) == head + 2 * (4 + 8 * n)
    assert a[:head] == b[:head]
    for k, ref in enumerate((A, B)):
        lo = head + k * (4 + 8 * n)
        assert a[lo:lo + 4] == b[lo:lo + 4]
        got = np.frombuffer(a[lo + 4:lo + 4 + 8 * n], ">f8")
        scale = np.abs(A).max() if mode == GLUON_PROPS else np.abs(ref).max()
        assert np.abs(got - ref).max() < 1e-12 * scale


@pytest.mark.parametrize("mode,definition", [(GLUON_PROPS, LOG_DEF), (NONEXCEPTIONAL, LINEAR_DEF)])
def test_cuts_struct_refusals(glu, tmp_path, mode, definition):
    dims = (4, 4, 4, 8)
    set_latt(glu, dims, 0.08)
    U = np.zeros((int(np.prod(dims)), 4, 18))
    U[..., [0, 8, 16]] = 1.0
    ci = CutInfo(dir=mode, type=1, max_mom=3, where=(str(tmp_path) + "/").encode(), definition=definition)
    assert glu.cuts_struct(Sites(U).p, ci) == -1  # GLU_FAILURE
    assert os.listdir(tmp_path) == []
