"""Error paths of the C ABI on a device: every refused request returns a
negative code with a message and leaves the library usable."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")


def create(dims, gftype=0, precision=0):
    h = C.c_void_p()
    rc = gl.lib().glu_gfix_create(C.byref(h), (gl.SZ * 4)(*dims), gftype, precision)
    return rc, h


@pytest.mark.parametrize("dims", [(1, 8, 8, 8), (8, 8, 8, 5000), (11, 8, 8, 8), (8, 8, 8, 13), (2560, 4, 4, 4)])
def test_bad_extents_are_refused(dims):
    rc, h = create(dims)
    assert rc < 0 and not h.value
    msg = gl.lib().glu_gfix_last_error().decode()
    assert "extent" in msg


def test_volume_beyond_int32_is_refused():
    rc, h = create((2048, 2048, 512, 2))
    assert rc < 0 and not h.value
    assert "volume" in gl.lib().glu_gfix_last_error().decode()


def test_bad_arguments_leave_the_context_usable():
    dims = (4, 4, 4, 4)
    U = np.zeros((256, 4, 18))
    U[..., [0, 8, 16]] = 1.0
    with gl.Gfix(dims) as g:
        g.set_links(U)
        lib = gl.lib()
        th, it = C.c_double(), gl.SZ()
        assert lib.glu_gfix_run(g._h, 7, 0.08, 1e-14, 100, C.byref(th), C.byref(it)) < 0
        assert lib.glu_gfix_run(g._h, gl.CG, 0.08, 1e-14, 0, C.byref(th), C.byref(it)) < 0
        assert lib.glu_gfix_slice_info(g._h, (gl.SZ * 4)(), np.zeros(4).ctypes.data_as(gl.DP)) < 0
        assert lib.glu_gfix_restore(g._h) < 0  # no snapshot yet
        with pytest.raises(gl.GfixError):
            g.luxury(copies=0)
        # still works: unit links are already fixed
        it2, th2 = g.run(gl.CG, 0.08, 1e-14, 100)
        assert it2 == 0 and th2 < 1e-14


def test_landau_only_entry_points_refuse_coulomb_contexts():
    with gl.Gfix((4, 4, 4, 4), gl.COULOMB) as g:
        with pytest.raises(gl.GfixError):
            g.landau()
        with pytest.raises(gl.GfixError):
            g.luxury(copies=2)
