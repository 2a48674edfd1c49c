"""NERSC output of the device links against the oracle's restatement of
GLU's writer (IO/writers.c:330-422, IO/write_headers.c:77-138).

The checksum is integer work and must agree bit for bit: the header word
the device computes equals the word the oracle's reader computes from the
bytes written.  The links written are the device links after the in-place
reunitarisation, exactly as glu_gfix_get_links returns them, and agree with
the oracle's reunitarisation of the same input to 1e-14 (the device
normalises with a Newton-refined reciprocal square root, GLU with
1/sqrt).  Header values are printed with 15 decimals."""
import os

import numpy as np
import pytest

from oracle import oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

CONF = os.path.join(GOLDEN, "conf_lat400.nersc")
VOLATILE = ("CREATOR_MACHINE", "CREATION_DATE")
NUMERIC = ("CHECKSUM", "LINK_TRACE", "PLAQUETTE")


def header(path, info):
    with open(path, "rb") as f:
        return [ln.split(" = ", 1) if " = " in ln else [ln, ""]
                for ln in f.read(info.data_offset).decode().splitlines()]


@pytest.mark.parametrize("two_rows,big", [(True, True), (False, False)])
def test_write_nersc_against_the_oracle_writer(tmp_path, two_rows, big):
    U, info = O.read_nersc(CONF)
    dims = tuple(info.dims)
    L = O.Lattice(dims, nth=4)
    p, q = str(tmp_path / "gpu.nersc"), str(tmp_path / "oracle.nersc")
    print(f"write_nersc fixture two_rows={two_rows} big={big}", flush=True)
    with gl.Gfix(dims) as g:
        g.set_links(U)
        g.write_nersc(p, two_rows=two_rows, big_endian=big, label="fixed", sequence=400)
        Ud = g.get_links()  # reunitarised in place
    Ug, ig = O.read_nersc(p)
    assert ig.checksum_header == ig.checksum_computed
    assert bool(ig.two_rows) == two_rows and bool(ig.big_endian) == big
    if two_rows:  # row 2 is completed by the reader
        assert np.array_equal(Ug[..., :12], Ud[..., :12])
    else:
        assert np.array_equal(Ug, Ud)
    R = U.copy()
    O.write_nersc(q, L, R, two_rows=two_rows, big_endian=big, label="fixed", sequence=400)
    assert np.abs(Ug - R).max() < 1e-14
    hg, ho = header(p, ig), header(q, O.read_nersc(q)[1])
    assert [k for k, _ in hg] == [k for k, _ in ho]
    for (k, vg), (_, vo) in zip(hg, ho):
        if k in ("LINK_TRACE", "PLAQUETTE"):
            assert abs(float(vg) - float(vo)) < 1e-14, k
        elif k not in VOLATILE + NUMERIC:
            assert vg == vo, k
    assert abs(ig.plaquette - L.plaquette(Ug)) < 1e-14
    assert abs(ig.link_trace - L.links(Ug)[0]) < 1e-14


def test_write_nersc_streams_several_chunks(tmp_path):
    dims = (16, 16, 16, 48)  # 196608 sites: two staging chunks
    p = str(tmp_path / "hb.nersc")
    print(f"write_nersc heatbath {dims}", flush=True)
    with gl.Gfix(dims) as g:
        g.heatbath(6.0, 2, 1, seed=11)
        g.write_nersc(p, two_rows=False, big_endian=True, sequence=2)
        Ud = g.get_links()
        m = g.measure()
    Ug, ig = O.read_nersc(p)
    assert list(ig.dims) == list(dims)
    assert ig.checksum_header == ig.checksum_computed
    assert np.array_equal(Ug, Ud)
    assert abs(ig.plaquette - m["plaquette"]) < 1e-14


def test_write_nersc_reports_an_unwritable_path(tmp_path):
    U, info = O.read_nersc(CONF)
    with gl.Gfix(tuple(info.dims)) as g:
        g.set_links(U)
        with pytest.raises(gl.GfixError, match="cannot open"):
            g.write_nersc(str(tmp_path / "missing_dir" / "x.nersc"))
