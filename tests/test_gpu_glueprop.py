"""Gluon propagators of the device links against the oracle's restatement of
GLU's GLUON_PROPS cuts (Cuts/cuts.c:114-184, Cuts/glueprop.c:95-151).

The momentum list is integer work and must be identical.  The propagators
agree to 1e-12 relative to the largest value: the device transforms 16
packed fields and unpacks them at p and -p, the oracle (like GLU)
transforms all 36 entries, so the two differ by rounding only.  After
Landau gauge fixing the momentum-space Landau condition holds: the
longitudinal propagator is rounding noise (compared on the scale of the
transverse one) and GLU's p-space error figure falls by orders of
magnitude."""
import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import heatbath_config

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")


def haar_links(dims, seed):
    """independent Haar-random SU(3) links in GLU's layout"""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims)) * 4
    z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    q = q / (np.linalg.det(q) ** (1 / 3))[:, None, None]
    out = np.empty((n, 18))
    out[:, 0::2], out[:, 1::2] = q.reshape(n, 9).real, q.reshape(n, 9).imag
    return out.reshape(-1, 4, 18)


@pytest.fixture(scope="module")
def fixed():
    dims = (8, 8, 8, 16)
    U, _ = heatbath_config(dims, sweeps=20)
    with gl.Gfix(dims) as g:
        g.set_links(U)
        _, _, _, err0 = g.gluon_props("psq", 9)
        it, th = g.run(gl.CG, 0.08, 1e-14, 2000)
        W = g.get_links()
    assert th < 1e-14
    return dims, W, err0


@pytest.mark.parametrize("cut,max_mom", [("psq", 9), ("hypercubic", 4), ("cylinder", 12), ("conical", 12)])
def test_gluon_props_match_the_oracle(fixed, cut, max_mom):
    dims, W, err0 = fixed
    print(f"gluon_props {dims} cut={cut} max_mom={max_mom}", flush=True)
    L = O.Lattice(dims, nth=4)
    mo, Go, Fo = O.gluon_prop(L, W, cut, max_mom, 1.0, 20.0)
    with gl.Gfix(dims) as g:
        g.set_links(W)
        mg, Gg, Fg, err = g.gluon_props(cut, max_mom, 1.0, 20.0)
    assert np.array_equal(mg, mo)
    assert np.array_equal(mg[::-1], -mg)
    assert np.abs(Gg - Go).max() < 1e-12 * np.abs(Go).max()
    # in Landau gauge p_mu A_mu(p) = 0, so F is rounding noise on the scale
    # of G: both sides are compared on that scale
    assert np.abs(Fg - Fo).max() < 1e-12 * np.abs(Go).max()
    assert np.abs(Fg).max() < 1e-12 * np.abs(Gg).max()
    # GLU's momentum-space Landau-condition figure; the unfixed field is far off
    assert err < 1e-5 * err0, (err, err0)


@pytest.mark.parametrize("dims,precision", [((6, 6, 8, 8), gl.FP64), ((6, 6, 8, 8), gl.MIXED),
                                            ((5, 6, 4, 8), gl.FP64)])
def test_gluon_props_of_an_unfixed_field(dims, precision):
    """Off Landau gauge the longitudinal part is of the size of the
    transverse one and is compared on its own scale; the links are FP64 in
    every context, and an odd extent goes through the radix-5 transform."""
    print(f"gluon_props unfixed {dims} precision={precision}", flush=True)
    U = haar_links(dims, seed=5)
    mo, Go, Fo = O.gluon_prop(O.Lattice(dims, nth=2), U, "psq", 3)
    with gl.Gfix(dims, gl.LANDAU, precision) as g:
        g.set_links(U)
        mg, Gg, Fg, _ = g.gluon_props("psq", 3)
    assert np.array_equal(mg, mo)
    assert np.abs(Fo).max() > 0.1 * np.abs(Go).max()
    assert np.abs(Gg - Go).max() < 1e-12 * np.abs(Go).max()
    assert np.abs(Fg - Fo).max() < 1e-12 * np.abs(Fo).max()


def test_gluon_props_file_matches_the_oracle(fixed, tmp_path):
    dims, W, _ = fixed
    p, q = str(tmp_path / "gpu.bin"), str(tmp_path / "oracle.bin")
    with gl.Gfix(dims) as g:
        g.set_links(W)
        mg, Gg, Fg, _ = g.gluon_props("psq", 6, path=p)
    mo, Go, Fo = O.gluon_prop(O.Lattice(dims, nth=2), W, "psq", 6)
    O.write_glueprop(q, mo, Go, Fo)
    a, b = open(p, "rb").read(), open(q, "rb").read()
    n = len(mo)
    head = 4 + 20 * n
    assert len(a) == len(b) == head + 2 * (4 + 8 * n)
    assert a[:head] == b[:head]  # count and momenta, big-endian ints
    ga = np.frombuffer(a[head + 4:head + 4 + 8 * n], ">f8")
    assert np.array_equal(ga, Gg)
    assert np.abs(ga - np.frombuffer(b[head + 4:head + 4 + 8 * n], ">f8")).max() < 1e-12 * np.abs(Go).max()


def test_gluon_props_refusals():
    with gl.Gfix((4, 4, 4, 8)) as g:
        g.heatbath(6.0, 1, 0, seed=2)
        with pytest.raises(gl.GfixError, match="unknown momentum cut"):
            g.gluon_props(7, 3)
    with gl.Gfix((4, 4, 4, 8), gl.COULOMB) as g:
        with pytest.raises(gl.GfixError, match="Landau context"):
            g.gluon_props("psq", 3)


@pytest.mark.parametrize("cut,max_mom", [("psq", 9), ("cylinder", 12)])
def test_gluon_exceptional_matches_the_oracle(fixed, cut, max_mom):
    """GLU's default cuts output (MOMgg exceptional g2, g3) on the Landau-
    fixed field: g3 to 1e-12 of its largest value, g2 likewise and equal to
    the propagators' G + F / 3 (their common trace)."""
    dims, W, _ = fixed
    print(f"gluon_exceptional {dims} cut={cut} max_mom={max_mom}", flush=True)
    mo, g2o, g3o = O.gluon_exceptional(O.Lattice(dims, nth=4), W, cut, max_mom, 1.0, 20.0)
    with gl.Gfix(dims) as g:
        g.set_links(W)
        mg, g2, g3, _ = g.gluon_exceptional(cut, max_mom, 1.0, 20.0)
        _, G, F, _ = g.gluon_props(cut, max_mom, 1.0, 20.0)
    assert np.array_equal(mg, mo)
    assert np.abs(g2 - g2o).max() < 1e-12 * np.abs(g2o).max()
    assert np.abs(g3 - g3o).max() < 1e-12 * np.abs(g3o).max()
    assert np.abs(g3o).max() > 0
    z = len(g2) // 2
    want = G + F / 3.0
    want[z] = G[z]
    assert np.abs(g2 - want).max() < 1e-12 * np.abs(want).max()


def test_gluon_exceptional_of_an_unfixed_field():
    dims = (6, 6, 8, 8)
    U = haar_links(dims, seed=6)
    mo, g2o, g3o = O.gluon_exceptional(O.Lattice(dims, nth=2), U, "psq", 3)
    with gl.Gfix(dims) as g:
        g.set_links(U)
        mg, g2, g3, _ = g.gluon_exceptional("psq", 3)
    assert np.array_equal(mg, mo)
    assert np.abs(g2 - g2o).max() < 1e-12 * np.abs(g2o).max()
    assert np.abs(g3 - g3o).max() < 1e-12 * np.abs(g3o).max()
