"""The device input generators behind the large benchmark lattices: the
quenched heatbath and the seeded random gauge transform.

They stand in for GLU's own generators (Update/KPHB.c, Geometry/
random_config.c), which are not restated bit for bit, so the checks are
physical and algebraic: the heatbath reaches the known beta = 6.0
plaquette, keeps the links on SU(3), depends only on its seed; the gauge
transform preserves the plaquette to rounding and moves the links; and the
device heatbath follows the oracle's chain (same counter-based draws, same
arithmetic) until rounding can flip an accept/reject decision.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")


def test_heatbath_plaquette_and_unitarity():
    dims = (8, 8, 8, 8)
    with gl.Gfix(dims) as g:
        g.heatbath(6.0, 100, 4, seed=3)
        m = g.measure()
        U = g.get_links()
    # <P> at beta = 6.0 is 0.5937 in large volumes; 8^4 sits within 0.005
    assert 0.585 < m["plaquette"] < 0.600, m["plaquette"]
    M = U.reshape(-1, 3, 3, 2)
    M = M[..., 0] + 1j * M[..., 1]
    eye = np.einsum("nij,nkj->nik", M, M.conj())
    assert np.abs(eye - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-12


def test_heatbath_depends_only_on_its_seed():
    dims = (8, 8, 8, 8)
    out = []
    for seed in (5, 5, 6):
        with gl.Gfix(dims) as g:
            g.heatbath(6.0, 5, 2, seed=seed)
            out.append(g.get_links())
    assert np.array_equal(out[0], out[1])
    assert not np.array_equal(out[0], out[2])


def test_heatbath_follows_the_oracle_chain():
    dims = (4, 4, 4, 8)
    L = O.Lattice(dims, nth=4)
    Uo = L.heatbath(6.0, 3, 2, seed=9)
    with gl.Gfix(dims) as g:
        g.heatbath(6.0, 3, 2, seed=9)
        Ug = g.get_links()
        pg = g.measure()["plaquette"]
    assert abs(pg - L.plaquette(Uo)) < 1e-10
    assert np.abs(Ug - Uo).max() < 1e-10


def test_random_transform_preserves_the_plaquette():
    dims = (8, 8, 8, 16)
    with gl.Gfix(dims) as g:
        g.heatbath(6.0, 20, 2, seed=1)
        m0 = g.measure()
        g.random_transform(seed=77)
        m1 = g.measure()
    assert abs(m1["plaquette"] - m0["plaquette"]) < 1e-14
    assert abs(m1["link_trace"] - m0["link_trace"]) > 1e-3
