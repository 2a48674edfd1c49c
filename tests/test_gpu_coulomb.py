"""GPU parity of the batched Coulomb path against the CPU oracle (which fixes
the slices one after another, as GLU does).

Tolerances: per-slice theta < accuracy; per-slice iteration counts equal
(or within 2 where rounding changes the last step); spatial link trace
1e-10 relative; plaquette preserved to 1e-13.  The Coulomb condition fixes
each slice's transform only up to a constant SU(3) rotation per slice, and
which rotation a solve ends on depends on rounding along its trajectory:
temporal link traces therefore agree only to 1e-8 relative between two
correct solves, and spatial links are compared by their per-link traces
(invariant under that rotation) to 1e-10."""
import os

import numpy as np
import pytest

from oracle import oracle as O

from conftest import GOLDEN
from gpu_common import check_trajectory, heatbath_config, rel

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")


def links_sp_t(U):
    tr = U[..., 0] + U[..., 8] + U[..., 16]
    return tr[:, :3].mean() / 3, tr[:, 3].mean() / 3


def run_both(U, dims, method, gf, acc, max_iters):
    L = O.Lattice(dims, nth=4)
    W = U.copy()
    r = L.coulomb(W, gf, acc, max_iters, facg=(method == gl.CG))
    with gl.Gfix(dims, gl.COULOMB) as g:
        g.set_links(U)
        it, th = g.run(method, gf, acc, max_iters)
        its, ths = g.slice_info()
        g.reunit()  # Gfix/Coulomb.c:97
        m = g.measure()
        Wg = g.get_links()
    return L, W, r, (it, th, its, ths, m, Wg)


def test_conf_fixture_matches_oracle():
    U, info = O.read_nersc(os.path.join(GOLDEN, "conf_lat400.nersc"))
    dims = tuple(info.dims)
    L, W, r, (it, th, its, ths, m, Wg) = run_both(U, dims, gl.CG, 0.09, 1e-14, 650)
    assert max(ths) < 1e-14
    d = np.abs(np.array(its) - np.array(r["slice_iters"]))
    assert d.max() <= 2, (its, r["slice_iters"])
    sp_o, t_o = links_sp_t(W)
    sp_g, t_g = links_sp_t(Wg)
    assert rel(sp_g, sp_o) < 1e-10 and rel(t_g, t_o) < 1e-8
    assert abs(m["plaquette"] - L.plaquette(U)) < 1e-13
    # per-link traces: a solve ending at theta < 1e-14 fixes them to about
    # 1e-10, the oracle's own spread between two summation orders
    W1 = U.copy()
    O.Lattice(dims, nth=1).coulomb(W1, 0.09, 1e-14, 650, facg=True)
    tr = lambda X: X[:, :3, 0] + X[:, :3, 8] + X[:, :3, 16]
    spread = np.abs(tr(W1) - tr(W)).max()
    assert np.abs(tr(Wg) - tr(W)).max() < max(1e-10, 20 * spread), spread


def test_heatbath_batched_slices():
    dims = (8, 8, 8, 12)
    U, plaq = heatbath_config(dims, sweeps=30)
    L, W, r, (it, th, its, ths, m, Wg) = run_both(U, dims, gl.CG, 0.09, 1e-14, 650)
    assert max(ths) < 1e-14
    assert it == sum(its)
    assert np.abs(np.array(its) - np.array(r["slice_iters"])).max() <= 2
    assert abs(m["plaquette"] - plaq) < 1e-13
    # Coulomb condition on every slice: spatial divergence vanishes
    with gl.Gfix(dims) as g:
        g.set_links(Wg)
    assert L.theta_lin(Wg, max_dir=3) < 1e-13
    sp_o, _ = links_sp_t(W)
    sp_g, _ = links_sp_t(Wg)
    assert rel(sp_g, sp_o) < 1e-10


def test_slice_trajectories_match_the_oracle():
    """Every slice's first 21 rows (theta_k on the accelerated field, step,
    beta_k) against the oracle, which fixes the slices one after another:
    1e-12 relative, or 20x the oracle's own spread over four other thread
    counts (summation orders).  On 8^3 slices the parabola fit of the line
    search amplifies rounding by about 10^3 over 20 iterations, so the
    later rows of some slices carry 1e-10 differences between any two
    summation orders, the oracle's included."""
    dims = (8, 8, 8, 12)
    U, _ = heatbath_config(dims, sweeps=30)
    tr = {}
    for nth in (4, 1, 2, 3, 8):
        W = U.copy()
        tr[nth] = O.Lattice(dims, nth=nth).coulomb(W, 0.09, 1e-14, 650, facg=True, trace_slice="all",
                                                   trace=21)["traces"]
    with gl.Gfix(dims, gl.COULOMB) as g:
        g.set_links(U)
        g.run(gl.CG, 0.09, 1e-14, 650)
        rows = [g.slice_trace(t) for t in range(dims[3])]
    for t in range(dims[3]):
        n = min(21, len(tr[4][t]["trace_theta"]))
        check_trajectory(f"coulomb slice {t}", rows[t], tr[4][t], [tr[k][t] for k in (1, 2, 3, 8)], n=n)


def test_odd_spatial_extent_and_fasd():
    dims = (6, 6, 6, 4)
    U, plaq = heatbath_config(dims, sweeps=20)
    L, W, r, (it, th, its, ths, m, Wg) = run_both(U, dims, gl.SD, 0.09, 1e-12, 4000)
    assert max(ths) < 1e-12
    assert np.abs(np.array(its) - np.array(r["slice_iters"])).max() <= 2
    assert abs(m["plaquette"] - plaq) < 1e-13
    dims = (5, 6, 4, 4)  # odd spatial extent: GLU's odd-dimension search
    U, plaq = heatbath_config(dims, sweeps=10)
    L, W, r, (it, th, its, ths, m, Wg) = run_both(U, dims, gl.CG, 0.09, 1e-14, 650)
    assert max(ths) < 1e-14
    assert abs(m["plaquette"] - plaq) < 1e-13


def test_pure_gauge_spatial_links_become_unit():
    dims = (4, 4, 4, 6)
    V = int(np.prod(dims))
    U = np.zeros((V, 4, 18))
    U[..., [0, 8, 16]] = 1.0
    O.Lattice(dims, nth=2).random_gtrans(U, seed=8)
    with gl.Gfix(dims, gl.COULOMB) as g:
        g.set_links(U)
        g.run(gl.CG, 0.09, 1e-14, 650)
        Wg = g.get_links()
    assert np.abs(Wg[:, :3, [0, 8, 16]] - 1).max() < 1e-12
