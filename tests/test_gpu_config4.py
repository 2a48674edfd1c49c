"""Batched Coulomb gauge fixing at the benchmark's size, 48^3x96: all 96
time-slice problems advance in the same launches, with uneven iteration
counts per slice.

The input is the benchmark's: a device heatbath config (beta 6.0, 200
sweeps from cold, 4 overrelaxation passes each, seed 2024) under a random
gauge transform (seed 2025).

Full-size checks: every slice converges (theta < 1e-14), the plaquette
is preserved to 1e-13 after GLU's final reunitarisation (Gfix/Coulomb.c:97),
and the Coulomb condition holds on the whole lattice (theta_test_lin over
the three spatial directions < 1e-13).

Oracle check on a cut: slices are independent problems in their spatial
links (Gfix/CFACG.c:593-614), so the first two time slices, cut out as a
48^3x2 lattice, start the same trajectories under the oracle as inside the
full batched run.  Their first 21 rows (theta_k measured on the accelerated
field, the step, beta_k) are held to 1e-12 relative, or 20x the oracle's own
spread between two thread counts where that is larger.  The complete solves
are not compared link by link: at this size the late trajectory is decided
by rounding (two oracle thread counts took 402 / 438 and 1098 / 666
iterations on these two slices and ended on different Gribov copies), so a
slice's final transform is a property of its summation order, not of the
algorithm.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import check_trajectory

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

DIMS = (48, 48, 48, 96)
NTH = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 4))


@pytest.fixture(scope="module")
def run():
    print(f"[config4] heatbath {DIMS} beta 6.0, 200 sweeps", flush=True)
    with gl.Gfix(DIMS, gl.COULOMB, gl.FP64) as g:
        g.heatbath(6.0, 200, 4, seed=2024, cold_start=True)
        plaq = g.measure()["plaquette"]
        g.random_transform(seed=2025)
        U = g.get_links()
        print("[config4] batched Coulomb FACG, 96 slices", flush=True)
        it, th = g.run(gl.CG, 0.09, 1e-14, 650)
        its, ths = g.slice_info()
        rows = [g.slice_trace(t) for t in (0, 1)]
        g.reunit()
        m = g.measure()
        W = g.get_links()
    print(f"[config4] {it} slice iterations in all (per slice {min(its)}..{max(its)}), largest slice theta "
          f"{max(ths):.3e}, plaquette drift {abs(m['plaquette'] - plaq):.1e}", flush=True)
    return U, plaq, it, th, its, ths, m, W, rows


def test_every_slice_converges_and_the_plaquette_is_kept(run):
    U, plaq, it, th, its, ths, m, W, _ = run
    assert len(its) == DIMS[3] and it == sum(its)
    assert max(ths) < 1e-14 and th == max(ths)
    assert min(its) < max(its)  # uneven slices in one batched run
    assert abs(m["plaquette"] - plaq) < 1e-13


def test_coulomb_condition_on_the_whole_lattice(run):
    W = run[7]
    th3 = O.Lattice(DIMS, nth=NTH).theta_lin(W, max_dir=3)
    print(f"[config4] theta_test_lin over x, y, z: {th3:.3e}", flush=True)
    assert th3 < 1e-13


def test_two_slices_cut_out_match_the_oracle(run):
    U, plaq, it, th, its, ths, m, W, rows = run
    Lcu = DIMS[0] * DIMS[1] * DIMS[2]
    cut = (DIMS[0], DIMS[1], DIMS[2], 2)
    Uc = np.ascontiguousarray(U[: 2 * Lcu])
    traces = {}
    for nth in (NTH, max(1, NTH // 2)):
        print(f"[config4] oracle on the 48^3x2 cut, {nth} threads, 20 iterations per slice", flush=True)
        r = O.Lattice(cut, nth=nth).coulomb(Uc.copy(), 0.09, 1e-14, 20, facg=True, trace_slice="all", trace=21)
        traces[nth] = r["traces"]
    print(f"[config4] gpu full solve: slices 0, 1 took {its[:2]} iterations", flush=True)
    for t in (0, 1):
        assert len(rows[t]) >= 21
        check_trajectory(f"config4 slice {t}", rows[t], traces[NTH][t], traces[max(1, NTH // 2)][t], n=21)
