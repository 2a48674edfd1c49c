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
48^3x2 lattice, must be fixed the same way by the oracle as inside the
full batched run: per-slice iteration counts equal to within 2 (or the
oracle's own spread between two thread counts), the gauge-invariant
per-link traces of the spatial links to 1e-10 (or 20x that spread).  The
temporal links of slice 0 pair slices 0 and 1 in both lattices; each slice's
transform is fixed only up to a constant rotation, so their mean trace is
compared to 1e-8 relative, as for the smaller lattices.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import rel

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
        g.reunit()
        m = g.measure()
        W = g.get_links()
    print(f"[config4] {it} slice iterations in all (per slice {min(its)}..{max(its)}), largest slice theta "
          f"{max(ths):.3e}, plaquette drift {abs(m['plaquette'] - plaq):.1e}", flush=True)
    return U, plaq, it, th, its, ths, m, W


def test_every_slice_converges_and_the_plaquette_is_kept(run):
    U, plaq, it, th, its, ths, m, W = run
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
    U, plaq, it, th, its, ths, m, W = run
    Lcu = DIMS[0] * DIMS[1] * DIMS[2]
    cut = (DIMS[0], DIMS[1], DIMS[2], 2)
    Uc = np.ascontiguousarray(U[: 2 * Lcu])
    res = {}
    for nth in (NTH, max(1, NTH // 2)):
        print(f"[config4] oracle on the 48^3x2 cut, {nth} threads", flush=True)
        L = O.Lattice(cut, nth=nth)
        Wc = Uc.copy()
        r = L.coulomb(Wc, 0.09, 1e-14, 650, facg=True)
        L.reunit(Wc)
        res[nth] = (r, Wc)
    (r1, W1), (r2, W2) = res[NTH], res[max(1, NTH // 2)]
    print(f"[config4] slices 0, 1: gpu iterations {its[:2]}, oracle {r1['slice_iters']} / {r2['slice_iters']}",
          flush=True)
    assert max(r1["slice_theta"]) < 1e-14
    dit = max(abs(a - b) for a, b in zip(r1["slice_iters"], r2["slice_iters"]))
    assert max(abs(a - b) for a, b in zip(its[:2], r1["slice_iters"])) <= max(2, dit)
    tr = lambda X: X[:, :3, 0] + X[:, :3, 8] + X[:, :3, 16]
    Wg = W[: 2 * Lcu]
    spread = np.abs(tr(W2) - tr(W1)).max()
    err = np.abs(tr(Wg) - tr(W1)).max()
    print(f"[config4] spatial per-link traces: max abs diff {err:.2e} (oracle spread {spread:.2e})", flush=True)
    assert err < max(1e-10, 20 * spread)
    t0 = lambda X: (X[:Lcu, 3, 0] + X[:Lcu, 3, 8] + X[:Lcu, 3, 16]).mean()
    assert rel(t0(Wg), t0(W1)) < 1e-8
