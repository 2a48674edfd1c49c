"""The 24^3x48 FP64 Landau FACG run on the GPU checked against the CPU oracle
on the same inputs: four gauge copies of one heatbath config (beta 6.0, 500
sweeps from cold, 4 overrelaxation passes each), each the config under its
own seeded random gauge transform.

Per copy: the first 20 iterations of the trajectory (theta_k, alpha_k,
beta_k) against the oracle with two thread counts (two summation orders),
to 1e-12 relative or 20x the oracle's own spread where that is larger (the
realised errors and the rows that needed the allowance are printed); then
the complete solve through glu_gfix_landau, Landau()'s restart policy
included, which must converge to theta < 1e-14 with the plaquette
preserved to 1e-14.

The late trajectory is chaotic: once theta is small the line search's
parabola fit is decided by rounding, and the PR+ direction can turn uphill
(the vertex falls below zero, the step is clamped to 0.08 and theta grows
for a while).  The oracle itself then depends on its summation order: with
different thread counts it can take different numbers of iterations, end
on different Gribov copies (functionals differing by 1e-6..1e-4 relative),
or fail where the other converges, and GLU resolves a failure by a random
restart.  The final functional is therefore compared with a complete oracle
solve on one copy (seed 49) against the oracle's own spread; the other
copies are compared side by side, one and several oracle threads, by
tests/report_config2.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import check_trajectory

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

DIMS = (24, 24, 24, 48)
SWEEPS = 500
NTH = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 4))
SEEDS = (46, 47, 48, 49)


@pytest.fixture(scope="module")
def ctx():
    print(f"[config2] heatbath {DIMS} beta 6.0, {SWEEPS} sweeps", flush=True)
    g = gl.Gfix(DIMS, gl.LANDAU, gl.FP64)
    g.heatbath(6.0, SWEEPS, 4, seed=24)
    plaq = g.measure()["plaquette"]
    g.snapshot()
    yield g, plaq
    g.close()


def gauge_copy(g, seed):
    g.restore()
    g.random_transform(seed)


@pytest.mark.parametrize("seed", SEEDS)
def test_copy_trajectory_and_solve(ctx, seed):
    g, plaq = ctx
    print(f"[config2] seed {seed}: trajectory", flush=True)
    gauge_copy(g, seed)
    U = g.get_links()
    g.run(gl.CG, 0.08, 1e-14, 20)
    tr = g.trace()
    r = {}
    for nth in (NTH, max(1, NTH // 2)):
        W = U.copy()
        r[nth] = O.Lattice(DIMS, nth=nth).facg(W, 0.08, 1e-14, 20, trace=64)
    check_trajectory(f"config2 seed {seed}", tr, r[NTH], r[max(1, NTH // 2)], n=21)
    gauge_copy(g, seed)
    it, th, restarts = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=seed)
    g.reunit()
    m = g.measure()
    print(f"[config2] seed {seed}: gpu {it} iterations, {restarts} restarts, theta {th:.3e}, "
          f"functional {m['functional']:.15f}", flush=True)
    assert th < 1e-14
    assert abs(m["plaquette"] - plaq) < 1e-14


def test_final_state_against_oracle(ctx):
    g, plaq = ctx
    gauge_copy(g, 49)
    U = g.get_links()
    it, th, restarts = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=49)
    g.reunit()
    m = g.measure()
    F = {}
    for nth in sorted({NTH, max(1, NTH // 2)}):
        print(f"[config2] oracle full solve, {nth} threads", flush=True)
        L = O.Lattice(DIMS, nth=nth)
        W = U.copy()
        r = L.facg(W, 0.08, 1e-14, 1000)
        if r["iters"] == 123456789:
            print(f"[config2] oracle with {nth} threads failed (GLU would restart)", flush=True)
            continue
        L.reunit(W)
        F[nth] = (L.functional(W), r["iters"])
    print(f"[config2] seed 49: gpu {it} iterations ({restarts} restarts) F {m['functional']:.15f}; "
          f"oracle {F}", flush=True)
    assert th < 1e-14 and abs(m["plaquette"] - plaq) < 1e-14
    assert F, "no oracle solve converged"
    if restarts:
        # a restarted GPU solve fixed another random copy of the input: its
        # functional is not comparable with the oracle's first attempt
        return
    ref = F[max(F)][0]
    spread = max(abs(a - ref) for a, _ in F.values())
    assert abs(m["functional"] - ref) <= max(1e-10 * abs(ref), spread)
