"""The 24^3x48 FP64 Landau FACG run on the GPU checked against the CPU oracle
on the same input: the first 20 iterations of the trajectory, then the final
theta, functional and plaquette of the complete solve.

The input is a device heatbath config (beta 6.0, 200 sweeps from cold) with
a seeded random gauge transform, downloaded so that the oracle starts from
the identical links.  Tolerances as in test_gpu_landau.py.

The late trajectory is chaotic: once theta is small the line search's
parabola fit is decided by rounding, and the PR+ direction can turn uphill
(the vertex falls below zero, the step is clamped to 0.08 and theta grows
for a while).  The oracle itself then depends on its summation order: run
with 8 and with 16 threads on the same input it takes different numbers of
iterations and can end on different Gribov copies (functionals differing by
1e-6..1e-4 relative), or fail where the other converges.  The final state is
therefore checked against the oracle's own spread: the GPU must converge,
preserve the plaquette, and end within 1e-10 of the oracle's functional or
no farther from it than the two oracle runs are from each other.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import rel

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

DIMS = (24, 24, 24, 48)
NTH = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 4))


@pytest.fixture(scope="module")
def run():
    with gl.Gfix(DIMS, gl.LANDAU, gl.FP64) as g:
        g.heatbath(6.0, 200, 4, seed=24)
        plaq = g.measure()["plaquette"]
        g.random_transform(49)
        U = g.get_links()
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        tr = g.trace()
        raw = g.measure()
        g.reunit()
        m = g.measure()
    return U, plaq, it, th, tr, raw, m


def test_trajectory_first_20(run):
    U, plaq, it, th, tr, raw, m = run
    r = {}
    for nth in (1, NTH):
        L = O.Lattice(DIMS, nth=nth)
        W = U.copy()
        r[nth] = L.facg(W, 0.08, 1e-14, 20, trace=64)
    n = 21
    for col, key in enumerate(("trace_theta", "trace_alpha", "trace_beta")):
        o1, o2, gg = r[1][key][:n], r[NTH][key][:n], tr[:n, col]
        spread = np.maximum.accumulate(rel(o2, o1)) * np.abs(o1)
        tol = np.maximum(1e-12 * np.abs(o1), 20 * spread)
        bad = np.abs(gg - o1) > tol
        assert not bad.any(), (key, np.nonzero(bad)[0], gg[bad], o1[bad])


def test_final_state_against_oracle(run):
    U, plaq, it, th, tr, raw, m = run
    F = {}
    iters = {}
    for nth in sorted({NTH, max(1, NTH // 2)}):
        L = O.Lattice(DIMS, nth=nth)
        W = U.copy()
        r = L.facg(W, 0.08, 1e-14, 1000)
        assert r["theta"] < 1e-14
        L.reunit(W)
        F[nth], iters[nth] = L.functional(W), r["iters"]
    ref = F[NTH]
    spread = max(abs(a - ref) for a in F.values())
    print(f"[config2] gpu iters {it} F {m['functional']:.15f}; oracle iters {iters} F {F}")
    assert it < 123456789 and th < 1e-14
    assert abs(m["plaquette"] - plaq) < 1e-14
    assert abs(m["functional"] - ref) <= max(1e-10 * abs(ref), spread)
