"""The 24^3x48 FP64 Landau FACG run on the GPU checked against the CPU oracle
on the same input: the first 20 iterations of the trajectory, then the final
theta, functional and plaquette of the complete solve.

The input is a device heatbath config (beta 6.0, 200 sweeps from cold) with
a seeded random gauge transform, downloaded so that the oracle starts from
the identical links.  Tolerances as in test_gpu_landau.py.
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
        g.random_transform(48)
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
        tol = np.maximum(1e-12 * np.abs(o1), 20 * np.abs(o1 - o2))
        bad = np.abs(gg - o1) > tol
        assert not bad.any(), (key, np.nonzero(bad)[0], gg[bad], o1[bad])


def test_final_state_against_oracle(run):
    U, plaq, it, th, tr, raw, m = run
    L = O.Lattice(DIMS, nth=NTH)
    W = U.copy()
    r = L.facg(W, 0.08, 1e-14, 1000)
    print(f"[config2] gpu iters {it}, oracle iters {r['iters']}, functional gpu "
          f"{m['functional']:.15f} oracle {L.functional(W):.15f}")
    assert th < 1e-14 and r["theta"] < 1e-14
    assert abs(m["plaquette"] - plaq) < 1e-14
    # the late trajectory is chaotic (rounding decides the line-search steps
    # once theta is small), so iteration counts differ; both runs must still
    # reach the same gauge-fixed configuration (Gribov copy)
    assert rel(m["functional"], L.functional(W)) < 1e-10
