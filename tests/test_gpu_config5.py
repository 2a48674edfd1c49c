"""The serving stream at its real request sizes: Landau requests of
16^3x32, 24^3x48 and 32^3x64 at beta 5.8, 6.0 and 6.2, mixed precision to
theta < 1e-14, admitted and run concurrently by the serving loop.

Each request is a device heatbath config of its (size, beta) (100 sweeps
from cold, 4 overrelaxation passes each) under a random gauge transform of
its own.  The bar is bitwise: every request served among the others gives
exactly the links, theta and iteration count of the same request fixed
alone on a context of its own with glu_gfix_landau (Landau()'s restart
policy, the server's restart seed and switch point)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")
torch = pytest.importorskip("torch")

REQUESTS = [  # dims, beta, gauge transform seed
    ((16, 16, 16, 32), 5.8, 501),
    ((32, 32, 32, 64), 6.0, 502),
    ((24, 24, 24, 48), 6.2, 503),
    ((16, 16, 16, 32), 6.0, 504),
    ((24, 24, 24, 48), 5.8, 505),
    ((32, 32, 32, 64), 6.2, 506),
]
RESTART_SEED = 1


@pytest.fixture(scope="module")
def inputs():
    out = []
    for dims, beta, seed in REQUESTS:
        print(f"[config5] request input {dims} beta {beta}", flush=True)
        with gl.Gfix(dims, gl.LANDAU, gl.FP64) as g:
            g.heatbath(beta, 100, 4, seed=int(10 * beta), cold_start=True)
            g.random_transform(seed)
            b = torch.empty(math.prod(dims) * 72, dtype=torch.float64, device="cuda")
            g.get_links_device(b.data_ptr())
        out.append(b)
    torch.cuda.synchronize()
    return out


def test_served_mix_matches_requests_alone(inputs):
    alone = []
    for (dims, beta, _), b in zip(REQUESTS, inputs):
        with gl.Gfix(dims, gl.LANDAU, gl.MIXED) as g:
            g.set_links_device(b.data_ptr())
            it, th, nr = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=RESTART_SEED)
            w = torch.empty_like(b)
            g.get_links_device(w.data_ptr())
        torch.cuda.synchronize()
        print(f"[config5] alone {dims} beta {beta}: {it} iterations, {nr} restarts, theta {th:.2e}", flush=True)
        alone.append((it, th, w))
    bufs = [b.clone() for b in inputs]
    torch.cuda.synchronize()
    with gl.Server(max_active=4, restart_seed=RESTART_SEED) as srv:
        tickets = [srv.submit(dims, b.data_ptr(), gl.LANDAU, gl.MIXED) for (dims, _, _), b in zip(REQUESTS, bufs)]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(f"[config5] served: {[(r[0], f'{r[2]:.0f} ms') for r in res]}, p50 {st['p50_ms']:.0f} ms, "
          f"p99 {st['p99_ms']:.0f} ms", flush=True)
    assert st["completed"] == len(REQUESTS)
    for k, ((it, th, w), (it2, th2, _), b) in enumerate(zip(alone, res, bufs)):
        assert th < 1e-14, k
        assert (it2, th2) == (it, th), k
        assert torch.equal(b, w), k
