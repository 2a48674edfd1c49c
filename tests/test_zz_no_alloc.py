"""glu_gfix_measure_noalloc (the allocation-free variant of glu_gfix_measure
and glu_gfix_theta) against those two entry points on the same links.

Three calls in a row on one device share its scratch array: a 24^3x48
lattice (the full 1024-block grid, several sites per thread), then two
smaller ones (fewer blocks: the finishing kernel must read only the rows
the new call wrote).  Links come from the device heatbath under a random
gauge transform, so theta is far from zero.  The two paths sum in different
orders, so they agree to rounding: 1e-13 relative.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")
torch = pytest.importorskip("torch")

KEYS = ("plaquette", "link_trace", "link_max", "functional", "tlink")


def _reference(dims, seed):
    with gl.Gfix(dims, gl.LANDAU, gl.FP64) as g:
        g.heatbath(6.0, 4, 1, seed=seed, cold_start=True)
        g.random_transform(seed=seed + 1)
        U = g.get_links()
        m = g.measure()
        th = g.theta()
    return U, [m[k] for k in KEYS] + [th]


def _noalloc(U, dims):
    links = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    out = torch.full((6,), float("nan"), dtype=torch.float64, device="cuda")
    d = (C.c_size_t * 4)(*dims)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    rc = gl.lib().glu_gfix_measure_noalloc(C.c_void_p(links.data_ptr()), d, C.c_void_p(out.data_ptr()),
                                           C.c_void_p(stream.cuda_stream))
    assert rc == 0, gl.lib().glu_gfix_last_error().decode()
    stream.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("calls", [[(24, 24, 24, 48), (8, 8, 8, 8), (6, 6, 6, 10)]])
def test_noalloc_matches_measure_and_theta(calls):
    for k, dims in enumerate(calls):
        print(f"[noalloc] call {k}: {dims}", flush=True)
        U, ref = _reference(dims, seed=11 + 7 * k)
        got = _noalloc(U, dims)
        for name, a, b in zip(KEYS + ("theta_lin",), got, ref):
            print(f"[noalloc]   {name}: {a:.16e} vs {b:.16e}", flush=True)
            assert abs(a - b) <= 1e-13 * max(abs(b), 1e-3), (dims, name, a, b)


def test_noalloc_refuses_bad_extents():
    d = (C.c_size_t * 4)(8, 8, 1, 8)
    rc = gl.lib().glu_gfix_measure_noalloc(C.c_void_p(1), d, C.c_void_p(1), None)
    assert rc == -1 and b"extent" in gl.lib().glu_gfix_last_error()
    d = (C.c_size_t * 4)(4096, 4096, 4096, 2)
    rc = gl.lib().glu_gfix_measure_noalloc(C.c_void_p(1), d, C.c_void_p(1), None)
    assert rc == -1 and b"2^31" in gl.lib().glu_gfix_last_error()
