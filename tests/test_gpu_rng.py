"""GLU's random numbers on the device, through the C ABI: the random gauge
transform drawn from a context's copy of GLU's MWC_4096 streams, and every
path that draws from them -- Landau()'s restarts (the device-resident
glu_gfix_landau and the drop-in Landau(), which re-reads its NERSC input
file), LUXURY_GAUGE copies and the Coulomb slice restarts -- against the
oracle replaying GLU's draws with the same SEED and thread count.

Tolerances: the transforms agree to 1e-14 (the streams consume the same
uint32 draws; log() and the Gram-Schmidt may differ in the last bit); a
restarted solve has GLU's iteration count and links to 1e-9 (as for the
continuation runs of test_gpu_landau.py)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle import oracle as O

from conftest import GOLDEN
from gpu_common import heatbath_config

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")

FIXTURE = os.path.join(GOLDEN, "conf_lat400.nersc")


@pytest.mark.parametrize("dims,nth,seed", [((6, 6, 6, 6), 4, 7), ((5, 4, 6, 3), 7, 123), ((8, 8, 8, 8), 256, 1)])
def test_random_transform_matches_glu_stream(dims, nth, seed):
    U, _ = heatbath_config(dims, sweeps=4)
    L = O.Lattice(dims, nth=nth)
    R = O.lib().or_rng_new(seed, nth)
    W = U.copy()
    with gl.Gfix(dims) as g:
        g.set_links(U)
        g.rng_seed(seed, nth)
        for k in range(2):  # the streams continue from one transform to the next
            O.lib().or_random_gtrans(L.ref, R, O._dp(W))
            g.glu_random_transform()
            d = np.abs(g.get_links() - W).max()
            print(f"[rng] {dims} nth {nth} seed {seed} transform {k}: max abs diff {d:.1e}", flush=True)
            assert d < 1e-14
    O.lib().or_rng_free(R)


def oracle_landau(U, dims, nth, seed, acc, max_iters):
    """GLU's Landau() restart loop (Gfix/Landau.c:301-345) on the oracle:
    attempt 0 on the input, attempt k on the input under random_transform
    drawn from the streams seeded (seed, nth)."""
    L = O.Lattice(dims, nth=nth)
    R = O.lib().or_rng_new(seed, nth)
    W = U.copy()
    r = L.facg(W, 0.08, acc, max_iters)
    total, fails = r["iters"], 0
    if total == 123456789:
        for fails in range(1, 7):
            W = U.copy()
            O.lib().or_random_gtrans(L.ref, R, O._dp(W))
            r = L.facg(W, 0.08, acc, max_iters)
            if r["iters"] != 123456789:
                total += r["iters"]
                break
            total += max_iters
    O.lib().or_rng_free(R)
    return total, fails, W


# the reference's CONF/lat..400 (4^3x8): with at most 52 iterations per
# attempt the first attempt fails (theta 3.5e-10 at the cap) and the first
# restart, drawn with SEED 3 and 4 threads, converges through a continuation
# (theta 3.6e-12 < 1e3 acc at the cap)
SEED, NTH, MAXIT = 3, 4, 52


def test_device_landau_restarts_draw_glu_stream():
    U, info = O.read_nersc(FIXTURE)
    dims = tuple(info.dims)
    total, fails, W = oracle_landau(U, dims, NTH, SEED, 1e-14, MAXIT)
    assert fails == 1 and total > 123456789
    with gl.Gfix(dims) as g:
        g.set_links(U)
        g.rng_seed(SEED, NTH)  # the thread count; glu_gfix_landau reseeds with SEED at its first restart
        it, th, nr = g.landau(gl.CG, 0.08, 1e-14, MAXIT, restart_seed=SEED)
        Wg = g.get_links()
    print(f"[rng] glu_gfix_landau: {it - 123456789} + sentinel, {nr} restarts; oracle {total - 123456789}, {fails}",
          flush=True)
    assert nr == fails and th < 1e-14
    assert abs(int(it) - total) <= 2
    assert np.abs(Wg - W).max() < 1e-9


class LattInfo(C.Structure):
    """struct latt_info, Headers/GLU_types.h:229-243"""
    _fields_ = [("dims", C.c_size_t * 4), ("Lsq", C.c_size_t), ("Lcu", C.c_size_t), ("Volume", C.c_size_t),
                ("flow", C.c_size_t), ("gf_alpha", C.c_double), ("sm_alpha", C.c_double * 3), ("head", C.c_int),
                ("Seed", C.c_uint32 * 1), ("twiddles", C.c_double * 4), ("Nthreads", C.c_uint32),
                ("su2_data", C.c_void_p), ("argc", C.c_int)]


def site_array(U):
    """GLU's struct site array over U[V, 4, 18] (in place)"""
    V = U.shape[0]
    ptr = (U.ctypes.data + 144 * np.arange(V * 4, dtype=np.uint64)).reshape(V, 4)
    dt = np.dtype([("O", "u8"), ("neighbor", "u8", (4,)), ("back", "u8", (4,))])
    site = np.zeros(V, dt)
    site["O"] = ptr.ctypes.data + 32 * np.arange(V, dtype=np.uint64)
    return site, ptr


def test_dropin_landau_rereads_the_file_and_draws_glu_stream():
    """Landau() as GLU calls it, with its input file: the restart re-reads
    CONF/lat..400 through the library's NERSC reader and applies the random
    transform of GLU's process stream (Latt.Seed, Latt.Nthreads)."""
    U, info = O.read_nersc(FIXTURE)
    dims = tuple(info.dims)
    total, fails, W = oracle_landau(U, dims, NTH, SEED, 1e-14, MAXIT)
    lib = C.CDLL(gl.LIB_PATH)
    lib.Landau.restype = C.c_size_t
    lib.Landau.argtypes = [C.c_void_p, C.c_double, C.c_size_t, C.c_char_p]
    L = LattInfo.in_dll(lib, "Latt")
    for mu in range(4):
        L.dims[mu] = dims[mu]
    L.Lsq = dims[0] * dims[1]
    L.Lcu = L.Lsq * dims[2]
    L.Volume = L.Lcu * dims[3]
    L.gf_alpha = 0.08
    L.Seed[0] = SEED
    L.Nthreads = NTH
    Ul = np.ascontiguousarray(U.copy())
    site, ptr = site_array(Ul)
    n = lib.Landau(site.ctypes.data, 1e-14, MAXIT, FIXTURE.encode())
    print(f"[rng] Landau(): {n - 123456789} + sentinel; oracle {total - 123456789}", flush=True)
    assert abs(int(n) - total) <= 2
    # Landau() reunitarises before its report (Gfix/Landau.c:48)
    Wr = W.copy()
    O.Lattice(dims, nth=NTH).reunit(Wr)
    assert np.abs(Ul - Wr).max() < 1e-9


def test_luxury_copies_draw_glu_stream():
    """LUXURY_GAUGE (Gfix/Landau.c:88-213): every copy is the input under a
    random transform from the streams; the chosen copy's functional and the
    final links follow the oracle replay."""
    dims, nth, seed, copies = (6, 6, 6, 6), 3, 9, 3
    U, _ = heatbath_config(dims, sweeps=10)
    L = O.Lattice(dims, nth=nth)
    R = O.lib().or_rng_new(seed, nth)
    best, chosen, Wb = 1.0, -1, U.copy()
    for k in range(copies):
        W = U.copy()
        O.lib().or_random_gtrans(L.ref, R, O._dp(W))
        r = L.facg(W, 0.08, 1e-6, 1000)
        f = L.functional(W)
        if f < best and r["iters"] not in (1000, 123456789):
            best, chosen, Wb = f, k, W.copy()
    O.lib().or_rng_free(R)
    r = L.facg(Wb, 0.08, 1e-14, 1000)
    with gl.Gfix(dims) as g:
        g.set_links(U)
        g.rng_seed(seed, nth)
        it, th, pick = g.luxury(gl.CG, 0.08, 1e-14, 1000, copies=copies, seed=seed)
        Wg = g.get_links()
    print(f"[rng] luxury: copy {pick} (oracle {chosen}), functional {L.functional(Wg):.15f} "
          f"(oracle {L.functional(Wb):.15f})", flush=True)
    assert pick == chosen and th < 1e-14
    assert abs(L.functional(Wg) - L.functional(Wb)) < 1e-10


def test_coulomb_restarts_draw_glu_stream_in_slice_order():
    """A cap of 3 iterations makes every slice fail and take its 8 random
    restarts (Gfix/CFACG.c:460-530).  GLU fixes the slices in turn, so the
    streams are consumed slice by slice; the batched device run must draw
    the same matrices for the same slices, and end on the oracle's links."""
    dims, nth, seed = (4, 4, 4, 4), 3, 5
    U, _ = heatbath_config(dims, sweeps=10)
    L = O.Lattice(dims, nth=nth)
    W = U.copy()
    r = L.coulomb(W, 0.09, 1e-14, 3, facg=True, seed=seed)
    with gl.Gfix(dims, gl.COULOMB) as g:
        g.set_links(U)
        g.rng_seed(seed, nth)
        it, th = g.run(gl.CG, 0.09, 1e-14, 3)
        its, ths = g.slice_info()
        g.reunit()
        Wg = g.get_links()
    print(f"[rng] coulomb restarts: slice iterations {its} (oracle {r['slice_iters']}), "
          f"max link diff {np.abs(Wg - W).max():.1e}", flush=True)
    assert list(its) == list(r["slice_iters"]) == [27] * dims[3]
    assert np.abs(Wg - W).max() < 1e-9
