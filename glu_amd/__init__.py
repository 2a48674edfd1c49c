"""glu_amd -- MI355X (gfx950) gauge fixing behind GLU's C entry points.

The product is the C-ABI shared library ``libglu_gfix.so`` (declared in
include/glu_gfix.h).  This module is a thin ctypes binding used by tests and
benchmarks; it loads the in-tree library and fails loudly when it is
missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libglu_gfix.so")

LANDAU, COULOMB = 0, 1
FP64, MIXED = 0, 1
CG, SD = 0, 1

SZ = C.c_size_t
DP = C.POINTER(C.c_double)
_lib = None

EXPORTS = [
    "Landau", "Coulomb", "FACG", "FASD", "Coulomb_FA", "cuts_struct",
    "glu_gfix_create", "glu_gfix_destroy", "glu_gfix_set_links", "glu_gfix_get_links",
    "glu_gfix_set_links_device", "glu_gfix_snapshot", "glu_gfix_restore", "glu_gfix_run",
    "glu_gfix_trace", "glu_gfix_measure", "glu_gfix_theta", "glu_gfix_reunit",
    "glu_gfix_mixed_split", "glu_gfix_slice_info", "glu_gfix_stage_deriv", "glu_gfix_stage_fa",
    "glu_gfix_stage_line_search", "glu_gfix_stage_egauge", "glu_gfix_footprint",
    "glu_gfix_last_error", "glu_gfix_heatbath", "glu_gfix_random_transform",
    "glu_gfix_profile", "glu_gfix_profile_read", "glu_gfix_set_switch_theta",
    "glu_gfix_get_links_device", "glu_gfix_set_graph", "glu_gfix_landau", "glu_gfix_autotune", "glu_gfix_luxury",
    "glu_gfix_server_create", "glu_gfix_server_configure", "glu_gfix_server_set_batching", "glu_gfix_server_submit",
    "glu_gfix_server_wait", "glu_gfix_server_stats", "glu_gfix_server_destroy", "glu_gfix_write_nersc",
    "glu_gfix_gluon_props", "glu_gfix_gluon_exceptional", "glu_gfix_rng_seed", "glu_gfix_glu_random_transform",
    "glu_gfix_rng_draws", "glu_gfix_phase_ms", "glu_gfix_server_submit_glu", "glu_gfix_measure_noalloc",
    "glu_gfix_slice_trace", "glu_gfix_glu_heatbath", "glu_gfix_facg_batch",
]

KERNELS = ["deriv", "fft_axis", "pr_sums", "cg_update", "line_search", "egauge", "egauge_deriv", "lie_pack",
           "gluon_prop", "c_deriv", "c_pr_sums", "c_cg_update", "c_line_search", "c_slice_update"]
CUTS = {"hypercubic": 0, "psq": 1, "cylinder": 2, "conical": 3}


class GfixError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GfixError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                            "or python -m glu_amd.build")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "glu_gfix_create": (C.c_int, [C.POINTER(P), C.POINTER(SZ), C.c_int, C.c_int]),
            "glu_gfix_destroy": (C.c_int, [P]),
            "glu_gfix_set_links": (C.c_int, [P, DP]),
            "glu_gfix_get_links": (C.c_int, [P, DP]),
            "glu_gfix_set_links_device": (C.c_int, [P, C.c_void_p]),
            "glu_gfix_snapshot": (C.c_int, [P]),
            "glu_gfix_restore": (C.c_int, [P]),
            "glu_gfix_run": (C.c_int, [P, C.c_int, C.c_double, C.c_double, SZ, DP, C.POINTER(SZ)]),
            "glu_gfix_trace": (C.c_int, [P, DP, SZ, C.POINTER(SZ)]),
            "glu_gfix_measure": (C.c_int, [P, DP, DP, DP, DP, DP]),
            "glu_gfix_theta": (C.c_int, [P, DP]),
            "glu_gfix_reunit": (C.c_int, [P]),
            "glu_gfix_write_nersc": (C.c_int, [P, C.c_char_p, C.c_int, C.c_int, C.c_char_p, SZ]),
            "glu_gfix_gluon_props": (C.c_int, [P, C.c_int, SZ, C.c_double, C.c_double, C.c_char_p,
                                               C.POINTER(SZ), C.POINTER(C.c_int), DP, DP, SZ, DP]),
            "glu_gfix_gluon_exceptional": (C.c_int, [P, C.c_int, SZ, C.c_double, C.c_double, C.c_char_p,
                                                     C.POINTER(SZ), C.POINTER(C.c_int), DP, DP, SZ, DP]),
            "glu_gfix_mixed_split": (C.c_int, [P, C.POINTER(SZ), C.POINTER(SZ)]),
            "glu_gfix_slice_info": (C.c_int, [P, C.POINTER(SZ), DP]),
            "glu_gfix_slice_trace": (C.c_int, [P, SZ, DP, SZ, C.POINTER(SZ)]),
            "glu_gfix_stage_deriv": (C.c_int, [P, DP, DP]),
            "glu_gfix_stage_fa": (C.c_int, [P, DP, DP]),
            "glu_gfix_stage_line_search": (C.c_int, [P, DP, DP, DP]),
            "glu_gfix_stage_egauge": (C.c_int, [P, DP, C.c_double]),
            "glu_gfix_footprint": (SZ, [C.POINTER(SZ), C.c_int, C.c_int]),
            "glu_gfix_last_error": (C.c_char_p, []),
            "glu_gfix_heatbath": (C.c_int, [P, C.c_double, C.c_int, C.c_int, C.c_uint64, C.c_int]),
            "glu_gfix_random_transform": (C.c_int, [P, C.c_uint64]),
            "glu_gfix_profile": (C.c_int, [P, C.c_int]),
            "glu_gfix_set_switch_theta": (C.c_int, [P, C.c_double]),
            "glu_gfix_profile_read": (C.c_int, [P, C.c_int, C.POINTER(SZ), DP, DP]),
            "glu_gfix_get_links_device": (C.c_int, [P, C.c_void_p]),
            "glu_gfix_set_graph": (C.c_int, [P, C.c_int]),
            "glu_gfix_autotune": (C.c_int, [P, C.c_char_p, SZ]),
            "glu_gfix_luxury": (C.c_int, [P, C.c_int, C.c_double, C.c_double, SZ, C.c_int, C.c_uint64, DP,
                                          C.POINTER(SZ), C.POINTER(C.c_int)]),
            "glu_gfix_server_create": (C.c_int, [C.POINTER(P), SZ, C.c_int]),
            "glu_gfix_server_configure": (C.c_int, [P, C.c_double, C.c_int, C.c_uint64]),
            "glu_gfix_server_set_batching": (C.c_int, [P, C.c_int]),
            "glu_gfix_landau": (C.c_int, [P, C.c_int, C.c_double, C.c_double, SZ, C.c_uint64, DP,
                                          C.POINTER(SZ), C.POINTER(C.c_int)]),
            "glu_gfix_server_submit": (C.c_int, [P, C.POINTER(SZ), C.c_int, C.c_int, C.c_int, C.c_double,
                                                 C.c_double, SZ, C.c_void_p, C.POINTER(C.c_int)]),
            "glu_gfix_server_wait": (C.c_int, [P, C.c_int, DP, C.POINTER(SZ), DP]),
            "glu_gfix_server_stats": (C.c_int, [P, C.POINTER(SZ), DP, DP, C.POINTER(SZ)]),
            "glu_gfix_server_destroy": (C.c_int, [P]),
            "glu_gfix_rng_seed": (C.c_int, [P, C.c_uint32, C.c_int]),
            "glu_gfix_glu_random_transform": (C.c_int, [P]),
            "glu_gfix_glu_heatbath": (C.c_int, [P, C.c_double, C.c_int, C.c_int, C.c_int]),
            "glu_gfix_rng_draws": (C.c_int, [C.c_uint32, C.c_int, C.c_int, SZ, DP]),
            "glu_gfix_phase_ms": (C.c_int, [P, DP]),
            "glu_gfix_server_submit_glu": (C.c_int, [P, C.POINTER(SZ), C.c_int, C.c_char_p, C.c_void_p,
                                                     C.POINTER(C.c_int)]),
            "glu_gfix_measure_noalloc": (C.c_int, [C.c_void_p, C.POINTER(SZ), C.c_void_p, C.c_void_p]),
            "glu_gfix_facg_batch": (C.c_int, [C.POINTER(SZ), SZ, C.c_void_p, C.c_double, C.c_double, SZ,
                                              C.POINTER(SZ), DP, DP, SZ]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        raise GfixError(f"{what} failed ({rc}): {lib().glu_gfix_last_error().decode()}")


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(DP)


class Gfix:
    """One lattice's device state (wraps glu_gfix_ctx)."""

    def __init__(self, dims, gftype: int = LANDAU, precision: int = FP64):
        self.dims = tuple(int(d) for d in dims)
        self.V = int(np.prod(self.dims))
        self.gftype = gftype
        self._h = C.c_void_p()
        d = (SZ * 4)(*self.dims)
        _check(lib().glu_gfix_create(C.byref(self._h), d, gftype, precision), "glu_gfix_create")

    def close(self):
        if self._h:
            lib().glu_gfix_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_links(self, U: np.ndarray):
        U = np.ascontiguousarray(U, np.float64)
        assert U.size == self.V * 72
        _check(lib().glu_gfix_set_links(self._h, _dp(U)), "set_links")

    def set_links_device(self, ptr: int):
        _check(lib().glu_gfix_set_links_device(self._h, C.c_void_p(ptr)), "set_links_device")

    def get_links_device(self, ptr: int):
        _check(lib().glu_gfix_get_links_device(self._h, C.c_void_p(ptr)), "get_links_device")

    def autotune(self) -> str:
        """Pick the fastest launch shapes for this context; returns the timings."""
        buf = C.create_string_buffer(4096)
        _check(lib().glu_gfix_autotune(self._h, buf, len(buf)), "autotune")
        return buf.value.decode()

    def set_graph(self, on: bool):
        _check(lib().glu_gfix_set_graph(self._h, int(bool(on))), "set_graph")

    def get_links(self) -> np.ndarray:
        U = np.empty((self.V, 4, 18))
        _check(lib().glu_gfix_get_links(self._h, _dp(U)), "get_links")
        return U

    def snapshot(self):
        _check(lib().glu_gfix_snapshot(self._h), "snapshot")

    def restore(self):
        _check(lib().glu_gfix_restore(self._h), "restore")

    def run(self, method=CG, gf_tune=0.08, accuracy=1e-14, max_iters=1000):
        th = C.c_double()
        it = SZ()
        _check(lib().glu_gfix_run(self._h, method, gf_tune, accuracy, max_iters,
                                  C.byref(th), C.byref(it)), "run")
        return int(it.value), th.value

    def landau(self, method=CG, gf_tune=0.08, accuracy=1e-14, max_iters=1000, restart_seed=1):
        """Landau() semantics: run, with GLU's random restarts on failure.
        Returns (iters, theta, restarts)."""
        th, it, nf = C.c_double(), SZ(), C.c_int()
        _check(lib().glu_gfix_landau(self._h, method, gf_tune, accuracy, max_iters, restart_seed,
                                     C.byref(th), C.byref(it), C.byref(nf)), "landau")
        return it.value, th.value, nf.value

    def luxury(self, method=CG, gf_tune=0.08, accuracy=1e-14, max_iters=1000, copies=20, seed=1):
        """LUXURY_GAUGE: best of `copies` Gribov copies.  Returns (iters, theta, chosen copy)."""
        th, it, best = C.c_double(), SZ(), C.c_int()
        _check(lib().glu_gfix_luxury(self._h, method, gf_tune, accuracy, max_iters, copies, seed,
                                     C.byref(th), C.byref(it), C.byref(best)), "luxury")
        return it.value, th.value, best.value

    def trace(self, cap=16384) -> np.ndarray:
        rows = np.zeros((cap, 3))
        n = SZ()
        _check(lib().glu_gfix_trace(self._h, _dp(rows), cap, C.byref(n)), "trace")
        return rows[: n.value].copy()

    def measure(self):
        v = [C.c_double() for _ in range(5)]
        _check(lib().glu_gfix_measure(self._h, *[C.byref(x) for x in v]), "measure")
        return dict(zip(("plaquette", "link_trace", "link_max", "functional", "tlink"),
                        (x.value for x in v)))

    def theta(self) -> float:
        t = C.c_double()
        _check(lib().glu_gfix_theta(self._h, C.byref(t)), "theta")
        return t.value

    def reunit(self):
        _check(lib().glu_gfix_reunit(self._h), "reunit")

    def _gluon_cuts(self, fn, cut, max_mom, cyl_width, angle, path):
        t = CUTS[cut] if isinstance(cut, str) else int(cut)
        n = SZ()
        _check(fn(self._h, t, max_mom, cyl_width, angle, None, C.byref(n), None, None, None, 0, None),
               fn.__name__)
        k = n.value
        mom = np.zeros((k, 4), np.int32)
        a, b, err = np.zeros(k), np.zeros(k), C.c_double()
        _check(fn(self._h, t, max_mom, cyl_width, angle, None if path is None else os.fspath(path).encode(),
                  C.byref(n), mom.ctypes.data_as(C.POINTER(C.c_int)), _dp(a), _dp(b), k, C.byref(err)),
               fn.__name__)
        return mom, a, b, err.value

    def gluon_props(self, cut="psq", max_mom=4, cyl_width=1.0, angle=20.0, path=None):
        """GLU's GLUON_PROPS cuts on the device links: (mom[n, 4], G[n], F[n],
        momentum-space Landau-condition figure)."""
        return self._gluon_cuts(lib().glu_gfix_gluon_props, cut, max_mom, cyl_width, angle, path)

    def gluon_exceptional(self, cut="psq", max_mom=4, cyl_width=1.0, angle=20.0, path=None):
        """GLU's default (EXCEPTIONAL) cuts output on the device links:
        (mom[n, 4], g2[n], g3[n], momentum-space Landau-condition figure)."""
        return self._gluon_cuts(lib().glu_gfix_gluon_exceptional, cut, max_mom, cyl_width, angle, path)

    def write_nersc(self, path: str, two_rows: bool = False, big_endian: bool = True, label: str = "",
                    sequence: int = 1):
        """GLU's write_lat, NERSC branch: reunitarises the links in place,
        then writes them with the header's checksum, plaquette and trace."""
        _check(lib().glu_gfix_write_nersc(self._h, os.fspath(path).encode(), int(two_rows), int(big_endian),
                                          label.encode(), sequence), "write_nersc")

    def set_switch_theta(self, theta: float):
        _check(lib().glu_gfix_set_switch_theta(self._h, theta), "set_switch_theta")

    def phase_ms(self):
        """(FP32 phase, switch, FP64 phase) host milliseconds of the last run."""
        ms = np.zeros(3)
        _check(lib().glu_gfix_phase_ms(self._h, _dp(ms)), "phase_ms")
        return tuple(float(x) for x in ms)

    def mixed_split(self):
        a, b = SZ(), SZ()
        _check(lib().glu_gfix_mixed_split(self._h, C.byref(a), C.byref(b)), "mixed_split")
        return int(a.value), int(b.value)

    def slice_trace(self, slice_: int, cap: int = 64) -> np.ndarray:
        rows = np.zeros((cap, 3))
        n = SZ()
        _check(lib().glu_gfix_slice_trace(self._h, slice_, _dp(rows), cap, C.byref(n)), "slice_trace")
        return rows[: n.value].copy()

    def slice_info(self):
        T = self.dims[3]
        its = (SZ * T)()
        th = np.zeros(T)
        _check(lib().glu_gfix_slice_info(self._h, its, _dp(th)), "slice_info")
        return list(its), th

    def heatbath(self, beta=6.0, sweeps=200, nor=4, seed=1, cold_start=True):
        _check(lib().glu_gfix_heatbath(self._h, beta, sweeps, nor, seed, int(cold_start)),
               "heatbath")

    def random_transform(self, seed=1):
        _check(lib().glu_gfix_random_transform(self._h, seed), "random_transform")

    def rng_seed(self, seed: int, nthreads: int):
        """GLU's initialise_par_rng on the context's device streams."""
        _check(lib().glu_gfix_rng_seed(self._h, seed, nthreads), "rng_seed")

    def glu_random_transform(self):
        """GLU's random_transform drawn from the context's streams."""
        _check(lib().glu_gfix_glu_random_transform(self._h), "glu_random_transform")

    def glu_heatbath(self, beta=6.0, sweeps=200, nor=4, cold_start=True):
        """GLU's own HB + OR update (Update/KPHB.c) from the context's
        streams (seed them with rng_seed first: SEED, Nthreads)."""
        _check(lib().glu_gfix_glu_heatbath(self._h, beta, sweeps, nor, int(cold_start)), "glu_heatbath")

    def profile(self, mask=-1):
        """Time kernel classes with HIP events (bit k = KERNELS[k], bit
        k + len(KERNELS) its FP32 variant); -1 all, 0 off."""
        _check(lib().glu_gfix_profile(self._h, int(mask)), "profile")

    def profile_read(self):
        """{(kernel, 'fp64'|'fp32'): (launches, mean_ms, total_ms)}"""
        out = {}
        for k in range(2 * len(KERNELS)):
            n, mean, tot = SZ(), C.c_double(), C.c_double()
            _check(lib().glu_gfix_profile_read(self._h, k, C.byref(n), C.byref(mean), C.byref(tot)),
                   "profile_read")
            if n.value:
                out[(KERNELS[k % len(KERNELS)], "fp32" if k >= len(KERNELS) else "fp64")] = (
                    int(n.value), mean.value, tot.value)
        return out

    # stage-level entry points
    def stage_deriv(self):
        F = np.empty((4, self.V), np.complex128)
        t = C.c_double()
        _check(lib().glu_gfix_stage_deriv(self._h, _dp(F), C.byref(t)), "stage_deriv")
        return t.value, F

    def stage_fa(self, F):
        F = np.ascontiguousarray(F, np.complex128)
        out = np.empty_like(F)
        _check(lib().glu_gfix_stage_fa(self._h, _dp(F), _dp(out)), "stage_fa")
        return out

    def stage_line_search(self, F):
        F = np.ascontiguousarray(F, np.complex128)
        pr = np.zeros(3)
        a = C.c_double()
        _check(lib().glu_gfix_stage_line_search(self._h, _dp(F), _dp(pr), C.byref(a)),
               "stage_line_search")
        return pr, a.value

    def stage_egauge(self, F, alpha):
        F = np.ascontiguousarray(F, np.complex128)
        _check(lib().glu_gfix_stage_egauge(self._h, _dp(F), alpha), "stage_egauge")


def rng_draws(seed: int, nthreads: int, thread: int, n: int) -> np.ndarray:
    """par_rng_dbl(thread) of GLU's streams seeded (seed, nthreads)."""
    out = np.zeros(n)
    _check(lib().glu_gfix_rng_draws(seed, nthreads, thread, n, _dp(out)), "rng_draws")
    return out


def facg_batch(dims, ptr: int, nlat: int, gf_tune=0.08, accuracy=1e-14, max_iters=1000, trace_cap=0):
    """FACG() of nlat lattices of one small shape in one launch (device
    buffer `ptr`: nlat x V x 72 doubles in GLU's layout, fixed in place).
    Returns (iters, theta[, trace rows per lattice])."""
    d = (SZ * 4)(*dims)
    it = (SZ * nlat)()
    th = np.zeros(nlat)
    tr = np.zeros((nlat, max(trace_cap, 1), 3))
    _check(lib().glu_gfix_facg_batch(d, nlat, C.c_void_p(ptr), gf_tune, accuracy, max_iters, it, _dp(th),
                                     _dp(tr) if trace_cap else None, trace_cap), "facg_batch")
    its = [int(x) for x in it]
    return (its, th, tr) if trace_cap else (its, th)


def footprint(dims, gftype=LANDAU, precision=FP64) -> int:
    return int(lib().glu_gfix_footprint((SZ * 4)(*dims), gftype, precision))


class Server:
    """Serving loop (wraps glu_gfix_server).  Request buffers are device
    pointers to GLU-layout links, fixed in place."""

    def __init__(self, hbm_budget: int = 0, max_active: int = 8, switch_theta: float = 0.0,
                 use_graph: bool = True, restart_seed: int = 1, batching: bool = False):
        h = C.c_void_p()
        _check(lib().glu_gfix_server_create(C.byref(h), hbm_budget, max_active), "server_create")
        self._h = h
        _check(lib().glu_gfix_server_configure(h, switch_theta, int(bool(use_graph)), restart_seed),
               "server_configure")
        if batching:  # small FP64 Landau lattices in shared launches
            _check(lib().glu_gfix_server_set_batching(h, 1), "server_set_batching")

    def submit(self, dims, ptr: int, gftype=LANDAU, precision=MIXED, method=CG, gf_tune=0.08,
               accuracy=1e-14, max_iters=1000) -> int:
        t = C.c_int()
        _check(lib().glu_gfix_server_submit(self._h, (SZ * 4)(*dims), gftype, precision, method, gf_tune,
                                            accuracy, max_iters, C.c_void_p(ptr), C.byref(t)), "server_submit")
        return t.value

    def submit_glu(self, dims, ptr: int, options: str, precision=MIXED) -> int:
        """Submit with GLU input-file keys (GFTYPE, ACCURACY, MAX_ITERS,
        GF_TUNE, SEED, RANDOM_TRANSFORM) given as KEY = VALUE lines."""
        t = C.c_int()
        _check(lib().glu_gfix_server_submit_glu(self._h, (SZ * 4)(*dims), precision, options.encode(),
                                                C.c_void_p(ptr), C.byref(t)), "server_submit_glu")
        return t.value

    def wait(self, ticket: int):
        th, it, lat = C.c_double(), SZ(), C.c_double()
        _check(lib().glu_gfix_server_wait(self._h, ticket, C.byref(th), C.byref(it), C.byref(lat)),
               "server_wait")
        return it.value, th.value, lat.value

    def stats(self):
        n, p50, p99, peak = SZ(), C.c_double(), C.c_double(), SZ()
        _check(lib().glu_gfix_server_stats(self._h, C.byref(n), C.byref(p50), C.byref(p99), C.byref(peak)),
               "server_stats")
        return {"completed": n.value, "p50_ms": p50.value, "p99_ms": p99.value, "peak_bytes": peak.value}

    def close(self):
        if getattr(self, "_h", None):
            lib().glu_gfix_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
