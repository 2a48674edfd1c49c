"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (glu_amd) never
imports it.

Arrays use GLU's layouts: links ``U[V, 4, 18]`` float64 (site-major, x
fastest, row-major 3x3 complex as re/im pairs) and packed fields
``F[4, V]`` complex128.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

SZ = C.c_size_t
DP = C.POINTER(C.c_double)
IP = C.POINTER(C.c_int)


class Lat(C.Structure):
    _fields_ = [("dims", SZ * 4), ("V", SZ), ("Lcu", SZ),
                ("fwd", C.POINTER(SZ)), ("bwd", C.POINTER(SZ)), ("nth", C.c_int)]


class Trace(C.Structure):
    _fields_ = [("cap", SZ), ("n", SZ), ("theta", DP), ("alpha", DP),
                ("beta", DP), ("func", DP)]


class NerscInfo(C.Structure):
    _fields_ = [("dims", SZ * 4), ("plaquette", C.c_double),
                ("link_trace", C.c_double), ("checksum_header", C.c_uint32),
                ("checksum_computed", C.c_uint32), ("two_rows", C.c_int),
                ("double_prec", C.c_int), ("big_endian", C.c_int),
                ("data_offset", C.c_long)]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        LP = C.POINTER(Lat)
        sig = {
            "or_lat_init": (C.c_int, [LP, C.POINTER(SZ), C.c_int]),
            "or_lat_free": (None, [LP]),
            "or_psq": (C.c_double, [LP, SZ, C.c_int]),
            "or_gram_reunit": (None, [DP]),
            "or_set_gauge_matrix": (None, [DP, DP, SZ, C.c_double, SZ]),
            "or_gtransform_local": (None, [DP, DP, DP]),
            "or_retrace_abc_dag": (C.c_double, [DP, DP, DP]),
            "or_herm_proj_short": (None, [DP, DP]),
            "or_approx_minimum": (C.c_double, [DP, DP]),
            "or_plaquette": (C.c_double, [LP, DP]),
            "or_links": (C.c_double, [LP, DP, DP]),
            "or_all_links": (None, [LP, DP, DP, DP]),
            "or_functional": (C.c_double, [LP, DP]),
            "or_theta_lin": (C.c_double, [LP, DP, C.c_int]),
            "or_latt_reunit": (None, [LP, DP]),
            "or_fast_deriv": (C.c_double, [LP, DP, SZ, DP]),
            "or_latt_deriv": (C.c_double, [LP, DP, SZ, C.c_int, DP]),
            "or_fa_deriv": (C.c_double, [LP, DP, DP]),
            "or_fft": (None, [DP, C.POINTER(SZ), C.c_int, C.c_int, SZ]),
            "or_set_fft_mode": (None, [C.c_int]),
            "or_dft_naive": (None, [DP, C.POINTER(SZ), C.c_int, C.c_int]),
            "or_psq_table": (None, [LP, C.c_int, DP]),
            "or_fourier_accelerate": (None, [LP, DP, DP, C.c_int, SZ]),
            "or_egauge_landau": (None, [LP, DP, DP, C.c_double]),
            "or_line_search_landau": (C.c_double, [LP, DP, DP, DP]),
            "or_facg": (SZ, [LP, DP, C.c_double, C.c_double, SZ, DP,
                             C.POINTER(Trace)]),
            "or_fasd": (SZ, [LP, DP, C.c_double, C.c_double, SZ, DP,
                             C.POINTER(Trace)]),
            "or_coulomb": (SZ, [LP, DP, C.c_double, C.c_double, SZ, C.c_int,
                                C.c_uint32, C.POINTER(SZ), DP]),
            "or_rng_new": (C.c_void_p, [C.c_uint32, C.c_int]),
            "or_rng_free": (None, [C.c_void_p]),
            "or_rng_dbl": (C.c_double, [C.c_void_p, C.c_int]),
            "or_sunitary_gen": (None, [C.c_void_p, C.c_int, DP]),
            "or_random_gtrans": (None, [LP, C.c_void_p, DP]),
            "or_gtransform": (None, [LP, DP, DP]),
            "or_read_nersc": (C.c_int, [C.c_char_p, C.POINTER(NerscInfo), DP]),
            "or_write_nersc": (C.c_int, [C.c_char_p, LP, DP, C.c_int, C.c_int, C.c_int, C.c_char_p, SZ,
                                         C.c_char_p, C.c_char_p]),
            "or_momentum_list": (C.c_int, [LP, C.c_int, SZ, C.c_double, C.c_double, IP,
                                           C.POINTER(SZ), SZ, C.POINTER(SZ)]),
            "or_gluon_prop": (C.c_int, [LP, DP, C.c_int, SZ, C.c_double, C.c_double, IP, DP, DP, SZ,
                                        C.POINTER(SZ)]),
            "or_write_glueprop": (C.c_int, [C.c_char_p, SZ, IP, DP, DP]),
            "or_gluon_exceptional": (C.c_int, [LP, DP, C.c_int, SZ, C.c_double, C.c_double, IP, DP, DP, SZ,
                                               C.POINTER(SZ)]),
            "or_heatbath": (None, [LP, DP, C.c_double, C.c_int, C.c_int,
                                   C.c_uint64, C.c_int]),
            "or_glu_heatbath": (None, [LP, DP, C.c_double, SZ, SZ, C.c_uint32, C.c_int]),
            "or_glu_overrelax": (None, [LP, DP]),
            "or_glu_staple": (None, [LP, DP, SZ, C.c_int, DP]),
            "or_glu_colour": (C.c_int, [LP, SZ]),
            "or_glu_kp_samples": (None, [C.c_uint32, C.c_double, SZ, DP]),
            "or_set_plant": (None, [C.c_int]),
            "or_coulomb_trace": (None, [C.POINTER(Trace), SZ]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _dp(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(DP)


class Lattice:
    """Geometry + neighbour tables for a 4D periodic lattice (x fastest)."""

    def __init__(self, dims, nth: int | None = None):
        self.dims = tuple(int(d) for d in dims)
        self._lat = Lat()
        d = (SZ * 4)(*self.dims)
        if nth is None:
            nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        if lib().or_lat_init(C.byref(self._lat), d, int(nth)) != 0:
            raise MemoryError("or_lat_init failed")
        self.V = int(self._lat.V)
        self.Lcu = int(self._lat.Lcu)
        self.nth = int(nth)

    def __del__(self):
        try:
            lib().or_lat_free(C.byref(self._lat))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._lat)

    def fwd(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._lat.fwd, (self.V, 4)).copy()

    def bwd(self) -> np.ndarray:
        return np.ctypeslib.as_array(self._lat.bwd, (self.V, 4)).copy()

    # measurements
    def plaquette(self, U):
        return lib().or_plaquette(self.ref, _dp(U))

    def links(self, U):
        mx = C.c_double()
        tr = lib().or_links(self.ref, _dp(U), C.byref(mx))
        return tr, mx.value

    def functional(self, U):
        return lib().or_functional(self.ref, _dp(U))

    def theta_lin(self, U, max_dir=4):
        return lib().or_theta_lin(self.ref, _dp(U), max_dir)

    def psq(self, D=4):
        n = self.V if D == 4 else self.Lcu
        out = np.empty(n)
        lib().or_psq_table(self.ref, D, _dp(out))
        return out

    def fa_deriv(self, U):
        F = np.empty((4, self.V), np.complex128)
        th = lib().or_fa_deriv(self.ref, _dp(U), _dp(F))
        return th, F

    def fourier_accelerate(self, F, D=4):
        psq = self.psq(D)
        F = np.ascontiguousarray(F, np.complex128).copy()
        lib().or_fourier_accelerate(self.ref, _dp(F), _dp(psq), D, 4)
        return F

    def egauge(self, U, F, alpha):
        lib().or_egauge_landau(self.ref, _dp(U), _dp(F), alpha)

    def reunit(self, U):
        lib().or_latt_reunit(self.ref, _dp(U))

    def random_gtrans(self, U, seed: int, nth: int | None = None):
        r = lib().or_rng_new(seed, nth or self.nth)
        lib().or_random_gtrans(self.ref, r, _dp(U))
        lib().or_rng_free(r)

    def gtransform(self, U, G):
        lib().or_gtransform(self.ref, _dp(U), _dp(G))

    def heatbath(self, beta, sweeps, nor=4, seed=1, U=None):
        cold = U is None
        if U is None:
            U = np.zeros((self.V, 4, 18))
        lib().or_heatbath(self.ref, _dp(U), beta, sweeps, nor, seed, int(cold))
        return U

    def glu_heatbath(self, beta, sweeps, nor=4, seed=1, U=None):
        """GLU's own HB + OR update from GLU's MWC streams (self.nth of
        them); from unit links unless U is given."""
        cold = U is None
        if U is None:
            U = np.zeros((self.V, 4, 18))
        lib().or_glu_heatbath(self.ref, _dp(U), beta, sweeps, nor, seed, int(cold))
        return U

    def glu_overrelax(self, U):
        lib().or_glu_overrelax(self.ref, _dp(U))
        return U

    def glu_staple(self, U, i, mu):
        S = np.zeros(18)
        lib().or_glu_staple(self.ref, _dp(U), i, mu, _dp(S))
        return S

    def glu_colours(self):
        return np.array([lib().or_glu_colour(self.ref, i) for i in range(self.V)])

    def facg(self, U, gf_alpha=0.08, acc=1e-14, max_iters=1000, trace=0,
             functional=False, sd=False):
        th = C.c_double()
        tr = None
        bufs = {}
        if trace:
            tr = Trace()
            tr.cap = trace
            tr.n = 0
            for k in ("theta", "alpha", "beta") + (("func",) if functional else ()):
                bufs[k] = np.zeros(trace)
                setattr(tr, k, _dp(bufs[k]))
        f = lib().or_fasd if sd else lib().or_facg
        it = f(self.ref, _dp(U), gf_alpha, acc, max_iters, C.byref(th),
               C.byref(tr) if tr is not None else None)
        out = {"iters": int(it), "theta": th.value}
        if tr is not None:
            for k, v in bufs.items():
                out["trace_" + k] = v[: tr.n].copy()
        return out

    def coulomb(self, U, gf_alpha=0.09, acc=1e-14, max_iters=650, facg=True,
                seed=1, trace_slice=None, trace=0):
        """Coulomb_FA; with trace_slice set, the trajectory (theta_k, step,
        beta_k) of that slice is returned as trace_theta/alpha/beta;
        trace_slice="all" returns one such dict per slice as "traces"."""
        T = self.dims[3]
        its = (SZ * T)()
        th = np.zeros(T)
        tr, bufs = None, {}
        if trace_slice == "all":
            tr = (Trace * T)()
            bufs = [{k: np.zeros(trace) for k in ("theta", "alpha", "beta")} for _ in range(T)]
            for t in range(T):
                tr[t].cap, tr[t].n = trace, 0
                for k, v in bufs[t].items():
                    setattr(tr[t], k, _dp(v))
            lib().or_coulomb_trace(tr, C.c_size_t(-1).value)
        elif trace_slice is not None:
            tr = Trace()
            tr.cap, tr.n = trace, 0
            for k in ("theta", "alpha", "beta"):
                bufs[k] = np.zeros(trace)
                setattr(tr, k, _dp(bufs[k]))
            lib().or_coulomb_trace(C.byref(tr), trace_slice)
        try:
            tot = lib().or_coulomb(self.ref, _dp(U), gf_alpha, acc, max_iters,
                                   int(facg), seed, its, _dp(th))
        finally:
            if tr is not None:
                lib().or_coulomb_trace(None, 0)
        out = {"iters": int(tot), "slice_iters": list(its), "slice_theta": th}
        if trace_slice == "all":
            out["traces"] = [{"trace_" + k: v[: tr[t].n].copy() for k, v in bufs[t].items()}
                             for t in range(T)]
        elif tr is not None:
            for k, v in bufs.items():
                out["trace_" + k] = v[: tr.n].copy()
        return out


def glu_kp_samples(seed, norm, n):
    """n SU(2) heatbath values a0 of GLU's Kennedy-Pendleton step."""
    a0 = np.zeros(n)
    lib().or_glu_kp_samples(seed, norm, n, _dp(a0))
    return a0


class planted:
    """Context manager: run the oracle with one planted mistake
    (OR_PLANT_* of glu_oracle.h), to show that a pin catches it."""
    FR_ALWAYS, MAX_FACTOR, THETA_AFTER, EGAUGE_SIGN, NO_CLAMP, NO_BETA, NO_STALE, OLD_KEPT = range(1, 9)
    STAPLE_NODAG, OR_SIGN, KP_NOCOS = range(9, 12)

    def __init__(self, code: int):
        self.code = code

    def __enter__(self):
        lib().or_set_plant(self.code)
        return self

    def __exit__(self, *a):
        lib().or_set_plant(0)


class direct_dft:
    """Context manager: the oracle's Fourier transforms as direct per-axis
    sums instead of its mixed-radix decimation (or_set_fft_mode) -- an
    exact DFT with other rounding, whose trajectories show how far the
    transform's arithmetic moves them."""

    def __enter__(self):
        lib().or_set_fft_mode(1)
        return self

    def __exit__(self, *a):
        lib().or_set_fft_mode(0)


def fft(data: np.ndarray, dims, nd: int, sign: int) -> np.ndarray:
    """Unnormalised DFT over the first `nd` axes of `dims` (x fastest) of
    every row of `data` (shape [nb, prod(dims[:nd])])."""
    a = np.ascontiguousarray(data, np.complex128).copy()
    nb = a.shape[0] if a.ndim > 1 else 1
    d = (SZ * 4)(*(list(dims) + [1] * (4 - len(dims))))
    lib().or_fft(_dp(a), d, nd, sign, nb)
    return a


def dft_naive(data: np.ndarray, dims, nd: int, sign: int) -> np.ndarray:
    a = np.ascontiguousarray(data, np.complex128).copy()
    d = (SZ * 4)(*(list(dims) + [1] * (4 - len(dims))))
    lib().or_dft_naive(_dp(a), d, nd, sign)
    return a


def read_nersc(path: str):
    info = NerscInfo()
    rc = lib().or_read_nersc(path.encode(), C.byref(info), None)
    if rc != 0:
        raise IOError(f"NERSC header read failed ({rc}): {path}")
    V = int(np.prod(list(info.dims)))
    U = np.zeros((V, 4, 18))
    rc = lib().or_read_nersc(path.encode(), C.byref(info), _dp(U))
    if rc != 0:
        raise IOError(f"NERSC data read failed ({rc}): {path}")
    return U, info


CUTS = {"hypercubic": 0, "psq": 1, "cylinder": 2, "conical": 3}


def momentum_list(lat: Lattice, cut="psq", max_mom=4, cyl_width=1.0, angle=20.0):
    """GLU's momentum list of a cut: (mom[n, 4] signed ints, sites[n])."""
    n = SZ()
    t = CUTS[cut] if isinstance(cut, str) else cut
    rc = lib().or_momentum_list(lat.ref, t, max_mom, cyl_width, angle, None, None, 0, C.byref(n))
    if rc == -2:
        raise ValueError(f"unknown cut {cut}")
    mom = np.zeros((max(n.value, 1), 4), np.int32)
    idx = np.zeros(max(n.value, 1), np.uintp)
    rc = lib().or_momentum_list(lat.ref, t, max_mom, cyl_width, angle, mom.ctypes.data_as(IP),
                                idx.ctypes.data_as(C.POINTER(SZ)), n.value, C.byref(n))
    assert rc == 0
    return mom[:n.value], idx[:n.value].astype(np.int64)


def gluon_prop(lat: Lattice, U: np.ndarray, cut="psq", max_mom=4, cyl_width=1.0, angle=20.0):
    """Transverse and longitudinal propagators of GLU's GLUON_PROPS cuts:
    (mom[n, 4], G[n], F[n])."""
    mom0, _ = momentum_list(lat, cut, max_mom, cyl_width, angle)
    n = SZ()
    cap = max(len(mom0), 1)
    mom = np.zeros((cap, 4), np.int32)
    G, F = np.zeros(cap), np.zeros(cap)
    rc = lib().or_gluon_prop(lat.ref, _dp(np.ascontiguousarray(U)), CUTS[cut] if isinstance(cut, str) else cut,
                             max_mom, cyl_width, angle, mom.ctypes.data_as(IP), _dp(G), _dp(F), cap,
                             C.byref(n))
    assert rc == 0
    k = n.value
    return mom[:k], G[:k], F[:k]


def gluon_exceptional(lat: Lattice, U: np.ndarray, cut="psq", max_mom=4, cyl_width=1.0, angle=20.0):
    """GLU's exceptional two- and three-point functions: (mom[n, 4], g2[n], g3[n])."""
    mom0, _ = momentum_list(lat, cut, max_mom, cyl_width, angle)
    n = SZ()
    cap = max(len(mom0), 1)
    mom = np.zeros((cap, 4), np.int32)
    g2, g3 = np.zeros(cap), np.zeros(cap)
    rc = lib().or_gluon_exceptional(lat.ref, _dp(np.ascontiguousarray(U)),
                                    CUTS[cut] if isinstance(cut, str) else cut, max_mom, cyl_width, angle,
                                    mom.ctypes.data_as(IP), _dp(g2), _dp(g3), cap, C.byref(n))
    assert rc == 0
    k = n.value
    return mom[:k], g2[:k], g3[:k]


def write_glueprop(path: str, mom: np.ndarray, G: np.ndarray, F: np.ndarray):
    mom = np.ascontiguousarray(mom, np.int32)
    rc = lib().or_write_glueprop(path.encode(), len(G), mom.ctypes.data_as(IP), _dp(np.ascontiguousarray(G)),
                                 _dp(np.ascontiguousarray(F)))
    if rc != 0:
        raise IOError(f"propagator write failed ({rc})")


def write_nersc(path: str, lat: Lattice, U: np.ndarray, two_rows=False, big_endian=True, reunit=True,
                label="", sequence=1, creator="oracle", date="Mon Jan  0 00:00:00 1970"):
    """GLU's write_lat, NERSC branch; with reunit (GLU always does) U is
    reunitarised in place, as GLU does to the lattice it writes."""
    assert U.flags.c_contiguous and U.dtype == np.float64
    rc = lib().or_write_nersc(path.encode(), lat.ref, _dp(U), int(two_rows), int(big_endian), int(reunit),
                              label.encode(), sequence, creator.encode(), date.encode())
    if rc != 0:
        raise IOError(f"NERSC write failed ({rc})")


def rng_draws(seed: int, nth: int, thread: int, n: int) -> np.ndarray:
    r = lib().or_rng_new(seed, nth)
    out = np.array([lib().or_rng_dbl(r, thread) for _ in range(n)])
    lib().or_rng_free(r)
    return out


# single-matrix helpers (links are 18-double vectors)
def gram_reunit(U18):
    u = np.ascontiguousarray(U18, np.float64).copy()
    lib().or_gram_reunit(_dp(u))
    return u


def gtransform_local(a, b, c):
    b = np.ascontiguousarray(b, np.float64).copy()
    lib().or_gtransform_local(_dp(np.ascontiguousarray(a, np.float64)), _dp(b),
                              _dp(np.ascontiguousarray(c, np.float64)))
    return b


def retrace_abc_dag(a, b, c):
    return lib().or_retrace_abc_dag(*[_dp(np.ascontiguousarray(x, np.float64))
                                      for x in (a, b, c)])


def herm_proj_short(U18):
    q = np.zeros(10)
    lib().or_herm_proj_short(_dp(q), _dp(np.ascontiguousarray(U18, np.float64)))
    return q


def set_gauge_matrix(F, alpha, i):
    F = np.ascontiguousarray(F, np.complex128)
    g = np.zeros(18)
    lib().or_set_gauge_matrix(_dp(g), _dp(F), F.shape[1], alpha, i)
    return g


def approx_minimum(alphas, f):
    a = np.ascontiguousarray(alphas, np.float64)
    b = np.ascontiguousarray(f, np.float64)
    return lib().or_approx_minimum(_dp(a), _dp(b))


def sunitary(seed: int, nth: int, thread: int, n: int):
    r = lib().or_rng_new(seed, nth)
    out = np.zeros((n, 18))
    for k in range(n):
        lib().or_sunitary_gen(r, thread, _dp(out[k]))
    lib().or_rng_free(r)
    return out
