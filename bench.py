"""Benchmark: Landau gauge fixing (FACG) of 32^3x64 configurations to
theta < 1e-14 on one MI355X, mixed precision (FP32 streaming until theta
falls below --switch-theta, then FP64) against the plain FP64 path on the
same input.

One step = restore the input links from their device copy + one complete
gauge fixing to theta < 1e-14 (both inside the timed region).  The input is
a device heatbath config (beta = 6.0, cold start, 4 overrelaxation passes
per sweep) with a seeded random gauge transform, generated once before the
timed region.

Prints one JSON line (see the keys below).  Four further workloads print a
line of their own: --workload coulomb (48^3x96 Coulomb FACG, the 96 time
slices batched), --workload serve (a Poisson stream of 64 Landau
requests of mixed sizes through the serving loop), --workload glueprop
(gluon propagators of a Landau-fixed 32^3x64 config, GLU's GLUON_PROPS
cuts), --workload batch (512 gauge copies of an 8^4 config fixed in one
launch, one workgroup per lattice) and --workload dropin (GLU's own
Landau() on host struct site links, FP64 and mixed).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Landau configs/s to theta<1e-14 at 32^3x64 (mixed vs FP64); "
          "p50/p99 latency; GB/s")
PEAK_HBM_GBS = 8000.0
SENTINEL = 123456789  # FACG()'s failure return
# algorithmic bytes per site per launch (FP64 / FP32 streaming) of each
# kernel: each logical step reads its inputs once and writes its outputs
# once; an FFT axis pass reads and writes the 4 packed fields
ALG_BYTES = {
    "deriv": (640, 320), "fft_axis": (128, 64), "pr_sums": (128, 64),
    # sn = in + beta sn; GLU's copy old = in is a buffer swap here
    "cg_update": (192, 96), "line_search": (640, 320), "egauge": (1216, 608),
    # the gauge update fused with the next iteration's derivative
    "egauge_deriv": (1216 + 640, 608 + 320),
}
# per Landau FACG iteration (FP64 / FP32): the sum over the logical steps
ALG_ITER = (3072, 1536)
# the least each kernel must move with the links stored as their first two
# columns (384 / 192 B per site for the 4 directions): links read (and
# written), the packed fields it reads and writes once; the fused update
# of the FP32 phase also reads and writes the FP64 site transform (2 x 96)
STORED_BYTES = {
    "deriv": (384 + 64, 192 + 32), "fft_axis": (128, 64), "pr_sums": (128, 64), "cg_update": (192, 96),
    "line_search": (384 + 64, 192 + 32), "egauge": (384 + 64 + 384, 192 + 32 + 192),
    "egauge_deriv": (384 + 64 + 384 + 64, 192 + 32 + 192 + 32 + 192),
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dims", default="32,32,32,64")
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=20,
                    help="CPU baseline sample: k and 2k oracle iterations")
    ap.add_argument("--workload", choices=("landau", "coulomb", "serve", "glueprop", "batch", "dropin"),
                    default="landau")
    ap.add_argument("--batch", type=int, default=512, help="batch: lattices per launch")
    ap.add_argument("--max-mom", type=int, default=64, help="glueprop: PSQ cut p^2 <= max_mom")
    ap.add_argument("--graph", action="store_true", help="replay iteration batches from HIP graphs")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--slots", type=lambda v: [int(x) for x in v.split(",")], default=[4, 8, 1],
                    help="serve: concurrent requests of the Poisson runs (the first gives p50/p99)")
    ap.add_argument("--rate", type=float, default=2.0,
                    help="serve: Poisson arrivals per second (about 70%% of the one-at-a-time capacity)")
    ap.add_argument("--switch-theta", type=float, default=1e-11,
                    help="mixed precision: FP32 until theta falls below this")
    return ap.parse_args()


_JSON_OUT = None


def route_stdout():
    """Keep the process's stdout for the one JSON line and send everything
    else written to file descriptor 1 -- the progress lines GLU's entry
    points and the oracle print from C included -- to stderr."""
    global _JSON_OUT
    sys.stdout.flush()
    _JSON_OUT = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)


def emit(line):
    out = _JSON_OUT or sys.stdout
    out.write(json.dumps(line) + "\n")
    out.flush()


def pct(v, q):
    v = sorted(v)
    k = (len(v) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def kernel_mask(name, prec):
    import glu_amd as gl
    return 1 << (gl.KERNELS.index(name) + (len(gl.KERNELS) if prec == "fp32" else 0))


def solve(g, seed):
    """One step: the input config under its own seeded random gauge
    transform, gauge fixed with Landau()'s semantics (restarts included)."""
    g.restore()
    g.random_transform(seed=seed)
    it, th, restarts = g.landau(0, 0.08, 1e-14, 1000, restart_seed=seed)
    if not th < 1e-14:
        raise RuntimeError(f"did not converge: theta {th}")
    return it, restarts


def timed_runs(g, steps, warmup, seed):
    """Warm-up steps (the first one times every kernel class, to find the
    dominant kernel), then the timed steps with HIP events around the
    dominant kernel's launches only.  Step k fixes gauge copy seed + k of
    the input, so every precision sees the same copies."""
    prof_all = None
    for w in range(max(warmup, 1)):
        if w == 0:
            g.profile(-1)
        solve(g, seed + 1000 + w)
        if w == 0:
            prof_all = g.profile_read()
            g.profile(0)
    dom = max(prof_all.items(), key=lambda kv: kv[1][2])[0]
    g.measure()  # synchronises the device
    lat, its, restarts, split, func, phases = [], [], 0, [0, 0], [], [0.0, 0.0, 0.0]
    g.profile(kernel_mask(*dom))
    for k in range(steps):
        t = time.perf_counter()
        it, nr = solve(g, seed + k)
        lat.append(time.perf_counter() - t)
        its.append(it)
        restarts += nr
        n32, n64 = g.mixed_split()
        split[0] += n32
        split[1] += n64
        for j, ms in enumerate(g.phase_ms()):  # the last attempt's phases
            phases[j] += ms
        func.append(g.measure()["functional"])  # outside the step's clock
    total = sum(lat)
    live = g.profile_read()
    g.profile(0)
    phases = {k: round(v / steps, 2) for k, v in zip(("fp32", "switch", "fp64"), phases)}
    return total, lat, its, restarts, split, func, prof_all, dom, live[dom], phases


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: every "
              "configuration fits one MI355X and the multi-GPU design (one config "
              "per GPU, no communication) has no figure to measure here.", file=sys.stderr)
        sys.exit(2)
    route_stdout()
    import glu_amd as gl

    if a.workload == "coulomb":
        return coulomb_main(a)
    if a.workload == "serve":
        return serve_main(a)
    if a.workload == "glueprop":
        return glueprop_main(a)
    if a.workload == "batch":
        return batch_main(a)
    if a.workload == "dropin":
        return dropin_main(a)
    dims = tuple(int(x) for x in a.dims.split(","))
    V = math.prod(dims)
    print(f"[bench] lattice {dims} V={V}: generating input "
          f"(device heatbath beta=6.0, {a.sweeps} sweeps)", file=sys.stderr, flush=True)
    gm = gl.Gfix(dims, gl.LANDAU, gl.MIXED)
    gm.set_switch_theta(a.switch_theta)
    t = time.perf_counter()
    gm.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    plaq = gm.measure()["plaquette"]
    gm.snapshot()
    print(f"[bench] input ready in {time.perf_counter() - t:.1f}s, plaquette {plaq:.12f}",
          file=sys.stderr, flush=True)
    tune = gm.autotune()  # launch shapes, before any timing
    gm.set_graph(a.graph)
    print("[bench] autotune:\n" + tune.rstrip(), file=sys.stderr, flush=True)

    total, lat, its, restarts, (n32, n64), func, prof, dom, dom_live, phases = timed_runs(
        gm, a.steps, a.warmup, a.seed + 1)
    mixed_value = a.steps / total
    print(f"[bench] mixed: {mixed_value:.3f} configs/s, iters {its} (fp32 {n32} + fp64 {n64} in all), "
          f"restarts {restarts}", file=sys.stderr, flush=True)
    host = host_boundary(gm, a.seed + 1)
    # the timed mean solve with one upload and one download added
    host["configs_per_s_with_transfers"] = round(
        1.0 / (total / a.steps + 1e-3 * (host["upload_ms"] + host["download_ms"])), 4)

    fp64 = None
    if not a.no_fp64:
        gm.restore()
        U = gm.get_links()
        g64 = gl.Gfix(dims, gl.LANDAU, gl.FP64)
        g64.set_links(U)
        g64.snapshot()
        g64.autotune()
        g64.set_graph(a.graph)
        t64, lat64, its64, rs64, _, func64, prof64, _, _, _ = timed_runs(g64, a.steps, a.warmup, a.seed + 1)
        rel = [abs(x - y) / abs(y) for x, y in zip(func, func64)]
        n64only = sum(i for i in its64 if i < 123456789)
        fp64 = {"configs_per_s": a.steps / t64, "ms_per_step": 1e3 * t64 / a.steps,
                "iters": its64, "restarts": rs64, "p50_ms": 1e3 * pct(lat64, 0.5),
                "ms_per_iter": round(1e3 * t64 / max(n64only, 1), 4),
                # error of the mixed result against the FP64 one, per gauge copy:
                # ~1e-13 on the same Gribov copy, 1e-6..1e-4 on a neighbouring one
                "functional_rel_diff_mixed": [float(f"{r:.2e}") for r in rel]}
        print(f"[bench] fp64: {fp64['configs_per_s']:.3f} configs/s, iters {its64}, restarts {rs64}; "
              f"functional rel. diff mixed vs fp64 {fp64['functional_rel_diff_mixed']}",
              file=sys.stderr, flush=True)
        prof.update({k: v for k, v in prof64.items() if k not in prof})

    # dominant kernel of the mixed run and its roofline, from the HIP events
    # recorded around its launches during the timed steps
    (kname, prec), (nl, mean_ms, tot_ms) = dom, dom_live
    roof = roofline(kname, prec, V, mean_ms, dims)
    roof.update({"launches": nl})
    # per-kernel breakdown from the fully profiled warm-up run (one config),
    # priced with the bytes the stored format must move; a rate above the
    # HBM peak (working sets served from the 256 MB Infinity Cache) is
    # reported as such, not as bandwidth
    def rate(k, p, m):
        r = STORED_BYTES[k][0 if p == "fp64" else 1] * V / (m * 1e-3) / 1e9
        return (round(r, 1), None) if r <= PEAK_HBM_GBS else (None, round(r, 1))

    kernels = {}
    for (k, p), (n, m, tt) in sorted(prof.items()):
        if k not in STORED_BYTES:
            continue
        gbs, above = rate(k, p, m)
        kernels[f"{k}<{p}>"] = {"launches": n, "mean_ms": round(m, 4), "total_ms": round(tt, 2), "GBps": gbs}
        if above is not None:
            kernels[f"{k}<{p}>"]["above_hbm_peak_GBps"] = above

    # streaming rate of the mixed solves from the algorithmic bytes
    alg_total = (n32 * ALG_ITER[1] + n64 * ALG_ITER[0]) * V
    gbs = alg_total / total / 1e9

    cpu = None
    if not a.no_cpu:
        ref_its = [i for i in (fp64["iters"] if fp64 else its) if i < 123456789]
        gm.restore()
        gm.random_transform(seed=a.seed + 1)
        cpu = cpu_baseline(gm, dims, sorted(ref_its)[len(ref_its) // 2], a.cpu_iters)

    line = {
        "metric": METRIC if dims == (32, 32, 32, 64) else METRIC.replace("32^3x64", "x".join(map(str, dims))),
        "value": round(mixed_value, 4), "unit": "configs/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, "
                 f"4 overrelaxation each, seed {a.seed}); step k fixes its own gauge copy "
                 f"(random gauge transform seed {a.seed + 1}+k), with Landau() restarts"),
        "config": {"workload": "landau_facg_mixed_" + "x".join(map(str, dims)),
                   "lattice": list(dims), "gf_tune": 0.08, "accuracy": 1e-14, "max_iters": 1000,
                   "switch_theta": a.switch_theta, "graphs": a.graph},
        "p50_ms": round(1e3 * pct(lat, 0.5), 2), "p99_ms": round(1e3 * pct(lat, 0.99), 2),
        "iters": its, "restarts": restarts, "iters_fp32": n32, "iters_fp64": n64,
        # kernel speed apart from the (chaotic) iteration counts
        "ms_per_iter": round(1e3 * total / max(n32 + n64, 1), 4),
        # host-clock split of a mixed solve (mean over the timed steps; a
        # restarted step counts its last attempt) and the per-iteration cost
        # of each precision
        "phase_ms_per_config": phases,
        "ms_per_iter_fp32": round(phases["fp32"] * a.steps / max(n32, 1), 4),
        "ms_per_iter_fp64_tail": round(phases["fp64"] * a.steps / max(n64, 1), 4),
        "stream_GBps": round(gbs, 1),
        "fp64": fp64,
        "mixed_speedup_vs_fp64": round(mixed_value / fp64["configs_per_s"], 3) if fp64 else None,
        "host_boundary": host,
        "roofline": roof, "kernels": kernels, "cpu_baseline": cpu,
    }
    emit(line)


def roofline(kname, prec, V, mean_ms, dims):
    """The roofline object of one kernel from its mean launch time:
    achieved = the algorithmic bytes credited per site (the steps it
    performs, each reading its inputs and writing its outputs once, links
    as 9 complex numbers) / time; achieved_stored prices the same work in
    the stored two-column format (the kernel's own minimum); traffic is the
    HBM bytes per launch read from the committed rocprofv3 counter summary
    of this kernel and lattice (FETCH_SIZE + WRITE_SIZE, calibrated on a
    copy kernel), frac_counter = traffic / time / peak."""
    p = 0 if prec == "fp64" else 1
    achieved = ALG_BYTES[kname][p] * V / (mean_ms * 1e-3) / 1e9
    stored = STORED_BYTES[kname][p] * V / (mean_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(f"{kname}_{prec}_{'x'.join(map(str, dims))}")
    out = {"kernel": f"{kname}<{prec}>", "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
           "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
           "achieved_stored": round(stored, 1), "frac_stored": round(stored / PEAK_HBM_GBS, 4),
           "mean_ms": round(mean_ms, 4)}
    if traffic:
        out["frac_counter"] = round(traffic / (mean_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
        out["traffic_per_site"] = round(traffic / V, 1)
    return out


def host_boundary(g, seed):
    """GLU's entry points hand over host links: the first gauge copy once
    more with the upload of its links from host memory and the download of
    the result inside the clock (outside `value`, whose inputs are resident
    in HBM)."""
    g.restore()
    g.random_transform(seed=seed)
    U = g.get_links()
    g.measure()
    t0 = time.perf_counter()
    g.set_links(U)
    g.measure()  # synchronises the device
    t1 = time.perf_counter()
    it, th, nr = g.landau(0, 0.08, 1e-14, 1000, restart_seed=seed)
    t2 = time.perf_counter()
    U = g.get_links()
    t3 = time.perf_counter()
    out = {"one_copy_configs_per_s": round(1.0 / (t3 - t0), 4), "upload_ms": round(1e3 * (t1 - t0), 1),
           "solve_ms": round(1e3 * (t2 - t1), 1), "download_ms": round(1e3 * (t3 - t2), 1),
           "bytes_each_way": int(U.nbytes)}
    print(f"[bench] with the host boundary: {out}", file=sys.stderr, flush=True)
    return out


def cpu_baseline(g, dims, iters_fp64, k):
    """The CPU oracle (C, OpenMP) on the context's current links (the first
    timed gauge copy): time k and 2k FACG iterations, take the
    per-iteration cost from the difference and scale it by the median FP64
    GPU iteration count of a full solve."""
    from oracle import oracle as O

    nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    U = g.get_links()
    L = O.Lattice(dims, nth=nth)
    ts = []
    for n in (k, 2 * k):
        W = U.copy()
        t = time.perf_counter()
        L.facg(W, 0.08, 1e-14, n)
        ts.append(time.perf_counter() - t)
        print(f"[bench] cpu oracle: {n} iterations in {ts[-1]:.1f}s", file=sys.stderr, flush=True)
    per_iter = max((ts[1] - ts[0]) / k, 1e-9)
    per_cfg = per_iter * (iters_fp64 + 1) + max(ts[0] - k * per_iter, 0.0)
    return {"value": round(1.0 / per_cfg, 6), "unit": "configs/s", "cores": nth, "kind": "port",
            "sample": (f"oracle FP64 FACG on the same {'x'.join(map(str, dims))} input: "
                       f"{k} and {2 * k} iterations timed ({ts[0]:.1f}s, {ts[1]:.1f}s), "
                       f"{per_iter:.2f}s/iteration scaled to the {iters_fp64} iterations of "
                       "the full FP64 solve")}


# per slice-site and slice iteration (FP64): derivative on rotated links,
# forward + backward FFT credited once each, PR sums, CG update, line
# search, slice-transform update
COULOMB_ALG = (576 + 144 + 64) + 2 * 128 + 128 + 256 + (576 + 144 + 64) + (144 + 64 + 144)
# per kernel, B per slice-site and launch: credited (links as 9 complex
# numbers, the CG update with GLU's in_old copy) and stored (the three
# spatial links' two columns 288 B, the slice transform's 96 B)
COULOMB_KERNEL_BYTES = {
    "c_deriv": (576 + 144 + 64, 288 + 96 + 64), "c_pr_sums": (128, 128), "c_cg_update": (256, 192),
    "c_line_search": (576 + 144 + 64, 288 + 96 + 64), "c_slice_update": (144 + 64 + 144, 96 + 64 + 96),
    "fft_axis": (128, 128),
}


def dropin_main(a):
    """GLU's own Landau() (Gfix/Landau.c:261-374 prototype): `struct site`
    links in host memory and the global Latt, as GLU calls it.  One call =
    the links gathered and uploaded, fixed (FP64, then mixed through
    GLU_GFIX_PRECISION), reunitarised, reported with GLU's [GF] lines
    (including its temporal-constancy logarithms) and written back into the
    host links.  The first call of each precision (context creation) is not
    timed."""
    import ctypes as C

    import glu_amd as gl

    dims = tuple(int(x) for x in a.dims.split(","))
    V = math.prod(dims)
    n = max(1, min(a.steps, 8))
    g = gl.Gfix(dims, gl.LANDAU, gl.FP64)
    g.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    plaq = g.measure()["plaquette"]
    g.snapshot()
    lib = C.CDLL(gl.LIB_PATH)
    lib.Landau.restype = C.c_size_t
    lib.Landau.argtypes = [C.c_void_p, C.c_double, C.c_size_t, C.c_char_p]

    class LattInfo(C.Structure):  # struct latt_info, Headers/GLU_types.h:229-243
        _fields_ = [("dims", C.c_size_t * 4), ("Lsq", C.c_size_t), ("Lcu", C.c_size_t), ("Volume", C.c_size_t),
                    ("flow", C.c_size_t), ("gf_alpha", C.c_double), ("sm_alpha", C.c_double * 3),
                    ("head", C.c_int), ("Seed", C.c_uint32 * 1), ("twiddles", C.c_double * 4),
                    ("Nthreads", C.c_uint32), ("su2_data", C.c_void_p), ("argc", C.c_int)]

    L = LattInfo.in_dll(lib, "Latt")
    for mu in range(4):
        L.dims[mu] = dims[mu]
    L.Lsq, L.Lcu, L.Volume = dims[0] * dims[1], dims[0] * dims[1] * dims[2], V
    L.gf_alpha, L.Seed[0], L.Nthreads = 0.08, 1, 1
    U = np.empty((V, 4, 18))
    # struct site: O -> 4 pointers, one per direction, to 9 complex numbers
    ptr = (U.ctypes.data + 144 * np.arange(V * 4, dtype=np.uint64)).reshape(V, 4)
    site = np.zeros(V, np.dtype([("O", "u8"), ("neighbor", "u8", (4,)), ("back", "u8", (4,))]))
    site["O"] = ptr.ctypes.data + 32 * np.arange(V, dtype=np.uint64)
    runs = {}
    for prec in ("fp64", "mixed"):
        os.environ["GLU_GFIX_PRECISION"] = prec
        times, its = [], []
        for k in range(n + 1):
            g.restore()
            g.random_transform(seed=a.seed + 1 + k)
            U[...] = g.get_links()
            t = time.perf_counter()
            it = lib.Landau(site.ctypes.data, 1e-14, 1000, b"")
            dt = time.perf_counter() - t
            if k:
                times.append(dt)
                its.append(int(it))
        runs[prec] = {"calls_per_s": round(len(times) / sum(times), 4), "s_per_call": round(sum(times) / len(times), 3),
                      "iters": its}
        print(f"[bench] Landau() {prec}: {runs[prec]}", file=sys.stderr, flush=True)
    os.environ.pop("GLU_GFIX_PRECISION", None)
    line = {
        "metric": f"GLU Landau() calls/s at {'x'.join(map(str, dims))} (host struct site links in and out, report "
                  "included)",
        "value": runs["mixed"]["calls_per_s"], "unit": "calls/s", "n_gpus": 1, "steps": n, "warmup": 1,
        "ms_per_step": round(1e3 * runs["mixed"]["s_per_call"], 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, seed {a.seed}); call k "
                 f"fixes gauge copy seed {a.seed + 1}+k, copied into GLU-layout host links"),
        "config": {"workload": "glu_landau_dropin_" + "x".join(map(str, dims)), "lattice": list(dims),
                   "accuracy": 1e-14, "max_iters": 1000, "gf_tune": 0.08},
        "plaquette": plaq, "runs": runs, "roofline": None, "cpu_baseline": None,
    }
    emit(line)


def batch_main(a):
    """Landau FACG of many small lattices of one shape, one workgroup per
    lattice, all in one launch (glu_gfix_facg_batch), FP64, against the
    one-lattice context path on the same gauge copies.  One step = restore
    the B inputs from their device copy + one launch over all of them, with
    Landau()'s random restarts (on the context path) for any that fail."""
    import torch

    import glu_amd as gl

    dims = (8, 8, 8, 8) if a.dims == "32,32,32,64" else tuple(int(x) for x in a.dims.split(","))
    V, B = math.prod(dims), a.batch
    g = gl.Gfix(dims, gl.LANDAU, gl.FP64)
    t = time.perf_counter()
    g.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    plaq = g.measure()["plaquette"]
    g.snapshot()
    src = torch.empty((B, V * 72), dtype=torch.float64, device="cuda")
    for k in range(B):  # gauge copy k: the input under random transform seed + 1 + k
        g.restore()
        g.random_transform(seed=a.seed + 1 + k)
        g.get_links_device(src[k].data_ptr())
    torch.cuda.synchronize()
    dev = torch.empty_like(src)
    print(f"[bench] batch {dims} x {B}: inputs ready in {time.perf_counter() - t:.1f}s, plaquette {plaq:.12f}",
          file=sys.stderr, flush=True)

    def step():
        dev.copy_(src)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        its, th = gl.facg_batch(dims, dev.data_ptr(), B, 0.08, 1e-14, 1000)
        t_launch = time.perf_counter() - t1
        failed = [k for k, x in enumerate(its) if x == SENTINEL or not th[k] < 1e-14]
        for k in failed:  # Landau()'s restarts, from the lattice's input
            g.set_links_device(src[k].data_ptr())
            it, thk, _ = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=a.seed + 1 + k)
            if not thk < 1e-14:
                raise RuntimeError(f"lattice {k} did not converge: theta {thk}")
            g.get_links_device(dev[k].data_ptr())
            its[k] = it
        torch.cuda.synchronize()
        return its, t_launch, len(failed)

    for _ in range(max(a.warmup, 1)):
        step()
    lat, launch, its_all, restarts = [], [], [], 0
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        its, tl, nf = step()
        lat.append(time.perf_counter() - t)
        launch.append(tl)
        its_all += its
        restarts += nf
    total = time.perf_counter() - t0
    g.set_links_device(dev[0].data_ptr())  # lattice 0 as fixed by the last step
    g.reunit()
    drift = abs(g.measure()["plaquette"] - plaq)
    # the context path on the first copies, one after another
    nctx = min(B, 16)
    ctx_its = []
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(nctx):
        g.set_links_device(src[k].data_ptr())
        it, thk, _ = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=a.seed + 1 + k)
        ctx_its.append(it)
    g.measure()
    t_ctx = time.perf_counter() - t
    # the serving loop on the same copies, every request submitted at once:
    # with shared launches, and one context per request (8 in flight)
    served = {}
    for batching in (True, False):
        n = B if batching else min(B, 128)
        bufs = src[:n].clone()
        torch.cuda.synchronize()
        with gl.Server(max_active=8, batching=batching) as srv:
            t = time.perf_counter()
            tickets = [srv.submit(dims, bufs[k].data_ptr(), gl.LANDAU, gl.FP64) for k in range(n)]
            for tk in tickets:
                srv.wait(tk)
            dt = time.perf_counter() - t
        served["shared_launches" if batching else "contexts"] = {"requests": n, "configs_per_s": round(n / dt, 1)}
        print(f"[bench] server, {'shared launches' if batching else 'contexts'}: {n} requests in {dt:.2f}s",
              file=sys.stderr, flush=True)
    n_ok = [x for x in its_all if x < SENTINEL]
    mean_it = sum(n_ok) / max(len(n_ok), 1)
    # each lattice iteration moves the logical bytes of one FACG iteration
    # of its V sites (links as GLU stores them, 9 complex numbers)
    launch_s = sum(launch) / len(launch)
    gbs = sum(its_all) / a.steps * V * ALG_ITER[0] / launch_s / 1e9 if not restarts else None
    cpu = None if a.no_cpu else batch_cpu_baseline(src, dims, B)
    line = {
        "metric": f"Landau configs/s to theta<1e-14 at {'x'.join(map(str, dims))}, {B} lattices per launch",
        "value": round(B * a.steps / total, 2), "unit": "configs/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, 4 overrelaxation "
                 f"each, seed {a.seed}); lattice k of a batch is its gauge copy under random transform "
                 f"seed {a.seed + 1}+k"),
        "config": {"workload": "landau_facg_batch_fp64_" + "x".join(map(str, dims)), "lattice": list(dims),
                   "lattices_per_launch": B, "gf_tune": 0.08, "accuracy": 1e-14, "max_iters": 1000},
        "p50_ms": round(1e3 * pct(lat, 0.5), 2), "launch_ms": round(1e3 * launch_s, 2),
        "iters_mean": round(mean_it, 1), "iters_max": max(n_ok) if n_ok else None, "restarts": restarts,
        "plaquette_drift_lattice0": drift,
        "context_path": {"configs_per_s": round(nctx / t_ctx, 2), "iters_mean": round(sum(ctx_its) / nctx, 1),
                         "copies": nctx},
        "speedup_vs_context_path": round((B * a.steps / total) / (nctx / t_ctx), 2),
        "server": served,
        "stream_GBps": round(gbs, 1) if gbs else None,
        "roofline": batch_roofline(gbs, launch_s, dims, B, sum(its_all) / a.steps),
        "cpu_baseline": cpu,
    }
    emit(line)


def batch_roofline(gbs, launch_s, dims, B, lattice_iters):
    """Roofline of the shared launch: the logical bytes of the FACG
    iterations it ran (3072 B per site per iteration) over its time, and
    beside them the HBM bytes the counters measured for the same launch
    (profiles/traffic.json, FETCH_SIZE x 2 + WRITE_SIZE as calibrated on a
    copy kernel of known size)."""
    out = {"kernel": "k_facg_batch (whole FACG per workgroup)", "bound": "hbm",
           "achieved": round(gbs, 1) if gbs else None, "peak": PEAK_HBM_GBS, "unit": "GB/s",
           "frac": round(gbs / PEAK_HBM_GBS, 4) if gbs else None, "traffic": None}
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    key = f"facg_batch_fp64_{'x'.join(map(str, dims))}_{B}"
    if os.path.exists(tfile) and gbs:
        t = json.load(open(tfile)).get(key)
        if t:
            out["traffic"] = t
            out["frac_counter"] = round(t / launch_s / 1e9 / PEAK_HBM_GBS, 4)
            out["traffic_per_site_iteration"] = round(t / (lattice_iters * math.prod(dims)), 1)
    return out


def batch_cpu_baseline(src, dims, B):
    """The CPU oracle (C, OpenMP) solving gauge copies of the batch one
    after another, until about 10 s have passed."""
    from oracle import oracle as O

    nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    L = O.Lattice(dims, nth=nth)
    t = time.perf_counter()
    n, its = 0, []
    while n < B and (time.perf_counter() - t < 10.0 or n == 0):
        W = src[n].cpu().numpy().reshape(-1, 4, 18).copy()
        its.append(L.facg(W, 0.08, 1e-14, 1000)["iters"])
        n += 1
    dt = time.perf_counter() - t
    return {"value": round(n / dt, 3), "unit": "configs/s", "cores": nth, "kind": "port",
            "sample": (f"oracle FP64 FACG on the first {n} gauge copies of the batch, one after another "
                       f"({dt:.1f}s, iterations {min(its)}..{max(its)})")}


def coulomb_main(a):
    """48^3x96 Coulomb FACG, FP64: all 96 slice problems advance together
    (one launch per kernel covers every slice still iterating)."""
    import glu_amd as gl

    dims = (48, 48, 48, 96) if a.dims == "32,32,32,64" else tuple(int(x) for x in a.dims.split(","))
    V = math.prod(dims)
    g = gl.Gfix(dims, gl.COULOMB, gl.FP64)
    t = time.perf_counter()
    g.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    plaq = g.measure()["plaquette"]
    g.random_transform(seed=a.seed + 1)
    g.snapshot()
    print(f"[bench] coulomb {dims}: input ready in {time.perf_counter() - t:.1f}s, plaquette {plaq:.12f}",
          file=sys.stderr, flush=True)
    for _ in range(max(a.warmup, 1)):
        g.restore()
        g.run(gl.CG, 0.09, 1e-14, 650)
    # one profiled solve: every slice kernel class timed with HIP events
    g.restore()
    g.profile(sum(kernel_mask(k, "fp64") for k in COULOMB_KERNEL_BYTES))
    g.run(gl.CG, 0.09, 1e-14, 650)
    prof = g.profile_read()
    g.profile(0)
    psits, _ = g.slice_info()
    g.measure()
    lat, its = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        g.restore()
        it, th = g.run(gl.CG, 0.09, 1e-14, 650)
        lat.append(time.perf_counter() - t)
        its.append(it)
        if not th < 1e-14:
            raise RuntimeError(f"did not converge: theta {th}")
    total = time.perf_counter() - t0
    sits, sth = g.slice_info()
    m = g.measure()
    # each slice iteration streams one slice's share of the Landau-like
    # step sequence; the batched launches cover the slices still running
    gbs = sum(sits) * (V // dims[3]) * COULOMB_ALG / (sum(lat) / len(lat)) / 1e9
    # per kernel: a launch covers the slices still iterating; the slice
    # stencils run once per slice iteration plus each slice's SD start, so
    # their slice-launches number sum(slice iterations) + T (the line search
    # is skipped for slices below CG_TOL and is priced by time only)
    Lcu = V // dims[3]
    slaunch = sum(psits) + dims[3]
    kernels = {}
    for (k, p), (n, mean, tt) in sorted(prof.items()):
        e = {"launches": n, "mean_ms": round(mean, 4), "total_ms": round(tt, 2)}
        if k in ("c_deriv", "c_pr_sums", "c_cg_update", "c_slice_update"):
            cred, stored = COULOMB_KERNEL_BYTES[k]
            e["GBps"] = round(cred * Lcu * slaunch / (tt * 1e-3) / 1e9, 1)
            e["GBps_stored"] = round(stored * Lcu * slaunch / (tt * 1e-3) / 1e9, 1)
        kernels[k] = e
    roof = None
    if "c_deriv" in kernels:
        d = kernels["c_deriv"]
        roof = {"kernel": "k_cderiv (slice derivative on rotated links)", "bound": "hbm",
                "achieved": d["GBps"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(d["GBps"] / PEAK_HBM_GBS, 4), "traffic": None,
                "achieved_stored": d["GBps_stored"], "frac_stored": round(d["GBps_stored"] / PEAK_HBM_GBS, 4),
                "slice_launches": slaunch, "total_ms": d["total_ms"]}
        # HBM bytes the counters measured per slice-site of a launch
        # (profiles/traffic.json; a launch covers the slices still running)
        tfile = os.path.join(ROOT, "profiles", "traffic.json")
        per = (json.load(open(tfile)).get(f"c_deriv_fp64_{'x'.join(map(str, dims))}_per_slice_site")
               if os.path.exists(tfile) else None)
        if per:
            roof["traffic"] = round(per * Lcu * slaunch / d["launches"], -5)
            roof["traffic_per_slice_site"] = per
            roof["frac_counter"] = round(per * Lcu * slaunch / (d["total_ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)
    cpu = None if a.no_cpu else coulomb_cpu_baseline(g, dims)
    line = {
        "metric": "Coulomb configs/s to theta<1e-14 at 48^3x96 (96 slices batched)",
        "value": round(a.steps / total, 4), "unit": "configs/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, "
                 f"seed {a.seed}) + seeded random gauge transform"),
        "config": {"workload": "coulomb_facg_fp64_" + "x".join(map(str, dims)), "lattice": list(dims),
                   "gf_tune": 0.09, "accuracy": 1e-14, "max_iters": 650},
        "p50_ms": round(1e3 * pct(lat, 0.5), 2), "total_iters": its,
        "slice_iters": {"min": min(sits), "max": max(sits), "mean": round(sum(sits) / len(sits), 1)},
        "max_slice_theta": float(max(sth)), "plaquette_drift": abs(m["plaquette"] - plaq),
        "stream_GBps": round(gbs, 1), "roofline": roof, "kernels": kernels, "cpu_baseline": cpu,
    }
    emit(line)


def coulomb_cpu_baseline(g, dims):
    """The CPU oracle (C, OpenMP, GLU's slice-by-slice Coulomb_FA) on the
    first two time slices of the same input, cut out as an L^3 x 2 lattice,
    scaled by T / 2 (the slices are independent problems)."""
    from oracle import oracle as O

    nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g.restore()
    U = g.get_links()
    Lcu = math.prod(dims[:3])
    cut = (dims[0], dims[1], dims[2], 2)
    W = np.ascontiguousarray(U[: 2 * Lcu])
    del U
    t = time.perf_counter()
    r = O.Lattice(cut, nth=nth).coulomb(W, 0.09, 1e-14, 650, facg=True)
    ts = time.perf_counter() - t
    per_cfg = ts * dims[3] / 2
    print(f"[bench] coulomb cpu oracle: 2 slices in {ts:.1f}s ({r['slice_iters']} iterations)", file=sys.stderr,
          flush=True)
    return {"value": round(1.0 / per_cfg, 6), "unit": "configs/s", "cores": nth, "kind": "port",
            "sample": (f"oracle Coulomb FACG (GLU's slice-by-slice loop) on time slices 0 and 1 of the same "
                       f"input cut out as {'x'.join(map(str, cut))}: {ts:.1f}s for {r['slice_iters']} slice "
                       f"iterations, scaled by {dims[3]}/2 slices")}


# gluon propagators, B/site per launch: the Lie-field packing reads the
# 2-column FP64 links (384) and writes 16 packed complex fields (256); an
# FFT axis pass reads and writes those 16 fields
GLUEPROP_ALG = {"lie_pack": 640, "fft_axis": 512}


def glueprop_main(a):
    """Transverse / longitudinal gluon propagators (GLU's GLUON_PROPS cuts,
    PSQ cut) of a Landau-fixed config: one step = one measurement on the
    device links (Lie fields, 16-field transform, Landau phase,
    projections over the momentum list, copy of the results to the host)."""
    import numpy as np

    import glu_amd as gl
    from oracle import oracle as O

    dims = tuple(int(x) for x in a.dims.split(","))
    V = math.prod(dims)
    g = gl.Gfix(dims, gl.LANDAU, gl.MIXED)
    t = time.perf_counter()
    g.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    g.random_transform(seed=a.seed + 1)
    g.autotune()
    g.set_switch_theta(a.switch_theta)
    it, th, _ = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=a.seed + 7)
    print(f"[bench] glueprop {dims}: input fixed in {time.perf_counter() - t:.1f}s ({it} iterations, "
          f"theta {th:.2e})", file=sys.stderr, flush=True)
    for _ in range(max(a.warmup, 1)):
        mom, G, F, err = g.gluon_props("psq", a.max_mom)
    g.profile(sum(kernel_mask(k, "fp64") for k in ("lie_pack", "fft_axis", "gluon_prop")))
    g.gluon_props("psq", a.max_mom)
    prof = g.profile_read()
    g.profile(0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        t = time.perf_counter()
        mom, G, F, err = g.gluon_props("psq", a.max_mom)
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    U = g.get_links()
    g.close()
    kern = {}
    for (name, prec), (n, mean, tot) in prof.items():
        kern[name] = {"launches": n, "mean_ms": round(mean, 4), "total_ms": round(tot, 3)}
        if name in GLUEPROP_ALG:
            kern[name]["GBps"] = round(GLUEPROP_ALG[name] * V / (mean * 1e-3) / 1e9, 1)
    dom = max(GLUEPROP_ALG, key=lambda k: kern.get(k, {}).get("total_ms", 0.0))
    ach = GLUEPROP_ALG[dom] * V / (kern[dom]["mean_ms"] * 1e-3) / 1e9
    # HBM bytes per launch from the calibrated counters, where measured
    tkey = {"fft_axis": "fft_axis_16f", "lie_pack": "lie_pack"}[dom] + "_fp64_" + "x".join(map(str, dims))
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    traffic = json.load(open(tfile)).get(tkey) if os.path.exists(tfile) else None
    # CPU: the oracle (GLU's 36-field transform, one thread) on a 16^3x32
    # sample of the same fixed field's shape, scaled by the volume
    cpu = None
    if not a.no_cpu:
        sd = (16, 16, 16, 32)
        Ls = O.Lattice(sd, nth=1)
        Us = np.ascontiguousarray(U.reshape(dims[::-1] + (4, 18))[:sd[3], :sd[2], :sd[1], :sd[0]]).reshape(-1, 4, 18)
        t = time.perf_counter()
        O.gluon_prop(Ls, Us, "psq", a.max_mom)
        ts = time.perf_counter() - t
        scale = V / math.prod(sd)
        cpu = {"value": round(1.0 / (ts * scale), 5), "unit": "measurements/s", "cores": 1, "kind": "port",
               "sample": f"oracle on a 16x16x16x32 corner of the same field ({ts:.2f}s), scaled by the volume "
                         f"ratio {scale:.0f} (FFT log factor ignored)"}
    line = {
        "metric": "gluon propagator measurements/s at 32^3x64 (GLU GLUON_PROPS, PSQ cut)",
        "value": round(a.steps / total, 3), "unit": "measurements/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, seed {a.seed}), "
                 f"random gauge transform, Landau-fixed to theta<1e-14 before the timed region"),
        "config": {"workload": "gluon_props_psq_" + "x".join(map(str, dims)), "lattice": list(dims),
                   "cut": "psq", "max_mom": a.max_mom, "momenta": int(len(G))},
        "p50_ms": round(1e3 * pct(lat, 0.5), 3), "pspace_landau_error": err,
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4), "traffic": traffic},
        "kernels": kern, "cpu_baseline": cpu,
    }
    emit(line)


SERVE_SIZES = ((16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64))
SERVE_BETAS = (5.8, 6.0, 6.2)


def serve_main(a):
    """A stream of Landau requests (mixed precision, theta < 1e-14) with
    Poisson arrivals through the serving loop.  Sizes and betas are drawn
    uniformly; each request is a heatbath config of its (size, beta) with a
    random gauge transform of its own, resident in device memory before
    the stream starts.  The arrival schedule is served with several slot
    counts (continuous batching) and one request at a time; then all
    requests are queued at once, which measures the capacity."""
    import threading

    import numpy as np
    import torch

    import glu_amd as gl

    rng = np.random.default_rng(a.seed)
    n = a.requests
    picks = [(SERVE_SIZES[rng.integers(3)], SERVE_BETAS[rng.integers(3)]) for _ in range(n)]
    gaps = rng.exponential(1.0 / a.rate, size=n)
    arrivals = np.cumsum(gaps) - gaps[0]
    t = time.perf_counter()
    base = {}
    for dims in SERVE_SIZES:
        for beta in SERVE_BETAS:
            if (dims, beta) not in {p for p in picks}:
                continue
            g = gl.Gfix(dims, gl.LANDAU, gl.FP64)
            g.heatbath(beta, a.sweeps, 4, seed=a.seed + int(10 * beta), cold_start=True)
            g.snapshot()
            base[(dims, beta)] = g
    bufs = []
    for k, (dims, beta) in enumerate(picks):
        g = base[(dims, beta)]
        g.restore()
        g.random_transform(seed=a.seed + 1000 + k)
        b = torch.empty(math.prod(dims) * 72, dtype=torch.float64, device="cuda")
        g.get_links_device(b.data_ptr())
        bufs.append(b)
    inputs = [b.clone() for b in bufs]
    for g in base.values():
        g.close()
    torch.cuda.synchronize()
    print(f"[bench] serve: {n} requests ready in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)

    def stream(slots, poisson=True):
        for b, x in zip(bufs, inputs):
            b.copy_(x)
        torch.cuda.synchronize()
        srv = gl.Server(max_active=slots, switch_theta=a.switch_theta)
        # warm the context pool: one request of each size, untimed
        for dims in SERVE_SIZES:
            k = next((i for i, p in enumerate(picks) if p[0] == dims), None)
            if k is not None:
                srv.wait(srv.submit(dims, bufs[k].data_ptr(), precision=gl.MIXED))
                bufs[k].copy_(inputs[k])
        torch.cuda.synchronize()
        tickets = [None] * n

        def feeder(t0):
            for k, (dims, _) in enumerate(picks):
                dt = t0 + arrivals[k] - time.perf_counter()
                if poisson and dt > 0:
                    time.sleep(dt)
                tickets[k] = srv.submit(dims, bufs[k].data_ptr(), precision=gl.MIXED)

        t0 = time.perf_counter()
        th = threading.Thread(target=feeder, args=(t0,))
        th.start()
        th.join()
        res = [srv.wait(tk) for tk in tickets]
        wall = time.perf_counter() - t0
        srv.close()
        lats = [r[2] for r in res]
        # a restarted request reports 123456789 + ..., as Landau() does
        restarted = sum(1 for r in res if r[0] >= 123456789)
        which = sorted("x".join(map(str, picks[k][0])) + f"@{picks[k][1]}" for k, r in enumerate(res)
                       if r[0] >= 123456789)
        its = [r[0] for r in res if r[0] < 123456789]
        if not all(r[1] < 1e-14 for r in res):
            raise RuntimeError("a request did not converge")
        return {"configs_per_s": round(n / wall, 3), "wall_s": round(wall, 2),
                "p50_ms": round(pct(lats, 0.5), 1), "p99_ms": round(pct(lats, 0.99), 1),
                "mean_iters": round(sum(its) / max(len(its), 1), 1), "restarted": restarted,
                "restarted_requests": which}

    runs = {}
    for slots in a.slots:
        runs[f"poisson_slots{slots}"] = r = stream(slots)
        print(f"[bench] serve Poisson {a.rate}/s, {slots} slot(s): {r}", file=sys.stderr, flush=True)
    for slots in (max(a.slots), 1):
        runs[f"saturated_slots{slots}"] = r = stream(slots, poisson=False)
        print(f"[bench] serve all at t=0, {slots} slot(s): {r}", file=sys.stderr, flush=True)
    mix = {}
    for dims, _ in picks:
        mix["x".join(map(str, dims))] = mix.get("x".join(map(str, dims)), 0) + 1
    sat = runs[f"saturated_slots{max(a.slots)}"]
    lat = runs[f"poisson_slots{a.slots[0]}"]
    line = {
        "metric": ("Landau serving stream: configs/s with all requests queued, p50/p99 latency "
                   "under Poisson arrivals (64 requests)"),
        "value": sat["configs_per_s"], "unit": "configs/s", "n_gpus": 1, "steps": n, "warmup": 3,
        "ms_per_step": round(1e3 * sat["wall_s"] / n, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32+fp64",
        "data": (f"synthetic: device heatbath ({a.sweeps} sweeps) per (size, beta), one random "
                 f"gauge transform per request, seed {a.seed}"),
        "config": {"workload": "landau_serve_stream", "requests": n, "rate_per_s": a.rate,
                   "sizes": mix, "betas": list(SERVE_BETAS), "accuracy": 1e-14,
                   "switch_theta": a.switch_theta, "slots": a.slots},
        "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"], "runs": runs,
    }
    emit(line)


if __name__ == "__main__":
    main()
