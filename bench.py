"""Benchmark: Landau gauge fixing (FACG) of 32^3x64 configurations to
theta < 1e-14 on one MI355X, mixed precision (FP32 streaming until
theta < 1e-7, then FP64) against the plain FP64 path on the same input.

One step = restore the input links from their device copy + one complete
gauge fixing to theta < 1e-14 (both inside the timed region).  The input is
a device heatbath config (beta = 6.0, cold start, 4 overrelaxation passes
per sweep) with a seeded random gauge transform, generated once before the
timed region.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Landau configs/s to theta<1e-14 at 32^3x64 (mixed vs FP64); "
          "p50/p99 latency; GB/s")
PEAK_HBM_GBS = 8000.0
# algorithmic bytes per site per launch (FP64 / FP32 streaming) of each
# kernel: each logical step reads its inputs once and writes its outputs
# once; an FFT axis pass reads and writes the 4 packed fields
ALG_BYTES = {
    "deriv": (640, 320), "fft_axis": (128, 64), "pr_sums": (128, 64),
    "cg_update": (256, 128), "line_search": (640, 320), "egauge": (1216, 608),
}
# per Landau FACG iteration (FP64 / FP32): the sum over the logical steps
ALG_ITER = (3136, 1568)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dims", default="32,32,32,64")
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--switch-theta", type=float, default=1e-11,
                    help="mixed precision: FP32 until theta falls below this")
    return ap.parse_args()


def pct(v, q):
    v = sorted(v)
    k = (len(v) - 1) * q
    lo, hi = math.floor(k), math.ceil(k)
    return v[lo] + (v[hi] - v[lo]) * (k - lo)


def kernel_mask(name, prec):
    import glu_amd as gl
    return 1 << (gl.KERNELS.index(name) + (len(gl.KERNELS) if prec == "fp32" else 0))


def timed_runs(g, steps, warmup):
    """Warm-up runs (the first one times every kernel class, to find the
    dominant kernel), then the timed steps with HIP events around the
    dominant kernel's launches only."""
    prof_all = None
    for w in range(max(warmup, 1)):
        g.restore()
        if w == 0:
            g.profile(-1)
        g.run(0, 0.08, 1e-14, 1000)
        if w == 0:
            prof_all = g.profile_read()
            g.profile(0)
    dom = max(prof_all.items(), key=lambda kv: kv[1][2])[0]
    g.measure()  # synchronises the device
    lat, its = [], []
    g.profile(kernel_mask(*dom))
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        g.restore()
        it, th = g.run(0, 0.08, 1e-14, 1000)
        lat.append(time.perf_counter() - t)
        its.append(it)
        if not th < 1e-14:
            raise RuntimeError(f"did not converge: theta {th}")
    total = time.perf_counter() - t0
    live = g.profile_read()
    g.profile(0)
    return total, lat, its, prof_all, dom, live[dom]


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: every "
              "configuration fits one MI355X and the multi-GPU design (one config "
              "per GPU, no communication) has no figure to measure here.", file=sys.stderr)
        sys.exit(2)
    import glu_amd as gl

    dims = tuple(int(x) for x in a.dims.split(","))
    V = math.prod(dims)
    print(f"[bench] lattice {dims} V={V}: generating input "
          f"(device heatbath beta=6.0, {a.sweeps} sweeps)", file=sys.stderr, flush=True)
    gm = gl.Gfix(dims, gl.LANDAU, gl.MIXED)
    gm.set_switch_theta(a.switch_theta)
    t = time.perf_counter()
    gm.heatbath(6.0, a.sweeps, 4, seed=a.seed, cold_start=True)
    plaq = gm.measure()["plaquette"]
    gm.random_transform(seed=a.seed + 1)
    gm.snapshot()
    print(f"[bench] input ready in {time.perf_counter() - t:.1f}s, plaquette {plaq:.12f}",
          file=sys.stderr, flush=True)

    total, lat, its, prof, dom, dom_live = timed_runs(gm, a.steps, a.warmup)
    n32, n64 = gm.mixed_split()
    gm.restore()
    mixed_value = a.steps / total
    print(f"[bench] mixed: {mixed_value:.3f} configs/s, iters {its} (fp32 {n32} + fp64 {n64})",
          file=sys.stderr, flush=True)

    fp64 = None
    if not a.no_fp64:
        U = gm.get_links()
        g64 = gl.Gfix(dims, gl.LANDAU, gl.FP64)
        g64.set_links(U)
        g64.snapshot()
        t64, lat64, its64, prof64, _, _ = timed_runs(g64, a.steps, a.warmup)
        fp64 = {"configs_per_s": a.steps / t64, "ms_per_step": 1e3 * t64 / a.steps,
                "iters": its64, "p50_ms": 1e3 * pct(lat64, 0.5)}
        print(f"[bench] fp64: {fp64['configs_per_s']:.3f} configs/s, iters {its64}",
              file=sys.stderr, flush=True)
        prof.update({k: v for k, v in prof64.items() if k not in prof})

    # dominant kernel of the mixed run and its roofline, from the HIP events
    # recorded around its launches during the timed steps
    (kname, prec), (nl, mean_ms, tot_ms) = dom, dom_live
    alg = ALG_BYTES[kname][0 if prec == "fp64" else 1] * V
    achieved = alg / (mean_ms * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get(f"{kname}_{prec}_{'x'.join(map(str, dims))}")
    roof = {"kernel": f"{kname}<{prec}>", "bound": "hbm", "achieved": round(achieved, 1),
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": traffic, "launches": nl, "mean_ms": round(mean_ms, 4)}
    # per-kernel breakdown from the fully profiled warm-up run (one config)
    kernels = {f"{k}<{p}>": {"launches": n, "mean_ms": round(m, 4), "total_ms": round(tt, 2),
                             "GBps": round(ALG_BYTES[k][0 if p == "fp64" else 1] * V / (m * 1e-3) / 1e9, 1)}
               for (k, p), (n, m, tt) in sorted(prof.items())}

    # streaming rate of the whole mixed solve from the algorithmic bytes
    alg_total = (n32 * ALG_ITER[1] + n64 * ALG_ITER[0]) * V
    gbs = alg_total / (sum(lat) / len(lat)) / 1e9

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(gm, dims, fp64["iters"][0] if fp64 else its[0], a.cpu_iters)

    line = {
        "metric": METRIC, "value": round(mixed_value, 4), "unit": "configs/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(1e3 * total / a.steps, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": (f"synthetic: device quenched heatbath beta=6.0 ({a.sweeps} sweeps from cold, "
                 f"4 overrelaxation each, seed {a.seed}) + seeded random gauge transform"),
        "config": {"workload": "landau_facg_mixed_" + "x".join(map(str, dims)),
                   "lattice": list(dims), "gf_tune": 0.08, "accuracy": 1e-14, "max_iters": 1000,
                   "switch_theta": a.switch_theta},
        "p50_ms": round(1e3 * pct(lat, 0.5), 2), "p99_ms": round(1e3 * pct(lat, 0.99), 2),
        "iters": its, "iters_fp32": n32, "iters_fp64": n64, "stream_GBps": round(gbs, 1),
        "fp64": fp64,
        "mixed_speedup_vs_fp64": round(mixed_value / fp64["configs_per_s"], 3) if fp64 else None,
        "roofline": roof, "kernels": kernels, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(g, dims, iters_fp64, k):
    """The CPU oracle (C, OpenMP) on the same input: time k and 2k FACG
    iterations, take the per-iteration cost from the difference and scale
    it by the FP64 GPU iteration count of the full solve."""
    from oracle import oracle as O

    nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g.restore()
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


if __name__ == "__main__":
    main()
