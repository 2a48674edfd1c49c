"""CPU checks of bench.py's contract: one GPU only, and the algorithmic
byte counts it prices the roofline with add up."""
import importlib.util
import os
import subprocess
import sys

from conftest import ROOT


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""  # no bench line


def test_iteration_bytes_are_the_sum_of_the_steps():
    """One FACG iteration: the fused update and next derivative, the line
    search, the CG update, the Polak-Ribiere sums and the Fourier
    acceleration, which is credited 256 / 128 B/site (a forward and a
    backward transform of the four fields) whatever number of passes it
    takes."""
    b = load_bench()
    fa = (256, 128)
    for p in (0, 1):
        steps = [b.ALG_BYTES[k][p] for k in ("egauge_deriv", "line_search", "cg_update", "pr_sums")]
        assert sum(steps) + fa[p] == b.ALG_ITER[p]
    # the fused kernel is credited with exactly the two steps it performs
    assert b.ALG_BYTES["egauge_deriv"] == tuple(x + y for x, y in zip(b.ALG_BYTES["egauge"], b.ALG_BYTES["deriv"]))
    # FP32 streaming moves half the bytes of FP64
    for k, (b64, b32) in b.ALG_BYTES.items():
        assert b64 == 2 * b32, k


def test_stdout_carries_only_the_json_line():
    """Progress lines written to file descriptor 1 from C (GLU's [GF]/[CG]
    lines of the library and the oracle) and from Python go to stderr once
    the bench has routed its output; stdout keeps the one JSON line."""
    code = (
        "import importlib.util, os, sys, ctypes\n"
        f"spec = importlib.util.spec_from_file_location('bench', {os.path.join(ROOT, 'bench.py')!r})\n"
        "b = importlib.util.module_from_spec(spec); spec.loader.exec_module(b)\n"
        "b.route_stdout()\n"
        "libc = ctypes.CDLL(None)\n"
        "libc.printf(b'[CG] random gauge transform 0 :: 1e-09 (650)\\n'); libc.fflush(None)\n"
        "print('[bench] progress', flush=True)\n"
        "b.emit({'metric': 'm', 'value': 1.0})\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == '{"metric": "m", "value": 1.0}\n'
    assert "[CG] random gauge transform" in r.stderr and "[bench] progress" in r.stderr
