"""GPU parity of the Landau path (HIP, through the C ABI) against the CPU
oracle on the same seeded inputs.

Tolerances (FP64 path):
  stage outputs (derivative, Fourier acceleration, probes, updated links)
      1e-13 relative to the field scale -- they differ only by rounding;
  trajectory theta_k / alpha_k / beta_k, first 20 iterations
      1e-12 relative, or 20x the oracle's own spread between two
      summation orders (1 vs 8 virtual threads) where that is larger,
      taking at iteration k the largest relative spread of iterations
      0..k: the parabola fit of the line search cancels O(12 V) probe sums,
      so its rounding floor grows as theta falls, for the reference as
      well, and rounding that entered the trajectory stays in it;
  final state: theta < accuracy, functional 1e-10 relative,
      plaquette preserved to 1e-14.
"""
import numpy as np
import pytest

from oracle import oracle as O

from gpu_common import check_trajectory, heatbath_config, rel

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")


def fresh(dims, U, precision=gl.FP64):
    g = gl.Gfix(dims, gl.LANDAU, precision)
    g.set_links(U)
    return g


DIMS = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def cfg():
    U, plaq = heatbath_config(DIMS)
    return U, plaq


def test_links_roundtrip(cfg):
    U, _ = cfg
    with fresh(DIMS, U) as g:
        back = g.get_links()
    # column 2 is rebuilt from columns 0 and 1 on the device
    assert np.abs(back - U).max() < 1e-14


def test_measure_matches_oracle(cfg):
    U, plaq = cfg
    L = O.Lattice(DIMS, nth=4)
    with fresh(DIMS, U) as g:
        m = g.measure()
        th = g.theta()
    assert abs(m["plaquette"] - L.plaquette(U)) < 1e-14
    tr, mx = L.links(U)
    assert abs(m["link_trace"] - tr) < 1e-14
    assert abs(m["link_max"] - mx / 3) < 1e-14
    assert abs(m["functional"] - L.functional(U)) < 1e-14
    assert rel(th, L.theta_lin(U)) < 1e-13


def test_stage_derivative(cfg):
    U, _ = cfg
    L = O.Lattice(DIMS, nth=4)
    th_o, F_o = L.fa_deriv(U)
    with fresh(DIMS, U) as g:
        th_g, F_g = g.stage_deriv()
    scale = np.abs(F_o).max()
    assert np.abs(F_g - F_o).max() < 1e-13 * scale
    assert rel(th_g, th_o) < 1e-13


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (6, 4, 12, 8), (4, 6, 8, 24), (16, 8, 4, 6), (5, 7, 10, 14),
                                  (32, 4, 2, 64), (96, 2, 48, 2), (16, 16, 4, 6), (24, 24, 2, 4),
                                  (32, 32, 2, 3), (48, 48, 2, 2),
                                  # the z-t column pass after the x-y plane pass
                                  (16, 16, 16, 16), (16, 16, 16, 32), (24, 24, 24, 48), (32, 32, 32, 64)])
def test_stage_fourier_accelerate(dims):
    rng = np.random.default_rng(5)
    V = int(np.prod(dims))
    F = rng.standard_normal((4, V)) + 1j * rng.standard_normal((4, V))
    L = O.Lattice(dims, nth=2)
    ref = L.fourier_accelerate(F)
    U = np.zeros((V, 4, 18))
    U[..., [0, 8, 16]] = 1.0
    with fresh(dims, U) as g:
        out = g.stage_fa(F)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("variant", [1, 2])
@pytest.mark.parametrize("dims", [(16, 16, 16, 32), (32, 32, 32, 64)])
def test_stage_fourier_accelerate_zt_variants(dims, variant, monkeypatch):
    """The z-t column pass with the other x-values-per-workgroup variants
    the autotuner may pick."""
    monkeypatch.setenv("GLU_GFIX_ZT_PV", str(variant))
    rng = np.random.default_rng(6)
    V = int(np.prod(dims))
    F = rng.standard_normal((4, V)) + 1j * rng.standard_normal((4, V))
    ref = O.Lattice(dims, nth=2).fourier_accelerate(F)
    U = np.zeros((V, 4, 18))
    U[..., [0, 8, 16]] = 1.0
    with fresh(dims, U) as g:
        out = g.stage_fa(F)
    assert np.abs(out - ref).max() < 1e-13 * np.abs(ref).max()


def test_stage_line_search_and_egauge(cfg):
    U, _ = cfg
    L = O.Lattice(DIMS, nth=4)
    _, F = L.fa_deriv(U)
    F = L.fourier_accelerate(F)
    red = np.zeros(3 + 8 * L.nth)
    W = U.copy()
    a_o = O.lib().or_line_search_landau(L.ref, O._dp(W), O._dp(F), O._dp(red))
    probes_o = np.array([sum(red[j + 8 * t] for t in range(L.nth)) for j in range(3)])
    with fresh(DIMS, U) as g:
        pr, a_g = g.stage_line_search(F)
        g.stage_egauge(F, a_g)
        W_g = g.get_links()
    assert np.all(rel(pr, probes_o) < 1e-13)
    assert rel(a_g, a_o) < 1e-10
    assert np.abs(W_g - W).max() < 1e-12


def oracle_run(U, dims, nth, **kw):
    L = O.Lattice(dims, nth=nth)
    W = U.copy()
    r = L.facg(W, trace=4000, **kw)
    return L, W, r


@pytest.mark.parametrize("dims", [(8, 8, 8, 8), (8, 8, 8, 16), (6, 10, 6, 12)])
def test_facg_trajectory_and_final(dims):
    U, plaq = heatbath_config(dims)
    L, W1, r1 = oracle_run(U, dims, 1)
    _, W8, r8 = oracle_run(U, dims, 8)
    with fresh(dims, U) as g:
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        tr = g.trace()
        raw = g.measure()
        g.reunit()  # as GLU's Landau() does before its report (Gfix/Landau.c:48)
        m = g.measure()
    # SD start + 20 iterations
    check_trajectory(f"facg {dims}", tr, r1, r8, n=21)
    assert th < 1e-14
    assert abs(it - r1["iters"]) <= max(2, r1["iters"] // 20)
    assert rel(m["functional"], L.functional(W1)) < 1e-10
    assert abs(m["plaquette"] - plaq) < 1e-14
    # before reunitarisation the drift is rounding of ~70 link updates, the
    # same size as the oracle's own
    drift_o = abs(L.plaquette(W1) - plaq)
    assert abs(raw["plaquette"] - plaq) < max(1e-14, 3 * drift_o)


def test_fasd_matches_oracle():
    dims = (6, 6, 6, 6)
    U, plaq = heatbath_config(dims, sweeps=20)
    L, W, r = oracle_run(U, dims, 4, gf_alpha=0.08, acc=1e-10, max_iters=3000, sd=True)
    with fresh(dims, U) as g:
        it, th = g.run(gl.SD, 0.08, 1e-10, 3000)
        tr = g.trace()
        m = g.measure()
    assert it == r["iters"]
    assert np.all(rel(tr[:20, 0], r["trace_theta"][:20]) < 1e-12)
    assert rel(m["functional"], L.functional(W)) < 1e-10


def test_continuation_and_failure():
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims)
    # pick max_iters so that the first segment ends with theta < 1e3 acc:
    # GLU restarts from an SD step with its stale reduction slots
    _, _, rfull = oracle_run(U, dims, 4)
    th = rfull["trace_theta"][1:]
    k = int(np.argmax(th < 1e-12)) + 1
    L, W, r = oracle_run(U, dims, 4, max_iters=k)
    with fresh(dims, U) as g:
        it, thg = g.run(gl.CG, 0.08, 1e-14, k)
        tr = g.trace()
        Wg = g.get_links()
    assert it == r["iters"] and thg < 1e-14
    assert len(tr) == len(r["trace_theta"])
    assert np.abs(Wg - W).max() < 1e-9
    # unreachable accuracy with few iterations: GLU's sentinel
    with fresh(dims, U) as g:
        it, thg = g.run(gl.CG, 0.08, 1e-14, 5)
    assert it == 123456789


def test_deterministic_bitwise(cfg):
    U, _ = cfg
    outs = []
    for _ in range(2):
        with fresh(DIMS, U) as g:
            g.run(gl.CG, 0.08, 1e-14, 1000)
            outs.append(g.get_links())
    assert np.array_equal(outs[0], outs[1])


def test_mixed_precision():
    dims = (8, 8, 8, 16)
    U, plaq = heatbath_config(dims)
    L, W, r = oracle_run(U, dims, 4)
    with fresh(dims, U, gl.MIXED) as g:
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        n32, n64 = g.mixed_split()
        g.reunit()
        m = g.measure()
    assert th < 1e-14 and n32 > 0 and n64 > 0
    assert abs(m["plaquette"] - plaq) < 1e-14
    # same Gribov copy as the FP64 oracle: functional agrees closely
    assert rel(m["functional"], L.functional(W)) < 1e-10


@pytest.mark.parametrize("stall", ["0", "30"])
def test_mixed_stall_exit(stall, monkeypatch):
    """The FP32 phase may hand over to FP64 at its first stall (the default
    on lattices up to 2^17 sites) or run to the switch point: both end
    converged, on FP64-preserved links."""
    monkeypatch.setenv("GLU_GFIX_STALL_FACTOR", stall)
    dims = (8, 8, 8, 8)
    U, plaq = heatbath_config(dims, sweeps=20)
    with fresh(dims, U, gl.MIXED) as g:
        g.set_switch_theta(1e-11)
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        n32, n64 = g.mixed_split()
        g.reunit()
        m = g.measure()
    assert th < 1e-14 and n32 > 0 and n64 > 0
    assert abs(m["plaquette"] - plaq) < 1e-14


@pytest.mark.parametrize("graph", ["0", "1"])
def test_batch_length_is_bitwise_neutral(graph, monkeypatch):
    """Iterations enqueued past a stop run as no-ops that keep the links and
    the CG fields (in, sn, in_old) in their buffers, so a solve does not
    depend on how many iterations the host enqueues between checks, with
    or without graph replay.  The mixed solve carries its CG state over
    the FP32 -> FP64 switch when the two thetas there agree, so a misplaced
    in_old would change its trajectory."""
    dims = (8, 8, 8, 16)
    U, _ = heatbath_config(dims, sweeps=10)
    monkeypatch.setenv("GLU_GFIX_GRAPH", graph)
    monkeypatch.setenv("GLU_GFIX_STALL_FACTOR", "0")  # hand over at the switch point, CG state kept
    out = []
    for batch in ("1", "3", "8"):
        monkeypatch.setenv("GLU_GFIX_BATCH", batch)
        print(f"batch {batch} graph {graph}", flush=True)
        with fresh(dims, U, gl.MIXED) as g:
            it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
            out.append((it, th, g.mixed_split(), g.trace(), g.get_links()))
    for o in out[1:]:
        assert out[0][:3] == o[:3]
        assert np.array_equal(out[0][3], o[3])
        assert np.array_equal(out[0][4], o[4])


def test_unit_and_pure_gauge():
    dims = (4, 4, 4, 8)
    V = int(np.prod(dims))
    U = np.zeros((V, 4, 18))
    U[..., [0, 8, 16]] = 1.0
    with fresh(dims, U) as g:
        it, th = g.run(gl.CG, 0.08, 1e-14, 100)
    assert it == 0 and th == 0.0  # FACG's early exit
    L = O.Lattice(dims, nth=2)
    L.random_gtrans(U, seed=3)
    with fresh(dims, U) as g:
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        m = g.measure()
    assert th < 1e-14 and abs(m["functional"]) < 1e-12


@pytest.mark.parametrize("precision", [gl.FP64, gl.MIXED])
def test_xcd_block_order_is_bitwise_neutral(precision, monkeypatch):
    """The workgroup orders (plain, XCD-contiguous, XCD time slabs) change
    which workgroup computes a block, not what it computes or where its
    partial sum goes: whole runs agree bitwise."""
    dims = (8, 8, 8, 16)
    U, _ = heatbath_config(dims, sweeps=10)
    out = []
    for xcd in ("0", "1", "2"):
        monkeypatch.setenv("GLU_GFIX_XCD", xcd)
        with fresh(dims, U, precision) as g:
            it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
            out.append((it, th, g.get_links()))
    for o in out[1:]:
        assert out[0][:2] == o[:2]
        assert np.array_equal(out[0][2], o[2])


def test_autotune_changes_speed_only():
    """The tuned launch shapes compute the same transforms (to rounding)
    and the tuned context still solves; a second context of the same shape
    reuses the cached choices."""
    dims = (8, 8, 16, 32)
    U, plaq = heatbath_config(dims, sweeps=10)
    with fresh(dims, U) as g:
        _, F = g.stage_deriv()
        before = g.stage_fa(F)
        rep = g.autotune()
        assert "fft axis" in rep and "stencil order" in rep
        after = g.stage_fa(F)
        assert np.abs(after - before).max() < 1e-14 * np.abs(before).max()
        it, th = g.run(gl.CG, 0.08, 1e-14, 1000)
        assert th < 1e-14 and it < 1000
        assert abs(g.measure()["plaquette"] - plaq) < 1e-14
    with fresh(dims, U) as g:
        assert g.autotune().startswith("(cached)")


def test_autotune_times_only_z_t_variants_that_exist():
    """At Lx = 8 the z-t variant with 16 x values per workgroup has no
    kernel; the tuner leaves it out instead of timing an empty launch (and
    then sending every Fourier acceleration to single-axis passes)."""
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims, sweeps=10)
    with fresh(dims, U, gl.MIXED) as g:
        _, F = g.stage_deriv()
        before = g.stage_fa(F)
        rep = g.autotune()
        assert "z-t column pass variant 0" in rep and "z-t column pass variant 2" not in rep, rep
        after = g.stage_fa(F)
        assert np.abs(after - before).max() < 1e-14 * np.abs(before).max()
