"""Graph replay and the serving loop, through the C ABI.

Both change only how the iteration kernels are launched and interleaved,
never what they compute, so the bar is bitwise: a batch replayed from a
captured HIP graph, and a request served among others, give exactly the
links, theta and iteration count of the eager, stand-alone
glu_gfix_run on the same input."""
import numpy as np
import pytest

from gpu_common import heatbath_config

pytestmark = pytest.mark.gpu

gl = pytest.importorskip("glu_amd")
torch = pytest.importorskip("torch")


def alone(dims, U, gftype=gl.LANDAU, precision=gl.FP64, method=gl.CG, acc=1e-14, graph=False,
          landau=False):
    with gl.Gfix(dims, gftype, precision) as g:
        g.set_graph(graph)
        g.set_links(U)
        if landau and gftype == gl.LANDAU:
            it, th, _ = g.landau(method, 0.08, acc, 1000, restart_seed=1)
        else:
            it, th = g.run(method, 0.08, acc, 1000)
        return g.get_links(), it, th


@pytest.mark.parametrize("precision,method", [(gl.FP64, gl.CG), (gl.MIXED, gl.CG), (gl.FP64, gl.SD)])
def test_graph_replays_match_eager(precision, method):
    dims = (8, 8, 8, 16)
    U, _ = heatbath_config(dims, sweeps=30)
    acc = 1e-14 if method == gl.CG else 1e-10
    W0, it0, th0 = alone(dims, U, precision=precision, method=method, acc=acc)
    with gl.Gfix(dims, gl.LANDAU, precision) as g:
        g.set_links(U)
        g.snapshot()
        for rep in range(3):
            # graph run, then an eager run in between: each replay must
            # reproduce the eager result
            g.restore()
            g.set_graph(True)
            it, th = g.run(method, 0.08, acc, 1000)
            assert (it, th) == (it0, th0), rep
            assert np.array_equal(g.get_links(), W0), rep
            g.restore()
            g.set_graph(False)
            it, th = g.run(method, 0.08, acc, 1000)
            assert (it, th) == (it0, th0)
            assert np.array_equal(g.get_links(), W0)


REQUESTS = [  # dims, gftype, precision, method, input seed
    ((8, 8, 8, 8), gl.LANDAU, gl.MIXED, gl.CG, 1),
    ((8, 8, 8, 16), gl.LANDAU, gl.FP64, gl.CG, 2),
    ((12, 12, 12, 12), gl.LANDAU, gl.MIXED, gl.CG, 3),
    ((8, 8, 8, 8), gl.LANDAU, gl.FP64, gl.SD, 4),
    ((6, 6, 6, 12), gl.COULOMB, gl.FP64, gl.CG, 5),
    ((8, 8, 8, 16), gl.LANDAU, gl.MIXED, gl.CG, 6),
    ((8, 8, 8, 8), gl.LANDAU, gl.MIXED, gl.CG, 7),
]


def inputs():
    out = []
    for dims, gftype, prec, meth, seed in REQUESTS:
        U, _ = heatbath_config(dims, sweeps=20, seed=seed, gseed=seed + 100)
        out.append(U)
    return out


def test_served_requests_match_standalone_runs():
    Us = inputs()
    ref = []
    for (dims, gftype, prec, meth, _), U in zip(REQUESTS, Us):
        acc = 1e-14 if meth == gl.CG else 1e-10
        ref.append(alone(dims, U, gftype, prec, meth, acc, landau=True))
    bufs = [torch.from_numpy(np.ascontiguousarray(U)).cuda() for U in Us]
    torch.cuda.synchronize()
    with gl.Server(max_active=4) as srv:
        tickets = []
        for (dims, gftype, prec, meth, _), b in zip(REQUESTS, bufs):
            acc = 1e-14 if meth == gl.CG else 1e-10
            tickets.append(srv.submit(dims, b.data_ptr(), gftype, prec, meth, 0.08, acc, 1000))
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(REQUESTS)
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    for k, ((W, it, th), (it2, th2, lat), b) in enumerate(zip(ref, res, bufs)):
        assert (it2, th2) == (it, th), k
        assert lat > 0
        assert np.array_equal(b.cpu().numpy(), W), k


def test_admission_respects_budget_and_rejects_oversize():
    dims_big, dims_small = (12, 12, 12, 12), (8, 8, 8, 8)
    fb = gl.footprint(dims_big, gl.LANDAU, gl.MIXED)
    fs = gl.footprint(dims_small, gl.LANDAU, gl.MIXED)
    budget = fb + fs // 2  # the big one runs alone; small ones two at a time
    Ub, _ = heatbath_config(dims_big, sweeps=20, seed=3, gseed=103)
    Us, _ = heatbath_config(dims_small, sweeps=20, seed=1, gseed=101)
    Wb = alone(dims_big, Ub, precision=gl.MIXED, landau=True)
    Ws = alone(dims_small, Us, precision=gl.MIXED, landau=True)
    order = [dims_small, dims_big, dims_small, dims_small, dims_big]
    bufs = [torch.from_numpy(np.ascontiguousarray(Ub if d == dims_big else Us)).cuda() for d in order]
    torch.cuda.synchronize()
    with gl.Server(hbm_budget=budget, max_active=8) as srv:
        with pytest.raises(gl.GfixError):
            srv.submit((16, 16, 16, 16), bufs[0].data_ptr())
        tickets = [srv.submit(d, b.data_ptr(), precision=gl.MIXED) for d, b in zip(order, bufs)]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["peak_bytes"] <= budget
    for d, b, (it, th, _) in zip(order, bufs, res):
        W, it0, th0 = Wb if d == dims_big else Ws
        assert (it, th) == (it0, th0)
        assert np.array_equal(b.cpu().numpy(), W)


def test_landau_restart_policy():
    """glu_gfix_landau against the policy of Gfix/Landau.c:301-345 replayed
    by hand: attempt 0 on the input, attempt k on the input after the k-th
    random_transform drawn from GLU's streams seeded SEED = restart_seed
    (256 streams, seeded once before the first restart); the first success
    ends it and the count is 123456789 + (k - 1) max_iters + its
    iterations."""
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims, sweeps=20)
    seed = 5
    found = False
    with gl.Gfix(dims, gl.LANDAU, gl.FP64) as g:
        g.set_links(U)
        g.snapshot()
        full, _ = g.run(gl.CG, 0.08, 1e-14, 1000)
        for M in range(full // 3, full, max(1, full // 12)):
            g.restore()
            it0, _ = g.run(gl.CG, 0.08, 1e-14, M)
            if it0 != 123456789:
                continue
            outcome = None  # (k, loc, theta) of the first successful restart
            g.rng_seed(seed, 256)
            for k in range(1, 7):
                g.restore()
                g.glu_random_transform()
                loc, th = g.run(gl.CG, 0.08, 1e-14, M)
                if loc != 123456789:
                    outcome = (k, loc, th, g.get_links())
                    break
            g.restore()
            if outcome is None:
                with pytest.raises(gl.GfixError):
                    g.landau(gl.CG, 0.08, 1e-14, M, restart_seed=seed)
                continue
            k, loc, th, W = outcome
            it, th2, nf = g.landau(gl.CG, 0.08, 1e-14, M, restart_seed=seed)
            assert nf == k and it == 123456789 + (k - 1) * M + loc and th2 == th
            assert np.array_equal(g.get_links(), W)
            found = True
            break
    assert found, "no max_iters gave a failure followed by a successful restart"


def test_server_reports_failure_and_keeps_serving():
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims, sweeps=20)
    W, it0, th0 = alone(dims, U, precision=gl.MIXED, landau=True)
    bad = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    good = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    torch.cuda.synchronize()
    with gl.Server(max_active=2) as srv:
        t_bad = srv.submit(dims, bad.data_ptr(), precision=gl.MIXED, max_iters=5)
        t_good = srv.submit(dims, good.data_ptr(), precision=gl.MIXED)
        with pytest.raises(gl.GfixError):
            srv.wait(t_bad)
        it, th, _ = srv.wait(t_good)
    assert (it, th) == (it0, th0)
    assert np.array_equal(good.cpu().numpy(), W)


def test_luxury_copies_pick_the_smallest_functional():
    """glu_gfix_luxury against luxury_copy_fast (Gfix/Landau.c:88-213,
    BEST_COPY) replayed by hand: copy k is the input under the k-th
    random_transform drawn from GLU's streams seeded SEED = seed (256
    streams), fixed to 1e-6 in at most 1000 iterations; the smallest
    functional among converged copies is then fixed to the accuracy."""
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims, sweeps=20)
    seed, copies = 11, 4
    with gl.Gfix(dims, gl.LANDAU, gl.FP64) as g:
        g.set_links(U)
        g.snapshot()
        funcs, its, fixed = [], [], []
        g.rng_seed(seed, 256)
        for k in range(copies):
            g.restore()
            g.glu_random_transform()
            it, _ = g.run(gl.CG, 0.08, 1e-6, 1000)
            funcs.append(g.measure()["functional"])
            its.append(it)
            fixed.append(g.get_links())
        ok = [k for k in range(copies) if its[k] not in (1000, 123456789)]
        best = min(ok, key=lambda k: funcs[k])
        g.set_links(fixed[best])
        fin, th_ref = g.run(gl.CG, 0.08, 1e-14, 1000)
        W_ref = g.get_links()
        g.restore()
        it, th, chosen = g.luxury(gl.CG, 0.08, 1e-14, 1000, copies=copies, seed=seed)
        assert chosen == best
        assert it == its[-1] + fin and th == th_ref
        assert np.array_equal(g.get_links(), W_ref)
        assert len(set(np.round(funcs, 12))) > 1  # the copies differ


def test_glu_input_keys_request():
    """A request given as GLU input-file keys: RANDOM_TRANSFORM = YES with
    SEED = 7 transforms the input with the streams seeded 7, then fixes it
    to 10^-ACCURACY -- bitwise the same as doing those steps alone."""
    dims = (8, 8, 8, 16)
    U, _ = heatbath_config(dims, sweeps=20, seed=8, gseed=108)
    opts = "MODE = GAUGE_FIXING\nGFTYPE = LANDAU\nACCURACY = 14\nMAX_ITERS = 1000\nGF_TUNE = 0.08\n" \
           "SEED = 7\nRANDOM_TRANSFORM = YES\n"
    with gl.Gfix(dims, gl.LANDAU, gl.MIXED) as g:
        g.set_links(U)
        g.rng_seed(7, 256)
        g.glu_random_transform()
        it0, th0, _ = g.landau(gl.CG, 0.08, 1e-14, 1000, restart_seed=0)
        W0 = g.get_links()
    b = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    torch.cuda.synchronize()
    with gl.Server(max_active=2) as srv:
        it, th, _ = srv.wait(srv.submit_glu(dims, b.data_ptr(), opts, gl.MIXED))
        with pytest.raises(gl.GfixError, match="GF_TUNE"):
            srv.submit_glu(dims, b.data_ptr(), "GFTYPE = LANDAU\nACCURACY = 14\nMAX_ITERS = 10\n")
        with pytest.raises(gl.GfixError, match="MAG"):
            srv.submit_glu(dims, b.data_ptr(), "GFTYPE = MAG\nACCURACY = 14\nMAX_ITERS = 10\nGF_TUNE = 0.1\n")
    assert (it, th) == (it0, th0) and th < 1e-14
    assert np.array_equal(b.cpu().numpy(), W0)


def test_coulomb_request_does_not_hold_up_landau_runs():
    """A Coulomb request runs on a host thread of its own: Landau requests
    submitted after it finish while it is still running, with the results
    they have alone."""
    dc, dl = (12, 12, 12, 16), (8, 8, 8, 8)
    Uc, _ = heatbath_config(dc, sweeps=20, seed=21, gseed=121)
    Ul, _ = heatbath_config(dl, sweeps=20, seed=22, gseed=122)
    with gl.Gfix(dc, gl.COULOMB, gl.FP64) as g:
        g.set_links(Uc)
        it, th = g.run(gl.CG, 0.09, 1e-14, 650)
        Wc = (g.get_links(), it, th)
    Wl = alone(dl, Ul, gl.LANDAU, gl.FP64, landau=True)
    bc = torch.from_numpy(np.ascontiguousarray(Uc)).cuda()
    bl = [torch.from_numpy(np.ascontiguousarray(Ul)).cuda() for _ in range(3)]
    torch.cuda.synchronize()
    with gl.Server(max_active=4) as srv:
        tc = srv.submit(dc, bc.data_ptr(), gl.COULOMB, gl.FP64, gl.CG, 0.09, 1e-14, 650)
        tl = [srv.submit(dl, b.data_ptr(), gl.LANDAU, gl.FP64, gl.CG, 0.08, 1e-14, 1000) for b in bl]
        rl = [srv.wait(t) for t in tl]
        rc = srv.wait(tc)
    print(f"[serve] coulomb {rc[2]:.0f} ms; landau {[round(r[2]) for r in rl]} ms", flush=True)
    assert max(r[2] for r in rl) < rc[2]  # the short runs overtook the long Coulomb run
    assert (rc[0], rc[1]) == (Wc[1], Wc[2]) and np.array_equal(bc.cpu().numpy(), Wc[0])
    for r, b in zip(rl, bl):
        assert (r[0], r[1]) == (Wl[1], Wl[2]) and np.array_equal(b.cpu().numpy(), Wl[0])


def batch_alone(dims, U, max_iters=1000):
    """glu_gfix_facg_batch on this lattice alone: (links, iterations, theta)."""
    b = torch.from_numpy(np.ascontiguousarray(U)).cuda()
    torch.cuda.synchronize()
    its, th = gl.facg_batch(dims, b.data_ptr(), 1, 0.08, 1e-14, max_iters)
    return b.cpu().numpy(), its[0], float(th[0])


BATCH_REQUESTS = [  # dims, precision, input seed: the FP64 8^4 and 6^4 requests share launches
    ((8, 8, 8, 8), gl.FP64, 11),
    ((8, 8, 8, 32), gl.FP64, 12),   # an extent above 16: a context
    ((8, 8, 8, 8), gl.FP64, 13),
    ((6, 6, 6, 6), gl.FP64, 14),
    ((8, 8, 8, 8), gl.MIXED, 15),   # mixed precision: a context
    ((8, 8, 8, 8), gl.FP64, 16),
    ((6, 6, 6, 6), gl.FP64, 17),
    ((8, 8, 8, 8), gl.FP64, 18),
]


def test_small_lattices_share_launches():
    """With batching on, the FP64 Landau requests of 8^4 and 6^4 lattices
    are fixed in shared launches: each result equals, bitwise, the one
    glu_gfix_facg_batch gives for that lattice alone, while the others
    are served on contexts and equal their stand-alone runs."""
    Us = [heatbath_config(d, sweeps=20, seed=s, gseed=s + 100)[0] for d, _, s in BATCH_REQUESTS]
    ref = []
    for (dims, prec, _), U in zip(BATCH_REQUESTS, Us):
        small = prec == gl.FP64 and max(dims) <= 16
        ref.append(batch_alone(dims, U) if small else alone(dims, U, precision=prec, landau=True))
    bufs = [torch.from_numpy(np.ascontiguousarray(U)).cuda() for U in Us]
    torch.cuda.synchronize()
    with gl.Server(max_active=3, batching=True) as srv:
        tickets = [srv.submit(d, b.data_ptr(), gl.LANDAU, p) for (d, p, _), b in zip(BATCH_REQUESTS, bufs)]
        res = [srv.wait(t) for t in tickets]
        st = srv.stats()
    assert st["completed"] == len(BATCH_REQUESTS)
    for k, ((W, it, th), (it2, th2, lat), b) in enumerate(zip(ref, res, bufs)):
        assert th < 1e-14 and (it2, th2) == (it, th), (k, it, it2, th, th2)
        assert np.array_equal(b.cpu().numpy(), W), k


def test_batched_request_restarts_on_a_context():
    """A small-lattice request whose FACG fails in the shared launch keeps
    its input and takes Landau()'s restarts on a context: the iteration
    count, theta and links of the policy replayed by hand (attempt k on the
    input under the k-th random_transform from GLU's streams seeded
    SEED = restart_seed)."""
    dims = (8, 8, 8, 8)
    U, _ = heatbath_config(dims, sweeps=20)
    seed = 5
    _, full, _ = batch_alone(dims, U)
    found = False
    with gl.Gfix(dims, gl.LANDAU, gl.FP64) as g:
        g.set_links(U)
        g.snapshot()
        for M in range(full // 3, full, max(1, full // 12)):
            _, it0, _ = batch_alone(dims, U, M)
            if it0 != 123456789:
                continue
            outcome = None
            g.rng_seed(seed, 256)
            for k in range(1, 7):
                g.restore()
                g.glu_random_transform()
                loc, th = g.run(gl.CG, 0.08, 1e-14, M)
                if loc != 123456789:
                    outcome = (k, loc, th, g.get_links())
                    break
            if outcome is None:
                continue
            k, loc, th, W = outcome
            b = torch.from_numpy(np.ascontiguousarray(U)).cuda()
            torch.cuda.synchronize()
            with gl.Server(max_active=2, batching=True, restart_seed=seed) as srv:
                it, th2, _ = srv.wait(srv.submit(dims, b.data_ptr(), gl.LANDAU, gl.FP64, max_iters=M))
            assert it == 123456789 + (k - 1) * M + loc and th2 == th
            assert np.array_equal(b.cpu().numpy(), W)
            found = True
            break
    assert found, "no max_iters gave a failure in the shared launch followed by a successful restart"
