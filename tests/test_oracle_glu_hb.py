"""Pins of the oracle's restatement of GLU's own quenched update
(Update/KPHB.c, hb.c, relax.c, SU2_rotate_SSE.c, Smear/staples.c,
Geometry/draughtboard.c) on GLU's per-thread MWC_4096 streams.

No output of GLU's heatbath exists in the reference or can be produced
here (GLU cannot be built in this image), so the bits are not pinned; what
the algorithm fixes is:

1. the draughtboard is a partition of the lattice into two colours on even
   extents with every link joining two colours (so a colour's links can be
   updated independently, in any order);
2. the staple sum: sum over links of Re tr(U_mu(x) staple^dag) counts each
   plaquette four times, 72 V <P> (the plaquette is pinned separately);
3. overrelaxation is microcanonical: a pass leaves the Wilson action, i.e.
   the average plaquette, unchanged to rounding, and the links on SU(3);
4. the Kennedy-Pendleton step samples a0 from the SU(2) heatbath density
   sqrt(1 - a0^2) exp(a0 / norm) on [-1, 1] (Kolmogorov-Smirnov bound);
5. the full update keeps SU(3), thermalises a cold start to the known
   beta = 6.0 plaquette, and depends on SEED and the stream count only.

Each of a wrong staple, a wrong reflection and a KP trial without its
cos^2 factor is planted in the oracle and caught by one of these pins.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _config(dims, sweeps=3, seed=7, nth=2):
    L = O.Lattice(dims, nth=nth)
    U = L.glu_heatbath(6.0, sweeps, 1, seed=seed)
    return L, U


def _mats(U):
    M = U.reshape(-1, 3, 3, 2)
    return M[..., 0] + 1j * M[..., 1]


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (6, 4, 2, 8)])
def test_draughtboard_separates_neighbours(dims):
    L = O.Lattice(dims, nth=1)
    col = L.glu_colours()
    assert set(col.tolist()) == {0, 1}
    assert np.count_nonzero(col == 0) == L.V // 2
    for i in range(L.V):
        x = np.unravel_index(i, dims[::-1])[::-1]
        for mu in range(4):
            y = list(x)
            y[mu] = (y[mu] + 1) % dims[mu]
            j = np.ravel_multi_index(tuple(y[::-1]), dims[::-1])
            assert col[i] != col[j]


def test_odd_extent_uses_three_colours():
    L = O.Lattice((3, 4, 4, 4), nth=1)
    assert set(L.glu_colours().tolist()) == {0, 1, 2}


def _staple_total(L, U):
    tot = 0.0
    for i in range(L.V):
        for mu in range(4):
            S = L.glu_staple(U, i, mu)
            s = S[0::2] + 1j * S[1::2]
            u = U[i, mu, 0::2] + 1j * U[i, mu, 1::2]
            tot += np.real(np.trace(u.reshape(3, 3) @ s.reshape(3, 3).conj().T))
    return tot


def test_staples_count_each_plaquette_four_times():
    dims = (4, 4, 4, 4)
    L, U = _config(dims)
    L.random_gtrans(U, seed=3, nth=2)  # generic links, same plaquette
    assert abs(_staple_total(L, U) - 72 * L.V * L.plaquette(U)) < 1e-10 * L.V
    with O.planted(O.planted.STAPLE_NODAG):
        bad = _staple_total(L, U)
    assert abs(bad - 72 * L.V * L.plaquette(U)) > 1e-3 * L.V


def test_overrelaxation_keeps_the_action_and_su3():
    dims = (6, 4, 4, 4)
    L, U = _config(dims)
    p0 = L.plaquette(U)
    W = L.glu_overrelax(U.copy())
    assert abs(L.plaquette(W) - p0) < 1e-13
    assert np.abs(W - U).max() > 1e-2  # it moved the links
    M = _mats(W)
    assert np.abs(M @ M.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-13
    with O.planted(O.planted.OR_SIGN):
        Wb = L.glu_overrelax(U.copy())
    assert abs(L.plaquette(Wb) - p0) > 1e-6


def _ks(a0, norm):
    """Kolmogorov-Smirnov distance of the samples from the SU(2) heatbath
    density sqrt(1 - a^2) exp(a / norm) on [-1, 1]."""
    x = np.linspace(-1.0, 1.0, 400001)
    p = np.sqrt(np.clip(1 - x * x, 0, None)) * np.exp((x - 1.0) / norm)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(x))])
    cdf /= cdf[-1]
    s = np.sort(a0)
    F = np.interp(s, x, cdf)
    n = len(s)
    return max(np.max(np.arange(1, n + 1) / n - F), np.max(F - np.arange(n) / n))


@pytest.mark.parametrize("norm", [0.5, 0.125])
def test_kennedy_pendleton_samples_the_su2_density(norm):
    n = 100000
    a0 = O.glu_kp_samples(11, norm, n)
    assert np.all(np.abs(a0) <= 1.0)
    d = _ks(a0, norm)
    assert d < 1.95 / np.sqrt(n), d  # 99.9% KS bound
    with O.planted(O.planted.KP_NOCOS):
        bad = O.glu_kp_samples(11, norm, n)
    assert _ks(bad, norm) > 5 / np.sqrt(n)


def test_heatbath_thermalises_on_su3():
    dims = (8, 8, 8, 8)
    L = O.Lattice(dims, nth=4)
    U = L.glu_heatbath(6.0, 40, 4, seed=2)
    M = _mats(U)
    assert np.abs(M @ M.conj().transpose(0, 2, 1) - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(M) - 1).max() < 1e-12
    # <P> at beta = 6.0 is 0.5937 in large volumes; 8^4 sits within 0.005,
    # and 40 sweeps from cold are past the approach to equilibrium
    assert 0.585 < L.plaquette(U) < 0.600, L.plaquette(U)


def test_streams_decide_the_result():
    dims = (4, 4, 4, 4)
    _, a = _config(dims, seed=7, nth=2)
    _, b = _config(dims, seed=7, nth=2)
    _, c = _config(dims, seed=8, nth=2)
    _, d = _config(dims, seed=7, nth=3)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert not np.array_equal(a, d)
