"""Pins of the oracle's gluon propagators (GLU's GLUON_PROPS cuts).

The momentum lists are checked against brute-force enumerations written
from the cut definitions (a sphere, a hypercube, the distance from the
nearest body diagonal, and that distance with an angle), the propagators
against numpy: the traceless Hermitian part (U - U^dag)/2i - tr/3, numpy's
inverse FFT times V for the unnormalised exp(+ip.x) transform, the
exp(i p_mu / 2) Landau phase and the transverse / longitudinal projectors
with p_mu = 2 sin(pi k_mu / L_mu), evaluated at every momentum of the list
(the oracle, like GLU, fills the second half by p -> -p).  A plane wave
pins the normalisation in closed form."""
import numpy as np
import pytest

from oracle import oracle as O


def haar(rng, n):
    z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    return q / (np.linalg.det(q) ** (1 / 3))[:, None, None]


def to_glu(M):
    """complex [V, 4, 3, 3] -> GLU's double [V, 4, 18]"""
    out = np.empty(M.shape[:2] + (18,))
    f = M.reshape(M.shape[:2] + (9,))
    out[..., 0::2], out[..., 1::2] = f.real, f.imag
    return np.ascontiguousarray(out)


def signed_momenta(dims):
    """GLU's enumeration: k_mu = c_mu - L_mu / 2, x fastest"""
    V = int(np.prod(dims))
    i = np.arange(V)
    k = np.empty((V, 4), np.int64)
    for mu in range(4):
        k[:, mu] = (i // int(np.prod(dims[:mu]))) % dims[mu] - dims[mu] // 2
    return k


def brute_list(dims, cut, max_mom, width=1.0, angle=20.0):
    k = signed_momenta(dims)
    small = min(dims)
    rats = small / np.array(dims, float)
    psq = ((rats * k) ** 2).sum(1)
    if cut == "hypercubic":
        keep = (np.abs(rats * k) <= np.sqrt(max_mom)).all(1)
    elif cut == "psq":
        keep = psq <= max_mom
    else:
        q = 2 * np.sin(np.pi * k / np.array(dims))
        diags = np.array(np.meshgrid(*[[-1, 1]] * 4, indexing="ij")).reshape(4, -1).T / 2.0  # unit
        proj = q @ diags.T                                        # q . n
        dist = np.sqrt(np.maximum((q ** 2).sum(1)[:, None] - proj ** 2, 0))
        w = 2 * np.pi * width / small
        if cut == "cylinder":
            ok = (dist <= w + 1e-12).any(1)
        else:
            qn = np.sqrt((q ** 2).sum(1))
            with np.errstate(invalid="ignore", divide="ignore"):
                th = np.arccos(np.clip(proj / qn[:, None], -1, 1))
            ok = ((dist <= w + 1e-12) & (th <= np.radians(angle) + 1e-12)).any(1) | (qn < 1e-7)
        keep = (psq <= max_mom) & ok
    return k[keep]


CASES = [((4, 4, 4, 8), "psq", 3), ((4, 4, 4, 8), "hypercubic", 1), ((6, 6, 6, 12), "psq", 6),
         ((6, 6, 6, 12), "cylinder", 8), ((6, 6, 6, 12), "conical", 8), ((4, 6, 4, 8), "psq", 3)]


@pytest.mark.parametrize("dims,cut,max_mom", CASES)
def test_momentum_list_against_brute_force(dims, cut, max_mom):
    L = O.Lattice(dims, nth=1)
    mom, idx = O.momentum_list(L, cut, max_mom, 1.0, 20.0)
    assert np.array_equal(mom, brute_list(dims, cut, max_mom))
    # sites of the unnormalised transform: negative components wrap
    site = np.zeros(len(mom), np.int64)
    for mu in range(4):
        site += (mom[:, mu] % dims[mu]) * int(np.prod(dims[:mu]))
    assert np.array_equal(idx, site)
    # these cuts stay inside (-L/2, L/2): the list is +-symmetric about 0
    assert np.array_equal(mom[::-1], -mom) and not mom[len(mom) // 2].any()


def test_unknown_cut_is_refused():
    with pytest.raises(ValueError):
        O.momentum_list(O.Lattice((4, 4, 4, 4), nth=1), 7, 2)


def numpy_props(U, dims, mom):
    V = int(np.prod(dims))
    M = U[..., 0::2] + 1j * U[..., 1::2]
    M = M.reshape(V, 4, 3, 3)
    A = (M - np.conj(np.swapaxes(M, -1, -2))) / 2j
    A = A - np.trace(A, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    F = A.reshape(dims[::-1] + (4, 3, 3))
    F = np.fft.ifftn(F, axes=(0, 1, 2, 3)) * V          # sum_x A(x) exp(+ip.x)
    F = F.reshape(V, 4, 3, 3)
    L = np.array(dims)
    G, Fl = [], []
    for k in mom:
        site = int(sum((k[mu] % dims[mu]) * np.prod(dims[:mu]) for mu in range(4)))
        Ap = F[site] * np.exp(1j * np.pi * k / L)[:, None, None]
        tr = np.einsum("mij,nij->mn", Ap, np.conj(Ap)).real
        p = 2 * np.sin(np.pi * k / L)
        p2 = (p ** 2).sum()
        P = np.outer(p, p) / (p2 if p2 >= 3e-14 else 1.0)
        zero = not k.any()
        G.append(((np.eye(4) - P) * tr).sum() * 2 / (8 * (4 if zero else 3) * V))
        Fl.append((P * tr).sum() * 2 / (8 * V))
    return np.array(G), np.array(Fl)


@pytest.mark.parametrize("dims,cut,max_mom", [((4, 4, 4, 6), "psq", 3), ((6, 4, 4, 8), "cylinder", 3),
                                              ((6, 4, 4, 8), "hypercubic", 1), ((5, 6, 4, 8), "psq", 3)])
def test_gluon_prop_against_numpy(dims, cut, max_mom):
    rng = np.random.default_rng(3)
    V = int(np.prod(dims))
    U = to_glu(haar(rng, V * 4).reshape(V, 4, 3, 3))
    L = O.Lattice(dims, nth=2)
    mom, G, F = O.gluon_prop(L, U, cut, max_mom, 1.0, 20.0)
    # GLU fills the second half by p -> -p, which needs a symmetric list
    # (it warns otherwise): these cuts stay clear of the -L/2 edge
    assert np.array_equal(mom[::-1], -mom)
    Gn, Fn = numpy_props(U, dims, mom)
    assert np.abs(G - Gn).max() < 1e-12 * np.abs(Gn).max()
    assert np.abs(F - Fn).max() < 1e-12 * np.abs(Fn).max()
    assert F[len(F) // 2] == 0.0  # no longitudinal part at p = 0


def test_gluon_prop_plane_wave():
    """sin(theta_mu(x)) = eps_mu cos(q.x) on the T3 diagonal gives
    A_mu = eps_mu cos(q.x) diag(1, -1, 0): at +-q, A_mu(p) = eps_mu V / 2
    diag(1, -1, 0) and the propagators follow from the projectors."""
    dims = (4, 4, 4, 8)
    V = int(np.prod(dims))
    k = signed_momenta(dims)  # enumeration order is also x fastest
    c = np.stack(np.unravel_index(np.arange(V), dims[::-1])[::-1], 1)  # lattice coordinates
    q = np.array([1, 0, 1, 2])
    eps = np.array([0.2, -0.1, 0.05, 0.3])
    ph = 2 * np.pi * (c * q / np.array(dims)).sum(1)
    th = np.arcsin(eps[None, :] * np.cos(ph)[:, None])  # [V, 4]
    M = np.zeros((V, 4, 3, 3), complex)
    M[..., 0, 0], M[..., 1, 1], M[..., 2, 2] = np.exp(1j * th), np.exp(-1j * th), 1
    L = O.Lattice(dims, nth=1)
    mom, G, F = O.gluon_prop(L, to_glu(M), "psq", 3)
    p = 2 * np.sin(np.pi * q / np.array(dims))
    P = np.outer(p, p) / (p ** 2).sum()
    ph2 = np.pi * q / np.array(dims)
    tr = np.outer(eps, eps) * V ** 2 / 2 * np.cos(ph2[:, None] - ph2[None, :])
    Gq, Fq = ((np.eye(4) - P) * tr).sum() * 2 / (24 * V), (P * tr).sum() * 2 / (8 * V)
    for i, m in enumerate(mom):
        if np.array_equal(m, q) or np.array_equal(m, -q):
            assert abs(G[i] - Gq) < 1e-12 * abs(Gq) and abs(F[i] - Fq) < 1e-12 * abs(Fq)
        else:
            assert abs(G[i]) < 1e-12 and abs(F[i]) < 1e-12
    assert sum(np.array_equal(m, q) or np.array_equal(m, -q) for m in mom) == 2
    del k


def test_write_glueprop_format(tmp_path):
    mom = np.array([[-1, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]], np.int32)
    G, F = np.array([1.5, 2.0, 1.5]), np.array([0.25, 0.0, 0.25])
    p = str(tmp_path / "g.bin")
    O.write_glueprop(p, mom, G, F)
    b = open(p, "rb").read()
    n = np.frombuffer(b[:4], ">u4")[0]
    assert n == 3
    ints = np.frombuffer(b[4:4 + 4 * 5 * 3], ">i4").reshape(3, 5)
    assert (ints[:, 0] == 4).all() and np.array_equal(ints[:, 1:], mom)
    o = 4 + 60
    assert np.frombuffer(b[o:o + 4], ">u4")[0] == 3
    assert np.array_equal(np.frombuffer(b[o + 4:o + 28], ">f8"), G)
    o += 28
    assert np.frombuffer(b[o:o + 4], ">u4")[0] == 3
    assert np.array_equal(np.frombuffer(b[o + 4:o + 28], ">f8"), F)
    assert len(b) == o + 28


# ------------------------------------------ exceptional two/three point ---
def test_exceptional_g2_is_the_propagator_trace():
    """With A(-p) = A(p)^dag, sum_mu tr A_mu(p) A_mu(-p) is the trace the
    projectors split: g2 = G + F / (ND - 1) at p != 0 and g2 = G at p = 0."""
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(7)
    V = int(np.prod(dims))
    U = to_glu(haar(rng, V * 4).reshape(V, 4, 3, 3))
    L = O.Lattice(dims, nth=2)
    m1, G, F = O.gluon_prop(L, U, "psq", 3)
    m2, g2, g3 = O.gluon_exceptional(L, U, "psq", 3)
    assert np.array_equal(m1, m2)
    z = len(g2) // 2
    want = G + F / 3.0
    want[z] = G[z]
    assert np.abs(g2 - want).max() < 1e-12 * np.abs(want).max()
    assert g3[z] == 0.0  # p = 0: the rho sum vanishes


def test_exceptional_is_colour_rotation_invariant():
    """U -> W U W^dag with one constant W in SU(3) rotates every A_mu(p) the
    same way; g2 and g3 are traces of products and do not change."""
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(8)
    V = int(np.prod(dims))
    M = haar(rng, V * 4).reshape(V, 4, 3, 3)
    W = haar(rng, 1)[0]
    R = W @ M @ W.conj().T
    L = O.Lattice(dims, nth=2)
    _, a2, a3 = O.gluon_exceptional(L, to_glu(M), "psq", 3)
    _, b2, b3 = O.gluon_exceptional(L, to_glu(R), "psq", 3)
    assert np.abs(a2 - b2).max() < 1e-12 * np.abs(a2).max()
    assert np.abs(a3 - b3).max() < 1e-12 * np.abs(a3).max()
    assert np.abs(a3).max() > 1e-6  # not trivially zero


def test_exceptional_g3_against_numpy():
    """The field pipeline (lists, transform, Landau phase, sites of -p and
    0) by numpy, and the contraction by GLU's TRACE_PROJECTION form
    (Cuts/MOMgg.c:67-91): the explicit triple sum
    sum_{mu nu rho} (delta_mu nu - p_mu p_nu / p^2) p_rho
    tr(A_mu(p) A_nu(-p) A_rho(0)) / p^2.  The oracle follows the default
    contraction_matrices (MOMgg.c:93-125), which groups the same sum into
    matrix products; the two agree by the cyclicity of the trace, so a
    reversed product, a lost sign on the -p/p^2 term or A(p) and A(-p)
    exchanged would show here."""
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(9)
    V = int(np.prod(dims))
    U = to_glu(haar(rng, V * 4).reshape(V, 4, 3, 3))
    mom, g2, g3 = O.gluon_exceptional(O.Lattice(dims, nth=2), U, "psq", 3)
    M = (U[..., 0::2] + 1j * U[..., 1::2]).reshape(V, 4, 3, 3)
    A = (M - np.conj(np.swapaxes(M, -1, -2))) / 2j
    A = A - np.trace(A, axis1=-2, axis2=-1)[..., None, None] * np.eye(3) / 3
    Ft = (np.fft.ifftn(A.reshape(dims[::-1] + (4, 3, 3)), axes=(0, 1, 2, 3)) * V).reshape(V, 4, 3, 3)
    Lv = np.array(dims)

    def at(k):
        site = int(sum((k[mu] % dims[mu]) * np.prod(dims[:mu]) for mu in range(4)))
        return Ft[site] * np.exp(1j * np.pi * k / Lv)[:, None, None]

    A0 = at(np.zeros(4, int))
    for i, k in enumerate(mom):
        Ap, Am = at(k), at(-k)
        p = 2 * np.sin(np.pi * k / Lv)
        p2 = (p ** 2).sum()
        p2 = p2 if p2 >= 3e-14 else 1.0
        T = np.einsum("aij,bjk,cki->abc", Ap, Am, A0)  # tr(A_mu(p) A_nu(-p) A_rho(0))
        proj = np.eye(4) - np.outer(p, p) / p2
        want = (np.einsum("ab,c,abc->", proj, p, T) / p2).real * 4 / (2 * 3 * 8 * 3 * V)
        assert abs(g3[i] - want) < 1e-12 * max(np.abs(g3).max(), 1e-300)
    assert np.abs(g3).max() > 1e-6
