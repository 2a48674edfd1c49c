"""Pins of the oracle's iteration loops: FACG and FASD (Gfix/FACG.c:181-446)
and the Coulomb slice loop (Gfix/CFACG.c:286-430), beyond the helpers that
tests/test_oracle_pins.py pins one by one.

1. Closed forms.  On a pure gauge of small amplitude over a constant
   abelian background, U_mu(x) = G(x) B_mu G(x + mu)^dag with
   G(x) = exp(i Phi(x) T), B_mu = diag(e^{i c_mu}, e^{-i c_mu}, 1) and
   Phi(x) = A cos(p.x), the gauge functional is, to second order in A, the
   quadratic form -(1/2) sum_x sum_mu w_mu (Phi(x + mu) - Phi(x))^2 (times
   tr T^2) with w_mu = cos c_mu (T = lambda_3, or any generator when c = 0).
   Then, with p^_mu^2 = 2 (1 - cos p_mu) and p^2 = sum_mu p^_mu^2:
     - the derivative is Delta(x) = -A (sum_mu w_mu p^_mu^2) cos(p.x) T, so
       theta_0 = (2/3) A^2 (sum_mu w_mu p^_mu^2)^2 (GLU's local trace
       2 tr Delta^2, GFNORM_LANDAU = 1/(NC V), <cos^2> = 1/2);
     - the Fourier acceleration multiplies the mode by MAX V / p^2 = 4 ND / p^2
       (MAX_LANDAU = 4 ND / V, the unnormalised transforms' factor V), so a
       step alpha multiplies Phi by (1 - 16 alpha W), W = sum_mu w_mu
       p^_mu^2 / p^2: FASD's theta falls by (1 - 16 alpha W)^2 per step, the
       three-probe parabola of the SD start is exact with its vertex at
       alpha = 1 / (16 W), and after that step theta is O(A^2) theta_0;
     - two modes of equal p^2 and different W make the accelerated problem
       a two-eigenvalue quadratic, on which GLU's CG (Polak-Ribiere on the
       accelerated field, exact line searches) is linear CG: theta falls
       to O(A^2) theta_0 after two steps, where steepest descent cannot.
   Coulomb, per time slice, with the three spatial directions:
   MAX_COULOMB = 4 (ND - 1) / Lcu gives the factor (1 - 12 alpha W) and the
   vertex 1 / (12 W); theta is measured on the accelerated field
   (CFACG.c:315), 2 q(z) / (NC Lcu) summed with GLU's packed metric q, so
   theta_0 = 48 A^2 W^2 for an off-diagonal generator and 144 A^2 W^2 for
   lambda_3 (q counts the packed diagonal re^2 + im^2 + re im, i.e. 3x).

2. Reconstruction.  On a generic heatbath input the first 12 FACG
   iterations are rebuilt from the pinned helpers (derivative, Fourier
   acceleration, line search, gauge update) and an independent inner
   product -- the Frobenius product of the 3x3 matrices the packed fields
   unpack to (diagonal Im z0, Re z0, -Im z0 - Re z0, off-diagonal z1, z2,
   z3; that is the matrix whose Frobenius norm GLU's sum_PR3 / sum_DER3,
   Gfix/FACG.c:99-178, compute) -- and must give the oracle's recorded
   theta_k, alpha_k, beta_k and final links.

3. Planted mistakes.  Each slip of glu_oracle.h's OR_PLANT_* list, planted
   in the oracle on the CPU, makes at least one pin above fail.

Parity unpinned: GLU's stale reduction slots at a continuation restart
(Gfix/FACG.c:249-265 reads the previous segment's red[] values); no closed
form fixes them (OR_PLANT_NO_STALE passes every pin here).
"""
import numpy as np
import pytest

from oracle import oracle as O

GEN = {
    "l3": np.diag([1.0, -1.0, 0.0]).astype(complex),
    "l1": np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], complex),
    "l4": np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], complex),
    "l6": np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0]], complex),
}
CG_TOL = 5e-11


def exp_i(T, phi):
    """exp(i phi T) for Hermitian T, one matrix per entry of phi"""
    w, v = np.linalg.eigh(T)
    return np.einsum("ij,nj,kj->nik", v, np.exp(1j * np.outer(phi, w)), v.conj())


def coords(dims):
    V = int(np.prod(dims))
    x = np.zeros((V, 4), int)
    r = np.arange(V)
    for mu in range(4):
        x[:, mu] = r % dims[mu]
        r //= dims[mu]
    return x


def phat2(dims, n):
    p = 2 * np.pi * np.array(n, float) / np.array(dims, float)
    return 2 * (1 - np.cos(p))


def pure_gauge(dims, gen, amp, moms, bg=(0.0, 0.0, 0.0, 0.0)):
    x = coords(dims)
    V = len(x)
    phi = np.zeros(V)
    for n in moms:
        phi += amp * np.cos(x @ (2 * np.pi * np.array(n) / np.array(dims)))
    G = exp_i(GEN[gen], phi)
    S = [1, dims[0], dims[0] * dims[1], dims[0] * dims[1] * dims[2]]
    U = np.zeros((V, 4, 18))
    for mu in range(4):
        B = np.diag([np.exp(1j * bg[mu]), np.exp(-1j * bg[mu]), 1.0])
        f = np.where(x[:, mu] < dims[mu] - 1, np.arange(V) + S[mu], np.arange(V) - (dims[mu] - 1) * S[mu])
        M = (G @ B @ np.conj(np.transpose(G[f], (0, 2, 1)))).reshape(V, 9)
        U[:, mu, 0::2] = M.real
        U[:, mu, 1::2] = M.imag
    return U


def weight(dims, n, bg, D=4):
    ph = phat2(dims, n)[:D]
    w = np.cos(np.array(bg, float))[:D]
    return (w * ph).sum() / ph.sum(), (w * ph).sum()


AMP = 1e-4


# ----------------------------------------------------------------- Landau --
def landau_theta0(dims, gen, n, bg=(0, 0, 0, 0)):
    U = pure_gauge(dims, gen, AMP, [n], bg)
    r = O.Lattice(dims, nth=2).facg(U, 0.08, 1e-30, 1, trace=4)
    _, s = weight(dims, n, bg)
    return r["trace_theta"][0] / ((2.0 / 3.0) * AMP**2 * s**2)


def landau_fasd_rate(dims, gen, n, alpha, bg=(0, 0, 0, 0)):
    U = pure_gauge(dims, gen, AMP, [n], bg)
    r = O.Lattice(dims, nth=2).facg(U, alpha, 1e-30, 5, trace=8, sd=True)
    th = r["trace_theta"]
    W, _ = weight(dims, n, bg)
    return th[1:5] / th[:4] / (1 - 16 * alpha * W) ** 2


def landau_sd_vertex(dims, gen, n, bg=(0, 0, 0, 0)):
    U = pure_gauge(dims, gen, AMP, [n], bg)
    r = O.Lattice(dims, nth=2).facg(U, 0.08, 1e-30, 2, trace=8)
    W, _ = weight(dims, n, bg)
    return r["trace_alpha"][0] * 16 * W, r["trace_theta"][1] / r["trace_theta"][0]


def landau_two_mode_cg(dims, bg):
    U = pure_gauge(dims, "l3", 1e-3, [(1, 0, 0, 0), (0, 1, 0, 0)], bg)
    r = O.Lattice(dims, nth=2).facg(U, 0.08, 1e-30, 3, trace=8)
    th = r["trace_theta"]
    return th[1] / th[0], th[2] / th[0]


@pytest.mark.parametrize("gen", sorted(GEN))
@pytest.mark.parametrize("n", [(1, 0, 0, 0), (1, 2, 0, 1)])
def test_landau_theta0_closed_form(gen, n):
    assert abs(landau_theta0((6, 6, 6, 8), gen, n) - 1) < 1e-6


@pytest.mark.parametrize("gen", ["l3", "l1", "l6"])
@pytest.mark.parametrize("alpha", [0.08, 0.05])
def test_landau_fasd_rate(gen, alpha):
    q = landau_fasd_rate((6, 6, 6, 8), gen, (1, 2, 0, 1), alpha)
    assert np.abs(q - 1).max() < 1e-5, q


@pytest.mark.parametrize("gen", ["l3", "l4"])
def test_landau_sd_start_vertex(gen):
    v, drop = landau_sd_vertex((6, 6, 6, 8), gen, (1, 0, 0, 1))
    assert abs(v - 1) < 1e-5 and drop < 1e-10


def test_landau_background_weights():
    """w_mu = cos c_mu weights the directions: the accelerated step no
    longer cancels the mode exactly, by the factor W."""
    dims, bg = (6, 6, 6, 8), (0.0, 1.0, 0.0, 0.5)
    assert abs(landau_theta0(dims, "l3", (0, 1, 0, 1), bg) - 1) < 1e-6
    q = landau_fasd_rate(dims, "l3", (0, 1, 0, 1), 0.08, bg)
    assert np.abs(q - 1).max() < 1e-5
    v, _ = landau_sd_vertex(dims, "l3", (0, 1, 0, 0), bg)
    assert abs(v - 1) < 1e-5


def test_landau_cg_terminates_on_two_eigenvalues():
    one, two = landau_two_mode_cg((6, 6, 6, 8), (0.0, 1.0, 0.0, 0.0))
    assert one > 1e-2  # one steepest-descent step leaves the slower mode
    assert two < 1e-8  # the conjugate step removes it


# ---------------------------------------------------------------- Coulomb --
def coulomb_run(dims, gen, n, bg, facg, gf=0.09, amp=AMP, moms=None):
    U = pure_gauge(dims, gen, amp, moms or [n], bg)
    return O.Lattice(dims, nth=2).coulomb(U, gf, 1e-22, 60, facg=facg, trace_slice=1, trace=64)


@pytest.mark.parametrize("gen,factor", [("l1", 48.0), ("l6", 48.0), ("l3", 144.0)])
def test_coulomb_theta0_on_the_accelerated_field(gen, factor):
    r = coulomb_run((6, 6, 6, 4), gen, (1, 2, 1, 0), (0, 0, 0, 0), True)
    assert abs(r["trace_theta"][0] / (factor * AMP**2) - 1) < 1e-6


@pytest.mark.parametrize("gen", ["l1", "l3"])
def test_coulomb_fasd_rate_and_vertex(gen):
    dims, bg, n = (6, 6, 6, 4), (0.0, 0.7, 0.0, 0.0), (1, 1, 0, 0)
    W, _ = weight(dims, n, bg, D=3)
    if gen != "l3":
        bg = (0, 0, 0, 0)
        W = 1.0
    r = coulomb_run(dims, gen, n, bg, False, gf=0.05)
    th = r["trace_theta"]
    assert np.abs(th[1:4] / th[:3] / (1 - 12 * 0.05 * W) ** 2 - 1).max() < 1e-5
    r = coulomb_run(dims, gen, n, bg, True)
    assert abs(r["trace_alpha"][0] * 12 * W - 1) < 1e-5
    assert r["trace_theta"][1] / r["trace_theta"][0] < 1e-10


def test_coulomb_cg_terminates_on_two_eigenvalues():
    r = coulomb_run((6, 6, 6, 4), "l3", None, (0.0, 1.0, 0.0, 0.0), True, amp=1e-3,
                    moms=[(1, 0, 0, 0), (0, 1, 0, 0)])
    th = r["trace_theta"]
    assert th[1] / th[0] > 1e-2 and th[2] / th[0] < 1e-8


# --------------------------------------------------------- reconstruction --
def unpack(F):
    """3x3 matrices of GLU's packed-field metric, one per site"""
    V = F.shape[1]
    M = np.zeros((V, 3, 3), complex)
    M[:, 0, 0] = F[0].imag
    M[:, 1, 1] = F[0].real
    M[:, 2, 2] = -F[0].imag - F[0].real
    for (a, b), f in (((0, 1), 1), ((0, 2), 2), ((1, 2), 3)):
        M[:, a, b] = F[f]
        M[:, b, a] = np.conj(F[f])
    return M


def frob(F, G):
    return float(np.sum(unpack(F) * np.conj(unpack(G))).real)


def landau_reconstruction(K=12):
    """Rebuild K FACG iterations from the helpers; returns the largest
    relative mismatch of theta, alpha, beta and of the final links."""
    dims = (4, 4, 4, 4)
    L = O.Lattice(dims, nth=2)
    U0 = L.heatbath(6.0, 10, 2, seed=3)
    L.random_gtrans(U0, seed=53, nth=2)
    Wo = U0.copy()
    r = L.facg(Wo, 0.08, 1e-14, K, trace=K + 2)
    W = U0.copy()
    red = np.zeros(3 + 8 * L.nth)
    th, F = L.fa_deriv(W)
    F = L.fourier_accelerate(F)
    a = O.lib().or_line_search_landau(L.ref, O._dp(W), O._dp(F), O._dp(red))
    rows = [(th, a, 0.0)]
    sn, old, inold = F.copy(), F.copy(), frob(F, F)
    for _ in range(K):
        th, F = L.fa_deriv(W)
        F = L.fourier_accelerate(F)
        insum, sc = frob(F, F), frob(F, F - old)
        beta = max(0.0, sc / inold)
        if th < CG_TOL:
            beta = insum / inold
        inold = insum
        sn = F + beta * sn
        old = F.copy()
        if th > CG_TOL:
            red[:] = 0
            a = O.lib().or_line_search_landau(L.ref, O._dp(W), O._dp(np.ascontiguousarray(sn)), O._dp(red))
        else:
            L.egauge(W, np.ascontiguousarray(sn), 0.08)
            a = 0.08
        rows.append((th, a, beta))
    rows = np.array(rows)
    n = min(len(rows), len(r["trace_theta"]))
    got = np.stack([r["trace_theta"][:n], r["trace_alpha"][:n], r["trace_beta"][:n]], 1)
    err = np.abs(got - rows[:n]) / np.maximum(np.abs(rows[:n]), 1e-300)
    err[rows[:n] == 0] = np.abs(got - rows[:n])[rows[:n] == 0]
    return err.max(axis=0), np.abs(W - Wo).max(), n


def test_landau_loop_reconstruction():
    err, dlinks, n = landau_reconstruction()
    assert n == 13
    assert err[0] < 1e-13 and err[1] < 1e-9 and err[2] < 1e-9, err
    assert dlinks < 1e-12


# ------------------------------------------------------- planted mistakes --
def failing_pins():
    """Run a compact set of the pins above; names of those that fail."""
    bad = []
    dims = (4, 4, 4, 6)
    try:
        if abs(landau_theta0(dims, "l1", (1, 0, 1, 1)) - 1) > 1e-5:
            bad.append("theta0")
        if np.abs(landau_fasd_rate(dims, "l3", (1, 0, 1, 1), 0.08) - 1).max() > 1e-4:
            bad.append("fasd_rate")
        v, drop = landau_sd_vertex(dims, "l3", (1, 0, 0, 1))
        if abs(v - 1) > 1e-4 or drop > 1e-8:
            bad.append("sd_vertex")
        one, two = landau_two_mode_cg((4, 4, 4, 6), (0.0, 1.0, 0.0, 0.0))
        if not (one > 1e-2 and two < 1e-7):
            bad.append("two_mode_cg")
        r = coulomb_run(dims, "l1", (1, 1, 0, 0), (0, 0, 0, 0), True)
        th = r["trace_theta"]
        if (abs(th[0] / (48 * AMP**2) - 1) > 1e-5 or abs(r["trace_alpha"][0] * 12 - 1) > 1e-4
                or th[1] / th[0] > 1e-8):
            bad.append("coulomb")
        err, dlinks, n = landau_reconstruction()
        if n != 13 or err.max() > 1e-9 or dlinks > 1e-12:
            bad.append("reconstruction")
    except (FloatingPointError, ValueError, IndexError, ZeroDivisionError):
        bad.append("crashed")
    return bad


@pytest.mark.parametrize("plant", ["FR_ALWAYS", "MAX_FACTOR", "THETA_AFTER", "EGAUGE_SIGN", "NO_CLAMP",
                                   "NO_BETA", "OLD_KEPT"])
def test_each_planted_mistake_fails_a_pin(plant):
    assert failing_pins() == []  # the faithful oracle passes them all
    with O.planted(getattr(O.planted, plant)):
        bad = failing_pins()
    print(f"[plant {plant}] fails: {bad}")
    assert bad, plant
