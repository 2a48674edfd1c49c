/*
 * or_heatbath.c -- oracle-side input generator: quenched Wilson-action
 * heatbath + overrelaxation (Cabibbo-Marinari over the three SU(2)
 * subgroups, Kennedy-Pendleton SU(2) heatbath, microcanonical reflections).
 * TEST INFRASTRUCTURE ONLY.  It produces thermalised-looking inputs for the
 * gauge-fixing tests; it is not a restatement of GLU's Update/ code and is
 * not bit-compatible with it.  Its random numbers come from a counter-based
 * splitmix64 keyed on (seed, sweep, direction, site, draw) so the result does
 * not depend on the thread count.
 */
#include "or_internal.h"

static inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

typedef struct {
  uint64_t s;
} hb_rng;

static inline double hb_u(hb_rng *r) /* uniform in (0, 1] */
{
  r->s += 0x9E3779B97F4A7C15ULL;
  return ((double)(mix64(r->s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

/* staple sum A with Re tr(U_mu(x) A) the local action: standard Wilson */
static void staple(const or_lat *L, const double *U, size_t i, int mu, cplx *A)
{
  for (int k = 0; k < 9; k++) A[k] = 0.0;
  const size_t xp = L->fwd[i * OR_ND + mu];
  for (int nu = 0; nu < OR_ND; nu++) {
    if (nu == mu) continue;
    const size_t xn = L->fwd[i * OR_ND + nu], xm = L->bwd[i * OR_ND + nu];
    const size_t xpm = L->bwd[xp * OR_ND + nu];
    const cplx *a = (const cplx *)LINK(U, xp, nu), *b = (const cplx *)LINK(U, xn, mu),
               *c = (const cplx *)LINK(U, i, nu);
    /* up: U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
    cplx t[9];
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        t[r * 3 + s] = a[r * 3] * conj(b[s * 3]) + a[r * 3 + 1] * conj(b[s * 3 + 1]) +
                       a[r * 3 + 2] * conj(b[s * 3 + 2]);
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        A[r * 3 + s] += t[r * 3] * conj(c[s * 3]) + t[r * 3 + 1] * conj(c[s * 3 + 1]) +
                        t[r * 3 + 2] * conj(c[s * 3 + 2]);
    /* down: U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) */
    const cplx *d = (const cplx *)LINK(U, xpm, nu), *e = (const cplx *)LINK(U, xm, mu),
               *f = (const cplx *)LINK(U, xm, nu);
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        t[r * 3 + s] = conj(d[0 * 3 + r]) * conj(e[s * 3 + 0]) +
                       conj(d[1 * 3 + r]) * conj(e[s * 3 + 1]) +
                       conj(d[2 * 3 + r]) * conj(e[s * 3 + 2]);
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++)
        A[r * 3 + s] += t[r * 3] * f[s] + t[r * 3 + 1] * f[3 + s] + t[r * 3 + 2] * f[6 + s];
  }
}

static const int SUB[3][2] = {{0, 1}, {0, 2}, {1, 2}};

/* one Cabibbo-Marinari update of link u with staple A */
static void cm_update(cplx *u, const cplx *A, double beta, int heat, hb_rng *r)
{
  for (int sg = 0; sg < 3; sg++) {
    const int p = SUB[sg][0], q = SUB[sg][1];
    /* w = (u A) restricted to the (p, q) block */
    cplx w[4];
    const int idx[2] = {p, q};
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        const int ra = idx[a], cb = idx[b];
        w[a * 2 + b] = u[ra * 3] * A[cb] + u[ra * 3 + 1] * A[3 + cb] + u[ra * 3 + 2] * A[6 + cb];
      }
    const double a0 = 0.5 * creal(w[0] + w[3]), a3 = 0.5 * cimag(w[0] - w[3]);
    const double a1 = 0.5 * cimag(w[1] + w[2]), a2 = 0.5 * creal(w[1] - w[2]);
    const double k = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
    if (k < 1e-14) continue;
    const double v0 = a0 / k, v1 = a1 / k, v2 = a2 / k, v3 = a3 / k;
    double x0, x1, x2, x3;
    if (heat) {
      const double al = 2.0 * beta * k / 3.0;
      double lam2;
      for (int tries = 0;; tries++) { /* Kennedy-Pendleton */
        const double r1 = hb_u(r), r2 = hb_u(r), r3 = hb_u(r), r4 = hb_u(r);
        const double c = cos(OR_TWOPI * r2);
        lam2 = -(log(r1) + c * c * log(r3)) / (2.0 * al);
        if (r4 * r4 <= 1.0 - lam2 || tries > 1000) break;
      }
      x0 = 1.0 - 2.0 * lam2;
      const double rad = sqrt(fmax(0.0, 1.0 - x0 * x0));
      const double cth = 2.0 * hb_u(r) - 1.0, phi = OR_TWOPI * hb_u(r);
      const double sth = sqrt(fmax(0.0, 1.0 - cth * cth));
      x1 = rad * sth * cos(phi);
      x2 = rad * sth * sin(phi);
      x3 = rad * cth;
    } else { /* reflection: X = V^dag V^dag V, i.e. R = V^dag V^dag */
      x0 = v0; x1 = -v1; x2 = -v2; x3 = -v3;
    }
    /* R = X V^dag as quaternions: (x0 + i x.s)(v0 - i v.s) */
    const double R0 = x0 * v0 + x1 * v1 + x2 * v2 + x3 * v3;
    const double R1 = -x0 * v1 + x1 * v0 + (x2 * v3 - x3 * v2);
    const double R2 = -x0 * v2 + x2 * v0 + (x3 * v1 - x1 * v3);
    const double R3 = -x0 * v3 + x3 * v0 + (x1 * v2 - x2 * v1);
    /* matrix [[R0 + i R3, R2 + i R1], [-R2 + i R1, R0 - i R3]] */
    const cplx m00 = R0 + I * R3, m01 = R2 + I * R1, m10 = -R2 + I * R1,
               m11 = R0 - I * R3;
    for (int s = 0; s < 3; s++) {
      const cplx up = u[p * 3 + s], uq = u[q * 3 + s];
      u[p * 3 + s] = m00 * up + m01 * uq;
      u[q * 3 + s] = m10 * up + m11 * uq;
    }
  }
}

void or_heatbath(const or_lat *L, double *U, double beta, int sweeps, int nor,
                 uint64_t seed, int cold_start)
{
  if (cold_start) {
    memset(U, 0, L->V * OR_ND * 18 * sizeof(double));
    for (size_t l = 0; l < L->V * OR_ND; l++)
      U[l * 18 + 0] = U[l * 18 + 8] = U[l * 18 + 16] = 1.0;
  }
  for (int sw = 0; sw < sweeps; sw++) {
    for (int pass = 0; pass <= nor; pass++) {
      const int heat = pass == 0;
      for (int mu = 0; mu < OR_ND; mu++)
        for (int par = 0; par < 2; par++) {
#pragma omp parallel for schedule(static)
          for (size_t i = 0; i < L->V; i++) {
            size_t s = 0, ii = i;
            for (int d = 0; d < OR_ND; d++) { s += ii % L->dims[d]; ii /= L->dims[d]; }
            if ((int)(s & 1) != par) continue;
            hb_rng r;
            r.s = mix64(seed ^ mix64(((uint64_t)sw << 42) ^ ((uint64_t)pass << 38) ^
                                     ((uint64_t)mu << 36) ^ (uint64_t)i));
            cplx A[9];
            staple(L, U, i, mu, A);
            cm_update((cplx *)LINK(U, i, mu), A, beta, heat, &r);
          }
        }
    }
    or_latt_reunit(L, U);
  }
}
