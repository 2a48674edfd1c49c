/*
 * or_glu_hb.c -- oracle: GLU's own quenched update, the input generator of
 * GLU's MODE = HEATBATH (Update/KPHB.c:121-257): per sweep a heatbath over
 * the whole lattice (Update/hb.c:156-194) and NOR overrelaxation passes
 * (Update/relax.c:76-114), with GLU's draughtboard colouring
 * (Geometry/draughtboard.c:48-123), Wilson staples (Smear/staples.c:32-79),
 * the SU(2) subgroup algebra (Update/SU2_rotate_SSE.c:75-394) and the
 * Kennedy-Pendleton SU(2) heatbath (Update/hb.c:53-127), drawing from GLU's
 * per-thread MWC_4096 streams (or_rng.c) in libgomp's static blocks.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * Where C leaves GLU's arithmetic unspecified this restatement fixes it:
 * the two draws inside KP's log expression (hb.c:58) are taken left to
 * right, and the SSE complex sums of only_subgroup / su2_rotate are grouped
 * as the intrinsics nest them (a + (b + c)), without fused multiply-adds
 * (an x86-64 build without -mfma).  No GLU output is available here to
 * check the bits against; the restatement is pinned by the identities in
 * tests/test_oracle_glu_hb.py instead.
 */
#include "or_internal.h"

#define GLU_NHBMAX 42 /* Update/hb.c:44 */

/* ---------------------------------------------------------------------- */
/* draughtboard: Geometry/draughtboard.c:48-123                           */
/* ---------------------------------------------------------------------- */
int or_glu_colour(const or_lat *L, size_t i)
{
  size_t even = 0, odd = 0, sub = 1;
  for (int mu = 0; mu < OR_ND; mu++) {
    const size_t x = (i / sub) % L->dims[mu];
    sub *= L->dims[mu];
    if ((L->dims[mu] & 1) == 0)
      even += x;
    else
      odd += x;
  }
  return (int)((even % 2 + odd) % 3);
}

typedef struct {
  int ncol;
  size_t n[3];
  size_t *sq[3];
} glu_db;

static void db_init(const or_lat *L, glu_db *db)
{
  for (int c = 0; c < 3; c++) db->n[c] = 0;
  for (size_t i = 0; i < L->V; i++) db->n[or_glu_colour(L, i)]++;
  db->ncol = db->n[2] ? 3 : 2;
  for (int c = 0; c < 3; c++) {
    db->sq[c] = db->n[c] ? malloc(db->n[c] * sizeof(size_t)) : NULL;
    db->n[c] = 0;
  }
  for (size_t i = 0; i < L->V; i++) {
    const int c = or_glu_colour(L, i);
    db->sq[c][db->n[c]++] = i;
  }
}

static void db_free(glu_db *db)
{
  for (int c = 0; c < 3; c++) free(db->sq[c]);
}

/* ---------------------------------------------------------------------- */
/* staples: Smear/staples.c:63-76                                         */
/* ---------------------------------------------------------------------- */
/* r = a b c^dag */
static void mul_abcdag(cplx *r, const cplx *a, const cplx *b, const cplx *c)
{
  cplx t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = a[3 * i] * b[j] + a[3 * i + 1] * b[3 + j] + a[3 * i + 2] * b[6 + j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r[3 * i + j] = t[3 * i] * conj(c[3 * j]) + t[3 * i + 1] * conj(c[3 * j + 1]) +
                     t[3 * i + 2] * conj(c[3 * j + 2]);
}

/* r = a^dag b c */
static void mul_adagbc(cplx *r, const cplx *a, const cplx *b, const cplx *c)
{
  cplx t[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      t[3 * i + j] = conj(a[i]) * b[j] + conj(a[3 + i]) * b[3 + j] + conj(a[6 + i]) * b[6 + j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      r[3 * i + j] = t[3 * i] * c[j] + t[3 * i + 1] * c[3 + j] + t[3 * i + 2] * c[6 + j];
}

/* sum over nu != mu of the upper staple U_nu(x) U_mu(x+nu) U_nu(x+mu)^dag
   and the lower one U_nu(x-nu)^dag U_mu(x-nu) U_nu(x-nu+mu), so that
   Re tr(U_mu(x) stap^dag) is the sum of the plaquettes holding U_mu(x) */
void or_glu_staple(const or_lat *L, const double *U, size_t i, int mu, double *stap18)
{
  cplx *stap = (cplx *)stap18;
  for (int k = 0; k < 9; k++) stap[k] = 0.0;
  const size_t t2 = L->fwd[i * OR_ND + mu];
  for (int j = 0; j < OR_ND - 1; j++) {
    const int nu = j < mu ? j : j + 1;
    const size_t t1 = L->fwd[i * OR_ND + nu], b1 = L->bwd[i * OR_ND + nu];
    const size_t b2 = L->fwd[b1 * OR_ND + mu];
    cplx a[9];
    mul_abcdag(a, (const cplx *)LINK(U, i, nu), (const cplx *)LINK(U, t1, mu), (const cplx *)LINK(U, t2, nu));
    for (int k = 0; k < 9; k++) stap[k] += a[k];
    if (or_plant == OR_PLANT_STAPLE_NODAG)
      mul_abcdag(a, (const cplx *)LINK(U, b1, nu), (const cplx *)LINK(U, b1, mu), (const cplx *)LINK(U, b2, nu));
    else
      mul_adagbc(a, (const cplx *)LINK(U, b1, nu), (const cplx *)LINK(U, b1, mu), (const cplx *)LINK(U, b2, nu));
    for (int k = 0; k < 9; k++) stap[k] += a[k];
  }
}

/* ---------------------------------------------------------------------- */
/* SU(2) subgroups: Update/SU2_rotate_SSE.c                               */
/* ---------------------------------------------------------------------- */
/* rows (r0, r1) of the subgroup: 0 -> (0, 1), 1 -> (1, 2), 2 -> (2, 0) */
static const int SG_A[3] = {0, 1, 2}, SG_B[3] = {1, 2, 0};

/* s0, s1 of the subgroup of U staple^dag, normalised; scale = 1 / |(s0, s1)|
   (SU2_rotate_SSE.c:75-200): with rows a = SG_A, b = SG_B,
   s0 = sum_k U[a][k] conj(S[a][k]) + sum_k conj(U[b][k]) S[b][k],
   s1 = sum_k U[a][k] conj(S[b][k]) - sum_k conj(U[b][k]) S[a][k],
   each three-term sum nested as x0 + (x1 + x2) */
static void only_subgroup(cplx *s0, cplx *s1, double *scale, const cplx *U, const cplx *S, int sg)
{
  const int a = 3 * SG_A[sg], b = 3 * SG_B[sg];
#define SUM3(e0, e1, e2) ((e0) + ((e1) + (e2)))
  const cplx t0 = SUM3(U[a] * conj(S[a]), U[a + 1] * conj(S[a + 1]), U[a + 2] * conj(S[a + 2]));
  const cplx t3c = SUM3(conj(U[b]) * S[b], conj(U[b + 1]) * S[b + 1], conj(U[b + 2]) * S[b + 2]);
  const cplx t1 = SUM3(U[a] * conj(S[b]), U[a + 1] * conj(S[b + 1]), U[a + 2] * conj(S[b + 2]));
  const cplx t2c = SUM3(conj(U[b]) * S[a], conj(U[b + 1]) * S[a + 1], conj(U[b + 2]) * S[a + 2]);
#undef SUM3
  cplx m0 = t0 + t3c, m1 = t1 - t2c;
  /* (re0^2 + re1^2) + (im0^2 + im1^2), then its inverse root */
  const double z = 1.0 / sqrt((creal(m0) * creal(m0) + creal(m1) * creal(m1)) +
                              (cimag(m0) * cimag(m0) + cimag(m1) * cimag(m1)));
  *s0 = creal(m0) * z + I * (cimag(m0) * z);
  *s1 = creal(m1) * z + I * (cimag(m1) * z);
  *scale = z;
}

/* U <- s U with s = [[s0, s1], [-conj(s1), conj(s0)]] acting on the rows
   (a, b) of subgroup sg (SU2_rotate_SSE.c:287-394; GLU writes the third
   subgroup over rows (0, 2) with s0, s1 in the conjugated places, which is
   this form over rows (2, 0)) */
static void su2_rotate(cplx *U, cplx s0, cplx s1, int sg)
{
  const int a = 3 * SG_A[sg], b = 3 * SG_B[sg];
  for (int k = 0; k < 3; k++) {
    const cplx ua = U[a + k], ub = U[b + k];
    U[a + k] = s0 * ua + s1 * ub;
    U[b + k] = conj(s0) * ub - conj(s1) * ua;
  }
}

/* ---------------------------------------------------------------------- */
/* Kennedy-Pendleton SU(2) heatbath: Update/hb.c:53-127                   */
/* ---------------------------------------------------------------------- */
static int kp(double *d, double norm, or_rng *R, int th)
{
  const double x4 = or_rng_dbl(R, th);
  const double x3 = cos(OR_TWOPI * or_rng_dbl(R, th));
  const double l1 = log(or_rng_dbl(R, th));
  const double l2 = log(or_rng_dbl(R, th));
  *d = -(l1 + l2 * (or_plant == OR_PLANT_KP_NOCOS ? 1.0 : x3 * x3)) * norm;
  return (1. - 0.5 * (*d)) < x4 * x4;
}

/* d = 1 - a0 with GLU's retry cap: the last trial is kept after NHBMAX
   rejections following the first */
static double kp_d(double norm, or_rng *R, int th)
{
  double d = 0.0;
  size_t iters = 1;
  if (kp(&d, norm, R, th))
    while (kp(&d, norm, R, th) && iters < GLU_NHBMAX) iters++;
  return d;
}

static void generate_su2(cplx *s0, cplx *s1, double norm, or_rng *R, int th)
{
  const double d = kp_d(norm, R, th);
  const double a0 = 1.0 - d;
  const double ar2 = fabs(1.0 - a0 * a0);
  const double a3 = sqrt(ar2) * (2.0 * or_rng_dbl(R, th) - 1.0);
  const double rho = sqrt(fabs(ar2 - a3 * a3));
  const double xr2 = OR_TWOPI * or_rng_dbl(R, th);
  double a1 = cos(xr2), a2 = sin(xr2);
  a1 *= rho;
  a2 *= rho;
  const double rS0 = creal(*s0), iS0 = cimag(*s0), rS1 = creal(*s1), iS1 = cimag(*s1);
  *s0 = a0 * rS0 + a3 * iS0 + a2 * rS1 + a1 * iS1 + I * (-a0 * iS0 + rS0 * a3 - a2 * iS1 + rS1 * a1);
  *s1 = -a0 * rS1 + a3 * iS1 + a2 * rS0 - a1 * iS0 + I * (-a0 * iS1 - a3 * rS1 + a2 * iS0 + a1 * rS0);
}

/* hb() of Update/hb.c:131-152: the three subgroups in turn */
static void hb_link(cplx *U, const cplx *stap, double invbeta, or_rng *R, int th)
{
  for (int sg = 0; sg < 3; sg++) {
    cplx s0, s1;
    double scale;
    only_subgroup(&s0, &s1, &scale, U, stap, sg);
    generate_su2(&s0, &s1, invbeta * scale, R, th);
    su2_rotate(U, s0, s1, sg);
  }
}

/* microcanonical reflection (Update/relax.c:33-44) over the subgroups */
static void or_link(cplx *U, const cplx *stap)
{
  for (int sg = 0; sg < 3; sg++) {
    cplx s0, s1;
    double scale;
    only_subgroup(&s0, &s1, &scale, U, stap, sg);
    const double z = creal(s0) * creal(s0) - cimag(s0) * cimag(s0) - creal(s1) * creal(s1) -
                     cimag(s1) * cimag(s1);
    s1 = (or_plant == OR_PLANT_OR_SIGN ? 2. : -2.) * creal(s0) * s1;
    s0 = z - I * (2. * creal(s0) * cimag(s0));
    su2_rotate(U, s0, s1, sg);
  }
}

/* ---------------------------------------------------------------------- */
/* lattice sweeps                                                         */
/* ---------------------------------------------------------------------- */
/* one heatbath (heat = 1, Update/hb.c:156-194, each link reunitarised) or
   overrelaxation (heat = 0, Update/relax.c:76-114) pass: directions in
   turn, each colour's sites split into static blocks per GLU thread */
static void pass(const or_lat *L, double *U, const glu_db *db, double invbeta, or_rng *R, int heat)
{
  for (int cmu = 0; cmu < OR_ND * db->ncol; cmu++) {
    const int mu = cmu / db->ncol, c = cmu % db->ncol;
    const size_t *sq = db->sq[c];
    OR_THREADS(L->nth, db->n[c], lo, hi, th) {
      for (size_t k = lo; k < hi; k++) {
        const size_t it = sq[k];
        cplx stap[9];
        or_glu_staple(L, U, it, mu, (double *)stap);
        cplx *u = (cplx *)LINK(U, it, mu);
        if (heat) {
          hb_link(u, stap, invbeta, R, th);
          or_gram_reunit((double *)u);
        } else {
          or_link(u, stap);
        }
      }
    }
  }
}

void or_glu_heatbath(const or_lat *L, double *U, double beta, size_t sweeps, size_t nor, uint32_t seed,
                     int cold_start)
{
  if (cold_start) {
    memset(U, 0, L->V * OR_ND * OR_LINK * sizeof(double));
    for (size_t l = 0; l < L->V * OR_ND; l++) U[l * OR_LINK + 0] = U[l * OR_LINK + 8] = U[l * OR_LINK + 16] = 1.0;
  }
  or_rng *R = or_rng_new(seed, L->nth);
  glu_db db;
  db_init(L, &db);
  const double invbeta = OR_NC / beta; /* Update/KPHB.c:173 */
  for (size_t s = 0; s < sweeps; s++) {
    pass(L, U, &db, invbeta, R, 1);
    for (size_t o = 0; o < nor; o++) pass(L, U, &db, invbeta, R, 0);
  }
  db_free(&db);
  or_rng_free(R);
}

void or_glu_overrelax(const or_lat *L, double *U)
{
  glu_db db;
  db_init(L, &db);
  pass(L, U, &db, 0.0, NULL, 0);
  db_free(&db);
}

/* n trial values a0 = 1 - d of the SU(2) heatbath at `norm` from stream 0
   of GLU's streams seeded `seed` (tests of the KP distribution) */
void or_glu_kp_samples(uint32_t seed, double norm, size_t n, double *a0)
{
  or_rng *R = or_rng_new(seed, 1);
  for (size_t k = 0; k < n; k++) a0[k] = 1.0 - kp_d(norm, R, 0);
  or_rng_free(R);
}
