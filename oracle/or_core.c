/*
 * or_core.c -- oracle: geometry, SU(3) algebra, measurements, derivative.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 */
#include "or_internal.h"

/* ---------------------------------------------------------------------- */
/* geometry                                                               */
/* ---------------------------------------------------------------------- */

/* lexicographic site, x fastest: Geometry/geometry.c:28-36 (gen_site) */
static size_t site_of(const size_t *dims, const int x[OR_ND])
{
  size_t r = (size_t)x[OR_ND - 1];
  for (int mu = OR_ND - 1; mu > 0; mu--) r = dims[mu - 1] * r + (size_t)x[mu - 1];
  return r;
}

/* coordinates of site i: Geometry/geometry.c:105-120 (get_mom_2piBZ) */
static void coords_of(const size_t *dims, size_t i, int x[OR_ND])
{
  for (int mu = 0; mu < OR_ND; mu++) {
    x[mu] = (int)(i % dims[mu]);
    i /= dims[mu];
  }
}

/* neighbour tables: Geometry/init.c:30-41 (init_navig) with
   Geometry/geometry.c:289-305 (gen_shift), periodic in every direction */
int or_lat_init(or_lat *L, const size_t dims[OR_ND], int nth)
{
  memset(L, 0, sizeof(*L));
  L->V = 1;
  for (int mu = 0; mu < OR_ND; mu++) {
    L->dims[mu] = dims[mu];
    L->V *= dims[mu];
  }
  L->Lcu = L->V / dims[OR_ND - 1];
  L->nth = nth > 0 ? nth : 1;
  L->fwd = malloc(L->V * OR_ND * sizeof(size_t));
  L->bwd = malloc(L->V * OR_ND * sizeof(size_t));
  if (!L->fwd || !L->bwd) return -1;
  for (size_t i = 0; i < L->V; i++) {
    int x[OR_ND];
    coords_of(L->dims, i, x);
    for (int mu = 0; mu < OR_ND; mu++) {
      int y[OR_ND];
      memcpy(y, x, sizeof(y));
      y[mu] = (int)((x[mu] + 1) % (int)dims[mu]);
      L->fwd[i * OR_ND + mu] = site_of(L->dims, y);
      y[mu] = x[mu] == 0 ? (int)dims[mu] - 1 : x[mu] - 1;
      L->bwd[i * OR_ND + mu] = site_of(L->dims, y);
    }
  }
  return 0;
}

void or_lat_free(or_lat *L)
{
  free(L->fwd);
  free(L->bwd);
  L->fwd = L->bwd = NULL;
}

/* lattice p^2 with the zero mode set to 1: Geometry/geometry.c:170-186.
   Momentum integers come from the 0..L-1 zone, twiddle 2pi/L
   (Geometry/init.c:69-71, TWOPI = 6.283185307179586 in
   Headers/GLU_definitions.h:139). */
double or_psq(const or_lat *L, size_t i, int D)
{
  int n[OR_ND];
  coords_of(L->dims, i, n);
  double kcos = 0.0;
  for (int mu = 0; mu < D; mu++)
    kcos += cos(n[mu] * (OR_TWOPI / (double)L->dims[mu]));
  if (kcos == (double)D) return 1.0;
  return 2.0 * ((double)D - kcos);
}

/* ---------------------------------------------------------------------- */
/* SU(3)                                                                  */
/* ---------------------------------------------------------------------- */

/* Gram-Schmidt on columns 0 and 1, column 2 = conj(col0 x col1):
   Matrix_Ops/gramschmidt.c:73-165 (generic form of gramschmidt_SSE.c:38-98) */
void or_gram_reunit(double *U18)
{
  cplx *U = (cplx *)U18;
  double n = 0.0;
  n += creal(U[0]) * creal(U[0]) + cimag(U[0]) * cimag(U[0]);
  n += creal(U[3]) * creal(U[3]) + cimag(U[3]) * cimag(U[3]);
  n += creal(U[6]) * creal(U[6]) + cimag(U[6]) * cimag(U[6]);
  n = 1.0 / sqrt(n);
  cplx c0 = U[0] * n, c1 = U[3] * n, c2 = U[6] * n;
  /* projection of column 1 onto column 0: sum conj(col0_k) col1_k */
  cplx p = conj(c0) * U[1] + conj(c1) * U[4] + conj(c2) * U[7];
  cplx d0 = U[1] - p * c0, d1 = U[4] - p * c1, d2 = U[7] - p * c2;
  double m = creal(d0) * creal(d0) + cimag(d0) * cimag(d0) +
             creal(d1) * creal(d1) + cimag(d1) * cimag(d1) +
             creal(d2) * creal(d2) + cimag(d2) * cimag(d2);
  m = 1.0 / sqrt(m);
  d0 *= m; d1 *= m; d2 *= m;
  U[0] = c0; U[3] = c1; U[6] = c2;
  U[1] = d0; U[4] = d1; U[7] = d2;
  U[2] = conj(c1 * d2 - c2 * d1);
  U[5] = conj(c2 * d0 - c0 * d2);
  U[8] = conj(c0 * d1 - c1 * d0);
}

/* g = P_SU(3)(1 + alpha X(F_i)), the exp_approx branch of set_gauge_matrix:
   Gfix/CG.c:47-100 (only columns 0 and 1 are set before reunitarising) */
void or_set_gauge_matrix(double *g18, const double *F, size_t V, double alpha,
                         size_t i)
{
  const cplx *f0 = (const cplx *)F, *f1 = f0 + V, *f2 = f1 + V, *f3 = f2 + V;
  cplx *g = (cplx *)g18;
  g[0] = 1.0 + I * alpha * cimag(f0[i]);
  g[1] = alpha * f1[i];
  g[3] = -conj(g[1]);
  g[4] = 1.0 - I * alpha * creal(f0[i]);
  g[6] = -alpha * conj(f2[i]);
  g[7] = -alpha * conj(f3[i]);
  g[2] = g[5] = g[8] = 0.0;
  or_gram_reunit(g18);
}

/* b <- a b c^dagger; columns 0 and 1 of the product are computed and
   column 2 completed as conj(col0 x col1), as the SSE variant does:
   Gfix/gtrans.c:32-141 */
void or_gtransform_local(const double *a18, double *b18, const double *c18)
{
  const cplx *a = (const cplx *)a18, *c = (const cplx *)c18;
  cplx *b = (cplx *)b18;
  cplx t[9]; /* t = b c^dagger, columns 0 and 1 */
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 2; s++)
      t[r * 3 + s] = b[r * 3 + 0] * conj(c[s * 3 + 0]) +
                     b[r * 3 + 1] * conj(c[s * 3 + 1]) +
                     b[r * 3 + 2] * conj(c[s * 3 + 2]);
  cplx o[9];
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 2; s++)
      o[r * 3 + s] = a[r * 3 + 0] * t[0 + s] + a[r * 3 + 1] * t[3 + s] +
                     a[r * 3 + 2] * t[6 + s];
  o[2] = conj(o[3] * o[7] - o[4] * o[6]);
  o[5] = conj(o[1] * o[6] - o[0] * o[7]);
  o[8] = conj(o[0] * o[4] - o[1] * o[3]);
  memcpy(b, o, sizeof(o));
}

/* Re tr(a b c^dagger) with rows 0,1 of a.b and row 2 from the cofactors:
   Matrix_Ops/trace_abc_SSE.c:34-120 */
double or_retrace_abc_dag(const double *a18, const double *b18,
                          const double *c18)
{
  const cplx *a = (const cplx *)a18, *b = (const cplx *)b18,
             *c = (const cplx *)c18;
  cplx ab[9];
  for (int r = 0; r < 2; r++)
    for (int s = 0; s < 3; s++)
      ab[r * 3 + s] = a[r * 3 + 0] * b[0 + s] + a[r * 3 + 1] * b[3 + s] +
                      a[r * 3 + 2] * b[6 + s];
  ab[6] = conj(ab[1] * ab[5] - ab[2] * ab[4]);
  ab[7] = conj(ab[2] * ab[3] - ab[0] * ab[5]);
  ab[8] = conj(ab[0] * ab[4] - ab[1] * ab[3]);
  double re = 0.0, im = 0.0;
  for (int k = 0; k < 9; k++) {
    re += creal(ab[k]) * creal(c[k]);
    im += cimag(ab[k]) * cimag(c[k]);
  }
  return re + im;
}

/* traceless Hermitian part of (U - U^dagger)/2i in the 5-entry short form
   { Q00 (real), Q01, Q02, Q11 (real), Q12 }: Matrix_Ops/exactQ.c:400-428 */
void or_herm_proj_short(double *Q10, const double *U18)
{
  const cplx *U = (const cplx *)U18;
  cplx *Q = (cplx *)Q10;
  const double i0 = cimag(U[0]), i4 = cimag(U[4]), i8 = cimag(U[8]);
  Q[0] = (2.0 * i0 - i4 - i8) / 3.0;
  Q[1] = I * 0.5 * (conj(U[3]) - U[1]);
  Q[2] = I * 0.5 * (conj(U[6]) - U[2]);
  Q[3] = (2.0 * i4 - i0 - i8) / 3.0;
  Q[4] = I * 0.5 * (conj(U[7]) - U[5]);
}

/* vertex of the Lagrange parabola through 3 probes, clamped:
   Gfix/line_search.c:38-57 */
double or_approx_minimum(const double al[3], const double f[3])
{
  const double A = f[0] / ((al[0] - al[1]) * (al[0] - al[2]));
  const double B = f[1] / ((al[1] - al[0]) * (al[1] - al[2]));
  const double C = f[2] / ((al[2] - al[0]) * (al[2] - al[1]));
  const double sol =
      (al[0] * (B + C) + al[1] * (A + C) + al[2] * (A + B)) / (2 * (A + B + C));
  if (sol > al[2]) return al[2];
  if (isnan(sol) || isinf(sol) || sol < al[0]) return 0.08;
  return sol;
}

/* ---------------------------------------------------------------------- */
/* measurements                                                           */
/* ---------------------------------------------------------------------- */

static double re_tr(const double *u) { return u[0] + u[8] + u[16]; }

/* Re tr(a b c^dagger d^dagger): Field/plaqs_links.c:34-... */
static double plaq_face(const double *a, const double *b, const double *c,
                        const double *d)
{
  const cplx *A = (const cplx *)a, *B = (const cplx *)b, *C = (const cplx *)c,
             *D = (const cplx *)d;
  cplx ab[9], dc[9];
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) {
      ab[r * 3 + s] = A[r * 3] * B[s] + A[r * 3 + 1] * B[3 + s] +
                      A[r * 3 + 2] * B[6 + s];
      /* (d c)^dagger's conjugate-transpose partner: dc = d c */
      dc[r * 3 + s] = D[r * 3] * C[s] + D[r * 3 + 1] * C[3 + s] +
                      D[r * 3 + 2] * C[6 + s];
    }
  /* Re tr(ab (dc)^dagger) = Re sum_ij ab_ij conj(dc_ij) */
  double t = 0.0;
  for (int k = 0; k < 9; k++)
    t += creal(ab[k]) * creal(dc[k]) + cimag(ab[k]) * cimag(dc[k]);
  return t;
}

/* average plaquette: Field/plaqs_links.c:243-265 (serial site loop) */
double or_plaquette(const or_lat *L, const double *U)
{
  double plaq = 0.0;
  for (size_t i = 0; i < L->V; i++) {
    double p = 0.0;
    for (int mu = 0; mu < OR_ND; mu++) {
      const size_t t = L->fwd[i * OR_ND + mu];
      for (int nu = 0; nu < mu; nu++) {
        const size_t s = L->fwd[i * OR_ND + nu];
        p += plaq_face(LINK(U, i, mu), LINK(U, t, nu), LINK(U, s, mu),
                       LINK(U, i, nu));
      }
    }
    plaq += p;
  }
  return 2.0 * plaq / (double)(OR_NC * OR_ND * (OR_ND - 1) * L->V);
}

/* average link trace and the largest single-link trace:
   Field/plaqs_links.c:424-448 (indivlinks) */
double or_links(const or_lat *L, const double *U, double *maxtr)
{
  double link = 0.0, mx = 0.0;
  for (size_t i = 0; i < L->V; i++) {
    double loc = 0.0;
    for (int mu = 0; mu < OR_ND; mu++) {
      const double r = re_tr(LINK(U, i, mu));
      loc += r;
      if (r > mx) mx = r;
    }
    link += loc;
  }
  if (maxtr) *maxtr = mx;
  return link / (double)(OR_ND * OR_NC * L->V);
}

/* spatial and temporal link traces: Field/plaqs_links.c:399-422 */
void or_all_links(const or_lat *L, const double *U, double *sp, double *t)
{
  double s = 0.0, tt = 0.0;
  for (size_t i = 0; i < L->V; i++) {
    double p = 0.0;
    for (int mu = 0; mu < OR_ND - 1; mu++) p += re_tr(LINK(U, i, mu));
    s += p;
    tt += re_tr(LINK(U, i, OR_ND - 1));
  }
  *sp = s / ((OR_ND - 1) * OR_NC * (double)L->V);
  *t = tt / (OR_NC * (double)L->V);
}

/* F = 1 - sum Re tr U / (ND NC V): Gfix/gftests.c:133-147 */
double or_functional(const or_lat *L, const double *U)
{
  double tr = 0.0;
  for (size_t i = 0; i < L->V; i++) {
    double loc = 0.0;
    for (int mu = 0; mu < OR_ND; mu++) loc += re_tr(LINK(U, i, mu));
    tr += loc;
  }
  return 1.0 - tr / (double)(OR_ND * OR_NC * L->V);
}

/* reunitarise every link: Geometry/random_config.c:43-56 */
void or_latt_reunit(const or_lat *L, double *U)
{
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < L->V; i++)
    for (int mu = 0; mu < OR_ND; mu++) or_gram_reunit(LINK(U, i, mu));
}

/* ---------------------------------------------------------------------- */
/* derivative                                                             */
/* ---------------------------------------------------------------------- */

/* 4 tr(sum^2) in the short Hermitian form: Gfix/lin_derivs.c:30-50 */
static double trace_deriv(const double *s)
{
  /* s = {Re Q00, 0, Re Q01, Im Q01, Re Q02, Im Q02, Re Q11, 0, Re Q12, Im Q12} */
  return 4.0 * (s[0] * s[0] + s[6] * s[6] + s[0] * s[6] + s[2] * s[2] +
                s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + s[8] * s[8] +
                s[9] * s[9]);
}

/* generic derivative sum_mu [A_mu(x - mu) - A_mu(x)] built from the
   Hermitian projection: Gfix/lin_derivs.c:56-70 */
double or_latt_deriv(const or_lat *L, const double *U, size_t i, int max_dir,
                     double *sum10)
{
  memset(sum10, 0, 10 * sizeof(double));
  for (int mu = 0; mu < max_dir; mu++) {
    double A[10], sA[10];
    or_herm_proj_short(A, LINK(U, i, mu));
    or_herm_proj_short(sA, LINK(U, L->bwd[i * OR_ND + mu], mu));
    for (int k = 0; k < 10; k++) sum10[k] += sA[k] - A[k];
  }
  return trace_deriv(sum10);
}

/* the SU(3)-specialised derivative reading 15 of the 18 reals of each
   link and returning the local tr(Delta^2): Gfix/lin_derivs.c:75-193 */
double or_fast_deriv(const or_lat *L, const double *U, size_t i, double *sum10)
{
  double R0 = 0, R1 = 0, I1 = 0, R2 = 0, I2 = 0, R3 = 0, R4 = 0, I4 = 0;
  for (int mu = 0; mu < OR_ND; mu++) {
    const double *q = LINK(U, i, mu);
    const double *s = LINK(U, L->bwd[i * OR_ND + mu], mu);
    R0 += 2 * s[1] - s[9] - s[17];
    R0 += -2 * q[1] + q[9] + q[17];
    R1 += s[2] - s[6] - q[2] + q[6];
    I1 += s[3] + s[7] - q[3] - q[7];
    R2 += s[4] - s[12] - q[4] + q[12];
    I2 += s[5] + s[13] - q[5] - q[13];
    R3 += 2 * s[9] - s[1] - s[17];
    R3 += -2 * q[9] + q[1] + q[17];
    R4 += s[10] - s[14] - q[10] + q[14];
    I4 += s[11] + s[15] - q[11] - q[15];
  }
  /* sum = { R0/3, -i R1/2 + I1/2, -i R2/2 + I2/2, R3/3, -i R4/2 + I4/2 } */
  sum10[0] = R0 / 3.0; sum10[1] = 0.0;
  sum10[2] = 0.5 * I1; sum10[3] = -R1 * 0.5;
  sum10[4] = 0.5 * I2; sum10[5] = -R2 * 0.5;
  sum10[6] = R3 / 3.0; sum10[7] = 0.0;
  sum10[8] = 0.5 * I4; sum10[9] = -R4 * 0.5;
  return 0.4444444444444444 * (R0 * R0 + R3 * (R3 + R0)) +
         (R1 * R1 + I1 * I1 + R2 * R2 + I2 * I2 + R4 * R4 + I4 * I4);
}

/* pack Delta into the 4 FFT fields and accumulate the per-thread plain sum
   of the local traces (GLU's red[5 + th*CLINE]): Gfix/FACG.c:38-73.
   Returns theta = sum / (NC V) (GFNORM_LANDAU, GLU_definitions.h:427). */
double or_fa_deriv_red(const or_lat *L, const double *U, double *F, double *red)
{
  const size_t V = L->V;
  cplx *f0 = (cplx *)F, *f1 = f0 + V, *f2 = f1 + V, *f3 = f2 + V;
  OR_THREADS(L->nth, V, lo, hi, th) {
    double acc = red[OR_SLOT_TH + th * OR_CLINE];
    for (size_t i = lo; i < hi; i++) {
      double s[10];
      acc += or_fast_deriv(L, U, i, s);
      f0[i] = I * s[0] - s[6];
      f1[i] = I * s[2] - s[3];
      f2[i] = I * s[4] - s[5];
      f3[i] = I * s[8] - s[9];
    }
    red[OR_SLOT_TH + th * OR_CLINE] = acc;
  }
  double tr = 0.0;
  for (int th = 0; th < L->nth; th++) tr += red[OR_SLOT_TH + th * OR_CLINE];
  return tr / (double)(OR_NC * V);
}

double or_fa_deriv(const or_lat *L, const double *U, double *F)
{
  double *red = calloc(OR_NSTEPS + OR_CLINE * (size_t)L->nth, sizeof(double));
  const double th = or_fa_deriv_red(L, U, F, red);
  free(red);
  return th;
}

int or_plant = OR_PLANT_NONE;
void or_set_plant(int code) { or_plant = code; }

/* theta_test_lin = 2 sum tr(Delta^2) / (NC V): Gfix/gftests.c:215-250 */
double or_theta_lin(const or_lat *L, const double *U, int max_dir)
{
  double tr = 0.0;
#pragma omp parallel for reduction(+ : tr) schedule(static)
  for (size_t i = 0; i < L->V; i++) {
    double s[10];
    tr += or_latt_deriv(L, U, i, max_dir, s);
  }
  return 2.0 * tr / (double)(OR_NC * L->V);
}
