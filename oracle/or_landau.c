/*
 * or_landau.c -- oracle: Landau FACG / FASD (Fourier-accelerated CG/SD).
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * Follows Gfix/FACG.c, Gfix/line_search.c and Gfix/CGalloc.c.  GLU's
 * reduction array `red` is mirrored slot for slot because GLU reads stale
 * slots at a continuation restart (see or_facg).
 */
#include "or_internal.h"

#define PL1 0.15 /* Gfix/line_search.c:34-35 */
#define PL2 0.30

/* U_mu(x) <- g(x) U_mu(x) g(x+mu)^dagger, g = P(1 + alpha X(F)):
   Gfix/line_search.c:60-97 */
void or_egauge_landau(const or_lat *L, double *U, const double *F, double alpha)
{
  const size_t V = L->V;
  if (or_plant == OR_PLANT_EGAUGE_SIGN) alpha = -alpha;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < V; i++) {
    double A[18], B[18];
    or_set_gauge_matrix(A, F, V, alpha, i);
    for (int mu = 0; mu < OR_ND; mu++) {
      or_set_gauge_matrix(B, F, V, alpha, L->fwd[i * OR_ND + mu]);
      or_gtransform_local(A, LINK(U, i, mu), B);
    }
  }
}

/* three-probe line search along F, Kahan-summed per virtual thread into
   red[0..2], then egauge with the parabola vertex:
   Gfix/line_search.c:274-341.  Returns the step taken. */
double or_line_search_landau(const or_lat *L, double *U, const double *F,
                             double *red)
{
  const size_t V = L->V;
  OR_THREADS(L->nth, V, lo, hi, th) {
    double c[OR_NSTEPS] = {0.0, 0.0, 0.0};
    double *r = red + th * OR_CLINE;
    for (size_t i = lo; i < hi; i++) {
      double A1[18], A2[18], B[18], v[OR_NSTEPS] = {0.0, 0.0, 0.0};
      or_set_gauge_matrix(A1, F, V, PL1, i);
      or_set_gauge_matrix(A2, F, V, PL2, i);
      for (int mu = 0; mu < OR_ND; mu++) {
        const double *u = LINK(U, i, mu);
        const size_t f = L->fwd[i * OR_ND + mu];
        v[0] += u[0] + u[8] + u[16];
        or_set_gauge_matrix(B, F, V, PL1, f);
        v[1] += or_retrace_abc_dag(A1, u, B);
        or_set_gauge_matrix(B, F, V, PL2, f);
        v[2] += or_retrace_abc_dag(A2, u, B);
      }
      for (int n = 0; n < OR_NSTEPS; n++) { /* Kahan, line_search.c:318-323 */
        const double y = v[n] - c[n];
        const double t = r[n] + y;
        c[n] = (t - r[n]) - y;
        r[n] = t;
      }
    }
  }
  double val[OR_NSTEPS];
  for (int j = 0; j < OR_NSTEPS; j++) {
    val[j] = 0.0;
    for (int t = 0; t < L->nth; t++) val[j] -= red[j + OR_CLINE * t];
  }
  const double al[OR_NSTEPS] = {0.0, PL1, PL2};
  const double mn = or_approx_minimum(al, val);
  or_egauge_landau(L, U, F, mn);
  return mn;
}

/* Polak-Ribiere sums in the trace metric of the packed field, plain
   per-thread sums into red[3], red[4]: Gfix/FACG.c:99-145 */
static void sum_pr3(const or_lat *L, const double *Fin, const double *Fold,
                    double *red)
{
  const size_t V = L->V;
  const cplx *in = (const cplx *)Fin, *old = (const cplx *)Fold;
  OR_THREADS(L->nth, V, lo, hi, th) {
    double a1 = red[OR_SLOT_PR0 + th * OR_CLINE];
    double a2 = red[OR_SLOT_PR1 + th * OR_CLINE];
    for (size_t i = lo; i < hi; i++) {
      double s1 = 0.0, s2 = 0.0;
      const cplx z0 = in[i];
      s1 += 2.0 * (creal(z0) * creal(z0) + cimag(z0) * cimag(z0) +
                   creal(z0) * cimag(z0));
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * V + i];
        s1 += 2.0 * (creal(z) * creal(z) + cimag(z) * cimag(z));
      }
      cplx t = z0 - old[i];
      s2 += 2.0 * (creal(z0) * creal(t) + cimag(z0) * cimag(t));
      s2 += creal(z0) * cimag(t) + cimag(z0) * creal(t);
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * V + i];
        t = z - old[f * V + i];
        s2 += 2.0 * (creal(z) * creal(t) + cimag(z) * cimag(t));
      }
      a1 += s1;
      a2 += s2;
    }
    red[OR_SLOT_PR0 + th * OR_CLINE] = a1;
    red[OR_SLOT_PR1 + th * OR_CLINE] = a2;
  }
}

/* tr(dA dA) of the packed field into red[3]: Gfix/FACG.c:148-178 */
static void sum_der3(const or_lat *L, const double *Fin, double *red)
{
  const size_t V = L->V;
  const cplx *in = (const cplx *)Fin;
  OR_THREADS(L->nth, V, lo, hi, th) {
    double a = red[OR_SLOT_PR0 + th * OR_CLINE];
    for (size_t i = lo; i < hi; i++) {
      double s = 0.0;
      const cplx z0 = in[i];
      s += creal(z0) * creal(z0) + cimag(z0) * cimag(z0) + creal(z0) * cimag(z0);
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * V + i];
        s += creal(z) * creal(z) + cimag(z) * cimag(z);
      }
      a += 2 * s;
    }
    red[OR_SLOT_PR0 + th * OR_CLINE] = a;
  }
}

static double red_sum(const or_lat *L, const double *red, int slot)
{
  double s = 0.0;
  for (int t = 0; t < L->nth; t++) s += red[slot + OR_CLINE * t];
  return s;
}

static void red_zero(const or_lat *L, double *red)
{
  /* GLU zeroes CLINE*Nthreads entries (Gfix/FACG.c:280-282) */
  memset(red, 0, OR_CLINE * (size_t)L->nth * sizeof(double));
}

static void rec(or_trace *tr, double theta, double alpha, double beta,
                const or_lat *L, const double *U)
{
  if (!tr || tr->n >= tr->cap) return;
  if (tr->theta) tr->theta[tr->n] = theta;
  if (tr->alpha) tr->alpha[tr->n] = alpha;
  if (tr->beta) tr->beta[tr->n] = beta;
  if (tr->func) tr->func[tr->n] = or_functional(L, U);
  tr->n++;
}

typedef struct {
  double *in, *sn, *in_old, *psq, *red;
} facg_tmp;

/* steep_Landau_FACG: Gfix/FACG.c:238-337.  The SD start is recorded as a
   trajectory row with beta = 0; the loop rows carry theta_k, the step and
   beta_k of update k. */
static size_t steep_landau_facg(const or_lat *L, double *U, facg_tmp *T,
                                double gf_alpha, double *tr, double acc,
                                size_t max_iters, or_trace *trc)
{
  const size_t V = L->V, n4 = 4 * V;
  size_t loc_iters = 0;
  /* SD start (FACG.c:249-254): the red slots are NOT cleared here, so a
     continuation run sees the previous segment's values */
  if (or_plant == OR_PLANT_NO_STALE) red_zero(L, T->red);
  *tr = or_fa_deriv_red(L, U, T->in, T->red);
  or_fourier_accelerate(L, T->in, T->psq, OR_ND, 4);
  const double a0 = or_line_search_landau(L, U, T->in, T->red);
  rec(trc, or_plant == OR_PLANT_THETA_AFTER ? 0.5 * or_theta_lin(L, U, OR_ND) : *tr, a0, 0.0, L, U);
  memcpy(T->sn, T->in, n4 * 2 * sizeof(double));
  memcpy(T->in_old, T->in, n4 * 2 * sizeof(double));
  sum_der3(L, T->in, T->red);
  double inold = red_sum(L, T->red, OR_SLOT_PR0);
  for (;;) {
    red_zero(L, T->red);
    double trAA = 0.0;
    or_fa_deriv_red(L, U, T->in, T->red);
    trAA = red_sum(L, T->red, OR_SLOT_TH);
    *tr = trAA / (double)(OR_NC * V);
    or_fourier_accelerate(L, T->in, T->psq, OR_ND, 4);
    sum_pr3(L, T->in, T->in_old, T->red);
    const double insum = red_sum(L, T->red, OR_SLOT_PR0);
    const double sum_conj = red_sum(L, T->red, OR_SLOT_PR1);
    double beta = sum_conj / inold;
    if (beta < 0.0 && or_plant != OR_PLANT_NO_CLAMP) beta = 0.0; /* PRfmax(0, b): Gfix/CG.c:38-42 */
    if (*tr < OR_CG_TOL || or_plant == OR_PLANT_FR_ALWAYS) beta = insum / inold; /* Fletcher-Reeves */
    if (or_plant == OR_PLANT_NO_BETA) beta = 0.0;
    inold = insum;
    cplx *in = (cplx *)T->in, *sn = (cplx *)T->sn, *old = (cplx *)T->in_old;
    const int keep_old = or_plant == OR_PLANT_OLD_KEPT;
#pragma omp parallel for schedule(static)
    for (size_t j = 0; j < n4; j++) {
      sn[j] = in[j] + beta * sn[j];
      if (!keep_old) old[j] = in[j];
    }
    double step;
    if (*tr > OR_CG_TOL) {
      step = or_line_search_landau(L, U, T->sn, T->red);
    } else {
      or_egauge_landau(L, U, T->sn, gf_alpha);
      step = gf_alpha;
    }
    loc_iters++;
    rec(trc, or_plant == OR_PLANT_THETA_AFTER ? 0.5 * or_theta_lin(L, U, OR_ND) : *tr, step, beta, L, U);
    if (!(*tr > acc && loc_iters < max_iters)) break;
  }
  return loc_iters;
}

/* FACG: Gfix/FACG.c:341-392.  Returns the iteration count of the last
   segment, or 123456789 on non-convergence. */
size_t or_facg(const or_lat *L, double *U, double gf_alpha, double acc,
               size_t max_iters, double *theta, or_trace *trc)
{
  *theta = or_theta_lin(L, U, OR_ND);
  if (*theta < acc) return 0;
  facg_tmp T;
  const size_t n4 = 4 * L->V;
  T.in = malloc(n4 * 2 * sizeof(double));
  T.sn = malloc(n4 * 2 * sizeof(double));
  T.in_old = malloc(n4 * 2 * sizeof(double));
  T.psq = malloc(L->V * sizeof(double));
  T.red = calloc(OR_NSTEPS + OR_CLINE * (size_t)L->nth, sizeof(double));
  or_psq_table(L, OR_ND, T.psq);
  size_t iters;
  for (;;) {
    const size_t loc = steep_landau_facg(L, U, &T, gf_alpha, theta, acc,
                                         max_iters, trc);
    iters = loc;
    if (loc >= max_iters) {
      if (*theta < 1e3 * acc) {
        fprintf(stdout, "[GF] Continuation run \n");
        continue;
      }
      iters = OR_GF_FAIL;
    }
    break;
  }
  free(T.in); free(T.sn); free(T.in_old); free(T.psq); free(T.red);
  return iters;
}

/* FASD with the fixed step GF_TUNE (the GLU_GFIX_SD build):
   Gfix/FACG.c:181-235, 395-446.  On non-convergence GLU returns
   123456789 + max_iters (the failure sentinel is set, then the last
   segment's count is added, FACG.c:431-441). */
size_t or_fasd(const or_lat *L, double *U, double gf_alpha, double acc,
               size_t max_iters, double *theta, or_trace *trc)
{
  *theta = or_theta_lin(L, U, OR_ND);
  if (*theta < acc) return 0;
  const size_t V = L->V;
  double *in = malloc(4 * V * 2 * sizeof(double));
  double *psq = malloc(V * sizeof(double));
  double *red = calloc(OR_NSTEPS + OR_CLINE * (size_t)L->nth, sizeof(double));
  or_psq_table(L, OR_ND, psq);
  size_t iters = 0, loc;
  for (;;) {
    loc = 0;
    do {
      red_zero(L, red);
      *theta = or_fa_deriv_red(L, U, in, red);
      or_fourier_accelerate(L, in, psq, OR_ND, 4);
      or_egauge_landau(L, U, in, gf_alpha);
      loc++;
      rec(trc, or_plant == OR_PLANT_THETA_AFTER ? 0.5 * or_theta_lin(L, U, OR_ND) : *theta, gf_alpha, 0.0,
          L, U);
    } while (*theta > acc && loc < max_iters);
    if (loc >= max_iters) {
      if (*theta < 1e3 * acc) {
        fprintf(stdout, "[GF] Continuation run \n");
        continue;
      }
      iters = OR_GF_FAIL;
    }
    iters += loc;
    break;
  }
  free(in); free(psq); free(red);
  return iters;
}
