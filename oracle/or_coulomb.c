/*
 * or_coulomb.c -- oracle: Coulomb gauge fixing, FACG (or FASD) per time
 * slice with Fourier acceleration over the 3 spatial dimensions.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * Follows Gfix/CFACG.c, Gfix/Coulomb.c, Gfix/line_search.c:101-271 and
 * Gfix/gtrans.c:242-264.
 */
#include "or_internal.h"

#define PC1 0.17 /* Gfix/line_search.c:30-31 */
#define PC2 0.34

typedef struct {
  const or_lat *L;
  double *in, *sn, *in_old, *psq, *red;
  size_t *col0; /* draughtboard colour-0 sites of a slice */
  size_t ncol0;
  int two_colour;
  double gf_alpha;
  or_rng *rng;
} cl_ctx;

/* multab_suNC: rows 0,1 of b c, row 2 completed (MMUL/MMUL_SUNC_SSE.c) */
static void multab(double *a18, const double *b18, const double *c18)
{
  const cplx *b = (const cplx *)b18, *c = (const cplx *)c18;
  cplx *a = (cplx *)a18;
  for (int r = 0; r < 2; r++)
    for (int s = 0; s < 3; s++)
      a[r * 3 + s] = b[r * 3] * c[s] + b[r * 3 + 1] * c[3 + s] + b[r * 3 + 2] * c[6 + s];
  a[6] = conj(a[1] * a[5] - a[2] * a[4]);
  a[7] = conj(a[2] * a[3] - a[0] * a[5]);
  a[8] = conj(a[0] * a[4] - a[1] * a[3]);
}

/* derivative of the gauge-rotated spatial links of slice t:
   Gfix/CFACG.c:92-163 */
static void steep_deriv_cg(cl_ctx *C, const double *U, const double *g, size_t t)
{
  const or_lat *L = C->L;
  const size_t Lcu = L->Lcu;
  cplx *f0 = (cplx *)C->in, *f1 = f0 + Lcu, *f2 = f1 + Lcu, *f3 = f2 + Lcu;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < Lcu; i++) {
    const size_t Uidx = i + Lcu * t;
    double sum[10] = {0};
    for (int mu = 0; mu < OR_ND - 1; mu++) {
      const size_t Ubck = L->bwd[Uidx * OR_ND + mu];
      const size_t bck = L->bwd[i * OR_ND + mu], it = L->fwd[i * OR_ND + mu];
      double Cm[18], Dm[18], A[10], B[10];
      memcpy(Cm, LINK(U, Uidx, mu), sizeof(Cm));
      or_gtransform_local(g + i * 18, Cm, g + it * 18);
      memcpy(Dm, LINK(U, Ubck, mu), sizeof(Dm));
      or_gtransform_local(g + bck * 18, Dm, g + i * 18);
      or_herm_proj_short(B, Cm);
      or_herm_proj_short(A, Dm);
      for (int k = 0; k < 10; k++) sum[k] += A[k] - B[k];
    }
    f0[i] = I * sum[0] - sum[6];
    f1[i] = I * sum[2] - sum[3];
    f2[i] = I * sum[4] - sum[5];
    f3[i] = I * sum[8] - sum[9];
  }
}

/* Kahan sum of 2 tr(dA dA) into red[3]: Gfix/CFACG.c:166-202 */
static void sum_der2(cl_ctx *C, const double *F)
{
  const or_lat *L = C->L;
  const size_t n = L->Lcu;
  const cplx *in = (const cplx *)F;
  OR_THREADS(L->nth, n, lo, hi, th) {
    double c = 0.0, *r = C->red + OR_SLOT_PR0 + th * OR_CLINE;
    for (size_t i = lo; i < hi; i++) {
      double s = 0.0;
      const cplx z0 = in[i];
      s += creal(z0) * creal(z0) + cimag(z0) * cimag(z0) + creal(z0) * cimag(z0);
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * n + i];
        s += creal(z) * creal(z) + cimag(z) * cimag(z);
      }
      const double y = 2 * s - c, tt = *r + y;
      c = (tt - *r) - y;
      *r = tt;
    }
  }
}

/* Kahan Polak-Ribiere sums into red[3], red[4]: Gfix/CFACG.c:205-260 */
static void sum_pr(cl_ctx *C, const double *Fin, const double *Fold)
{
  const or_lat *L = C->L;
  const size_t n = L->Lcu;
  const cplx *in = (const cplx *)Fin, *old = (const cplx *)Fold;
  OR_THREADS(L->nth, n, lo, hi, th) {
    double c0 = 0.0, c1 = 0.0;
    double *r0 = C->red + OR_SLOT_PR0 + th * OR_CLINE;
    double *r1 = C->red + OR_SLOT_PR1 + th * OR_CLINE;
    for (size_t i = lo; i < hi; i++) {
      double s1 = 0.0, s2 = 0.0;
      const cplx z0 = in[i];
      s1 += creal(z0) * creal(z0) + cimag(z0) * cimag(z0) + creal(z0) * cimag(z0);
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * n + i];
        s1 += creal(z) * creal(z) + cimag(z) * cimag(z);
      }
      cplx t = z0 - old[i];
      s2 += 2.0 * (creal(z0) * creal(t) + cimag(z0) * cimag(t));
      s2 += creal(z0) * cimag(t) + cimag(z0) * creal(t);
      for (int f = 1; f < 4; f++) {
        const cplx z = in[f * n + i];
        t = z - old[f * n + i];
        s2 += 2.0 * (creal(z) * creal(t) + cimag(z) * cimag(t));
      }
      const double y = 2 * s1 - c0, tt = *r0 + y;
      c0 = (tt - *r0) - y;
      *r0 = tt;
      const double y2 = s2 - c1, t2 = *r1 + y2;
      c1 = (t2 - *r1) - y2;
      *r1 = t2;
    }
  }
}

/* trajectory of one slice, or of every slice (tests): or_coulomb_trace
   with slice = (size_t)-1 takes an array of one trace per time slice */
static or_trace *cl_trace = NULL;
static size_t cl_trace_slice = 0;
void or_coulomb_trace(or_trace *tr, size_t slice)
{
  cl_trace = tr;
  cl_trace_slice = slice;
}
static void cl_rec(size_t t, double theta, double alpha, double beta)
{
  or_trace *tr = cl_trace;
  if (!tr) return;
  if (cl_trace_slice == (size_t)-1)
    tr += t;
  else if (t != cl_trace_slice)
    return;
  if (tr->n >= tr->cap) return;
  if (tr->theta) tr->theta[tr->n] = theta;
  if (tr->alpha) tr->alpha[tr->n] = alpha;
  if (tr->beta) tr->beta[tr->n] = beta;
  tr->n++;
}

/* gauge(x) <- P(1 + alpha X(F(x))) gauge(x): Gfix/line_search.c:101-115 */
static void exponentiate_gauge_cg(cl_ctx *C, double *g, const double *F,
                                  double alpha)
{
  const size_t n = C->L->Lcu;
  if (or_plant == OR_PLANT_EGAUGE_SIGN) alpha = -alpha;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    double tmp[18], old[18];
    or_set_gauge_matrix(tmp, F, n, alpha, i);
    memcpy(old, g + i * 18, sizeof(old));
    multab(g + i * 18, tmp, old);
  }
}

/* Coulomb line search: colour-0 sites with forward and backward links when
   every spatial extent is even (Gfix/line_search.c:194-271), every site
   with forward links otherwise (:141-190); then exponentiate with the
   parabola vertex (:118-138). Returns the step. */
static double line_search_coulomb(cl_ctx *C, const double *U, double *g,
                                  const double *F, size_t t)
{
  const or_lat *L = C->L;
  const size_t n = L->Lcu;
  const size_t nloop = C->two_colour ? C->ncol0 : n;
  OR_THREADS(L->nth, nloop, lo, hi, th) {
    double c[OR_NSTEPS] = {0, 0, 0}, *r = C->red + th * OR_CLINE;
    for (size_t ii = lo; ii < hi; ii++) {
      const size_t idx = C->two_colour ? C->col0[ii] : ii;
      double A1[18], A2[18], B[18], Cm[18], Dm[18];
      double v[OR_NSTEPS] = {0, 0, 0};
      or_set_gauge_matrix(A1, F, n, PC1, idx);
      or_set_gauge_matrix(A2, F, n, PC2, idx);
      for (int mu = 0; mu < OR_ND - 1; mu++) {
        const size_t fwd = L->fwd[idx * OR_ND + mu];
        memcpy(Cm, LINK(U, idx + n * t, mu), sizeof(Cm));
        or_gtransform_local(g + idx * 18, Cm, g + fwd * 18);
        v[0] += Cm[0] + Cm[8] + Cm[16];
        if (C->two_colour) {
          const size_t bck = L->bwd[idx * OR_ND + mu];
          memcpy(Dm, LINK(U, bck + n * t, mu), sizeof(Dm));
          or_gtransform_local(g + bck * 18, Dm, g + idx * 18);
          v[0] += Dm[0] + Dm[8] + Dm[16];
          or_set_gauge_matrix(B, F, n, PC1, fwd);
          v[1] += or_retrace_abc_dag(A1, Cm, B);
          or_set_gauge_matrix(B, F, n, PC1, bck);
          v[1] += or_retrace_abc_dag(B, Dm, A1);
          or_set_gauge_matrix(B, F, n, PC2, fwd);
          v[2] += or_retrace_abc_dag(A2, Cm, B);
          or_set_gauge_matrix(B, F, n, PC2, bck);
          v[2] += or_retrace_abc_dag(B, Dm, A2);
        } else {
          or_set_gauge_matrix(B, F, n, PC1, fwd);
          v[1] += or_retrace_abc_dag(A1, Cm, B);
          or_set_gauge_matrix(B, F, n, PC2, fwd);
          v[2] += or_retrace_abc_dag(A2, Cm, B);
        }
      }
      for (int k = 0; k < OR_NSTEPS; k++) {
        const double y = v[k] - c[k], tt = r[k] + y;
        c[k] = (tt - r[k]) - y;
        r[k] = tt;
      }
    }
  }
  double val[OR_NSTEPS];
  for (int j = 0; j < OR_NSTEPS; j++) {
    val[j] = 0.0;
    for (int th = 0; th < L->nth; th++) val[j] -= C->red[j + OR_CLINE * th];
  }
  const double al[OR_NSTEPS] = {0.0, PC1, PC2};
  const double mn = or_approx_minimum(al, val);
  exponentiate_gauge_cg(C, g, F, mn);
  return mn;
}

static double redsum(const cl_ctx *C, int slot)
{
  double s = 0.0;
  for (int th = 0; th < C->L->nth; th++) s += C->red[slot + OR_CLINE * th];
  return s;
}

static void red_zero_all(cl_ctx *C)
{
  memset(C->red, 0, (OR_NSTEPS + OR_CLINE * (size_t)C->L->nth) * sizeof(double));
}

/* Gfix/CFACG.c:286-382 */
static size_t steep_step_facg(cl_ctx *C, const double *U, double *g, double *tr,
                              size_t t, double acc, size_t max_iters)
{
  const or_lat *L = C->L;
  const size_t n = L->Lcu, n4 = 4 * n;
  const double norm = 1.0 / (double)(OR_NC * n); /* GFNORM_COULOMB */
  red_zero_all(C);
  steep_deriv_cg(C, U, g, t);
  or_fourier_accelerate(L, C->in, C->psq, OR_ND - 1, 4);
  sum_der2(C, C->in);
  double inold = redsum(C, OR_SLOT_PR0);
  *tr = inold * norm;
  cl_rec(t, *tr, line_search_coulomb(C, U, g, C->in, t), 0.0);
  memcpy(C->sn, C->in, n4 * 2 * sizeof(double));
  memcpy(C->in_old, C->in, n4 * 2 * sizeof(double));
  size_t loc = 0;
  for (;;) {
    red_zero_all(C);
    steep_deriv_cg(C, U, g, t);
    or_fourier_accelerate(L, C->in, C->psq, OR_ND - 1, 4);
    sum_pr(C, C->in, C->in_old);
    const double insum = redsum(C, OR_SLOT_PR0), sc = redsum(C, OR_SLOT_PR1);
    *tr = insum * norm;
    double beta = sc / inold;
    if (beta < 0.0 && or_plant != OR_PLANT_NO_CLAMP) beta = 0.0;
    if (*tr < OR_CG_TOL || or_plant == OR_PLANT_FR_ALWAYS) beta = insum / inold;
    if (or_plant == OR_PLANT_NO_BETA) beta = 0.0;
    inold = insum;
    cplx *in = (cplx *)C->in, *sn = (cplx *)C->sn, *old = (cplx *)C->in_old;
    for (size_t j = 0; j < n4; j++) {
      sn[j] = in[j] + beta * sn[j];
      if (or_plant != OR_PLANT_OLD_KEPT) old[j] = in[j];
    }
    double step = C->gf_alpha;
    if (*tr > OR_CG_TOL)
      step = line_search_coulomb(C, U, g, C->sn, t);
    else
      exponentiate_gauge_cg(C, g, C->sn, C->gf_alpha);
    cl_rec(t, *tr, step, beta);
    loc++;
    if (!(*tr > acc && loc < max_iters)) break;
  }
  return loc;
}

/* Gfix/CFACG.c:385-430 (the GLU_GFIX_SD step: fixed GF_TUNE) */
static size_t steep_step_fasd(cl_ctx *C, const double *U, double *g, double *tr,
                              size_t t, double acc, size_t max_iters)
{
  const double norm = 1.0 / (double)(OR_NC * C->L->Lcu);
  size_t loc = 0;
  do {
    red_zero_all(C);
    steep_deriv_cg(C, U, g, t);
    or_fourier_accelerate(C->L, C->in, C->psq, OR_ND - 1, 4);
    exponentiate_gauge_cg(C, g, C->in, C->gf_alpha);
    sum_der2(C, C->in);
    *tr = redsum(C, OR_SLOT_PR0) * norm;
    cl_rec(t, *tr, C->gf_alpha, 0.0);
    loc++;
  } while (*tr > acc && loc < max_iters);
  return loc;
}

/* reunitarise the slice gauge and test unitarity: Gfix/CFACG.c:433-457 */
static int adequate_solution(cl_ctx *C, double *g, size_t iters)
{
  int ok = 1;
  for (size_t i = 0; i < C->L->Lcu; i++) {
    or_gram_reunit(g + i * 18);
    double s = 0.0;
    for (int k = 0; k < 18; k++) s += g[i * 18 + k] * g[i * 18 + k];
    if (fabs(s - OR_NC) > (double)iters * OR_PREC_TOL) ok = 0;
  }
  return ok;
}

/* Gfix/CFACG.c:460-530 */
static size_t steep_fix(cl_ctx *C, const double *U, double *g, size_t t,
                        double acc, size_t max_iters, int facg, double *tr_out)
{
  double tr = 1.0;
  size_t loc, tot = 0, control = 0;
  int trans = 0;
  for (;;) {
    loc = facg ? steep_step_facg(C, U, g, &tr, t, acc, max_iters)
               : steep_step_fasd(C, U, g, &tr, t, acc, max_iters);
    if (tr < acc) {
      if (!adequate_solution(C, g, loc > 0 ? loc : 1)) trans = 1;
    }
    if (loc >= max_iters) {
      if (tr < 1e4 * acc && !isnan(tr) && !isinf(tr)) {
        tot += loc;
        continue;
      }
      if (tr > acc || isnan(tr) || isinf(tr)) trans = 1;
    }
    if (control < 8 && trans) {
      fprintf(stdout, "[CG] random gauge transform %zu :: %e (%zu)\n", control,
              tr, loc);
      OR_THREADS(C->L->nth, C->L->Lcu, lo, hi, th) {
        for (size_t i = lo; i < hi; i++) or_sunitary_gen(C->rng, th, g + i * 18);
      }
      tot += loc;
      control++;
      trans = 0;
      continue;
    }
    tot += loc;
    break;
  }
  *tr_out = tr;
  return tot;
}

/* slice t: spatial U_i <- g U_i g(x+i)^dagger, temporal U_t <- g U_t g_up^dagger:
   Gfix/gtrans.c:242-264 */
static void gtransform_slice(const or_lat *L, double *U, const double *g,
                             const double *gup, size_t t)
{
  const size_t n = L->Lcu;
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < n; i++) {
    const size_t j = n * t + i;
    for (int mu = 0; mu < OR_ND - 1; mu++)
      or_gtransform_local(g + i * 18, LINK(U, j, mu), g + L->fwd[i * OR_ND + mu] * 18);
    or_gtransform_local(g + i * 18, LINK(U, j, OR_ND - 1), gup + i * 18);
  }
}

static void set_identity(double *g, size_t n)
{
  memset(g, 0, n * 18 * sizeof(double));
  for (size_t i = 0; i < n; i++) g[i * 18 + 0] = g[i * 18 + 8] = g[i * 18 + 16] = 1.0;
}

/* Coulomb_FA: Gfix/CFACG.c:532-636; returns the summed iteration count.
   slice_iters / slice_theta (optional, T entries) receive per-slice
   results. */
size_t or_coulomb(const or_lat *L, double *U, double gf_alpha, double acc,
                  size_t max_iters, int facg, uint32_t seed,
                  size_t *slice_iters, double *slice_theta)
{
  cl_ctx C;
  const size_t n = L->Lcu, T = L->dims[OR_ND - 1];
  C.L = L;
  C.gf_alpha = gf_alpha;
  C.in = malloc(4 * n * 2 * sizeof(double));
  C.sn = calloc(4 * n * 2, sizeof(double));
  C.in_old = calloc(4 * n * 2, sizeof(double));
  C.psq = malloc(n * sizeof(double));
  C.red = calloc(OR_NSTEPS + OR_CLINE * (size_t)L->nth, sizeof(double));
  C.rng = or_rng_new(seed, L->nth);
  or_psq_table(L, OR_ND - 1, C.psq);
  /* draughtboard of the spatial slice: Geometry/draughtboard.c:76-112 */
  C.two_colour = 1;
  for (int mu = 0; mu < OR_ND - 1; mu++)
    if (L->dims[mu] & 1) C.two_colour = 0;
  C.col0 = malloc(n * sizeof(size_t));
  C.ncol0 = 0;
  for (size_t i = 0; i < n; i++) {
    size_t s = 0, ii = i;
    for (int mu = 0; mu < OR_ND - 1; mu++) { s += ii % L->dims[mu]; ii /= L->dims[mu]; }
    if (s % 2 == 0) C.col0[C.ncol0++] = i;
  }
  double *g = malloc(n * 18 * sizeof(double)), *g_end = malloc(n * 18 * sizeof(double)),
         *g_up = malloc(n * 18 * sizeof(double));
  set_identity(g, n);
  set_identity(g_end, n);
  set_identity(g_up, n);
  double th;
  size_t its, tot = 0;
  its = steep_fix(&C, U, g_end, 0, acc, max_iters, facg, &th);
  if (slice_iters) slice_iters[0] = its;
  if (slice_theta) slice_theta[0] = th;
  tot += its;
  its = steep_fix(&C, U, g, 1, acc, max_iters, facg, &th);
  if (slice_iters) slice_iters[1] = its;
  if (slice_theta) slice_theta[1] = th;
  tot += its;
  gtransform_slice(L, U, g_end, g, 0);
  size_t t;
  for (t = 2; t < T; t++) {
    set_identity(g_up, n);
    its = steep_fix(&C, U, g_up, t, acc, max_iters, facg, &th);
    if (slice_iters) slice_iters[t] = its;
    if (slice_theta) slice_theta[t] = th;
    tot += its;
    gtransform_slice(L, U, g, g_up, t - 1);
    double *p = g;
    g = g_up;
    g_up = p;
  }
  gtransform_slice(L, U, g, g_end, t - 1);
  /* Gfix/Coulomb.c:97 */
  or_latt_reunit(L, U);
  free(C.in); free(C.sn); free(C.in_old); free(C.psq); free(C.red); free(C.col0);
  or_rng_free(C.rng);
  free(g); free(g_end); free(g_up);
  return tot;
}
