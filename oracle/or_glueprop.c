/*
 * or_glueprop.c -- oracle: transverse and longitudinal gluon propagators of
 * a gauge-fixed field, the GLUON_PROPS branch of cuts_struct
 * (Cuts/cuts.c:114-184).  TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * Each step follows the reference directly, with all nine entries of every
 * field (the device packs them; this does not):
 *   - A_mu(x) = Hermitian_proj(U_mu(x)), the LINEAR definition
 *     (Matrix_Ops/exactQ.c:358-373);
 *   - the unnormalised backward DFT of the 4 x 9 complex fields
 *     (Cuts/cuts.c:36-100 with CUT_FORWARD undefined: exp(+i p.x));
 *   - the momentum list of get_mom_veclist (Cuts/cut_routines.c:383-455):
 *     momenta enumerated in (-L/2 .. L/2-1) with x fastest, kept by
 *     safe_momenta (:333-380) with the anisotropy ratios of simorb_ratios
 *     (:73-95) and the cylinder tests cylinder_DF / cylinder_CJD
 *     (:191-298, CYLINDER_AS undefined, SIN_MOM momenta of
 *     Geometry/geometry.c:150-166);
 *   - correct_pspace_landau (Cuts/pspace_landau.c:28-86): A_mu(p) times
 *     exp(i p_mu / 2);
 *   - compute_gluon_prop and project_trans_long (Cuts/glueprop.c:32-151):
 *     the first half of the list and the zero, the rest by p -> -p;
 *   - the exceptional two- and three-point functions of
 *     write_exceptional_g2g3_MOMgg (Cuts/MOMgg.c:94-215), GLU's default
 *     cuts output;
 *   - the output of write_mom_veclist and write_g2g3_to_list
 *     (Cuts/cut_output.c:30-52, 311-336, 371-400), big-endian.
 */
#include "or_internal.h"

#include <stdint.h>

#define OR_CUT_HYPERCUBIC 0
#define OR_CUT_PSQ 1
#define OR_CUT_CYLINDER 2
#define OR_CUT_CONICAL 3

/* traceless Hermitian part, entry by entry as exactQ.c:361-373 */
static void hermitian_proj(cplx Q[9], const cplx U[9])
{
  const double c0 = cimag(U[0]), c4 = cimag(U[4]), c8 = cimag(U[8]);
  Q[0] = (2 * c0 - c4 - c8) / 3.;
  Q[1] = -I * (U[1] - conj(U[3])) / 2.;
  Q[2] = -I * (U[2] - conj(U[6])) / 2.;
  Q[3] = conj(Q[1]);
  Q[4] = (2 * c4 - c0 - c8) / 3.;
  Q[5] = -I * (U[5] - conj(U[7])) / 2.;
  Q[6] = conj(Q[2]);
  Q[7] = conj(Q[5]);
  Q[8] = -creal(Q[0]) - creal(Q[4]);
}

/* the diagonal of index i: entries -1 / +1 (cut_routines.c:97-129) */
static void get_diagonal(double n[OR_ND], size_t i)
{
  size_t subvol = 1;
  for (int mu = 0; mu < OR_ND; mu++) {
    n[mu] = (double)(((i - i % subvol) / subvol) % 2);
    if (n[mu] == 0) n[mu] = -1;
    subvol *= 2;
  }
}

static int cylinder_df(const double q[OR_ND], double width)
{
  const double norm = 1. / (2. * sqrt(OR_ND));
  const int diagonals = 2 << (OR_ND - 1);
  for (int d = 0; d < diagonals; d++) {
    double x[OR_ND];
    get_diagonal(x, (size_t)d);
    double sp = 0.;
    for (int nu = 0; nu < OR_ND; nu++) sp += q[nu] * x[nu];
    sp *= norm;
    double mod = 0.0;
    for (int nu = 0; nu < OR_ND; nu++) {
      const double t = q[nu] - sp * x[nu];
      mod += t * t;
    }
    if (sqrt(mod) <= width) return 1;
  }
  return 0;
}

static int cylinder_cjd(const double q[OR_ND], double width, double arc)
{
  const int diagonals = 2 << (OR_ND - 1);
  double qnorm = 0.;
  for (int mu = 0; mu < OR_ND; mu++) qnorm += q[mu] * q[mu];
  if (qnorm < OR_PREC_TOL) return 1;
  qnorm = 1.0 / sqrt(qnorm);
  for (int d = 0; d < diagonals; d++) {
    double x[OR_ND];
    get_diagonal(x, (size_t)d);
    double nnorm = 0.;
    for (int nu = 0; nu < OR_ND; nu++) nnorm += x[nu] * x[nu];
    nnorm = qnorm / sqrt(nnorm);
    double sp = 0.;
    for (int nu = 0; nu < OR_ND; nu++) sp += (q[nu] * x[nu] * nnorm);
    const double theta = acos(sp);
    const double delta_q = sin(theta) / qnorm;
    if ((delta_q <= width) && (theta <= arc * 0.01745329251994330)) return 1;
  }
  return 0;
}

int or_momentum_list(const or_lat *L, int type, size_t max_mom,
                     double cyl_width, double angle, int *mom, size_t *idx,
                     size_t cap, size_t *n)
{
  if (type < OR_CUT_HYPERCUBIC || type > OR_CUT_CONICAL) return -2; /* check_psq */
  size_t small = L->dims[0];
  for (int mu = 1; mu < OR_ND; mu++)
    if (L->dims[mu] < small) small = L->dims[mu];
  double rats[OR_ND];
  for (int mu = 0; mu < OR_ND; mu++) rats[mu] = (double)small / (double)L->dims[mu];
  const double latt_width = OR_TWOPI * cyl_width / small;
  size_t in = 0;
  for (size_t i = 0; i < L->V; i++) {
    int k[OR_ND];
    size_t subvol = 1;
    for (int mu = 0; mu < OR_ND; mu++) {
      k[mu] = (int)(((i - i % subvol) / subvol) % L->dims[mu]) -
              (int)(L->dims[mu] >> 1);
      subvol *= L->dims[mu];
    }
    double psq = 0.;
    for (int mu = 0; mu < OR_ND; mu++) {
      const double t = rats[mu] * k[mu];
      psq += t * t;
    }
    int keep = 0;
    if (type == OR_CUT_HYPERCUBIC) {
      const double root = sqrt((double)max_mom);
      int flag = 0;
      for (int mu = 0; mu < OR_ND; mu++)
        if (fabs((double)k[mu] * rats[mu]) <= root) flag++;
      keep = flag == OR_ND;
    } else if (type == OR_CUT_CYLINDER || type == OR_CUT_CONICAL) {
      if (psq <= (double)max_mom) {
        double q[OR_ND];
        for (int mu = 0; mu < OR_ND; mu++)
          q[mu] = 2. * sin((k[mu] * (OR_TWOPI / (double)L->dims[mu])) * 0.5);
        keep = type == OR_CUT_CYLINDER ? cylinder_df(q, latt_width)
                                       : cylinder_cjd(q, latt_width, angle);
      }
    } else { /* PSQ_CUT */
      keep = psq <= (double)max_mom;
    }
    if (!keep) continue;
    if (in < cap) {
      size_t site = 0, s = 1;
      for (int mu = 0; mu < OR_ND; mu++) {
        mom[in * OR_ND + mu] = k[mu];
        site += (size_t)(k[mu] < 0 ? k[mu] + (int)L->dims[mu] : k[mu]) * s;
        s *= L->dims[mu];
      }
      idx[in] = site;
    }
    in++;
  }
  *n = in;
  return in <= cap ? 0 : -1;
}

/* the Lie fields of cuts_struct up to the momentum-space Landau phase:
   A[(mu * 9 + j) * V + x], transformed and corrected at the listed sites */
static cplx *lie_fields(const or_lat *L, const double *U, size_t nm,
                        const int *mom, const size_t *idx, double tw[OR_ND])
{
  const size_t V = L->V;
  cplx *A = malloc(sizeof(cplx) * 36 * V);
  for (size_t i = 0; i < V; i++)
    for (int mu = 0; mu < OR_ND; mu++) {
      cplx Q[9];
      hermitian_proj(Q, (const cplx *)LINK(U, i, mu));
      for (int j = 0; j < 9; j++) A[(size_t)(mu * 9 + j) * V + i] = Q[j];
    }
  or_fft((double *)A, L->dims, OR_ND, +1, 36);
  for (int mu = 0; mu < OR_ND; mu++) tw[mu] = OR_TWOPI / (double)L->dims[mu];
  /* correct_pspace_landau */
  for (size_t i = 0; i < nm; i++)
    for (int mu = 0; mu < OR_ND; mu++) {
      const double cache = 0.5 * mom[i * OR_ND + mu] * tw[mu];
      const double p_mu = 2.0 * sin(cache);
      const cplx e = cos(cache) + I * 0.5 * p_mu;
      for (int j = 0; j < 9; j++) A[(size_t)(mu * 9 + j) * V + idx[i]] *= e;
    }
  return A;
}

/* psq_calc from the listed momentum.  GLU recovers it from the site
   through get_site_2piBZ (Geometry/geometry.c:42-61), the same map for
   even extents; for an odd extent L it sends the site of k = (L-1)/2
   through a negative coordinate and uses a neighbouring momentum. */
static double psq_of(const int *k, const double tw[OR_ND], double p[OR_ND])
{
  double tspsq = 0.;
  for (int mu = 0; mu < OR_ND; mu++) {
    p[mu] = 2.0 * sin(k[mu] * tw[mu] * 0.5);
    tspsq += p[mu] * p[mu];
  }
  return (tspsq < OR_PREC_TOL) ? 1.0 : tspsq;
}

#define AF(A, V, mu, j, site) (A)[(size_t)((mu) * 9 + (j)) * (V) + (site)]

int or_gluon_prop(const or_lat *L, const double *U, int type, size_t max_mom,
                  double cyl_width, double angle, int *mom, double *trans,
                  double *longi, size_t cap, size_t *n)
{
  const size_t V = L->V;
  size_t *idx = malloc(sizeof(size_t) * (cap ? cap : 1));
  const int rc = or_momentum_list(L, type, max_mom, cyl_width, angle, mom, idx, cap, n);
  if (rc) {
    free(idx);
    return rc;
  }
  const size_t nm = *n;
  double tw[OR_ND];
  cplx *A = lie_fields(L, U, nm, mom, idx, tw);
  const double g2_norm = 2.0 / ((OR_NC * OR_NC - 1) * (OR_ND - 1) * (double)V);
  const double g0_norm = 2.0 / ((OR_NC * OR_NC - 1) * (OR_ND) * (double)V);
  const double long_norm = 2.0 / ((OR_NC * OR_NC - 1) * (double)V);
  const size_t zero_pos = nm >> 1;
  for (size_t i = 0; i <= zero_pos && i < nm; i++) {
    const size_t posit = idx[i];
    double p[OR_ND];
    const double spsq = 1.0 / psq_of(mom + i * OR_ND, tw, p);
    double tr_ = 0.0, lo = 0.0;
    for (int mu = 0; mu < OR_ND; mu++)
      for (int nu = 0; nu < OR_ND; nu++) {
        cplx tr = 0.0;
        for (int j = 0; j < 9; j++)
          tr += AF(A, V, mu, j, posit) * conj(AF(A, V, nu, j, posit));
        const double common = p[mu] * p[nu] * spsq * creal(tr);
        tr_ += (mu != nu) ? -common : creal(tr) - common;
        lo += common;
      }
    trans[i] = tr_ * (i == zero_pos ? g0_norm : g2_norm);
    longi[i] = lo * long_norm;
  }
  for (size_t i = zero_pos + 1; i < nm; i++) {
    trans[i] = trans[nm - i - 1];
    longi[i] = longi[nm - i - 1];
  }
  free(A);
  free(idx);
  return 0;
}

/* 3x3 helpers of Matrix_Ops/U_Nops.c and trace_abc.c: c = a b, tr(a b) */
static void mul33(cplx *c, const cplx *a, const cplx *b)
{
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++)
      c[r * 3 + s] = a[r * 3] * b[s] + a[r * 3 + 1] * b[3 + s] + a[r * 3 + 2] * b[6 + s];
}
static cplx tr_ab(const cplx *a, const cplx *b)
{
  cplx t = 0.0;
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) t += a[r * 3 + s] * b[s * 3 + r];
  return t;
}

/* write_exceptional_g2g3_MOMgg (Cuts/MOMgg.c:141-215) with the default
   contraction_matrices (:94-121): for every listed p, with -p the mirror
   entry of the list and 0 its middle,
     g2 = sum_mu Re tr(A_mu(p) A_mu(-p)),
     g3 = Re [tr(AA Arho) + tr(Arho AP AM)] / p^2 with AA = sum_mu A_mu(p)
     A_mu(-p), Arho = sum_mu p_mu A_mu(0), AP = sum_mu p_mu A_mu(p),
     AM = -sum_mu p_mu A_mu(-p) / p^2 */
int or_gluon_exceptional(const or_lat *L, const double *U, int type,
                         size_t max_mom, double cyl_width, double angle,
                         int *mom, double *g2, double *g3, size_t cap, size_t *n)
{
  const size_t V = L->V;
  size_t *idx = malloc(sizeof(size_t) * (cap ? cap : 1));
  const int rc = or_momentum_list(L, type, max_mom, cyl_width, angle, mom, idx, cap, n);
  if (rc) {
    free(idx);
    return rc;
  }
  const size_t nm = *n;
  double tw[OR_ND];
  cplx *A = lie_fields(L, U, nm, mom, idx, tw);
  const double g2_norm = 2.0 / ((OR_NC * OR_NC - 1) * (OR_ND - 1) * (double)V);
  const double g0_norm = 2.0 / ((OR_NC * OR_NC - 1) * (OR_ND) * (double)V);
  const double g3_norm = 4.0 / (2.0 * OR_NC * (OR_NC * OR_NC - 1) * (OR_ND - 1) * (double)V);
  const size_t zero = nm ? (nm - 1) / 2 : 0;
  const size_t lzero = nm ? idx[zero] : 0;
  for (size_t i = 0; i < nm; i++) {
    const size_t posit = idx[i], cj = idx[nm - i - 1];
    double s2 = 0.0;
    cplx AA[9] = {0}, AM[9] = {0}, Arho[9] = {0}, AP[9] = {0};
    double p[OR_ND];
    const double spsq = psq_of(mom + i * OR_ND, tw, p);
    for (int mu = 0; mu < OR_ND; mu++) {
      cplx ap[9], am[9], a0[9], t[9];
      for (int j = 0; j < 9; j++) {
        ap[j] = AF(A, V, mu, j, posit);
        am[j] = AF(A, V, mu, j, cj);
        a0[j] = AF(A, V, mu, j, lzero);
      }
      s2 += creal(tr_ab(ap, am));
      mul33(t, ap, am);
      for (int j = 0; j < 9; j++) {
        AA[j] += t[j];
        Arho[j] += a0[j] * p[mu];
        AP[j] += ap[j] * p[mu];
        AM[j] += am[j] * (-p[mu] / spsq);
      }
    }
    g2[i] = s2 * (i != zero ? g2_norm : g0_norm);
    cplx t[9];
    mul33(t, Arho, AP);
    const cplx trans = tr_ab(AA, Arho), wiv = tr_ab(t, AM);
    g3[i] = creal((trans + wiv) / spsq) * g3_norm;
  }
  free(A);
  free(idx);
  return 0;
}

static void put32(FILE *f, uint32_t v)
{
  v = __builtin_bswap32(v);
  fwrite(&v, 4, 1, f);
}
static void put64(FILE *f, double d)
{
  uint64_t v;
  memcpy(&v, &d, 8);
  v = __builtin_bswap64(v);
  fwrite(&v, 8, 1, f);
}

int or_write_glueprop(const char *path, size_t n, const int *mom,
                      const double *trans, const double *longi)
{
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  put32(f, (uint32_t)n);
  for (size_t i = 0; i < n; i++) {
    put32(f, OR_ND);
    for (int mu = 0; mu < OR_ND; mu++) put32(f, (uint32_t)mom[i * OR_ND + mu]);
  }
  put32(f, (uint32_t)n);
  for (size_t i = 0; i < n; i++) put64(f, trans[i]);
  put32(f, (uint32_t)n);
  for (size_t i = 0; i < n; i++) put64(f, longi[i]);
  return fclose(f) == 0 ? 0 : -2;
}
