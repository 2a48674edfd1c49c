/*
 * or_rng.c -- oracle: GLU's parallel MWC_4096 generator, Haar-like random
 * SU(3) matrices and the random gauge transform.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * GLU keeps one generator per OpenMP thread (Utils/par_rng.c:43-132,
 * Utils/par_MWC_4096.c).  Its index is reset to 4095 on every call
 * (par_MWC_4096.c:35-36) so only table[0] is ever used: a lag-1
 * complementary multiply-with-carry.  Thread t is seeded with SEED + t,
 * warmed up by 10000 draws, and consumes its static block of sites in order.
 */
#include "or_internal.h"

struct or_rng {
  int nth;
  uint32_t *x;     /* table[t][0] */
  uint32_t *carry; /* carry[t]    */
};

/* Utils/par_MWC_4096.c:33-46 */
static uint32_t mwc(or_rng *R, int t)
{
  const uint64_t prod = 18782ULL * (uint64_t)R->x[t] + R->carry[t];
  R->carry[t] = (uint32_t)(prod >> 32);
  uint32_t x = (uint32_t)prod + R->carry[t];
  if (x < R->carry[t]) {
    x++;
    R->carry[t]++;
  }
  R->x[t] = 0xfffffffeu - x;
  return R->x[t];
}

/* x * 2^-32: Utils/par_MWC_4096.c:93-112 */
double or_rng_dbl(or_rng *R, int t)
{
  return (double)(mwc(R, t) * 2.3283064365386963e-10);
}

/* seeding (par_MWC_4096.c:72-89) and warm-up (par_rng.c:113-120) */
or_rng *or_rng_new(uint32_t seed, int nth)
{
  or_rng *R = malloc(sizeof(*R));
  R->nth = nth;
  R->x = malloc(nth * sizeof(uint32_t));
  R->carry = malloc(nth * sizeof(uint32_t));
  uint32_t *tab = malloc(4096 * sizeof(uint32_t));
  for (int t = 0; t < nth; t++) {
    tab[0] = seed + (uint32_t)t;
    for (uint32_t j = 1; j < 4096; j++)
      tab[j] = (uint32_t)(1812433253UL * (tab[j - 1] ^ (tab[j - 1] >> 30)) + j);
    R->x[t] = tab[0];
    R->carry[t] = tab[4095] % 809430660UL;
    for (int j = 0; j < 10000; j++) or_rng_dbl(R, t);
  }
  free(tab);
  return R;
}

void or_rng_free(or_rng *R)
{
  if (!R) return;
  free(R->x);
  free(R->carry);
  free(R);
}

/* Marsaglia polar with rejection, no factor 2 under the root:
   Utils/par_rng.c:180-186 */
static cplx polar_box(or_rng *R, int t)
{
  for (;;) {
    const double u = 2.0 * or_rng_dbl(R, t) - 1.0;
    const double v = 2.0 * or_rng_dbl(R, t) - 1.0;
    const double s = u * u + v * v;
    if (s < 1.0) return sqrt(-log(s) / s) * (u + I * v);
  }
}

/* 9 complex Gaussians then Gram-Schmidt: Matrix_Ops/gramschmidt_SSE.c:175-181,
   Utils/par_rng.c:159-168 */
void or_sunitary_gen(or_rng *R, int t, double *Z18)
{
  cplx *Z = (cplx *)Z18;
  for (int k = 0; k < 9; k++) Z[k] = polar_box(R, t);
  or_gram_reunit(Z18);
}

/* U_mu(x) <- G(x) U_mu(x) G(x+mu)^dagger: Gfix/gtrans.c:169-189 */
void or_gtransform(const or_lat *L, double *U, const double *G)
{
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < L->V; i++)
    for (int mu = 0; mu < OR_ND; mu++)
      or_gtransform_local(G + i * 18, LINK(U, i, mu),
                          G + L->fwd[i * OR_ND + mu] * 18);
}

/* random_gtrans: Geometry/random_config.c:61-95 (thread t draws its
   static block of sites in order) */
void or_random_gtrans(const or_lat *L, or_rng *R, double *U)
{
  double *G = malloc(L->V * 18 * sizeof(double));
  const int nth = R->nth;
  OR_THREADS(nth, L->V, lo, hi, th) {
    for (size_t i = lo; i < hi; i++) or_sunitary_gen(R, th, G + i * 18);
  }
  or_gtransform(L, U, G);
  free(G);
}
