/*
 * or_internal.h -- private helpers of the CPU oracle (test infrastructure).
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "glu_oracle.h"

typedef double complex cplx;

#define OR_TWOPI 6.283185307179586 /* Headers/GLU_definitions.h:139 */

/* reduction-array layout of GLU (Headers/GLU_definitions.h:460-471):
   per virtual thread th a stride of CLINE doubles; slots 0..2 hold the
   three line-search probes, 3 and 4 the Polak-Ribiere sums, 5 theta */
#define OR_NSTEPS 3
#define OR_CLINE 8
#define OR_SLOT_PR0 (OR_NSTEPS + 0)
#define OR_SLOT_PR1 (OR_NSTEPS + 1)
#define OR_SLOT_TH (OR_NSTEPS + 2)

/* Landau/Coulomb constants: Headers/GLU_definitions.h:401-454 */
#define OR_CG_TOL 5e-11
#define OR_GF_FAILURES 7
#define OR_PREC_TOL (3 * 1.0e-14)

#define LINK(U, i, mu) ((U) + ((size_t)(i) * OR_ND + (size_t)(mu)) * OR_LINK)

/* libgomp's static schedule without a chunk size: the first n % nth threads
   take one extra iteration */
static inline size_t or_blk_lo(size_t n, int nth, int t)
{
  size_t q = n / (size_t)nth, r = n % (size_t)nth;
  if ((size_t)t < r) { q++; r = 0; }
  return q * (size_t)t + r;
}
static inline size_t or_blk_hi(size_t n, int nth, int t)
{
  size_t q = n / (size_t)nth, r = n % (size_t)nth;
  if ((size_t)t < r) { q++; r = 0; }
  return q * (size_t)t + r + q;
}

/* run the body once per virtual thread `th` over its static block
   [lo, hi) of n iterations; blocks run concurrently under OpenMP */
#define OR_THREADS(nth, n, lo, hi, th)                                      \
  _Pragma("omp parallel for schedule(static, 1)")                          \
  for (int th = 0; th < (nth); th++)                                       \
    for (size_t lo = or_blk_lo((n), (nth), th), hi = or_blk_hi((n), (nth), th), \
                _once_##th = 1;                                             \
         _once_##th; _once_##th = 0)

double or_fa_deriv_red(const or_lat *L, const double *U, double *F,
                       double *red);

/* the planted mistake in force (OR_PLANT_*; 0 in every faithful run) */
extern int or_plant;

#endif
