/*
 * or_fft.c -- oracle: complex DFTs and the Fourier-acceleration step.
 * TEST INFRASTRUCTURE ONLY (see glu_oracle.h).
 *
 * GLU calls FFTW here (Gfix/FACG.c:77-95, plans in
 * Geometry/plan_ffts.c:103-168: unnormalised complex DFTs over the whole
 * lattice, FFTW_FORWARD = exp(-2 pi i k.n / L) and FFTW_BACKWARD with +).
 * FFTW is absent from this image; this file restates the mathematical DFT
 * FFTW computes: a recursive mixed-radix decimation in time whose twiddles
 * come straight from cos/sin, plus an O(V^2) definition-level DFT used only
 * to check it.
 */
#include "or_internal.h"

static size_t smallest_factor(size_t n)
{
  if (n % 4 == 0 && n != 4) return 4;
  for (size_t p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

/* out[k] = sum_n in[n*is] w^(n k) ; w[j] = exp(sign 2 pi i j / Ntot) and
   this level works on N = Ntot / ws points */
static void fft_rec(const cplx *in, size_t is, cplx *out, size_t N,
                    const cplx *w, size_t ws, size_t Ntot, cplx *tmp)
{
  if (N == 1) {
    out[0] = in[0];
    return;
  }
  const size_t p = smallest_factor(N), m = N / p;
  for (size_t r = 0; r < p; r++)
    fft_rec(in + r * is, is * p, out + r * m, m, w, ws * p, Ntot, tmp);
  cplx t[64];
  cplx *tt = p <= 64 ? t : tmp;
  for (size_t k = 0; k < m; k++) {
    for (size_t r = 0; r < p; r++)
      tt[r] = out[r * m + k] * w[(r * k * ws) % Ntot];
    for (size_t q = 0; q < p; q++) {
      cplx s = 0.0;
      for (size_t r = 0; r < p; r++) s += tt[r] * w[(r * q * m * ws) % Ntot];
      tmp[p + q] = s; /* staging after tt (tt may alias tmp[0..p)) */
    }
    for (size_t q = 0; q < p; q++) out[k + q * m] = tmp[p + q];
  }
}

/* 0: the mixed-radix decimation in time above; 1: the direct sum
   out[k] = sum_n in[n] w[(n k) mod N], n ascending, along each axis --
   another exact DFT with rounding of its own, to measure how far a
   trajectory depends on the arithmetic of the transform (FFTW's codelets,
   which GLU calls, are a third such arithmetic) */
static int or_fft_mode = 0;
void or_set_fft_mode(int mode) { or_fft_mode = mode; }

/* in-place unnormalised multi-dimensional DFT over nd axes */
void or_fft(double *data, const size_t *dims, int nd, int sign, size_t nb)
{
  size_t vol = 1;
  for (int a = 0; a < nd; a++) vol *= dims[a];
  cplx *D = (cplx *)data;
  size_t stride = 1;
  for (int a = 0; a < nd; a++) {
    const size_t N = dims[a];
    cplx *w = malloc(N * sizeof(cplx));
    for (size_t j = 0; j < N; j++) {
      const double ang = OR_TWOPI * (double)j / (double)N;
      w[j] = cos(ang) + I * (sign * sin(ang));
    }
    const size_t npencil = vol / N;
#pragma omp parallel
    {
      cplx *buf = malloc(N * sizeof(cplx)), *res = malloc(N * sizeof(cplx)),
           *tmp = malloc(2 * (N + 1) * sizeof(cplx));
#pragma omp for schedule(static) collapse(2)
      for (size_t b = 0; b < nb; b++)
        for (size_t pc = 0; pc < npencil; pc++) {
          const size_t lo = pc % stride, hi = pc / stride;
          cplx *base = D + b * vol + lo + hi * stride * N;
          for (size_t n = 0; n < N; n++) buf[n] = base[n * stride];
          if (or_fft_mode == 1) {
            for (size_t k = 0; k < N; k++) {
              cplx acc = 0.0;
              for (size_t n = 0; n < N; n++) acc += buf[n] * w[(n * k) % N];
              res[k] = acc;
            }
          } else {
            fft_rec(buf, 1, res, N, w, 1, N, tmp);
          }
          for (size_t n = 0; n < N; n++) base[n * stride] = res[n];
        }
      free(buf);
      free(res);
      free(tmp);
    }
    free(w);
    stride *= N;
  }
}

/* definition-level DFT, O(vol^2): for checking or_fft only */
void or_dft_naive(double *data, const size_t *dims, int nd, int sign)
{
  size_t vol = 1;
  for (int a = 0; a < nd; a++) vol *= dims[a];
  cplx *D = (cplx *)data, *out = calloc(vol, sizeof(cplx));
  for (size_t k = 0; k < vol; k++) {
    size_t kc[8], kk = k;
    for (int a = 0; a < nd; a++) { kc[a] = kk % dims[a]; kk /= dims[a]; }
    cplx s = 0.0;
    for (size_t n = 0; n < vol; n++) {
      size_t nn = n;
      double ph = 0.0;
      for (int a = 0; a < nd; a++) {
        const size_t na = nn % dims[a];
        nn /= dims[a];
        ph += (double)((na * kc[a]) % dims[a]) / (double)dims[a];
      }
      s += D[n] * (cos(OR_TWOPI * ph) + I * (sign * sin(OR_TWOPI * ph)));
    }
    out[k] = s;
  }
  memcpy(D, out, vol * sizeof(cplx));
  free(out);
}

/* psq[i] = MAX / p^2(i), MAX = 4 D / vol (the 1/vol of the unnormalised
   transforms folded in): Gfix/Landau.c:278-284, Gfix/Coulomb.c:55-59,
   MAX_LANDAU / MAX_COULOMB at Headers/GLU_definitions.h:414-420 */
void or_psq_table(const or_lat *L, int D, double *psq)
{
  const size_t vol = D == OR_ND ? L->V : L->Lcu;
  const double mx = 4.0 * (or_plant == OR_PLANT_MAX_FACTOR ? D - 1 : D) / (double)vol;
  for (size_t i = 0; i < vol; i++) psq[i] = mx / or_psq(L, i, D);
}

/* forward DFT, times psq, backward DFT, for each packed field:
   Gfix/FACG.c:77-95 (FOURIER_ACCELERATE3), Gfix/CFACG.c:70-89 (...2) */
void or_fourier_accelerate(const or_lat *L, double *F, const double *psq,
                           int D, size_t nb)
{
  const size_t vol = D == OR_ND ? L->V : L->Lcu;
  or_fft(F, L->dims, D, -1, nb);
  cplx *c = (cplx *)F;
#pragma omp parallel for schedule(static)
  for (size_t j = 0; j < nb * vol; j++) c[j] *= psq[j % vol];
  or_fft(F, L->dims, D, +1, nb);
}
