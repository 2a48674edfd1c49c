// glu_compat.hip -- GLU drop-in entry points: Landau(), Coulomb(), FACG(),
// FASD() and Coulomb_FA() with GLU's prototypes, reading GLU's global
// `Latt` and overwriting the links of `lat` in place.
//
// The host side mirrors GLU's drivers (Gfix/Landau.c:261-374,
// Gfix/Coulomb.c:33-100, Gfix/FACG.c:341-446): the same theta test,
// continuation, random restarts and [GF] report lines.  The work itself
// runs on the device through the glu_gfix_* context API.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

// GLU defines `Latt` (Run/Mainfile.c); standalone users get this weak one.
extern "C" {
__attribute__((weak)) struct latt_info Latt;
}

namespace {

using cd = std::complex<double>;

size_t volume() { return Latt.dims[0] * Latt.dims[1] * Latt.dims[2] * Latt.dims[3]; }

int precision_from_env() {
  const char *p = getenv("GLU_GFIX_PRECISION");
  return (p && (!strcmp(p, "mixed") || !strcmp(p, "MIXED"))) ? GLU_GFIX_MIXED : GLU_GFIX_FP64;
}

// one cached device context per (dims, gftype, precision)
glu_gfix_ctx *g_ctx = nullptr;
size_t g_dims[4];
int g_type = -1, g_prec = -1;

glu_gfix_ctx *get_ctx(int gftype) {
  const int prec = precision_from_env();
  if (g_ctx && g_type == gftype && g_prec == prec && !memcmp(g_dims, Latt.dims, sizeof(g_dims))) return g_ctx;
  if (g_ctx) glu_gfix_destroy(g_ctx);
  g_ctx = nullptr;
  if (glu_gfix_create(&g_ctx, Latt.dims, gftype, prec) != 0) {
    fprintf(stderr, "[GF] device context creation failed: %s\n", glu_gfix_last_error());
    g_ctx = nullptr;
    return nullptr;
  }
  g_ctx->verbose = 1;  // GLU's [GF] / [CG] progress lines
  // launch shapes timed once per shape and process (GLU_GFIX_AUTOTUNE=1):
  // worth it when a process fixes several configurations, as a GLU run
  // over many files does; the defaults serve a single call
  if (const char *e = getenv("GLU_GFIX_AUTOTUNE"))
    if (atoi(e) > 0 && glu_gfix_autotune(g_ctx, nullptr, 0) != 0)
      fprintf(stderr, "[GF] autotune failed (defaults kept): %s\n", glu_gfix_last_error());
  memcpy(g_dims, Latt.dims, sizeof(g_dims));
  g_type = gftype;
  g_prec = prec;
  return g_ctx;
}

void unpack(struct site *lat, const std::vector<double> &h) {
  const size_t V = volume();
  for (size_t i = 0; i < V; i++)
    for (int mu = 0; mu < 4; mu++) memcpy(lat[i].O[mu], &h[(i * 4 + mu) * 18], 18 * sizeof(double));
}

// struct site links straight to and from the device, through the pinned,
// double-buffered host transfers (several host threads gather or scatter
// each chunk while the device copies the previous one)
int sites_to_device(glu_gfix_ctx *c, const struct site *lat) {
  return upload_links(
      c,
      [](void *arg, double *dst, long long s0, long long n) {
        const struct site *l = (const struct site *)arg;
        for (long long i = s0; i < s0 + n; i++)
          for (int mu = 0; mu < 4; mu++) memcpy(dst + ((i - s0) * 4 + mu) * 18, l[i].O[mu], 18 * sizeof(double));
      },
      (void *)lat);
}
int device_to_sites(glu_gfix_ctx *c, struct site *lat) {
  return download_links(
      c,
      [](void *arg, const double *src, long long s0, long long n) {
        struct site *l = (struct site *)arg;
        for (long long i = s0; i < s0 + n; i++)
          for (int mu = 0; mu < 4; mu++) memcpy(l[i].O[mu], src + ((i - s0) * 4 + mu) * 18, 18 * sizeof(double));
      },
      lat);
}

// ------------------------------------------------------------------------
// The random gauge transform of the Landau restarts (Utils/par_rng.c:180-186,
// Geometry/random_config.c:108-120), drawn from GLU's process stream
// (mwc.hip: one MWC_4096 stream per GLU thread, seeded SEED + thread, each
// thread drawing its static block of sites).
// ------------------------------------------------------------------------
int rng_init() { return mwc_host_init(Latt.Seed[0], Latt.Nthreads ? (int)Latt.Nthreads : 1); }
double mwc_dbl(int t) { return mwc_host_dbl(t); }

cd polar_box(int t) {
  for (;;) {
    const double u = 2.0 * mwc_dbl(t) - 1.0, v = 2.0 * mwc_dbl(t) - 1.0;
    const double s = u * u + v * v;
    if (s < 1.0) return sqrt(-log(s) / s) * cd(u, v);
  }
}

void reunit_host(cd *U) {
  double n = std::norm(U[0]) + std::norm(U[3]) + std::norm(U[6]);
  n = 1.0 / sqrt(n);
  U[0] *= n;
  U[3] *= n;
  U[6] *= n;
  const cd p = std::conj(U[0]) * U[1] + std::conj(U[3]) * U[4] + std::conj(U[6]) * U[7];
  U[1] -= p * U[0];
  U[4] -= p * U[3];
  U[7] -= p * U[6];
  double m = std::norm(U[1]) + std::norm(U[4]) + std::norm(U[7]);
  m = 1.0 / sqrt(m);
  U[1] *= m;
  U[4] *= m;
  U[7] *= m;
  U[2] = std::conj(U[3] * U[7] - U[6] * U[4]);
  U[5] = std::conj(U[6] * U[1] - U[0] * U[7]);
  U[8] = std::conj(U[0] * U[4] - U[3] * U[1]);
}

// b <- a b c^dagger (columns 0,1 then completion), Gfix/gtrans.c:32-141
void gtrans_host(const cd *a, cd *b, const cd *c) {
  cd t[9], o[9];
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 2; s++)
      t[r * 3 + s] = b[r * 3] * std::conj(c[s * 3]) + b[r * 3 + 1] * std::conj(c[s * 3 + 1]) +
                     b[r * 3 + 2] * std::conj(c[s * 3 + 2]);
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 2; s++) o[r * 3 + s] = a[r * 3] * t[s] + a[r * 3 + 1] * t[3 + s] + a[r * 3 + 2] * t[6 + s];
  o[2] = std::conj(o[3] * o[7] - o[4] * o[6]);
  o[5] = std::conj(o[1] * o[6] - o[0] * o[7]);
  o[8] = std::conj(o[0] * o[4] - o[1] * o[3]);
  memcpy(b, o, sizeof(o));
}

int random_transform(std::vector<double> &h) {
  if (rng_init()) return -1;
  const size_t V = volume();
  const int nth = (int)g_mwc.x.size();
  std::vector<cd> G(V * 9);
  for (int t = 0; t < nth; t++) {  // libgomp static blocks
    size_t q = V / nth, r = V % nth;
    if ((size_t)t < r) {
      q++;
      r = 0;
    }
    const size_t lo = q * t + r, hi = lo + q;
    for (size_t i = lo; i < hi; i++) {
      for (int k = 0; k < 9; k++) G[i * 9 + k] = polar_box(t);
      reunit_host(&G[i * 9]);
    }
  }
  size_t S[4] = {1, Latt.dims[0], Latt.dims[0] * Latt.dims[1], Latt.dims[0] * Latt.dims[1] * Latt.dims[2]};
  for (size_t i = 0; i < V; i++) {
    size_t x[4], r = i;
    for (int mu = 0; mu < 4; mu++) {
      x[mu] = r % Latt.dims[mu];
      r /= Latt.dims[mu];
    }
    for (int mu = 0; mu < 4; mu++) {
      const size_t f = x[mu] + 1 < Latt.dims[mu] ? i + S[mu] : i - (Latt.dims[mu] - 1) * S[mu];
      gtrans_host(&G[i * 9], (cd *)&h[(i * 4 + mu) * 18], &G[f * 9]);
    }
  }
  return 0;
}

// ------------------------------------------------------------------------
// NERSC reader for the restarts (IO/read_headers.c:229-362,
// IO/readers.c:51-58, 170-302): 3x2 or 3x3, IEEE32/64, either endianness
// ------------------------------------------------------------------------
int read_nersc(const char *path, std::vector<double> &h) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  char line[1024];
  std::string hdr;
  bool end = false;
  while (fgets(line, sizeof(line), f)) {
    hdr += line;
    if (!strncmp(line, "END_HEADER", 10)) {
      end = true;
      break;
    }
  }
  auto value = [&](const char *key) -> std::string {
    size_t p = 0;
    const std::string k(key);
    while ((p = hdr.find(k, p)) != std::string::npos) {
      if (p == 0 || hdr[p - 1] == '\n') {
        size_t q = p + k.size();
        while (q < hdr.size() && (hdr[q] == ' ' || hdr[q] == '\t')) q++;
        if (q < hdr.size() && hdr[q] == '=') {
          q++;
          while (q < hdr.size() && (hdr[q] == ' ' || hdr[q] == '\t')) q++;
          size_t e = hdr.find('\n', q);
          std::string v = hdr.substr(q, e - q);
          while (!v.empty() && (v.back() == ' ' || v.back() == '\r' || v.back() == '\t')) v.pop_back();
          return v;
        }
      }
      p += k.size();
    }
    return "";
  };
  if (!end) {
    fclose(f);
    return -2;
  }
  const std::string fp = value("FLOATING_POINT"), dt = value("DATATYPE");
  const bool dbl = fp.find("64") != std::string::npos;
  const bool big = fp.find("BIG") != std::string::npos || fp == "IEEE32";
  const bool two = dt == "4D_SU3_GAUGE";
  if (!two && dt != "4D_SU3_GAUGE_3x3") {
    fclose(f);
    return -3;
  }
  const size_t V = volume(), per = two ? 12 : 18, n = V * 4 * per, w = dbl ? 8 : 4;
  std::vector<unsigned char> buf(n * w);
  if (fread(buf.data(), w, n, f) != n) {
    fclose(f);
    return -4;
  }
  fclose(f);
  h.assign(V * 72, 0.0);
  for (size_t k = 0; k < n; k++) {
    double d;
    if (dbl) {
      uint64_t u;
      memcpy(&u, &buf[8 * k], 8);
      if (big) u = __builtin_bswap64(u);
      memcpy(&d, &u, 8);
    } else {
      uint32_t u;
      float fl;
      memcpy(&u, &buf[4 * k], 4);
      if (big) u = __builtin_bswap32(u);
      memcpy(&fl, &u, 4);
      d = fl;
    }
    h[(k / per) * 18 + k % per] = d;
  }
  if (two)
    for (size_t l = 0; l < V * 4; l++) {
      cd *O = (cd *)&h[l * 18];
      O[6] = std::conj(O[1] * O[5] - O[2] * O[4]);
      O[7] = std::conj(O[2] * O[3] - O[0] * O[5]);
      O[8] = std::conj(O[0] * O[4] - O[1] * O[3]);
    }
  return 0;
}

// ------------------------------------------------------------------------
// report helpers
// ------------------------------------------------------------------------
// complex arithmetic of the report's logarithm, written out: the library
// multiply and divide of std::complex take the C99 Annex G slow path for
// non-finite values on every call, and 2 x 10^6 logarithms per report add
// up; the values are finite here and the formulas the usual ones
struct rc {
  double re, im;
};
inline rc operator+(rc a, rc b) { return {a.re + b.re, a.im + b.im}; }
inline rc operator-(rc a, rc b) { return {a.re - b.re, a.im - b.im}; }
inline rc operator*(rc a, rc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline rc operator*(double s, rc a) { return {s * a.re, s * a.im}; }
inline rc conjr(rc a) { return {a.re, -a.im}; }
inline double normr(rc a) { return a.re * a.re + a.im * a.im; }
inline rc operator/(rc a, rc b) {
  const double n = normr(b);
  return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}

// -i log(U) of a unitary 3x3 matrix from its eigen-decomposition (the
// "Log" column of the temporal-constancy line; GLU uses exact_log_slow)
void log_herm(const cd *Uc, cd *Qc) {
  rc U[9];
  for (int e = 0; e < 9; e++) U[e] = {Uc[e].real(), Uc[e].imag()};
  // characteristic polynomial l^3 - a l^2 + b l - d
  const rc a = U[0] + U[4] + U[8];
  const rc b = U[0] * U[4] - U[1] * U[3] + U[0] * U[8] - U[2] * U[6] + U[4] * U[8] - U[5] * U[7];
  const rc d = U[0] * (U[4] * U[8] - U[5] * U[7]) - U[1] * (U[3] * U[8] - U[5] * U[6]) +
               U[2] * (U[3] * U[7] - U[4] * U[6]);
  rc l[3] = {{1, 0}, {-0.5, 0.8660254037844386}, {-0.5, -0.8660254037844386}};
  // Durand-Kerner, until every correction is below 1e-15 (the eigenvalues
  // lie on the unit circle; convergence is quadratic, a dozen sweeps)
  for (int it = 0; it < 200; it++) {
    double step = 0.0;
    for (int k = 0; k < 3; k++) {
      const rc num = ((l[k] - a) * l[k] + b) * l[k] - d;
      rc den = {1, 0};
      for (int j = 0; j < 3; j++)
        if (j != k) den = den * (l[k] - l[j]);
      if (normr(den) > 0) {
        const rc c = num / den;
        l[k] = l[k] - c;
        step = std::max(step, std::sqrt(normr(c)));
      }
    }
    if (step < 1e-15) break;
  }
  rc Q[9];
  for (int k = 0; k < 9; k++) Q[k] = {0, 0};
  for (int k = 0; k < 3; k++) {
    // eigenvector: cross product of two rows of (U - l I)
    rc M[9];
    for (int e = 0; e < 9; e++) M[e] = U[e];
    M[0] = M[0] - l[k];
    M[4] = M[4] - l[k];
    M[8] = M[8] - l[k];
    rc best[3] = {{0, 0}, {0, 0}, {0, 0}};
    double bn = -1;
    for (int r1 = 0; r1 < 3; r1++)
      for (int r2 = r1 + 1; r2 < 3; r2++) {
        const rc *p = M + 3 * r1, *q = M + 3 * r2;
        const rc v[3] = {conjr(p[1] * q[2] - p[2] * q[1]), conjr(p[2] * q[0] - p[0] * q[2]),
                         conjr(p[0] * q[1] - p[1] * q[0])};
        const double nn = normr(v[0]) + normr(v[1]) + normr(v[2]);
        if (nn > bn) {
          bn = nn;
          best[0] = v[0];
          best[1] = v[1];
          best[2] = v[2];
        }
      }
    if (bn <= 0) continue;
    const double s = 1.0 / sqrt(bn);
    const double th = std::atan2(l[k].im, l[k].re);
    for (int r = 0; r < 3; r++)
      for (int c2 = 0; c2 < 3; c2++) Q[r * 3 + c2] = Q[r * 3 + c2] + th * ((s * best[r]) * conjr(s * best[c2]));
  }
  for (int e = 0; e < 9; e++) Qc[e] = cd(Q[e].re, Q[e].im);
}

// Gfix/gftests.c:26-130: average difference of the slice sums of the
// traceless Hermitian projection of the temporal links from slice 0
void const_time(const struct site *lat, double *lin, double *lg) {
  const size_t Lcu = Latt.dims[0] * Latt.dims[1] * Latt.dims[2], T = Latt.dims[3];
  auto slice_sum = [&](size_t t, cd *s, cd *sl) {
    for (int k = 0; k < 9; k++) s[k] = sl[k] = 0.0;
    for (size_t j = Lcu * t; j < Lcu * t + Lcu; j++) {
      const cd *U = (const cd *)lat[j].O[3];
      cd A[9];
      const double i0 = U[0].imag(), i4 = U[4].imag(), i8 = U[8].imag();
      A[0] = (2 * i0 - i4 - i8) / 3.0;
      A[1] = cd(0, -1) * (U[1] - std::conj(U[3])) / 2.0;
      A[2] = cd(0, -1) * (U[2] - std::conj(U[6])) / 2.0;
      A[4] = (2 * i4 - i0 - i8) / 3.0;
      A[5] = cd(0, -1) * (U[5] - std::conj(U[7])) / 2.0;
      for (int k : {0, 1, 2, 4, 5}) s[k] += A[k];
      cd Q[9];
      log_herm(U, Q);
      for (int k = 0; k < 9; k++) sl[k] += Q[k];
    }
  };
  // each slice's sums in its own order, the slices on host threads (an
  // exact logarithm per temporal link is the cost); then combined in slice
  // order, so the result does not depend on the thread count
  std::vector<cd> S(T * 9), SL(T * 9);
  {
    const size_t nt = std::min<size_t>(T, std::max(1u, std::min(std::thread::hardware_concurrency(), 16u)));
    std::vector<std::thread> th;
    for (size_t w = 0; w < nt; w++)
      th.emplace_back([&, w] {
        for (size_t t = w; t < T; t += nt) slice_sum(t, &S[t * 9], &SL[t * 9]);
      });
    for (std::thread &x : th) x.join();
  }
  const cd *s1 = &S[0], *l1 = &SL[0];
  double av = 0, avl = 0;
  for (size_t t = 1; t < T; t++) {
    const cd *s2 = &S[t * 9], *l2 = &SL[t * 9];
    av += fabs(s1[0].real() - s2[0].real()) + std::abs(s1[1] - s2[1]) + std::abs(s1[2] - s2[2]) +
          fabs(s1[4].real() - s2[4].real()) + std::abs(s1[5] - s2[5]);
    avl += fabs(l1[0].real() - l2[0].real()) + std::abs(l1[1] - l2[1]) + std::abs(l1[2] - l2[2]) +
           fabs(l1[4].real() - l2[4].real()) + std::abs(l1[5] - l2[5]);
  }
  const double check = 1.0 / (9.0 * (double)(T - 1));
  *lin = av * check;
  *lg = avl * check;
}

// Gfix/Landau.c:43-66
void output_fixing_info(glu_gfix_ctx *c, struct site *lat, double theta, size_t iters) {
  glu_gfix_reunit(c);
  double plaq = 0, tr = 0, mx = 0, func = 0, tl = 0;
  glu_gfix_measure(c, &plaq, &tr, &mx, &func, &tl);
  if (device_to_sites(c, lat)) fprintf(stderr, "[GF] link download failed: %s\n", glu_gfix_last_error());
  printf("[GF] Plaquette :: %1.15f \n[GF] Accuracy :: %1.4e\n", plaq, theta);
  printf("[GF] Iters :: %zu\n[GF] Link trace :: %1.15f || Maximum :: %1.15f\n", iters, tr, mx);
  double lin = 0, lg = 0;
  const_time(lat, &lin, &lg);
  printf("[GF] Temporal constance || Lin %e || Log %e \n", lin, lg);
  printf("[GF] Functional :: %1.15f\n", func);
}

// LUXURY_GAUGE copies requested at run time (GLU_GFIX_LUXURY=N; a value
// that is not a positive number means GLU's default of 20, configure.ac:493)
int luxury_copies() {
  const char *e = getenv("GLU_GFIX_LUXURY");
  if (!e) return 0;
  const int n = atoi(e);
  return n > 0 ? n : 20;
}

// FA_callback of Gfix/Landau.c:225-237: FACG, FASD, or luxury_copy_fast.
// The fixed links stay on the device, and in lat too unless the caller
// takes them later (download false: Landau() reunitarises them first)
size_t fa_call(struct site *lat, double *th, double acc, size_t max_iters, int method, bool callback,
               bool download = true) {
  glu_gfix_ctx *c = get_ctx(GLU_GFIX_LANDAU);
  if (!c) return 123456789;
  size_t iters = 0;
  const int lux = callback ? luxury_copies() : 0;
  // the copies draw from GLU's process stream, continued on the device
  if (lux && (rng_init() || mwc_from_host(c))) {
    fprintf(stderr, "[GF] random stream unavailable: %s\n", glu_gfix_last_error());
    return 123456789;
  }
  if (sites_to_device(c, lat) ||
      (lux ? glu_gfix_luxury(c, method, Latt.gf_alpha, acc, max_iters, lux, 0, th, &iters, nullptr) ||
                 mwc_to_host(c)
           : glu_gfix_run(c, method, Latt.gf_alpha, acc, max_iters, th, &iters))) {
    fprintf(stderr, "[GF] device gauge fixing failed: %s\n", glu_gfix_last_error());
    return 123456789;
  }
  if (download && device_to_sites(c, lat)) {
    fprintf(stderr, "[GF] link download failed: %s\n", glu_gfix_last_error());
    return 123456789;
  }
  return iters;
}

}  // namespace

extern "C" size_t FACG(struct site *lat, struct fftw_stuff *FFTW, double *th, const double acc,
                       const size_t max_iters) {
  return fa_call(lat, th, acc, max_iters, GLU_GFIX_CG, false);
}

extern "C" size_t FASD(struct site *lat, struct fftw_stuff *FFTW, double *th, const double acc,
                       const size_t max_iters) {
  return fa_call(lat, th, acc, max_iters, GLU_GFIX_SD, false);
}

// Gfix/Landau.c:261-374 (the FACG callback; FASD with GLU_GFIX_SD)
extern "C" size_t Landau(struct site *lat, const double accuracy, const size_t max_iter, const char *infile) {
  const int method = getenv("GLU_GFIX_SD") ? GLU_GFIX_SD : GLU_GFIX_CG;
  double theta = 0.0;
  size_t iters = fa_call(lat, &theta, accuracy, max_iter, method, true, false);
  size_t failure = 0;
  if (iters == 123456789) {
    fprintf(stderr, "\n[GF] Non-convergence ... Randomly-restarting \n\n[GF] Failure :: %zu || Iters :: %zu || ACC :: %e \n",
            failure, max_iter, theta);
    for (failure = 1; failure < 7; failure++) {
      std::vector<double> h;
      if (!infile || read_nersc(infile, h) != 0) {
        printf("[IO] Something funky happened when trying to read in config again!\n");
        break;
      }
      if (random_transform(h)) {
        fprintf(stderr, "[GF] random transform failed: %s\n", glu_gfix_last_error());
        break;
      }
      unpack(lat, h);
      const size_t loc = fa_call(lat, &theta, accuracy, max_iter, method, true, false);
      if (loc != 123456789) {
        iters += loc;
        break;
      }
      iters += max_iter;
      fprintf(stderr, "\n[GF] Non-convergence ... Randomly-restarting \n\n[GF] Failure :: %zu || Iters :: %zu || ACC :: %e \n",
              failure, loc, theta);
    }
  }
  glu_gfix_ctx *c = get_ctx(GLU_GFIX_LANDAU);
  if (failure == 7) {
    printf("\n[GF] Failure to converge to a sufficient solution \n[GF] Please lower the GF_TUNE parameter from %g \n",
           Latt.gf_alpha);
    if (c) (void)device_to_sites(c, lat);  // lat holds the last attempt, as GLU leaves it
    return (size_t)-1;
  }
  // the context holds the fixed links: reunitarised, reported and
  // written back into lat
  if (c) output_fixing_info(c, lat, theta, iters);
  return iters;
}

// Gfix/CFACG.c:532-636 with the slices batched on the device
extern "C" size_t Coulomb_FA(struct site *lat, struct fftw_stuff *FFTW, const double accuracy, const size_t max_iter,
                             const GLU_bool facg) {
  glu_gfix_ctx *c = get_ctx(GLU_GFIX_COULOMB);
  if (!c) return 0;
  size_t iters = 0;
  double th = 0.0;
  // initialise_par_rng on entry, free_par_rng on return (Gfix/CFACG.c:539,
  // 630): the slice restarts continue GLU's process stream on the device
  const int rc = rng_init() || mwc_from_host(c) || sites_to_device(c, lat) ||
                 glu_gfix_run(c, facg == GLU_TRUE ? GLU_GFIX_CG : GLU_GFIX_SD, Latt.gf_alpha, accuracy, max_iter,
                              &th, &iters);
  mwc_host_free();
  if (rc) {
    fprintf(stderr, "[GF] device Coulomb gauge fixing failed: %s\n", glu_gfix_last_error());
    return 0;
  }
  if (device_to_sites(c, lat)) fprintf(stderr, "[GF] link download failed: %s\n", glu_gfix_last_error());
  return iters;
}

// Gfix/Coulomb.c:33-100
extern "C" size_t Coulomb(struct site *lat, const double accuracy, const size_t iter) {
  glu_gfix_ctx *c = get_ctx(GLU_GFIX_COULOMB);
  if (!c) return 0;
  if (sites_to_device(c, lat)) fprintf(stderr, "[GF] link upload failed: %s\n", glu_gfix_last_error());
  double plaq = 0, tr = 0, mx = 0, func = 0, tl = 0;
  glu_gfix_measure(c, &plaq, &tr, &mx, &func, &tl);
  double sl = (4 * tr - tl) / 3.0;
  fprintf(stdout, "[GF] Initial Tlink :: %1.15f || Slink :: %1.15f \n[GF] Plaquette :: %1.15f \n", tl, sl, plaq);
  const size_t iters =
      Coulomb_FA(lat, nullptr, accuracy, iter, getenv("GLU_GFIX_SD") ? GLU_FALSE : GLU_TRUE);
  // the context still holds the fixed links Coulomb_FA wrote into lat
  glu_gfix_reunit(c);
  if (device_to_sites(c, lat)) fprintf(stderr, "[GF] link download failed: %s\n", glu_gfix_last_error());
  glu_gfix_measure(c, &plaq, &tr, &mx, &func, &tl);
  sl = (4 * tr - tl) / 3.0;
  fprintf(stdout,
          "[GF] Tuning :: %f || Iterations :: %zu ||\n[GF] Final Tlink :: %1.15f || Slink :: %1.15f \n[GF] Plaquette :: "
          "%1.15f \n",
          Latt.gf_alpha, iters, tl, sl, plaq);
  return iters;
}

namespace {

// the file name of Cuts/cut_output.c:115-258 for the momentum-space modes
// built here (SIN_MOM is GLU's default, Headers/GLU_definitions.h:348-352)
std::string cut_file(const struct cut_info &ci) {
  char tmp[256];
  std::string s(ci.where, strnlen(ci.where, sizeof(ci.where)));
  s += ci.dir == GLUON_PROPS ? "G2F2" : "MOMgg";
  snprintf(tmp, sizeof(tmp), "_SU3_%zux%zux%zux%zu_%zu_", Latt.dims[0], Latt.dims[1], Latt.dims[2], Latt.dims[3],
           ci.max_mom);
  s += tmp;
  switch (ci.type) {
    case HYPERCUBIC_CUT: s += "cutHYP"; break;
    case PSQ_CUT: s += "cutPSQ"; break;
    case CYLINDER_CUT: snprintf(tmp, sizeof(tmp), "cutCYL_w%g", ci.cyl_width); s += tmp; break;
    case CYLINDER_AND_CONICAL_CUT: snprintf(tmp, sizeof(tmp), "cutCON_%zu", ci.angle); s += tmp; break;
  }
  snprintf(tmp, sizeof(tmp), "_SINMOM_LIN.%zu.bin", Latt.flow);
  return s + tmp;
}

}  // namespace

// Cuts/cuts.c:114-184 for GLUON_PROPS and EXCEPTIONAL with the LINEAR
// definition; the transform, momentum list and contraction run on the device
extern "C" int cuts_struct(struct site *A, const struct cut_info CUTINFO) {
  const int kFailure = -1, kSuccess = 0;  // GLU_FAILURE, GLU_SUCCESS
  if (!A) {
    fprintf(stderr, "[CUTS] no lattice given\n");
    return kFailure;
  }
  if (CUTINFO.dir != GLUON_PROPS && CUTINFO.dir != EXCEPTIONAL) {
    fprintf(stderr, "[CUTS] cut mode %d is not provided by this library\n", (int)CUTINFO.dir);
    return kFailure;
  }
  if (CUTINFO.definition != LINEAR_DEF) {
    fprintf(stderr, "[CUTS] only the LINEAR field definition is provided by this library\n");
    return kFailure;
  }
  if ((int)CUTINFO.type < HYPERCUBIC_CUT || (int)CUTINFO.type > CYLINDER_AND_CONICAL_CUT) {
    fprintf(stderr, "I don't recognise the type.. Leaving \n");
    return kFailure;
  }
  glu_gfix_ctx *c = get_ctx(GLU_GFIX_LANDAU);
  if (!c) return kFailure;
  const std::string path = cut_file(CUTINFO);
  fprintf(stdout, "[CUTS] Our file is located @ %s \n", CUTINFO.where);
  size_t nmom = 0;
  const double angle = (double)CUTINFO.angle;
  int rc = sites_to_device(c, A);
  if (!rc)
    rc = CUTINFO.dir == GLUON_PROPS
             ? glu_gfix_gluon_props(c, (int)CUTINFO.type, CUTINFO.max_mom, CUTINFO.cyl_width, angle, path.c_str(),
                                    &nmom, nullptr, nullptr, nullptr, 0, nullptr)
             : glu_gfix_gluon_exceptional(c, (int)CUTINFO.type, CUTINFO.max_mom, CUTINFO.cyl_width, angle,
                                          path.c_str(), &nmom, nullptr, nullptr, nullptr, 0, nullptr);
  if (rc) {
    fprintf(stderr, "[CUTS] device cuts failed: %s\n", glu_gfix_last_error());
    return kFailure;
  }
  fprintf(stdout, "[CUTS] Writing finished...\n[CUTS] Outputting to %s\n", path.c_str());
  return kSuccess;
}
