/*
 * glu_oracle.h -- CPU restatement of GLU's Fourier-accelerated gauge fixing.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity checker for the HIP path in
 * glu_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it, and only to check or to time a CPU baseline.  The
 * product library never links or calls it.
 *
 * Conventions (all as in GLU, see the citations next to each function):
 *   - lattice sites are lexicographic with x fastest, t slowest
 *     (Geometry/geometry.c:28-36);
 *   - a link is 9 complex doubles, row-major, O[a*3+b] = U_ab, stored as 18
 *     interleaved re/im doubles (Headers/GLU_definitions.h:102);
 *   - a gauge field is double U[V][ND][18] (site-major, then direction);
 *   - the packed FFT "4-field" is 4 complex arrays of V entries each,
 *     double F[4][V][2] (Gfix/FACG.c:57-60, TRUE_HERM = 4).
 *
 * Reductions follow GLU's OpenMP structure: the site loop is cut into
 * "virtual threads" exactly like libgomp's static schedule, each virtual
 * thread accumulates its own partial in site order (plain or Kahan, as GLU
 * does), and partials are added in thread order.  With nth equal to GLU's
 * OMP_NUM_THREADS the summation order is GLU's.
 *
 * Each function is pinned by a check that a plausible mistake would fail
 * (tests/test_oracle_pins.py lists them): the NERSC reader, plaquette and
 * link trace by the CONF/lat..400 header; the NERSC writer by reproducing
 * that file's data bytes and checksum from its links; p^2 by gen_p_sq(0) == 1 and its
 * closed form; the FFT by a definition-level DFT and closed forms; the
 * derivative by the generic Hermitian projection, theta_lin == 2 theta_FA,
 * sum_x Delta(x) == 0 and an abelian closed form; reunitarisation,
 * set_gauge_matrix, gtransform_local and the triple trace by unitarity and
 * matrix products; the line search by the functional after trial updates;
 * the loops by pure-gauge inputs and invariants; the gluon propagators
 * (tests/test_oracle_glueprop.py) by brute-force momentum lists, numpy and
 * a plane wave; the exceptional g2 by the propagators' trace, g3 by
 * colour-rotation invariance and by GLU's alternative triple-sum
 * contraction (TRACE_PROJECTION, Cuts/MOMgg.c:67-91) in numpy; the
 * MWC_4096 stream by big-integer restatements of MT19937's init_genrand
 * (with its published state for seed 5489) and Marsaglia's complementary
 * multiply-with-carry.  Parity unpinned: the exact FACG/Coulomb
 * trajectories, and or_heatbath, which is an independent generator (a
 * counter-based Cabibbo-Marinari heatbath), not a restatement of GLU's.
 */
#ifndef GLU_ORACLE_H
#define GLU_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_ND 4
#define OR_NC 3
#define OR_LINK 18 /* doubles per link */

/* GLU's failure sentinel for FACG/FASD (Gfix/FACG.c:381) */
#define OR_GF_FAIL 123456789u

typedef struct {
  size_t dims[OR_ND];
  size_t V, Lcu;
  size_t *fwd; /* V*ND : fwd[i*ND+mu] = site i + mu-hat */
  size_t *bwd; /* V*ND : bwd[i*ND+mu] = site i - mu-hat */
  int nth;     /* virtual threads, GLU's Latt.Nthreads */
} or_lat;

/* per-iteration trajectory record (optional, any pointer may be NULL) */
typedef struct {
  size_t cap;      /* capacity of each array */
  size_t n;        /* entries written */
  double *theta;   /* theta_k measured before update k (GLU's *tr)  */
  double *alpha;   /* step applied at update k (line-search or fixed) */
  double *beta;    /* CG beta used at update k (0 for the SD start row) */
  double *func;    /* gauge functional after update k (costly)       */
} or_trace;

/* ---------------- geometry (Geometry/geometry.c, init.c) ---------------- */
int    or_lat_init(or_lat *L, const size_t dims[OR_ND], int nth);
void   or_lat_free(or_lat *L);
double or_psq(const or_lat *L, size_t i, int D); /* gen_p_sq geometry.c:170 */

/* ---------------- SU(3) kernels (Matrix_Ops, Gfix/CG.c, gtrans.c) ------- */
void   or_gram_reunit(double *U18);
void   or_set_gauge_matrix(double *g18, const double *F, size_t V,
                           double alpha, size_t i);
void   or_gtransform_local(const double *a, double *b, const double *c);
double or_retrace_abc_dag(const double *a, const double *b, const double *c);
void   or_herm_proj_short(double *Q10, const double *U18);
double or_approx_minimum(const double alphas[3], const double f[3]);

/* ---------------- measurements (Field/plaqs_links.c, gftests.c) --------- */
double or_plaquette(const or_lat *L, const double *U);
double or_links(const or_lat *L, const double *U, double *maxtr);
void   or_all_links(const or_lat *L, const double *U, double *sp, double *t);
double or_functional(const or_lat *L, const double *U);
double or_theta_lin(const or_lat *L, const double *U, int max_dir);
void   or_latt_reunit(const or_lat *L, double *U);

/* ---------------- derivative (Gfix/lin_derivs.c, FACG.c) ---------------- */
double or_fast_deriv(const or_lat *L, const double *U, size_t i, double *sum10);
double or_latt_deriv(const or_lat *L, const double *U, size_t i, int max_dir,
                     double *sum10);
double or_fa_deriv(const or_lat *L, const double *U, double *F); /* theta */

/* ---------------- FFT ---------------------------------------------------- */
/* unnormalised complex DFT over the first nd dims (x fastest) of `dims`,
   sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD, applied to `nb` arrays of
   prod(dims[0..nd-1]) complex entries laid out back to back */
void   or_fft(double *data, const size_t *dims, int nd, int sign, size_t nb);
void   or_set_fft_mode(int mode); /* 0 mixed radix (default), 1 direct per-axis sums */
void   or_dft_naive(double *data, const size_t *dims, int nd, int sign);
void   or_psq_table(const or_lat *L, int D, double *psq); /* MAX/p^2 */
void   or_fourier_accelerate(const or_lat *L, double *F, const double *psq,
                             int D, size_t nb);

/* ---------------- Landau (Gfix/FACG.c, line_search.c, Landau.c) --------- */
void   or_egauge_landau(const or_lat *L, double *U, const double *F,
                        double alpha);
double or_line_search_landau(const or_lat *L, double *U, const double *F,
                             double *red);
size_t or_facg(const or_lat *L, double *U, double gf_alpha, double acc,
               size_t max_iters, double *theta, or_trace *tr);
size_t or_fasd(const or_lat *L, double *U, double gf_alpha, double acc,
               size_t max_iters, double *theta, or_trace *tr);

/* ---------------- Coulomb (Gfix/CFACG.c, Coulomb.c) --------------------- */
size_t or_coulomb(const or_lat *L, double *U, double gf_alpha, double acc,
                  size_t max_iters, int facg, uint32_t seed,
                  size_t *slice_iters, double *slice_theta);

/* ---------------- RNG + random gauge transform (Utils/par_*.c) ---------- */
typedef struct or_rng or_rng;
or_rng *or_rng_new(uint32_t seed, int nth);
void    or_rng_free(or_rng *R);
double  or_rng_dbl(or_rng *R, int thread);
void    or_sunitary_gen(or_rng *R, int thread, double *Z18);
void    or_random_gtrans(const or_lat *L, or_rng *R, double *U);
void    or_gtransform(const or_lat *L, double *U, const double *G);

/* ---------------- NERSC I/O (IO/read_headers.c, readers.c) -------------- */
/* reads a NERSC file; returns 0 on success.  U is allocated by the caller
   with V*ND*18 doubles once dims are known: call with U == NULL first to
   get dims/header values, then again with storage. */
typedef struct {
  size_t dims[OR_ND];
  double plaquette, link_trace;
  uint32_t checksum_header, checksum_computed;
  int two_rows;     /* 4D_SU3_GAUGE (3x2) vs 4D_SU3_GAUGE_3x3 */
  int double_prec;  /* IEEE64 vs IEEE32 */
  int big_endian;
  long data_offset;
} or_nersc_info;
int or_read_nersc(const char *path, or_nersc_info *info, double *U);
int or_write_nersc(const char *path, const or_lat *L, double *U, int two_rows,
                   int big_endian, int reunit, const char *label,
                   size_t sequence, const char *creator, const char *date);

/* ---------------- gluon propagators (Cuts/, downstream of the path) ------ */
/* the momentum list of a cut (type 0 hypercubic, 1 psq, 2 cylinder,
   3 cylinder and cone): *n momenta, mom[n][4] signed, idx[n] their sites in
   the DFT output; returns -1 if n > cap, -2 for an unknown type */
int or_momentum_list(const or_lat *L, int type, size_t max_mom,
                     double cyl_width, double angle, int *mom, size_t *idx,
                     size_t cap, size_t *n);
/* transverse G(p^2) and longitudinal F(p^2) over that list */
int or_gluon_prop(const or_lat *L, const double *U, int type, size_t max_mom,
                  double cyl_width, double angle, int *mom, double *trans,
                  double *longi, size_t cap, size_t *n);
/* exceptional-momentum two- and three-point functions g2(p), g3(p) */
int or_gluon_exceptional(const or_lat *L, const double *U, int type,
                         size_t max_mom, double cyl_width, double angle,
                         int *mom, double *g2, double *g3, size_t cap, size_t *n);
int or_write_glueprop(const char *path, size_t n, const int *mom,
                      const double *trans, const double *longi);

/* ---------------- planted mistakes (tests of the pins) ------------------- */
/* The loop pins of tests/test_oracle_loops.py are shown to catch plausible
   slips by planting each one here, on the CPU, and checking that a pin
   fails.  0 = the faithful restatement (the default). */
enum {
  OR_PLANT_NONE = 0,
  OR_PLANT_FR_ALWAYS = 1,    /* Fletcher-Reeves beta also above CG_TOL     */
  OR_PLANT_MAX_FACTOR = 2,   /* psq with 4 (D - 1) / vol instead of 4 D    */
  OR_PLANT_THETA_AFTER = 3,  /* theta_k recorded after update k           */
  OR_PLANT_EGAUGE_SIGN = 4,  /* gauge update along -alpha                 */
  OR_PLANT_NO_CLAMP = 5,     /* Polak-Ribiere beta not clipped at 0       */
  OR_PLANT_NO_BETA = 6,      /* beta = 0: steepest descent directions     */
  OR_PLANT_NO_STALE = 7,     /* continuation without GLU's stale red[] slots */
  OR_PLANT_OLD_KEPT = 8,     /* in_old not refreshed after an iteration   */
  OR_PLANT_STAPLE_NODAG = 9, /* lower staple without the dagger on U_nu(x-nu) */
  OR_PLANT_OR_SIGN = 10,     /* overrelaxation reflection with s1 unflipped */
  OR_PLANT_KP_NOCOS = 11     /* Kennedy-Pendleton without the cos^2 factor */
};
void or_set_plant(int code);

/* Coulomb: record the trajectory of one slice (theta_k, step, beta) of the
   next or_coulomb calls into *tr (NULL: off); slice = (size_t)-1 records
   every slice, tr then pointing to one trace per time slice */
void or_coulomb_trace(or_trace *tr, size_t slice);

/* ---------------- input generation (not on the hot path) ---------------- */
/* quenched Wilson heatbath + overrelaxation, Cabibbo-Marinari with
   Kennedy-Pendleton SU(2) heatbath; own seeded generator */
void or_heatbath(const or_lat *L, double *U, double beta, int sweeps,
                 int nor, uint64_t seed, int cold_start);

/* GLU's own quenched update (Update/KPHB.c:121-257): `sweeps` times a
   Kennedy-Pendleton heatbath pass over GLU's draughtboard colouring and
   `nor` overrelaxation passes, from GLU's MWC_4096 streams seeded
   SEED = seed with L->nth of them, in libgomp's static blocks; from unit
   links when cold_start */
void or_glu_heatbath(const or_lat *L, double *U, double beta, size_t sweeps, size_t nor, uint32_t seed,
                     int cold_start);
/* one overrelaxation pass (Update/relax.c:76-114) */
void or_glu_overrelax(const or_lat *L, double *U);
/* GLU's Wilson staple sum of link (i, mu) (Smear/staples.c:63-76) */
void or_glu_staple(const or_lat *L, const double *U, size_t i, int mu, double *stap18);
/* draughtboard colour of site i (Geometry/draughtboard.c:48-68) */
int or_glu_colour(const or_lat *L, size_t i);
/* n values a0 = 1 - d of the SU(2) heatbath at `norm` (Update/hb.c:53-96) */
void or_glu_kp_samples(uint32_t seed, double norm, size_t n, double *a0);

#ifdef __cplusplus
}
#endif
#endif
