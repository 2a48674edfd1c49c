/*
 * glu_gfix.h -- C ABI of the MI355X gauge-fixing library (libglu_gfix.so).
 *
 * Two layers:
 *
 *  1. GLU drop-in.  Landau(), Coulomb(), FACG(), FASD() and Coulomb_FA()
 *     with GLU's exact prototypes.  Linked ahead of a libGLU.a built
 *     without Gfix/{Landau,FACG,CFACG,Coulomb}.o they replace GLU's CPU
 *     gauge fixing; they read GLU's global `Latt` (dims, gf_alpha, Seed,
 *     Nthreads) and overwrite the links of `lat` in place, as GLU does.
 *
 *  2. Plain-pointer API (glu_gfix_*): a context owning one lattice's
 *     device state.  Links cross the boundary in GLU's own memory layout,
 *     double[V][4][18] (site-major with x fastest, then direction, then the
 *     row-major 3x3 complex matrix as re/im pairs).
 *
 * Every glu_gfix_* call returns 0 on success and a negative code on
 * failure; glu_gfix_last_error() describes the last failure.  No call
 * falls back to the CPU: without a usable HIP device they fail.
 */
#ifndef GLU_GFIX_H
#define GLU_GFIX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* GLU types, laid out as GLU's Headers/GLU_types.h and GLU_enum.h     */
/* (skipped when GLU's own headers were included first)                */
/* ------------------------------------------------------------------ */
#ifndef GLU_ENUM_H
typedef enum { GLU_FALSE, GLU_TRUE } GLU_bool; /* Headers/GLU_enum.h:83-85 */
/* Headers/GLU_enum.h:159-196 */
typedef enum {
  EXCEPTIONAL, NONEXCEPTIONAL, FIELDS, SMEARED_GLUONS, INSTANTANEOUS_GLUONS, CONFIGSPACE_GLUONS,
  GLUON_PROPS, STATIC_POTENTIAL, TOPOLOGICAL_CORRELATOR, TOPOLOGICAL_MOMENTS, TOPOLOGICAL_SUSCEPTIBILITY,
  GLU_NOCUTS
} cut_mode;
typedef enum { LINEAR_DEF, LOG_DEF } lie_field_def;
typedef enum { HYPERCUBIC_CUT, PSQ_CUT, CYLINDER_CUT, CYLINDER_AND_CONICAL_CUT } momentum_cut_def;
#endif
#ifndef GLU_TYPES_H
/* Headers/GLU_types.h:269-273; O[mu] points at 9 complex doubles */
struct site {
  void **O;
  size_t neighbor[4];
  size_t back[4];
};
/* Headers/GLU_types.h:229-243 (the enum members are int-sized) */
struct latt_info {
  size_t dims[4];
  size_t Lsq;
  size_t Lcu;
  size_t Volume;
  size_t flow;
  double gf_alpha;
  double sm_alpha[3];
  int head;
  uint32_t Seed[1];
  double twiddles[4];
  uint32_t Nthreads;
  void *su2_data;
  int argc;
};
struct fftw_stuff; /* unused by this backend; callers may pass NULL */
/* Headers/GLU_types.h:94-103 */
struct cut_info {
  cut_mode dir;
  momentum_cut_def type;
  size_t max_mom;
  size_t max_t;
  char where[256];
  lie_field_def definition;
  size_t angle;
  double cyl_width;
};
#endif

/* Landau gauge fixing with restarts and GLU's [GF] report lines.
   Replaces Gfix/Landau.c:261 (prototype Headers/Landau.h:51).
   Returns the iteration count, or (size_t)-1 (GLU_FAILURE) when every
   restart failed. */
size_t Landau(struct site *lat, const double accuracy, const size_t max_iter,
              const char *infile);

/* Coulomb gauge fixing of every time slice plus the report lines.
   Replaces Gfix/Coulomb.c:33 (prototype Headers/Coulomb.h:40).
   Returns the summed iteration count over slices. */
size_t Coulomb(struct site *lat, const double accuracy, const size_t iter);

/* Fourier-accelerated Polak-Ribiere CG.  Replaces Gfix/FACG.c:341
   (prototype Headers/FACG.h:37).  *th receives the final theta; returns the
   iterations of the last segment or 123456789 on non-convergence. */
size_t FACG(struct site *lat, struct fftw_stuff *FFTW, double *th,
            const double acc, const size_t max_iters);

/* Fourier-accelerated steepest descent with the fixed step GF_TUNE (the
   GLU_GFIX_SD build).  Replaces Gfix/FACG.c:395 (Headers/FACG.h:54). */
size_t FASD(struct site *lat, struct fftw_stuff *FFTW, double *th,
            const double acc, const size_t max_iters);

/* Coulomb FACG (FACG == GLU_TRUE) or FASD per time slice, all slices
   batched on the device.  Replaces Gfix/CFACG.c:532 (Headers/CFACG.h:37). */
size_t Coulomb_FA(struct site *lat, struct fftw_stuff *FFTW,
                  const double accuracy, const size_t max_iter,
                  const GLU_bool FACG);

/* ------------------------------------------------------------------ */
/* plain-pointer API                                                   */
/* ------------------------------------------------------------------ */
typedef struct glu_gfix_ctx glu_gfix_ctx;

enum { GLU_GFIX_LANDAU = 0, GLU_GFIX_COULOMB = 1 };
enum { GLU_GFIX_FP64 = 0, GLU_GFIX_MIXED = 1 };
enum { GLU_GFIX_CG = 0, GLU_GFIX_SD = 1 };

/* Context for one lattice of extent dims[0..3] (x, y, z, t).  Allocates
   the device state of the whole run (links, 3 packed fields, partials).
   Mirrors the allocations of Gfix/CGalloc.c:98-129 and the FFTW plans of
   Geometry/plan_ffts.c:103-168, made once per context. */
int glu_gfix_create(glu_gfix_ctx **ctx, const size_t dims[4], int gftype,
                    int precision);
int glu_gfix_destroy(glu_gfix_ctx *ctx);

/* Upload / download links in GLU's layout double[V][4][18] (host). */
int glu_gfix_set_links(glu_gfix_ctx *ctx, const double *links);
int glu_gfix_get_links(glu_gfix_ctx *ctx, double *links);
/* Same, with the GLU-layout buffer already in device memory. */
int glu_gfix_set_links_device(glu_gfix_ctx *ctx, const double *links_dev);
int glu_gfix_get_links_device(glu_gfix_ctx *ctx, double *links_dev);

/* Keep a device copy of the current links / restore it (device to device). */
int glu_gfix_snapshot(glu_gfix_ctx *ctx);
int glu_gfix_restore(glu_gfix_ctx *ctx);

/* Run the gauge fixing on the device-resident links.  method: GLU_GFIX_CG
   (FACG) or GLU_GFIX_SD (FASD).  Landau: FACG()/FASD() semantics (theta
   test, continuation, 123456789 on failure) written to *iters.  Coulomb:
   Coulomb_FA() semantics, total iterations.  *theta: final accuracy
   (Coulomb: largest slice theta).  Synchronous. */
int glu_gfix_run(glu_gfix_ctx *ctx, int method, double gf_tune,
                 double accuracy, size_t max_iters, double *theta,
                 size_t *iters);

/* Landau() of Gfix/Landau.c:261-374 on the device links: glu_gfix_run
   with GLU's restart policy.  A run returning exactly 123456789 (FACG
   failure) is repeated from the input links after a random gauge transform,
   at most 6 times (GF_GLU_FAILURES = 7, Headers/GLU_definitions.h:400-402);
   failed attempts add max_iters to *iters, as GLU counts them.  GLU
   re-reads the config file and applies random_transform
   (Geometry/random_config.c:108-120) drawn from its MWC_4096 streams; here
   each restart applies the same transform, drawn from the context's copy
   of those streams (glu_gfix_rng_seed), to a device copy of the input.
   restart_seed != 0 seeds the streams with SEED = restart_seed (its low
   32 bits) at the first restart, as a GLU process starting with that SEED
   would; 0 continues the context's streams.  *failures (may be NULL):
   restarts taken.  Returns -5 when all 7 attempts fail (GLU: GLU_FAILURE). */
int glu_gfix_landau(glu_gfix_ctx *ctx, int method, double gf_tune,
                    double accuracy, size_t max_iters, uint64_t restart_seed,
                    double *theta, size_t *iters, int *failures);

/* luxury_copy_fast of Gfix/Landau.c:88-213 (the LUXURY_GAUGE build, with
   BEST_COPY): `copies` random gauge copies of the device links, each
   fixed to 1e-6 in at most 1000 iterations; the copy with the smallest
   functional among those that converged is then fixed to `accuracy`.
   *iters = last copy's iterations + the final run's, as GLU counts.  Each
   copy's random transform draws from the context's MWC_4096 streams,
   seeded with SEED = seed at the start (seed 0: continuing them), as GLU's
   copies draw from its streams.  *best_copy (may be NULL): the chosen copy,
   -1 if none qualified (the final run then starts from the input). */
int glu_gfix_luxury(glu_gfix_ctx *ctx, int method, double gf_tune,
                    double accuracy, size_t max_iters, int copies,
                    uint64_t seed, double *theta, size_t *iters,
                    int *best_copy);

/* FACG() (Gfix/FACG.c:341-392, prototype Headers/FACG.h:37) for nlat small
   lattices of one shape in one launch, one workgroup per lattice, FP64:
   links_dev holds nlat configurations back to back in GLU's layout
   (nlat x V x 72 doubles, device memory), each gauge fixed in place.  Per
   lattice: iters[k] is FACG()'s return value (0 when theta_test_lin is
   already below accuracy, the last segment's count after continuations
   below 1e3 accuracy, 123456789 when it fails) and theta[k] the final
   theta; with trace non-NULL, rows (theta_k, alpha_k, beta_k) as
   glu_gfix_trace gives them, trace_cap rows per lattice (host memory).
   Extents 2..16 (so at most 65536 sites), at most 65535 lattices; after
   100 continuation segments a lattice returns 123456789 (GLU does not
   bound them).  Landau()'s random restarts are left to the caller
   (glu_gfix_landau on the failed ones).  Waits for the device's pending
   work, runs on a stream of its own and returns when every lattice has
   stopped; allocates 640 B per site per lattice of device scratch for the
   call.  For lattices whose iterations are bound by launch latency (8^4:
   one iteration of the context path takes about 75 us). */
int glu_gfix_facg_batch(const size_t dims[4], size_t nlat, double *links_dev, double gf_tune,
                        double accuracy, size_t max_iters, size_t *iters, double *theta,
                        double *trace, size_t trace_cap);

/* Trajectory of the last run: rows (theta_k, alpha_k, beta_k); row 0 is the
   steepest-descent start.  Landau only.  Recording costs nothing extra. */
int glu_gfix_trace(glu_gfix_ctx *ctx, double *rows, size_t cap, size_t *n);

/* Observables of the device links: average plaquette
   (Field/plaqs_links.c:243-265), link trace and its maximum single-link
   value (:424-448), functional 1 - link trace (Gfix/gftests.c:133-147),
   temporal link trace (:399-422). */
int glu_gfix_measure(glu_gfix_ctx *ctx, double *plaquette, double *link_trace,
                     double *link_max, double *functional, double *tlink);

/* theta_test_lin of the device links (Gfix/gftests.c:215-250). */
int glu_gfix_theta(glu_gfix_ctx *ctx, double *theta_lin);

/* Reunitarise every device link (Geometry/random_config.c:43-56). */
int glu_gfix_reunit(glu_gfix_ctx *ctx);

/* Allocation-free variant of glu_gfix_measure + glu_gfix_theta.  It has no
   counterpart in GLU; it is for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops).  It reads caller-owned
   device links in GLU's layout (double[V][4][18], x fastest, as
   glu_gfix_set_links_device takes them) and writes 6 doubles to the device
   buffer out_dev, enqueued on `stream` (a hipStream_t; NULL: the default
   stream) without waiting: plaquette, link trace, largest link trace,
   functional, temporal link trace, theta_test_lin.  Extents 2..4096 and
   fewer than 2^31 sites; larger lattices are refused (-1).  Scratch: a fixed
   64 KiB device array per GPU, held until the process ends.  Calls that
   overlap on one GPU share it and race: serialise them.  Call it once
   outside graph capture on each device first (that call looks the array up). */
int glu_gfix_measure_noalloc(const double *links_dev, const size_t dims[4],
                             double *out_dev, void *stream);

/* GLU's momentum-space cuts of gauge-fixed links, replacing
   Cuts/cuts.c:114-184 (prototype Headers/cuts.h:41) for dir = GLUON_PROPS
   and EXCEPTIONAL (GLU's default) with the LINEAR definition: the file
   CUTINFO.where + GLU's name (Cuts/cut_output.c:115-258, from Latt.dims and
   Latt.flow) receives the momentum list and G/F or g2/g3.  Other modes
   and the LOG definition return GLU_FAILURE.  Unlike GLU, the links of A
   are left as they were (GLU overwrites them with the Lie fields). */
int cuts_struct(struct site *A, const struct cut_info CUTINFO);

/* Transverse G(p^2) and longitudinal F(p^2) gluon propagators of the device
   links: GLU's cuts_struct with dir = GLUON_PROPS (Cuts/cuts.c:114-184),
   the LINEAR field definition (Hermitian_proj, Matrix_Ops/exactQ.c:358-373),
   correct_pspace_landau (Cuts/pspace_landau.c:28-86) and
   compute_gluon_prop (Cuts/glueprop.c:95-151).  The momenta are those of
   GLU's momentum list for the cut (Cuts/cut_routines.c:300-455; cut_width
   in lattice units, angle in degrees for the cone), enumerated in GLU's
   order; *nmom receives their number (with every output NULL, the call
   only counts).  mom (nmom x 4 signed ints),
   transverse and longitudinal (nmom doubles each) may be NULL, else hold
   cap >= nmom entries; pspace_err (or NULL) receives GLU's momentum-space
   Landau-condition figure.  path (or NULL) receives GLU's big-endian
   output file (write_mom_veclist + write_g2g3_to_list,
   Cuts/cut_output.c:311-400).  Landau contexts only. */
enum { GLU_GFIX_CUT_HYPERCUBIC = 0, GLU_GFIX_CUT_PSQ = 1, GLU_GFIX_CUT_CYLINDER = 2, GLU_GFIX_CUT_CONICAL = 3 };
int glu_gfix_gluon_props(glu_gfix_ctx *ctx, int cut, size_t max_mom, double cyl_width, double angle,
                         const char *path, size_t *nmom, int *mom, double *transverse,
                         double *longitudinal, size_t cap, double *pspace_err);

/* GLU's default cuts output, the EXCEPTIONAL branch of cuts_struct: the
   gluon two- and three-point functions g2(p), g3(p) at exceptional
   momenta (p, -p, 0) of write_exceptional_g2g3_MOMgg (Cuts/MOMgg.c:141-215)
   with its default contraction_matrices (:94-121).  Arguments, list and
   file as for glu_gfix_gluon_props, with g2 / g3 in place of G / F. */
int glu_gfix_gluon_exceptional(glu_gfix_ctx *ctx, int cut, size_t max_mom, double cyl_width, double angle,
                               const char *path, size_t *nmom, int *mom, double *g2, double *g3,
                               size_t cap, double *pspace_err);

/* Write the device links as a NERSC file: the NERSC branch of write_lat
   (IO/writers.c:330-422) with write_header_NERSC (IO/write_headers.c:77-138).
   As in GLU the links are reunitarised in place first.  two_rows selects
   4D_SU3_GAUGE (rows 0 and 1 of each link) over 4D_SU3_GAUGE_3x3;
   big_endian IEEE64BIG (GLU's default) over IEEE64LITTLE; label and
   sequence fill ENSEMBLE_LABEL and SEQUENCE_NUMBER (GLU's reader refuses
   0).  Returns -3 when the file cannot be written. */
int glu_gfix_write_nersc(glu_gfix_ctx *ctx, const char *path, int two_rows, int big_endian,
                         const char *label, size_t sequence);

/* Mixed precision: run FP32 until theta < switch_theta, then FP64 (default
   1e-11, the fastest switch point measured on 32^3x64 solves, or the
   GLU_GFIX_SWITCH_THETA environment variable at creation).  A mixed
   result satisfies the same theta and plaquette bars as an FP64 one but
   usually lies on a different Gribov copy (a neighbouring local maximum
   of the gauge functional, 1e-6..1e-3 away in relative functional): the
   FP32 trajectory leaves the FP64 one early.  Use GLU_GFIX_FP64 where the
   copy GLU's own FP64 iteration reaches is required.
   The FP32 phase caps the Polak-Ribiere beta at 3 (GLU_GFIX_BETA_CAP;
   GLU_GFIX_SAFEGUARD=0 turns it off): near its rounding floor the FP32
   gradient makes PR+ directions run uphill for hundreds of iterations,
   and the cap keeps beta bounded there.  The FP64 phase follows GLU. */
int glu_gfix_set_switch_theta(glu_gfix_ctx *ctx, double switch_theta);

/* Host-clock milliseconds of the last Landau run (of a glu_gfix_landau or
   luxury call: its last attempt): ms[0] the FP32 phase (0 for an FP64 run),
   ms[1] the switch (the FP64 transform of the links, the CG state
   promoted), ms[2] the FP64 phase.  Each phase ends at the batch in which
   it stopped, so ms[0] includes up to a batch of no-op iterations. */
int glu_gfix_phase_ms(glu_gfix_ctx *ctx, double ms[3]);

/* Iterations of the last mixed-precision run spent in FP32 / FP64. */
int glu_gfix_mixed_split(glu_gfix_ctx *ctx, size_t *fp32_iters,
                         size_t *fp64_iters);

/* Per-slice results of the last Coulomb run (T entries each). */
int glu_gfix_slice_info(glu_gfix_ctx *ctx, size_t *iters, double *theta);

/* Coulomb: trajectory of one time slice in the last run, rows
   (theta_k, step, beta_k) as the slice's FACG steps took them
   (Gfix/CFACG.c:286-382): row 0 the steepest-descent start, then one row per
   iteration, continuations and restarts appended; at most 64 rows kept. */
int glu_gfix_slice_trace(glu_gfix_ctx *ctx, size_t slice, double *rows, size_t cap, size_t *n);

/* Stage-level entry points (used by the parity tests and the profiler):
   the derivative + packing of Gfix/FACG.c:38-73 (fields as 4*V complex =
   8*V doubles, field-major) with its theta; the Fourier acceleration of
   Gfix/FACG.c:77-95 applied to a host field; the three line-search probe
   sums and step of Gfix/line_search.c:274-341 for a host field; and the
   gauge update of Gfix/line_search.c:60-97 with a given step. */
int glu_gfix_stage_deriv(glu_gfix_ctx *ctx, double *fields, double *theta);
int glu_gfix_stage_fa(glu_gfix_ctx *ctx, const double *fields_in,
                      double *fields_out);
int glu_gfix_stage_line_search(glu_gfix_ctx *ctx, const double *fields,
                               double probes[3], double *alpha);
int glu_gfix_stage_egauge(glu_gfix_ctx *ctx, const double *fields,
                          double alpha);

/* Input generation (not part of GLU's gauge-fixing path; stands in for
   GLU's own generators, Update/KPHB.c:124-257 and
   Geometry/random_config.c:61-95, for lattices too large for the host):
   quenched Wilson heatbath (Cabibbo-Marinari, Kennedy-Pendleton) with `nor`
   overrelaxation passes per sweep on the device links, and a seeded
   Haar-like random gauge transform of them.  Counter-based random numbers:
   the result depends only on the seed. */
int glu_gfix_heatbath(glu_gfix_ctx *ctx, double beta, int sweeps, int nor,
                      uint64_t seed, int cold_start);
int glu_gfix_random_transform(glu_gfix_ctx *ctx, uint64_t seed);

/* GLU's random numbers.  GLU draws from one MWC_4096 stream per OpenMP
   thread (Utils/par_rng.c:43-186, Utils/par_MWC_4096.c:33-112): stream t is
   seeded with SEED + t and warmed up by 10000 draws, and a loop over n sites
   gives stream t libgomp's static block of them.  A context keeps a device
   copy of those streams; glu_gfix_rng_seed is initialise_par_rng with
   Latt.Seed = seed (nonzero) and Latt.Nthreads = nthreads (the result of
   every draw depends on both, as GLU's does).  The default, used when a
   run needs random numbers before any seeding, is seed 1 and 256 streams.
   Coulomb runs reseed from that seed on every call (GLU frees its streams
   at the end of Coulomb_FA, Gfix/CFACG.c:630). */
int glu_gfix_rng_seed(glu_gfix_ctx *ctx, uint32_t seed, int nthreads);
/* random_transform of Geometry/random_config.c:108-120 on the device links:
   Sunitary_gen (Matrix_Ops/gramschmidt_SSE.c:175-181: nine complex
   Gaussians of the polar method, Gram-Schmidt) per site from the context's
   streams, then U_mu(x) <- g(x) U_mu(x) g(x + mu)^dag.  The streams must be
   seeded. */
int glu_gfix_glu_random_transform(glu_gfix_ctx *ctx);
/* GLU's own quenched update, hb_update of Update/KPHB.c:121-257 with
   HEADER = UNIT (cold_start) or from the device links: `sweeps` times a
   Kennedy-Pendleton heatbath pass (Update/hb.c:156-194) over GLU's
   draughtboard colouring and `nor` overrelaxation passes
   (Update/relax.c:76-114), at BETA = beta, drawing from the context's
   copy of GLU's streams (glu_gfix_rng_seed: SEED, Nthreads) in GLU's
   order -- stream t takes libgomp's static block of each colour's sites.
   A heatbath pass runs one lane per stream. */
int glu_gfix_glu_heatbath(glu_gfix_ctx *ctx, double beta, int sweeps, int nor, int cold_start);
/* The first n doubles par_rng_dbl(thread) returns after GLU's seeding with
   (seed, nthreads): the host restatement of the streams the device copies,
   for checking them against GLU's. */
int glu_gfix_rng_draws(uint32_t seed, int nthreads, int thread, size_t n, double *out);

/* Per-kernel timing with HIP events on the context's stream.  Kernel
   classes: 0 derivative, 1 FFT axis pass, 2 Polak-Ribiere sums, 3 CG
   update, 4 line search, 5 gauge update (FASD), 6 gauge update fused with
   the next derivative (FACG), 7 Lie-field packing and 8 momentum
   projection of the gluon propagators; the Coulomb slice kernels 9
   derivative, 10 Polak-Ribiere sums, 11 CG update, 12 line search,
   13 slice-transform update; +14 for the FP32 variants.
   `mask` selects the classes timed (bit k = class k; -1 all, 0 off).
   While timing is on, a Landau run enqueues one iteration per batch, so
   no timed launch falls past the stop, and line searches skipped below
   CG_TOL are not counted: profile_read reports launches that did their
   work, their mean and total. */
int glu_gfix_profile(glu_gfix_ctx *ctx, int mask);
int glu_gfix_profile_read(glu_gfix_ctx *ctx, int kernel, size_t *launches,
                          double *mean_ms, double *total_ms);

/* Replay each batch of iterations from a HIP graph captured on first use
   (on = 1) instead of launching its kernels one by one.  Results are
   identical; only the launch cost changes.  Also set by the
   GLU_GFIX_GRAPH environment variable at creation.  Ignored while
   per-kernel profiling is on. */
int glu_gfix_set_graph(glu_gfix_ctx *ctx, int on);

/* Launch-shape autotuner: times the candidate shapes of this context's
   kernels on its own scratch fields -- pencils per workgroup of each FFT
   axis pass, the fused x-y plane pass against two axis passes, the
   XCD-contiguous workgroup order of the stencil kernels -- and keeps the
   fastest.  Results are cached per (dims, gftype, precision) in the
   process.  The choices change speed only.  `report` (may be NULL)
   receives the timings, one line per candidate.  Call between runs. */
int glu_gfix_autotune(glu_gfix_ctx *ctx, char *report, size_t cap);

/* Serving loop for streams of requests (GLU fixes one configuration per
   call, Wrappers/GF_wrap.c:179-212; this keeps its memory check,
   Utils/GLU_memcheck.c:72-95, as the admission rule for a device).
   Requests queue FIFO and are admitted while the device memory their
   contexts need (glu_gfix_footprint) fits under hbm_budget bytes (0: 90%
   of the free memory) and fewer than max_active run.  Admitted requests run
   concurrently, one context and stream each, driven batch by batch by one
   worker thread, so short runs overtake long ones.  Contexts are pooled
   and reused per shape.  A request's links are a device buffer in GLU's
   layout double[V][4][18], fixed in place (Landau requests: glu_gfix_landau
   semantics, restarts included; Coulomb: glu_gfix_run, on a host thread of
   its own so that the Landau runs in flight keep advancing); the caller
   keeps it alive until glu_gfix_server_wait returns.  Results are bitwise
   those of the same call on a context of their own.  Latency runs from
   submit to the output copy's completion. */
typedef struct glu_gfix_server glu_gfix_server;
int glu_gfix_server_create(glu_gfix_server **srv, size_t hbm_budget,
                           int max_active);
/* switch_theta for MIXED requests (0: the default); use_graph as
   glu_gfix_set_graph (default on); restart_seed as glu_gfix_landau
   (default 1).  Applies to requests admitted later. */
int glu_gfix_server_configure(glu_gfix_server *srv, double switch_theta,
                              int use_graph, uint64_t restart_seed);
/* Small lattices in shared launches (enable != 0; default off): FP64
   Landau CG requests whose extents are all 2..16 and that ask for no
   random transform skip the contexts.  Every queued request of one shape,
   gf_tune, accuracy and max_iters is fixed in one glu_gfix_facg_batch
   launch (one workgroup per lattice) on a host thread of the server; the
   launch takes one slot and its scratch (640 B per site per lattice)
   counts against the budget, and while it runs, later requests of that
   key gather for the next.  The other requests are admitted FIFO among
   themselves past the ones gathering.  A lattice whose
   FACG returns 123456789 keeps its input and goes back to the queue head
   for Landau()'s restarts on a context, as glu_gfix_landau takes them.  A
   request's result is then glu_gfix_facg_batch's for that lattice, which
   does not depend on the other lattices of the launch.  Applies to
   requests admitted later.  No counterpart in GLU, which fixes one
   configuration per call (Wrappers/GF_wrap.c:179-212). */
int glu_gfix_server_set_batching(glu_gfix_server *srv, int enable);
int glu_gfix_server_submit(glu_gfix_server *srv, const size_t dims[4],
                           int gftype, int precision, int method,
                           double gf_tune, double accuracy, size_t max_iters,
                           double *links_dev, int *ticket);
/* Submit with GLU's own input-file keys (read_gf_struct and the SEED and
   RANDOM_TRANSFORM tags of IO/input_reader.c:188-193, 444-481, 510-517):
   `options` holds KEY = VALUE lines (other keys are ignored).  GFTYPE
   (LANDAU or COULOMB), ACCURACY (the exponent: 10^-ACCURACY), MAX_ITERS and
   GF_TUNE are required.  SEED seeds the request's MWC_4096 streams (0: the
   entropy pool, as GLU); RANDOM_TRANSFORM = YES applies GLU's
   random_transform to the input first, and the Landau restarts then
   continue those streams; without it they start from SEED (default: the
   server's restart seed).  Method: FACG (CG). */
int glu_gfix_server_submit_glu(glu_gfix_server *srv, const size_t dims[4], int precision,
                               const char *options, double *links_dev, int *ticket);
/* Blocks until the request ends; returns its error if it failed. */
int glu_gfix_server_wait(glu_gfix_server *srv, int ticket, double *theta,
                         size_t *iters, double *latency_ms);
/* Completed requests, nearest-rank p50/p99 latency, peak bytes held. */
int glu_gfix_server_stats(glu_gfix_server *srv, size_t *completed,
                          double *p50_ms, double *p99_ms, size_t *peak_bytes);
/* Finishes every submitted request, then frees the pooled contexts.  Call
   it once no thread is inside glu_gfix_server_wait. */
int glu_gfix_server_destroy(glu_gfix_server *srv);

/* Bytes of device memory a context of these dims needs. */
size_t glu_gfix_footprint(const size_t dims[4], int gftype, int precision);

const char *glu_gfix_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
