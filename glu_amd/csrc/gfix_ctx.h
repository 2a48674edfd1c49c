// gfix_ctx.h -- the device context behind glu_gfix_ctx (internal).
#pragma once

#include <hip/hip_runtime.h>

#include "gfix_kernels.h"

#include <vector>

struct CoulombState;

// per-kernel event timing (glu_gfix_profile): kernel classes
enum ProfKernel {
  PK_DERIV = 0, PK_FFT, PK_PR, PK_CGUPD, PK_LSEARCH, PK_EGAUGE, PK_EGDERIV, PK_LIE, PK_PROP,
  // Coulomb slice kernels
  PK_CDERIV, PK_CPR, PK_CUPD, PK_CLINE, PK_CEXP, PK_COUNT
};
struct ProfPending {
  int id;
  hipEvent_t a, b;
};

// state of a Landau run driven batch by batch (see landau_begin/advance)
struct RunState {
  int method = 0, phase = 0, start = 0, mixed = 0;
  double gf = 0, acc = 0;
  size_t max_iters = 0, max_phase = 0;
  int segments = 0;
  int inflight = 0, finished = 0;
  // fused gauge updates enqueued in the batch in flight, and the device's
  // count of updates that had run when it was enqueued (Ctrl::upd)
  long long nenq = 0, upd0 = 0;
  hipEvent_t ev = nullptr;
  size_t iters = 0;
  double theta = 0.0;
  // host clock (ms since the run began) when the FP32 phase was seen to
  // end, when the switch's work was done, and when the run finished
  double t_begin = 0, t_fp32_end = 0, t_switched = 0, t_end = 0;
};

struct glu_gfix_ctx {
  size_t dims[4] = {0, 0, 0, 0};
  int device = 0;   // HIP device the context's memory lives on
  int verbose = 0;  // print GLU's progress lines (the drop-in entry points)
  gfix::Geo g{};
  int gftype = 0;
  int precision = 0;
  hipStream_t stream = nullptr;
  // links: FP64 column planes (4 dirs x 6 entries x V), FP32 copy and the
  // FP64 accumulated site transform for the mixed schedule, snapshot
  void *U64 = nullptr, *U32 = nullptr, *G64 = nullptr, *Usnap = nullptr;
  void *Uorig = nullptr;  // input links kept for the Landau() restarts
  void *Ubest = nullptr;  // best Gribov copy so far (glu_gfix_luxury)
  // second link buffers: the CG gauge update writes out of place and the
  // host swaps U64 <-> U64b (U32 <-> U32b) as it enqueues it; upar counts
  // the swaps mod 2 (FP64, FP32) so captured graphs know their start
  void *U64b = nullptr, *U32b = nullptr;
  int upar[2] = {0, 0};
  // the fused update also swaps the field buffers F <-> old (F32 <-> old32);
  // fpar counts those swaps mod 2
  int fpar[2] = {0, 0};
  // the fused update left the derivative of the current links (fields and
  // theta partials) in place: the next CG iteration needs no k_deriv
  int deriv_ready[2] = {0, 0};
  // FP32 packed fields of the mixed schedule (in, sn, in_old)
  void *F32 = nullptr, *sn32 = nullptr, *old32 = nullptr;
  // FP32 -> FP64 switch: 1e-11 was the fastest of the switch points
  // measured on 32^3x64 solves
  double switch_theta = 1e-11;
  // FP32-phase safeguards (Ctrl::sg_mode): beta capped at 3, chosen over
  // 20 gauge copies of a 32^3x64 beta = 6.0 config; the FP64 phase and
  // FP64 runs keep GLU's beta
  int sg_mode = 2;
  double beta_cap = 3.0;
  // FP32-phase stall exit (Ctrl::stall_factor, 0: off); set at creation
  // from the lattice size
  double stall_theta = 1e-6, stall_factor = 0.0;
  int use_xy = 1;     // x-y plane FFT passes where the plane fits a workgroup
  int use_zt = 1;     // Landau: z and t axes with p^2 in one pass after the x-y pass
  int zt_pv = 0;      // x values per workgroup of that pass (variant, tuned)
  // workgroup order (Geo::xcd) of the derivative / fused update and of the
  // line search; set by the autotuner
  int ord_eg = 1, ord_ls = 1;
  // line search on 256-site tiles (kLsTileShape[ls_tile - 1]; 0: one site
  // per thread); set at creation, then by the autotuner
  int ls_tile = 0;
  int ord_cl = 1;  // Coulomb slice stencils: XCD-contiguous (1) or plain (0)
  int use_graph = 0;  // replay iteration batches from captured HIP graphs
  // iterations enqueued between host checks; the iterations enqueued past
  // a stop run as no-ops, so the result does not depend on it
  int batch = 8;
  hipGraphExec_t graph[16] = {};  // CG/SD x FP64/FP32 x link- and field-buffer parity
  hipEvent_t run_ev = nullptr;
  RunState run;
  // packed fields (4 x V complex each): in, sn (search direction), in_old;
  // F and old trade places with every CG gauge update
  void *F = nullptr, *sn = nullptr, *old = nullptr;
  double *part = nullptr;  // per-block partial sums
  long long nblk = 0;
  gfix::Ctrl *ctrl = nullptr;   // device control block
  gfix::Ctrl *hctrl = nullptr;  // pinned host mirror
  double *trace = nullptr;
  long long trace_cap = 1 << 14;
  long long trace_n = 0;
  double2 *tw[4] = {nullptr, nullptr, nullptr, nullptr};
  double *cosv = nullptr;
  gfix::FftAxis ax[4]{};
  gfix::PsqInfo ps{};
  void *stage = nullptr;  // upload / download staging
  // pinned host halves of the host transfers (kHostChunk sites each) and
  // the events that free them (first host transfer)
  double *hstage[2] = {nullptr, nullptr};
  hipEvent_t hev[2] = {nullptr, nullptr};
  void *lie = nullptr;    // 16 packed Lie fields of the gluon propagators (first use)
  void *glue = nullptr;   // device copy of the last momentum list measured
  size_t last_fp32_iters = 0, last_fp64_iters = 0;
  CoulombState *cs = nullptr;
  // GLU's MWC_4096 streams on the device (x[mwc_nth], carry[mwc_nth]):
  // restarts, luxury copies and glu_gfix_glu_random_transform draw from
  // them; live once seeded (GLU's RNG_inited).  rng_seed / rng_threads
  // seed them when a run needs them unseeded (GLU's Latt.Seed, Nthreads).
  void *mwc = nullptr;
  int mwc_nth = 0, mwc_live = 0;
  uint32_t rng_seed = 1;
  int rng_threads = 256;
  // profiling: id = kernel class + PK_COUNT for the FP32 variants; prof is
  // a bit mask of the ids being timed
  int prof = 0;
  std::vector<ProfPending> prof_pending;
  std::vector<hipEvent_t> prof_pool;
  std::vector<float> prof_ms[2 * PK_COUNT];
};

// time one launch statement with HIP events on the context stream; a
// launch whose events cannot be made or recorded runs untimed
#define GFIX_PROF(c, id, stmt)                                                          \
  do {                                                                                 \
    ProfPending p_{(id), nullptr, nullptr};                                            \
    if (((c)->prof & (1 << (id))) && (p_.a = prof_event(c)) && (p_.b = prof_event(c)) && \
        hipEventRecord(p_.a, (c)->stream) == hipSuccess) {                             \
      stmt;                                                                            \
      if (hipEventRecord(p_.b, (c)->stream) == hipSuccess) {                           \
        (c)->prof_pending.push_back(p_);                                               \
      } else {                                                                         \
        (c)->prof_pool.push_back(p_.a);                                                \
        (c)->prof_pool.push_back(p_.b);                                                \
      }                                                                                \
    } else {                                                                           \
      if (p_.a) (c)->prof_pool.push_back(p_.a);                                        \
      if (p_.b) (c)->prof_pool.push_back(p_.b);                                        \
      stmt;                                                                            \
    }                                                                                  \
  } while (0)
hipEvent_t prof_event(glu_gfix_ctx *c);
void prof_harvest(glu_gfix_ctx *c);

int gfix_fail(int code, const char *fmt, ...);
size_t gfix_footprint(const size_t dims[4], int gftype, int precision);
template <typename T>
void enq_fa(glu_gfix_ctx *c, void *F, const gfix::Ctrl *ctrl, const int *slice_done, const int *slice_map = nullptr,
            int nmap = 0);

int coulomb_alloc(glu_gfix_ctx *c);
void coulomb_free(glu_gfix_ctx *c);
int coulomb_fix(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters, double *theta, size_t *iters);
int landau_fix(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters, double *theta, size_t *iters);
int landau_begin(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters);
int landau_advance(glu_gfix_ctx *c, int block);
int landau_restart_policy(size_t first, size_t loc, size_t max_iters, size_t *iters);
// FACG of nlat small lattices of one shape in one launch (links at the
// device pointers lat[0..nlat)), as glu_gfix_facg_batch; with keep_failed
// a lattice that fails keeps its input links
int facg_batch_run(const size_t dims[4], size_t nlat, double *const *lat, double gf_tune, double accuracy,
                   size_t max_iters, size_t *iters, double *theta, double *trace, size_t trace_cap,
                   int keep_failed);
gfix::Geo geo_order(const glu_gfix_ctx *c, int order);
extern "C" int glu_gfix_random_transform(glu_gfix_ctx *c, uint64_t seed);
extern "C" int glu_gfix_measure(glu_gfix_ctx *c, double *plaq, double *tr, double *mx, double *func, double *tlink);
extern "C" int glu_gfix_reunit(glu_gfix_ctx *c);

// links cross the host boundary in chunks of this many sites through the
// staging buffer (kChunk * 72 doubles, allocated on first use)
constexpr long long kChunk = 1 << 17;
int ensure_stage(glu_gfix_ctx *c);
// Host transfers of GLU-layout links, kHostChunk sites at a time through
// two pinned buffers: host threads fill one (gather) or empty one
// (scatter) while the device copies the other.  gather(dst, s0, n) writes
// sites [s0, s0 + n) to dst (n x 72 doubles); scatter(src, s0, n) reads
// them from src.  Both are called from several host threads at once, on
// disjoint site ranges.
constexpr long long kHostChunk = 1 << 16;
using SiteGather = void (*)(void *arg, double *dst, long long s0, long long n);
using SiteScatter = void (*)(void *arg, const double *src, long long s0, long long n);
int upload_links(glu_gfix_ctx *c, SiteGather gather, void *arg);
int download_links(glu_gfix_ctx *c, SiteScatter scatter, void *arg);
int check_extents(const size_t dims[4], int gftype);
void fft_plan(glu_gfix_ctx *c, double2 *const tw[4], double *cosv);
void glueprop_free(glu_gfix_ctx *c);

// GLU's MWC_4096 streams (mwc.hip)
struct MwcHost {
  std::vector<uint32_t> x, carry;
  bool live = false;
};
extern MwcHost g_mwc;  // the process stream of the drop-in entry points
void mwc_seed_states(uint32_t seed, int nth, uint32_t *x, uint32_t *carry);
int mwc_host_init(uint32_t seed, int nth);
void mwc_host_free();
double mwc_host_dbl(int thread);
int mwc_seed(glu_gfix_ctx *c, uint32_t seed, int nth);
int mwc_ensure(glu_gfix_ctx *c);
int mwc_from_host(glu_gfix_ctx *c);
int mwc_to_host(glu_gfix_ctx *c);
int mwc_sunitary(glu_gfix_ctx *c, void *G, long long Vp, long long n, const std::vector<long long> &first,
                 bool reunit);
int mwc_random_transform(glu_gfix_ctx *c);
