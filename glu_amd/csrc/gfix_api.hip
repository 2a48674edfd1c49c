// gfix_api.hip -- host side of the plain-pointer C ABI: device contexts,
// the FACG / FASD iteration drivers (Gfix/FACG.c:181-446) with device-side
// scalars, and the mixed-precision FP32 -> FP64 schedule.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

static thread_local std::string g_err;

int gfix_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" const char *glu_gfix_last_error(void) { return g_err.c_str(); }

namespace {

constexpr double kCgTol = 5e-11;       // CG_TOL, Headers/GLU_definitions.h:449-455
// mixed precision: lattices up to this many sites stop the FP32 phase at
// its first stall (theta 30x above its minimum once below 1e-6)
constexpr long long kStallSites = 1 << 17;
constexpr int kGluFailures = 7;        // GF_GLU_FAILURES, Headers/GLU_definitions.h:400-402

int factor_radices(int N, int *rad) {
  int n = 0;
  while (N % 4 == 0 && n < 8) {
    rad[n++] = 4;
    N /= 4;
  }
  if (N % 2 == 0 && n < 8) {
    rad[n++] = 2;
    N /= 2;
  }
  while (N % 3 == 0 && n < 8) {
    rad[n++] = 3;
    N /= 3;
  }
  while (N % 5 == 0 && n < 8) {
    rad[n++] = 5;
    N /= 5;
  }
  while (N % 7 == 0 && n < 8) {
    rad[n++] = 7;
    N /= 7;
  }
  return N == 1 ? n : -1;
}

size_t pow2_floor(size_t x) {
  size_t p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

}  // namespace

size_t gfix_footprint(const size_t dims[4], int gftype, int precision) {
  const size_t V = dims[0] * dims[1] * dims[2] * dims[3];
  size_t b = 2 * V * 4 * 6 * sizeof(double2);  // FP64 links, current and next
  b += 3 * V * 4 * sizeof(double2);           // in, sn, in_old
  if (precision == GLU_GFIX_MIXED)
    b += 2 * V * 4 * 6 * sizeof(float2) + V * 6 * sizeof(double2) + 3 * V * 4 * sizeof(float2);
  if (gftype == GLU_GFIX_COULOMB) b += 2 * V * 6 * sizeof(double2);  // slice transforms + scratch
  b += ((V + kBlock - 1) / kBlock) * 4 * sizeof(double) + (64 << 10);
  return b;
}

extern "C" size_t glu_gfix_footprint(const size_t dims[4], int gftype, int precision) {
  return gfix_footprint(dims, gftype, precision);
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// The FFT axis descriptors and p^2 data of a context whose dims, geometry
// and gftype are set: fills c->ax[mu] for the transformed axes and c->ps,
// leaving the device table pointers to the caller, and writes the host
// tables: tw[mu] (dims[mu] twiddles exp(-2 pi i k / N) for mu < D) and cosv
// (the cos(2 pi n / L_mu) of every axis, back to back: Latt.twiddles,
// Geometry/init.c:69-71).  The extents were checked by check_extents.
void fft_plan(glu_gfix_ctx *c, double2 *const tw[4], double *cosv) {
  const int D = c->gftype == GLU_GFIX_LANDAU ? 4 : 3;
  const long long V = c->g.V;
  int off = 0;
  for (int mu = 0; mu < 4; mu++) {
    const int N = c->g.L[mu];
    c->ps.off[mu] = off;
    c->ps.L[mu] = N;
    const double twd = 6.283185307179586 / (double)N;
    for (int n = 0; n < N; n++) cosv[off + n] = cos(n * twd);
    off += N;
    if (mu >= D) continue;
    FftAxis ax{};
    ax.N = N;
    ax.nst = factor_radices(N, ax.rad);
    ax.S = c->g.S[mu];
    ax.npencil = (D == 4 ? V : c->g.Lcu) / N;  // per transform (Coulomb: per slice)
    ax.P = (int)std::min<size_t>(512, pow2_floor(2048 / (size_t)N));
    for (int k = 0; k < N; k++) {
      const double a = 6.283185307179586 * (double)k / (double)N;
      tw[mu][k] = make_double2(cos(a), -sin(a));
    }
    c->ax[mu] = ax;
  }
  c->ps.D = D;
  c->ps.max = D == 4 ? 4.0 * 4 / (double)V : 4.0 * 3 / (double)c->g.Lcu;
}

// The transformed axes (x, y, z, and t for Landau) go through the LDS FFT:
// a workgroup holds P >= 1 pencils of N points (N P <= 2048) and a stage
// runs at most 1024 butterflies, so N is limited to 2048 and must factor
// into the radices 2, 3, 5 and 7.
int check_extents(const size_t dims[4], int gftype) {
  const int D = gftype == GLU_GFIX_LANDAU ? 4 : 3;
  long double vol = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (dims[mu] < 2 || dims[mu] > 4096) return gfix_fail(-1, "lattice extent %zu out of range", dims[mu]);
    vol *= (long double)dims[mu];
    if (mu >= D) continue;
    if (dims[mu] > 2048)
      return gfix_fail(-1, "lattice extent %zu along a transformed axis exceeds the FFT's 2048 points", dims[mu]);
    int rad[8];
    if (factor_radices((int)dims[mu], rad) < 0)
      return gfix_fail(-1, "lattice extent %zu: extents must factor into 2, 3, 5 and 7", dims[mu]);
  }
  if (vol >= (long double)(1ll << 31))
    return gfix_fail(-1, "lattice volume %.0Lf is beyond the supported 2^31 - 1 sites", vol);
  return 0;
}

extern "C" int glu_gfix_create(glu_gfix_ctx **out, const size_t dims[4], int gftype, int precision) {
  if (!out || !dims) return gfix_fail(-1, "glu_gfix_create: null argument");
  *out = nullptr;
  if (gftype != GLU_GFIX_LANDAU && gftype != GLU_GFIX_COULOMB) return gfix_fail(-1, "unknown gftype %d", gftype);
  if (precision != GLU_GFIX_FP64 && precision != GLU_GFIX_MIXED)
    return gfix_fail(-1, "unknown precision %d", precision);
  if (int rc = check_extents(dims, gftype)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return gfix_fail(-3, "no HIP device: the gauge-fixing library has no CPU path");
  glu_gfix_ctx *c = new glu_gfix_ctx();
  for (int mu = 0; mu < 4; mu++) {
    c->dims[mu] = dims[mu];
    c->g.L[mu] = (int)dims[mu];
  }
  c->g.S[0] = 1;
  for (int mu = 1; mu < 4; mu++) c->g.S[mu] = c->g.S[mu - 1] * (long long)dims[mu - 1];
  c->g.V = c->g.S[3] * (long long)dims[3];
  c->g.Lcu = c->g.S[3];
  geo_set_div(c->g);
  c->g.xcd = 0;
  // untuned line search: tile 5 where the lattice takes it.  The tile
  // fixes the summation order of the probes, and late in a solve the
  // trajectory follows rounding: at 24^3x48 the autotuner's 32^3x64 pick
  // (tile 2, 7% faster per FP64 iteration) ends one gauge copy on another
  // Gribov copy than the CPU reference does in both of its orders
  c->ls_tile = ls_tile_fits(c->g, 5) ? 5 : 0;
  if (const char *e = getenv("GLU_GFIX_LS_TILE")) c->ls_tile = ls_tile_fits(c->g, atoi(e)) ? atoi(e) : 0;
  if (const char *e = getenv("GLU_GFIX_XCD")) {
    c->ord_eg = c->ord_ls = std::min(std::max(atoi(e), 0), 2);
    c->ord_cl = c->ord_eg ? 1 : 0;
  }
  c->gftype = gftype;
  c->precision = precision;
  const long long V = c->g.V;
  c->nblk = (V + kBlock - 1) / kBlock;
  int rc = 0;
#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      rc = gfix_fail(-2, "%s failed: %s", #x, hipGetErrorString(e_));          \
      goto bad;                                                                \
    }                                                                          \
  } while (0)
  {
    CK(hipGetDevice(&c->device));
    CK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CK(hipEventCreateWithFlags(&c->run_ev, hipEventDisableTiming));
    if (const char *e = getenv("GLU_GFIX_GRAPH")) c->use_graph = atoi(e);
    if (const char *e = getenv("GLU_GFIX_BATCH")) c->batch = std::max(1, std::min(64, atoi(e)));
    if (const char *e = getenv("GLU_GFIX_ZT")) c->use_zt = atoi(e);
    if (const char *e = getenv("GLU_GFIX_ZT_PV")) c->zt_pv = std::min(std::max(atoi(e), 0), 2);
    CK(hipMalloc(&c->U64, (size_t)V * 24 * sizeof(double2)));
    CK(hipMalloc(&c->U64b, (size_t)V * 24 * sizeof(double2)));
    CK(hipMalloc(&c->F, (size_t)V * 4 * sizeof(double2)));
    CK(hipMalloc(&c->sn, (size_t)V * 4 * sizeof(double2)));
    CK(hipMalloc(&c->old, (size_t)V * 4 * sizeof(double2)));
    CK(hipMalloc(&c->part, (size_t)c->nblk * 4 * sizeof(double) + 4096 * sizeof(double)));
    CK(hipMalloc(&c->ctrl, sizeof(Ctrl)));
    CK(hipHostMalloc(&c->hctrl, sizeof(Ctrl), hipHostMallocDefault));
    CK(hipMalloc(&c->trace, (size_t)c->trace_cap * 3 * sizeof(double)));
    if (precision == GLU_GFIX_MIXED) {
      CK(hipMalloc(&c->U32, (size_t)V * 24 * sizeof(float2)));
      CK(hipMalloc(&c->U32b, (size_t)V * 24 * sizeof(float2)));
      CK(hipMalloc(&c->G64, (size_t)V * 6 * sizeof(double2)));
      CK(hipMalloc(&c->F32, (size_t)V * 4 * sizeof(float2)));
      CK(hipMalloc(&c->sn32, (size_t)V * 4 * sizeof(float2)));
      CK(hipMalloc(&c->old32, (size_t)V * 4 * sizeof(float2)));
      if (const char *e = getenv("GLU_GFIX_SWITCH_THETA")) c->switch_theta = atof(e);
      if (const char *e = getenv("GLU_GFIX_SAFEGUARD")) c->sg_mode = atoi(e);
      if (const char *e = getenv("GLU_GFIX_BETA_CAP")) c->beta_cap = atof(e);
      // launch-bound lattices: an FP32 iteration costs about what an FP64
      // one does, so the FP32 phase hands over at its first stall
      c->stall_factor = V <= kStallSites ? 30.0 : 0.0;
      if (const char *e = getenv("GLU_GFIX_STALL_FACTOR")) c->stall_factor = atof(e);
      if (const char *e = getenv("GLU_GFIX_STALL_THETA")) c->stall_theta = atof(e);
    }
    // FFT axes, twiddles and the p^2 cos tables
    const int D = gftype == GLU_GFIX_LANDAU ? 4 : 3;
    size_t ncos = 0;
    for (int mu = 0; mu < 4; mu++) ncos += dims[mu];
    std::vector<double> cosv(ncos);
    std::vector<double2> tw[4];
    double2 *twp[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int mu = 0; mu < D; mu++) {
      tw[mu].resize(dims[mu]);
      twp[mu] = tw[mu].data();
    }
    fft_plan(c, twp, cosv.data());
    for (int mu = 0; mu < D; mu++) {
      CK(hipMalloc(&c->tw[mu], dims[mu] * sizeof(double2)));
      CK(hipMemcpy(c->tw[mu], twp[mu], dims[mu] * sizeof(double2), hipMemcpyHostToDevice));
      c->ax[mu].tw = c->tw[mu];
    }
    CK(hipMalloc(&c->cosv, ncos * sizeof(double)));
    CK(hipMemcpy(c->cosv, cosv.data(), ncos * sizeof(double), hipMemcpyHostToDevice));
    c->ps.cosv = c->cosv;
    if (gftype == GLU_GFIX_COULOMB) {
      rc = coulomb_alloc(c);
      if (rc) goto bad;
    }
  }
#undef CK
  *out = c;
  return 0;
bad:
  glu_gfix_destroy(c);
  return rc;
}

extern "C" int glu_gfix_destroy(glu_gfix_ctx *c) {
  if (!c) return 0;
  // teardown releases everything it can; a failing release has no remedy
  // here, so its status is dropped
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *bufs[] = {c->U64, c->U32, c->U64b, c->U32b, c->G64, c->Usnap, c->Uorig, c->Ubest, c->F, c->sn, c->old, c->part, c->ctrl, c->trace,
                  c->cosv, c->tw[0], c->tw[1], c->tw[2], c->tw[3], c->stage, c->lie, c->F32, c->sn32, c->old32, c->mwc};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  coulomb_free(c);
  glueprop_free(c);
  prof_harvest(c);
  for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
  for (hipGraphExec_t &gx : c->graph)
    if (gx) (void)hipGraphExecDestroy(gx);
  if (c->run_ev) (void)hipEventDestroy(c->run_ev);
  if (c->hctrl) (void)hipHostFree(c->hctrl);
  for (int b = 0; b < 2; b++) {
    if (c->hstage[b]) (void)hipHostFree(c->hstage[b]);
    if (c->hev[b]) (void)hipEventDestroy(c->hev[b]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

// ---------------------------------------------------------------------------
// links in / out
// ---------------------------------------------------------------------------
int ensure_stage(glu_gfix_ctx *c) {
  if (c->stage) return 0;
  if (hipMalloc(&c->stage, (size_t)kChunk * 72 * sizeof(double)) != hipSuccess)
    return gfix_fail(-2, "staging allocation failed");
  return 0;
}

// host threads for the gather / scatter of one chunk
static int host_threads() {
  static const int n = [] {
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
  }();
  return n;
}

// f(s0, n) over [s0, s0 + n) split among the host threads
template <typename F> static void par_sites(long long s0, long long n, F f) {
  const int nt = (int)std::min<long long>(host_threads(), std::max<long long>(1, n / 2048));
  if (nt <= 1) {
    f(s0, n);
    return;
  }
  std::vector<std::thread> th;
  const long long per = (n + nt - 1) / nt;
  for (int t = 1; t < nt; t++) {
    const long long a = s0 + t * per, m = std::min(per, s0 + n - a);
    if (m > 0) th.emplace_back([=] { f(a, m); });
  }
  f(s0, std::min(per, n));
  for (std::thread &x : th) x.join();
}

static int ensure_host_stage(glu_gfix_ctx *c) {
  if (int rc = ensure_stage(c)) return rc;
  for (int b = 0; b < 2; b++) {
    if (!c->hstage[b] && hipHostMalloc((void **)&c->hstage[b], (size_t)kHostChunk * 72 * sizeof(double)) != hipSuccess) {
      c->hstage[b] = nullptr;
      return gfix_fail(-2, "pinned staging allocation failed");
    }
    if (!c->hev[b] && hipEventCreateWithFlags(&c->hev[b], hipEventDisableTiming) != hipSuccess) {
      c->hev[b] = nullptr;
      return gfix_fail(-2, "staging event creation failed");
    }
  }
  return 0;
}

int upload_links(glu_gfix_ctx *c, SiteGather gather, void *arg) {
  if (int rc = ensure_host_stage(c)) return rc;
  const long long V = c->g.V;
  int k = 0;
  for (long long s0 = 0; s0 < V; s0 += kHostChunk, k++) {
    const long long n = std::min(kHostChunk, V - s0);
    const int b = k & 1;
    // the copy that read this half two chunks ago has finished
    if (k >= 2 && hipEventSynchronize(c->hev[b]) != hipSuccess) return gfix_fail(-2, "link upload sync failed");
    double *h = c->hstage[b];
    par_sites(s0, n, [&](long long a, long long m) { gather(arg, h + (a - s0) * 72, a, m); });
    if (hipMemcpyAsync(c->stage, h, (size_t)n * 72 * sizeof(double), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipEventRecord(c->hev[b], c->stream) != hipSuccess)
      return gfix_fail(-2, "link upload failed");
    launch_pack<double>((const double *)c->stage, c->U64, V, 4, s0, n, c->stream);
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "link upload: %s", "sync failed");
  return 0;
}

int download_links(glu_gfix_ctx *c, SiteScatter scatter, void *arg) {
  if (int rc = ensure_host_stage(c)) return rc;
  const long long V = c->g.V;
  // empty the half that chunk (s0, n) was copied into, once its copy is done
  auto drain = [&](int b, long long s0, long long n) {
    if (hipEventSynchronize(c->hev[b]) != hipSuccess) return gfix_fail(-2, "link download sync failed");
    const double *h = c->hstage[b];
    par_sites(s0, n, [&](long long a, long long m) { scatter(arg, h + (a - s0) * 72, a, m); });
    return 0;
  };
  int k = 0;
  long long prev0 = 0, prevn = 0;
  for (long long s0 = 0; s0 < V; s0 += kHostChunk, k++) {
    const long long n = std::min(kHostChunk, V - s0);
    const int b = k & 1;
    launch_unpack<double>(c->U64, (double *)c->stage, V, 4, s0, n, c->stream);
    if (hipMemcpyAsync(c->hstage[b], c->stage, (size_t)n * 72 * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipEventRecord(c->hev[b], c->stream) != hipSuccess)
      return gfix_fail(-2, "link download failed");
    // the previous chunk's half empties while this chunk is copied
    if (k >= 1)
      if (int rc = drain(b ^ 1, prev0, prevn)) return rc;
    prev0 = s0;
    prevn = n;
  }
  if (k >= 1)
    if (int rc = drain((k - 1) & 1, prev0, prevn)) return rc;
  return 0;
}

extern "C" int glu_gfix_set_links(glu_gfix_ctx *c, const double *links) {
  if (!c || !links) return gfix_fail(-1, "glu_gfix_set_links: null argument");
  return upload_links(
      c,
      [](void *arg, double *dst, long long s0, long long n) {
        memcpy(dst, (const double *)arg + s0 * 72, (size_t)n * 72 * sizeof(double));
      },
      (void *)links);
}

extern "C" int glu_gfix_get_links(glu_gfix_ctx *c, double *links) {
  if (!c || !links) return gfix_fail(-1, "glu_gfix_get_links: null argument");
  return download_links(
      c,
      [](void *arg, const double *src, long long s0, long long n) {
        memcpy((double *)arg + s0 * 72, src, (size_t)n * 72 * sizeof(double));
      },
      links);
}

extern "C" int glu_gfix_set_links_device(glu_gfix_ctx *c, const double *links_dev) {
  if (!c || !links_dev) return gfix_fail(-1, "glu_gfix_set_links_device: null argument");
  launch_pack<double>(links_dev, c->U64, c->g.V, 4, 0, c->g.V, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "pack failed");
  return 0;
}

extern "C" int glu_gfix_get_links_device(glu_gfix_ctx *c, double *links_dev) {
  if (!c || !links_dev) return gfix_fail(-1, "glu_gfix_get_links_device: null argument");
  launch_unpack<double>(c->U64, links_dev, c->g.V, 4, 0, c->g.V, c->stream);
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "unpack failed");
  return 0;
}

extern "C" int glu_gfix_set_graph(glu_gfix_ctx *c, int on) {
  if (!c) return gfix_fail(-1, "null context");
  c->use_graph = on ? 1 : 0;
  return 0;
}

extern "C" int glu_gfix_snapshot(glu_gfix_ctx *c) {
  if (!c) return gfix_fail(-1, "null context");
  const size_t bytes = (size_t)c->g.V * 24 * sizeof(double2);
  if (!c->Usnap && hipMalloc(&c->Usnap, bytes) != hipSuccess) return gfix_fail(-2, "snapshot allocation failed");
  if (hipMemcpyAsync(c->Usnap, c->U64, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "snapshot copy failed");
  return 0;
}

extern "C" int glu_gfix_restore(glu_gfix_ctx *c) {
  if (!c || !c->Usnap) return gfix_fail(-1, "no snapshot to restore");
  const size_t bytes = (size_t)c->g.V * 24 * sizeof(double2);
  if (hipMemcpyAsync(c->U64, c->Usnap, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    return gfix_fail(-2, "restore copy failed");
  return 0;
}

// ---------------------------------------------------------------------------
// iteration building blocks
// ---------------------------------------------------------------------------
template <typename T> static int pk(int id) { return id + (sizeof(T) == 4 ? PK_COUNT : 0); }

// the geometry with the workgroup order chosen for a kernel class
Geo geo_order(const glu_gfix_ctx *c, int order) {
  Geo g = c->g;
  g.xcd = (order == 2 && !geo_slab_order_ok(g)) ? 1 : order;
  return g;
}

// forward transform, p^2 preconditioner, backward transform of the 4 packed
// fields: over the whole lattice (Landau) or per time slice (Coulomb)
template <typename T>
void enq_fa(glu_gfix_ctx *c, void *F, const Ctrl *ctrl, const int *slice_done, const int *slice_map, int nmap) {
  const int D = c->ps.D;
  const long long V = c->g.V;
  const int ns = D == 4 ? 1 : (slice_map ? nmap : c->g.L[3]);
  const long long sl = D == 4 ? 0 : c->g.Lcu;
  // number of x-y planes per transform
  const long long nplane = (D == 4 ? V : c->g.Lcu) / c->g.S[2];
  auto xy = [&](int mode) {
    bool ok = false;
    GFIX_PROF(c, pk<T>(PK_FFT),
              ok = c->use_xy && launch_fft_xy<T>(F, V, 4, c->ax[0], c->ax[1], mode, ctrl, c->stream, nplane, ns, sl,
                                                  slice_done, slice_map));
    return ok;
  };
  const bool fused = xy(0);
  if (fused && D == 4 && c->use_zt &&
      launch_fft_zt<T>(F, V, 4, c->ax[0], c->ax[1], c->ax[2], c->ax[3], c->ps, ctrl, c->stream, false, c->zt_pv)) {
    GFIX_PROF(c, pk<T>(PK_FFT),
              launch_fft_zt<T>(F, V, 4, c->ax[0], c->ax[1], c->ax[2], c->ax[3], c->ps, ctrl, c->stream, true,
                               c->zt_pv));
    xy(1);
    return;
  }
  for (int a = fused ? 2 : 0; a < D - 1; a++)
    GFIX_PROF(c, pk<T>(PK_FFT), launch_fft_axis<T>(F, V, 4, c->ax[a], 0, c->ps, ctrl, c->stream, ns, sl, slice_done,
                                                             slice_map));
  GFIX_PROF(c, pk<T>(PK_FFT),
            launch_fft_axis<T>(F, V, 4, c->ax[D - 1], 2, c->ps, ctrl, c->stream, ns, sl, slice_done, slice_map));
  for (int a = D - 2; a >= (fused ? 2 : 0); a--)
    GFIX_PROF(c, pk<T>(PK_FFT), launch_fft_axis<T>(F, V, 4, c->ax[a], 1, c->ps, ctrl, c->stream, ns, sl, slice_done,
                                                             slice_map));
  if (fused) xy(1);
}
hipEvent_t prof_event(glu_gfix_ctx *c) {
  if (!c->prof_pool.empty()) {
    hipEvent_t e = c->prof_pool.back();
    c->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// collect the durations of recorded launches (call after a stream sync)
void prof_harvest(glu_gfix_ctx *c) {
  for (const ProfPending &p : c->prof_pending) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) c->prof_ms[p.id].push_back(ms);
    c->prof_pool.push_back(p.a);
    c->prof_pool.push_back(p.b);
  }
  c->prof_pending.clear();
}
template void enq_fa<double>(glu_gfix_ctx *, void *, const Ctrl *, const int *, const int *, int);
template void enq_fa<float>(glu_gfix_ctx *, void *, const Ctrl *, const int *, const int *, int);

static double landau_norm(const glu_gfix_ctx *c) { return 1.0 / (3.0 * (double)c->g.V); }

// the buffers one precision phase works on
struct Bufs {
  void *U, *G, *F, *sn, *old;
};
template <typename T> static Bufs bufs_for(glu_gfix_ctx *c) {
  if (sizeof(T) == 4) return {c->U32, c->G64, c->F32, c->sn32, c->old32};
  return {c->U64, nullptr, c->F, c->sn, c->old};
}

// index of a precision in the per-precision state (0 FP64, 1 FP32)
template <typename T> static constexpr int tix() { return sizeof(T) == 4 ? 1 : 0; }

// the CG gauge update, out of place and fused with the next derivative;
// the host's link pointer moves to the buffer it writes.  The derivative
// goes into the field buffer of in_old, and the two field pointers swap:
// the `in` of this iteration becomes the in_old of the next (GLU's copy
// old = in, Gfix/FACG.c:320-322, without moving the data)
template <typename T> static void swap_cg_buffers(glu_gfix_ctx *c) {
  std::swap(sizeof(T) == 4 ? c->U32 : c->U64, sizeof(T) == 4 ? c->U32b : c->U64b);
  std::swap(sizeof(T) == 4 ? c->F32 : c->F, sizeof(T) == 4 ? c->old32 : c->old);
  c->upar[tix<T>()] ^= 1;
  c->fpar[tix<T>()] ^= 1;
}
template <typename T> static void enq_update_deriv(glu_gfix_ctx *c) {
  const Bufs b = bufs_for<T>(c);
  void *nxt = sizeof(T) == 4 ? c->U32b : c->U64b;
  GFIX_PROF(c, pk<T>(PK_EGDERIV),
            launch_egauge_deriv<T>(b.U, nxt, b.sn, b.G, b.F, b.old, c->part, geo_order(c, c->ord_eg), c->ctrl,
                                   c->stream));
  swap_cg_buffers<T>(c);
  c->deriv_ready[tix<T>()] = 1;
  c->run.nenq++;
}

// SD start of steep_Landau_FACG, Gfix/FACG.c:249-265.  sn = in_old = in and
// inold = <in, in> depend on `in` only: sn is copied before the update,
// which reads the direction from it, and `in` becomes in_old when the
// update's field swap puts the next derivative in its place.
template <typename T> static void enq_sd_start(glu_gfix_ctx *c, const Bufs &b) {
  hipStream_t s = c->stream;
  launch_begin_segment(c->ctrl, 1, s);
  launch_deriv<T>(b.U, b.F, c->part, geo_order(c, c->ord_eg), c->ctrl, s);
  launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), s);
  enq_fa<T>(c, b.F, c->ctrl, nullptr);
  launch_line_search<T>(b.U, b.F, c->part, geo_order(c, c->ord_ls), c->ctrl, 1, s, c->ls_tile);
  launch_fin_ls(c->part, c->nblk, c->ctrl, c->trace, 1, 0.15, 0.30, s);
  launch_copy<T>(b.F, b.sn, c->g.V, c->ctrl, s);
  launch_pr_sums<T>(b.F, b.F, c->part, c->g.V, c->ctrl, s);
  launch_fin_sd(c->part, c->nblk, c->ctrl, s);
  enq_update_deriv<T>(c);
  launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), s);
}

// one FACG iteration (Gfix/FACG.c:272-334).  On entry with a pending
// derivative, theta of the current links is already in the control block
// (set with the end of the previous iteration or the SD start).
template <typename T> static void enq_cg_iter(glu_gfix_ctx *c, const Bufs &b) {
  hipStream_t s = c->stream;
  if (!c->deriv_ready[tix<T>()]) {
    GFIX_PROF(c, pk<T>(PK_DERIV), launch_deriv<T>(b.U, b.F, c->part, geo_order(c, c->ord_eg), c->ctrl, s));
    launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), s);
    c->deriv_ready[tix<T>()] = 1;
  }
  enq_fa<T>(c, b.F, c->ctrl, nullptr);
  GFIX_PROF(c, pk<T>(PK_PR), launch_pr_sums<T>(b.F, b.old, c->part, c->g.V, c->ctrl, s));
  launch_fin_pr(c->part, c->nblk, c->ctrl, s);
  GFIX_PROF(c, pk<T>(PK_CGUPD), launch_cg_update<T>(b.F, b.sn, c->g.V, c->ctrl, s));
  GFIX_PROF(c, pk<T>(PK_LSEARCH),
            launch_line_search<T>(b.U, b.sn, c->part, geo_order(c, c->ord_ls), c->ctrl, 0, s, c->ls_tile));
  launch_fin_ls(c->part, c->nblk, c->ctrl, c->trace, 0, 0.15, 0.30, s);
  enq_update_deriv<T>(c);
  launch_fin_iter(c->part, c->nblk, c->ctrl, c->trace, landau_norm(c), s);
}

// one FASD iteration (Gfix/FACG.c:202-235, fixed step GF_TUNE)
template <typename T> static void enq_sd_iter(glu_gfix_ctx *c, const Bufs &b) {
  hipStream_t s = c->stream;
  GFIX_PROF(c, pk<T>(PK_DERIV), launch_deriv<T>(b.U, b.F, c->part, geo_order(c, c->ord_eg), c->ctrl, s));
  launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), s);
  enq_fa<T>(c, b.F, c->ctrl, nullptr);
  GFIX_PROF(c, pk<T>(PK_EGAUGE), launch_egauge<T>(b.U, b.F, b.G, c->g, c->ctrl, 1, s));
  launch_iter_end(c->ctrl, c->trace, 1, s);
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int read_ctrl(glu_gfix_ctx *c) {
  if (hipMemcpyAsync(c->hctrl, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "control block read failed: %s", hipGetErrorString(hipGetLastError()));
  if (c->prof) prof_harvest(c);
  return 0;
}
static int write_ctrl(glu_gfix_ctx *c) {
  if (hipMemcpyAsync(c->ctrl, c->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "control block write failed");
  return 0;
}

// theta_test_lin = 2 * (FA theta) on the links U (Gfix/gftests.c:215-250)
template <typename T> static int theta_lin(glu_gfix_ctx *c, void *U, double *out) {
  c->deriv_ready[0] = c->deriv_ready[1] = 0;  // F and the partials are reused here
  launch_deriv<T>(U, c->F, c->part, c->g, nullptr, c->stream);
  // the partials' sum lands in the spare slot after the 4 nblk partials
  double *sum = c->part + 4 * c->nblk;
  launch_sum_part(c->part, c->nblk, sum, c->stream);
  double s = 0.0;
  if (hipMemcpyAsync(&s, sum, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "theta readback failed");
  *out = 2.0 * s * landau_norm(c);
  return 0;
}

// ---------------------------------------------------------------------------
// FACG() / FASD() of Gfix/FACG.c:341-446 as a non-blocking state machine.
// landau_begin() checks theta (FACG.c:349-353), initialises the control
// block and enqueues the first batch; landau_advance() inspects a finished
// batch and enqueues the next, or closes a segment: continuation below
// 1e3 acc (FACG.c:372-378), the 123456789 sentinel, or -- mixed precision --
// the FP32 -> FP64 switch.  The synchronous glu_gfix_run() and the serving
// loop drive the same machine, so a request's result does not depend on how
// it was driven.
// ---------------------------------------------------------------------------
namespace {
enum { PH_FP32 = 0, PH_FP64 = 1 };
enum { START_NONE = 0, START_SD = 1, START_PLAIN = 2 };
}  // namespace

// (re)initialise the control block for a phase; resume keeps the CG state
static int phase_ctrl_init(glu_gfix_ctx *c, int resume, double switch_theta, size_t max_iters) {
  RunState &r = c->run;
  Ctrl &h = *c->hctrl;
  if (resume) {
    h.switched = 0;
    h.total = 0;
  } else {
    memset(&h, 0, sizeof(Ctrl));
    h.acc = r.acc;
    h.gf_alpha = r.gf;
    h.cg_tol = kCgTol;
    h.trace_cap = c->trace_cap;
    h.trace_n = c->trace_n;
  }
  h.switch_theta = switch_theta;
  h.sg_mode = c->sg_mode;
  h.beta_cap = c->beta_cap;
  h.sg_uphill = 0;
  h.theta_min = 1e300;
  h.stall_theta = c->stall_theta;
  h.stall_factor = c->stall_factor;
  h.max_iters = (long long)max_iters;
  h.done = 1;
  r.max_phase = max_iters;
  r.segments = 0;
  return write_ctrl(c);
}

template <typename T> static void enq_iterations(glu_gfix_ctx *c) {
  // per-kernel timing enqueues one iteration per batch: the iteration that
  // stops a run is then the last one enqueued, and every timed launch did
  // its work
  const int nb = c->prof ? 1 : c->batch;
  for (int k = 0; k < nb; k++) {
    if (c->run.method == GLU_GFIX_CG)
      enq_cg_iter<T>(c, bufs_for<T>(c));
    else
      enq_sd_iter<T>(c, bufs_for<T>(c));
  }
}

// one batch of iterations, replayed from a captured graph when enabled
template <typename T> static int enq_batch_body(glu_gfix_ctx *c) {
  if (!c->use_graph || c->prof) {
    enq_iterations<T>(c);
    return 0;
  }
  // a pending derivative is enqueued outside the graph, so every captured
  // batch starts and ends with it in place (and, 8 swaps later, on the
  // same link and field buffers it started on)
  if (c->run.method == GLU_GFIX_CG && !c->deriv_ready[tix<T>()]) {
    const Bufs b = bufs_for<T>(c);
    launch_deriv<T>(b.U, b.F, c->part, geo_order(c, c->ord_eg), c->ctrl, c->stream);
    launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), c->stream);
    c->deriv_ready[tix<T>()] = 1;
  }
  const int kind = (sizeof(T) == 4 ? 2 : 0) + (c->run.method == GLU_GFIX_CG ? 0 : 1) + 4 * c->upar[tix<T>()] +
                   8 * c->fpar[tix<T>()];
  if (!c->graph[kind]) {
    hipGraph_t gr = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return gfix_fail(-2, "graph capture failed to start");
    enq_iterations<T>(c);  // moves the host's buffer pointers as the batch will
    if (hipStreamEndCapture(c->stream, &gr) != hipSuccess || !gr) return gfix_fail(-2, "graph capture failed");
    const hipError_t e = hipGraphInstantiate(&c->graph[kind], gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);  // the instantiated graph no longer needs it
    if (e != hipSuccess) return gfix_fail(-2, "graph instantiation failed: %s", hipGetErrorString(e));
  } else if (c->run.method == GLU_GFIX_CG) {
    if (c->batch & 1) swap_cg_buffers<T>(c);  // a replayed batch of odd length ends on the other buffers
    c->run.nenq += c->batch;
  }
  return hipGraphLaunch(c->graph[kind], c->stream) == hipSuccess ? 0 : gfix_fail(-2, "graph launch failed");
}

static int enq_batch(glu_gfix_ctx *c) {
  RunState &r = c->run;
  const bool f32 = r.phase == PH_FP32;
  // the host mirror holds the control block as of the last batch's end (or
  // as just written), so its update count is the device's at this point
  r.upd0 = c->hctrl->upd;
  r.nenq = 0;
  if (r.start == START_SD) {
    if (f32)
      enq_sd_start<float>(c, bufs_for<float>(c));
    else
      enq_sd_start<double>(c, bufs_for<double>(c));
  } else if (r.start == START_PLAIN) {
    launch_begin_segment(c->ctrl, 0, c->stream);
  }
  r.start = START_NONE;
  if (int rc = f32 ? enq_batch_body<float>(c) : enq_batch_body<double>(c)) return rc;
  // a launch the runtime refused (bad configuration, resources) is
  // reported here rather than leaving the iteration silently undone
  if (const hipError_t e = hipGetLastError(); e != hipSuccess)
    return gfix_fail(-2, "iteration batch launch failed: %s", hipGetErrorString(e));
  // without an event (a context that must not create one) the host waits
  // on the stream instead
  if (hipMemcpyAsync(c->hctrl, c->ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      (r.ev && hipEventRecord(r.ev, c->stream) != hipSuccess))
    return gfix_fail(-2, "batch enqueue failed: %s", hipGetErrorString(hipGetLastError()));
  r.inflight = 1;
  return 0;
}

// After a batch: the host swapped its link and field pointers once per
// fused update it enqueued, the device ran only those before the stop
// (Ctrl::upd counts them).  Undo the swaps of the updates that did not run.
static int settle_buffers(glu_gfix_ctx *c) {
  RunState &r = c->run;
  const long long live = c->hctrl->upd - r.upd0;
  if (live < 0 || live > r.nenq || (live < r.nenq && !c->hctrl->done))
    return gfix_fail(-2, "gauge update count %lld of %lld enqueued is inconsistent", live, r.nenq);
  if ((r.nenq - live) & 1) {
    if (r.phase == PH_FP32)
      swap_cg_buffers<float>(c);
    else
      swap_cg_buffers<double>(c);
  }
  r.nenq = 0;
  return 0;
}

// per-kernel timing: a line search that found theta below CG_TOL returned
// at once; its timing is dropped (one iteration per batch while timing)
static void drop_skipped_timings(glu_gfix_ctx *c) {
  if (c->hctrl->did_ls) return;
  std::vector<ProfPending> keep;
  for (const ProfPending &p : c->prof_pending) {
    if (p.id % PK_COUNT == PK_LSEARCH) {
      c->prof_pool.push_back(p.a);
      c->prof_pool.push_back(p.b);
    } else {
      keep.push_back(p);
    }
  }
  c->prof_pending.swap(keep);
}

int landau_begin(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters) {
  RunState &r = c->run;
  if (r.inflight) return gfix_fail(-1, "a run is still in flight on this context");
  r = RunState{};
  r.t_begin = now_ms();
  r.ev = c->run_ev;
  r.method = method;
  r.gf = gf;
  r.acc = acc;
  r.max_iters = max_iters;
  c->trace_n = 0;
  c->last_fp32_iters = c->last_fp64_iters = 0;
  c->deriv_ready[0] = c->deriv_ready[1] = 0;
  double th0 = 0.0;
  if (int rc = theta_lin<double>(c, c->U64, &th0)) return rc;
  if (th0 < acc) {  // Gfix/FACG.c:349-353
    r.finished = 1;
    r.theta = th0;
    r.iters = 0;
    r.t_end = r.t_fp32_end = r.t_switched = now_ms();
    return 0;
  }
  r.mixed = c->precision == GLU_GFIX_MIXED && c->switch_theta > acc;
  if (r.mixed) {
    // FP32 phase on a demoted copy; the site transform accumulates in FP64
    launch_demote(c->U64, c->U32, c->g.V * 24, c->stream);
    launch_identity<double>(c->G64, c->g.V, c->stream);
    r.phase = PH_FP32;
    if (int rc = phase_ctrl_init(c, 0, c->switch_theta, max_iters)) return rc;
  } else {
    r.phase = PH_FP64;
    if (int rc = phase_ctrl_init(c, 0, 0.0, max_iters)) return rc;
  }
  r.start = method == GLU_GFIX_CG ? START_SD : START_PLAIN;
  return enq_batch(c);
}

// 1: finished; 2: a batch ended and the next one is enqueued; 0: the batch
// in flight has not ended (never with block); < 0: error
int landau_advance(glu_gfix_ctx *c, int block) {
  RunState &r = c->run;
  if (r.finished) return 1;
  if (r.inflight) {
    const hipError_t e = r.ev ? (block ? hipEventSynchronize(r.ev) : hipEventQuery(r.ev))
                              : (block ? hipStreamSynchronize(c->stream) : hipStreamQuery(c->stream));
    if (e == hipErrorNotReady) return 0;
    if (e != hipSuccess) return gfix_fail(-2, "iteration batch failed: %s", hipGetErrorString(e));
    r.inflight = 0;
    if (int rc = settle_buffers(c)) return rc;
    if (c->prof) {
      drop_skipped_timings(c);
      prof_harvest(c);
    }
  }
  Ctrl &h = *c->hctrl;
  if (!h.done) {
    if (int rc = enq_batch(c)) return rc;
    return 2;
  }
  // a segment ended; the iteration that stopped it did not finalise the
  // theta of the derivative it left, so a CG segment that resumes from
  // here takes the derivative again
  c->deriv_ready[0] = c->deriv_ready[1] = 0;
  c->trace_n = h.trace_n;
  const double theta = h.theta;
  const long long loc = h.iters;
  if (h.switched) {
    // FP32 -> FP64: replay the accumulated transform on the FP64 links and
    // carry the CG state when the links agree with where FP32 stopped
    c->last_fp32_iters = (size_t)h.total;
    r.t_fp32_end = now_ms();
    launch_apply_gauge(c->U64, c->G64, c->g, c->stream);
    launch_promote(c->sn32, c->sn, c->g.V * 4, c->stream);
    launch_promote(c->old32, c->old, c->g.V * 4, c->stream);
    double thl = 0.0;
    if (int rc = theta_lin<double>(c, c->U64, &thl)) return rc;  // waits for the switch's work
    r.t_switched = now_ms();
    // a stalled FP32 phase left an uphill direction behind: restart CG
    const int keep = h.switched == 1 && 0.5 * thl < 10.0 * theta;
    const size_t left = r.max_iters > c->last_fp32_iters ? r.max_iters - c->last_fp32_iters : 1;
    r.phase = PH_FP64;
    if (int rc = phase_ctrl_init(c, keep, 0.0, left)) return rc;
    r.start = keep ? START_PLAIN : (r.method == GLU_GFIX_CG ? START_SD : START_PLAIN);
    if (int rc = enq_batch(c)) return rc;
    return 2;
  }
  size_t out = (size_t)loc;
  if ((size_t)loc >= r.max_phase) {
    // GLU repeats continuation segments without bound (Gfix/FACG.c:363-379);
    // a cap keeps a pathological input from spinning forever
    if (theta < 1e3 * r.acc && r.segments < 63) {
      if (c->verbose) fprintf(stdout, "[GF] Continuation run \n");
      r.segments++;
      r.start = r.method == GLU_GFIX_CG ? START_SD : START_PLAIN;
      if (int rc = enq_batch(c)) return rc;
      return 2;
    }
    out = r.method == GLU_GFIX_CG ? (size_t)123456789 : (size_t)123456789 + (size_t)loc;
  }
  if (r.phase == PH_FP32) {  // mixed run that stopped in FP32 (converged or failed)
    c->last_fp32_iters = (size_t)h.total;
    launch_apply_gauge(c->U64, c->G64, c->g, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "sync failed");
    r.iters = out;
  } else {
    c->last_fp64_iters = (size_t)h.total;
    r.iters = (r.mixed && out < 123456789) ? c->last_fp32_iters + c->last_fp64_iters : out;
  }
  r.theta = theta;
  r.finished = 1;
  r.t_end = now_ms();
  if (r.t_fp32_end == 0) r.t_fp32_end = r.t_switched = r.phase == PH_FP32 ? r.t_end : r.t_begin;
  return 1;
}

int landau_fix(glu_gfix_ctx *c, int method, double gf, double acc, size_t max_iters, double *theta, size_t *iters) {
  if (int rc = landau_begin(c, method, gf, acc, max_iters)) return rc;
  for (;;) {
    const int st = landau_advance(c, 1);
    if (st < 0) return st;
    if (st == 1) break;
  }
  *theta = c->run.theta;
  *iters = c->run.iters;
  return 0;
}

// The restart policy of Landau(), Gfix/Landau.c:301-345: a run returning
// exactly 123456789 (an FACG failure; FASD's 123456789 + loc does not
// qualify, as in GLU) is repeated from the input links after a random gauge
// transform, at most GF_GLU_FAILURES - 1 = 6 times
// (Headers/GLU_definitions.h:400-402).  Failed attempts add max_iters to
// the count, the successful one its own iterations, so a restarted run
// reports 123456789 + ... as GLU does.  GLU re-reads the config file and
// applies random_transform (Geometry/random_config.c:108-120) drawn from
// its MWC_4096 streams; here the input comes from a device copy and the
// transform from the context's copy of those streams: seeded with
// restart_seed at the first restart (GLU's fresh process with SEED =
// restart_seed), or continuing the context's streams when restart_seed is
// 0.  *failures: restarts taken (7 = gave up).
int landau_restart_policy(size_t first, size_t loc, size_t max_iters, size_t *iters) {
  if (loc != 123456789) {
    *iters = first + loc;
    return 1;
  }
  *iters = first + max_iters;
  return 0;
}

extern "C" int glu_gfix_landau(glu_gfix_ctx *c, int method, double gf_tune, double accuracy, size_t max_iters,
                               uint64_t restart_seed, double *theta, size_t *iters, int *failures) {
  if (!c || !theta || !iters) return gfix_fail(-1, "glu_gfix_landau: null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "glu_gfix_landau: not a Landau context");
  if (method != GLU_GFIX_CG && method != GLU_GFIX_SD) return gfix_fail(-1, "unknown method %d", method);
  if (max_iters == 0) return gfix_fail(-1, "max_iters must be positive");
  const size_t bytes = (size_t)c->g.V * 24 * sizeof(double2);
  if (!c->Uorig && hipMalloc(&c->Uorig, bytes) != hipSuccess) return gfix_fail(-2, "input copy allocation failed");
  if (hipMemcpyAsync(c->Uorig, c->U64, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    return gfix_fail(-2, "input copy failed");
  size_t it = 0;
  if (int rc = landau_fix(c, method, gf_tune, accuracy, max_iters, theta, &it)) return rc;
  int fail = 0;
  if (it == 123456789) {
    size_t total = it;
    for (fail = 1; fail < kGluFailures; fail++) {
      if (hipMemcpyAsync(c->U64, c->Uorig, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return gfix_fail(-2, "input reload failed");
      if (fail == 1 && restart_seed) {
        if (int rc = mwc_seed(c, (uint32_t)restart_seed, c->rng_threads)) return rc;
      }
      if (int rc = mwc_random_transform(c)) return rc;
      size_t loc = 0;
      if (int rc = landau_fix(c, method, gf_tune, accuracy, max_iters, theta, &loc)) return rc;
      if (landau_restart_policy(total, loc, max_iters, &total)) break;
    }
    it = total;
  }
  if (failures) *failures = fail;
  *iters = it;
  if (fail == kGluFailures) return gfix_fail(-5, "no convergence after %d random restarts (GF_TUNE %g)", fail - 1, gf_tune);
  return 0;
}

// luxury_copy_fast of Gfix/Landau.c:88-213 (LUXURY_GAUGE, with BEST_COPY
// as Headers/GLU_definitions.h:441-443 sets it): each copy is the input
// under a random gauge transform, fixed to 1e-6 (at most 1000 iterations);
// the copy with the smallest functional among those that neither failed
// nor used all 1000 iterations is fixed to `accuracy`.  The count is the
// last copy's iterations plus the final run's, as in GLU.  Each copy's
// random_transform draws from the context's MWC_4096 streams, seeded with
// `seed` at the start (or continuing when seed is 0), as GLU's copies draw
// from its process streams.  If no copy qualifies GLU continues from an
// unset buffer; here the final run starts from the input.
extern "C" int glu_gfix_luxury(glu_gfix_ctx *c, int method, double gf_tune, double accuracy, size_t max_iters,
                               int copies, uint64_t seed, double *theta, size_t *iters, int *best_copy) {
  if (!c || !theta || !iters) return gfix_fail(-1, "glu_gfix_luxury: null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "glu_gfix_luxury: not a Landau context");
  if (method != GLU_GFIX_CG && method != GLU_GFIX_SD) return gfix_fail(-1, "unknown method %d", method);
  if (max_iters == 0 || copies < 1) return gfix_fail(-1, "max_iters and copies must be positive");
  const size_t bytes = (size_t)c->g.V * 24 * sizeof(double2);
  if (!c->Uorig && hipMalloc(&c->Uorig, bytes) != hipSuccess) return gfix_fail(-2, "input copy allocation failed");
  if (!c->Ubest && hipMalloc(&c->Ubest, bytes) != hipSuccess) return gfix_fail(-2, "best copy allocation failed");
  if (hipMemcpyAsync(c->Uorig, c->U64, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->Ubest, c->U64, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    return gfix_fail(-2, "input copy failed");
  double best = 1.0;  // GLU's maxlink under BEST_COPY
  int chosen = -1;
  size_t it = 0;
  if (int rc = seed ? mwc_seed(c, (uint32_t)seed, c->rng_threads) : mwc_ensure(c)) return rc;
  for (int k = 0; k < copies; k++) {
    if (hipMemcpyAsync(c->U64, c->Uorig, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      return gfix_fail(-2, "input reload failed");
    if (int rc = mwc_random_transform(c)) return rc;
    double th = 0.0, f = 0.0;
    if (int rc = landau_fix(c, method, gf_tune, 1e-6, 1000, &th, &it)) return rc;
    if (int rc = glu_gfix_measure(c, nullptr, nullptr, nullptr, &f, nullptr)) return rc;
    if (c->verbose) fprintf(stdout, "  [COPY] %d [FUNCTIONAL] %1.15f [ITER] %zu ", k, f, it);
    if (f < best && it != 1000 && it != 123456789) {
      best = f;
      chosen = k;
      if (hipMemcpyAsync(c->Ubest, c->U64, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
        return gfix_fail(-2, "best copy failed");
      if (c->verbose) fprintf(stdout, " -> Copy accepted\n");
    } else if (c->verbose) {
      fprintf(stdout, " -> Copy rejected %e\n", f - best);
    }
  }
  if (hipMemcpyAsync(c->U64, c->Ubest, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
    return gfix_fail(-2, "best copy reload failed");
  size_t fin = 0;
  if (int rc = landau_fix(c, method, gf_tune, accuracy, max_iters, theta, &fin)) return rc;
  *iters = it + fin;
  if (best_copy) *best_copy = chosen;
  return 0;
}

extern "C" int glu_gfix_run(glu_gfix_ctx *c, int method, double gf_tune, double accuracy, size_t max_iters,
                            double *theta, size_t *iters) {
  if (!c || !theta || !iters) return gfix_fail(-1, "glu_gfix_run: null argument");
  if (method != GLU_GFIX_CG && method != GLU_GFIX_SD) return gfix_fail(-1, "unknown method %d", method);
  if (max_iters == 0) return gfix_fail(-1, "max_iters must be positive");
  if (c->gftype == GLU_GFIX_COULOMB) return coulomb_fix(c, method, gf_tune, accuracy, max_iters, theta, iters);
  return landau_fix(c, method, gf_tune, accuracy, max_iters, theta, iters);
}

extern "C" int glu_gfix_trace(glu_gfix_ctx *c, double *rows, size_t cap, size_t *n) {
  if (!c || !n) return gfix_fail(-1, "glu_gfix_trace: null argument");
  const size_t m = std::min<size_t>(cap, (size_t)c->trace_n);
  if (m && (hipMemcpyAsync(rows, c->trace, m * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess))
    return gfix_fail(-2, "trace readback failed");
  *n = m;
  return 0;
}

extern "C" int glu_gfix_phase_ms(glu_gfix_ctx *c, double ms[3]) {
  if (!c || !ms) return gfix_fail(-1, "null argument");
  const RunState &r = c->run;
  if (!r.finished) return gfix_fail(-1, "no finished run on this context");
  ms[0] = r.t_fp32_end - r.t_begin;
  ms[1] = r.t_switched - r.t_fp32_end;
  ms[2] = r.t_end - r.t_switched;
  return 0;
}

extern "C" int glu_gfix_mixed_split(glu_gfix_ctx *c, size_t *a, size_t *b) {
  if (!c || !a || !b) return gfix_fail(-1, "null argument");
  *a = c->last_fp32_iters;
  *b = c->last_fp64_iters;
  return 0;
}

// ---------------------------------------------------------------------------
// observables
// ---------------------------------------------------------------------------
extern "C" int glu_gfix_measure(glu_gfix_ctx *c, double *plaq, double *tr, double *mx, double *func, double *tlink) {
  if (!c) return gfix_fail(-1, "null context");
  launch_measure(c->U64, c->part, c->g, c->stream);
  std::vector<double> h(c->nblk * 4);
  if (hipMemcpyAsync(h.data(), c->part, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "measure readback failed");
  double p = 0, t = 0, tt = 0, m = -1e300;
  for (long long b = 0; b < c->nblk; b++) {
    p += h[4 * b];
    t += h[4 * b + 1];
    tt += h[4 * b + 2];
    m = h[4 * b + 3] > m ? h[4 * b + 3] : m;
  }
  const double V = (double)c->g.V;
  if (plaq) *plaq = 2.0 * p / (3.0 * 4 * 3 * V);
  if (tr) *tr = t / (4 * 3 * V);
  if (mx) *mx = m / 3.0;
  if (func) *func = 1.0 - t / (4 * 3 * V);
  if (tlink) *tlink = tt / (3 * V);
  return 0;
}

extern "C" int glu_gfix_theta(glu_gfix_ctx *c, double *th) {
  if (!c || !th) return gfix_fail(-1, "null argument");
  return theta_lin<double>(c, c->U64, th);
}

extern "C" int glu_gfix_reunit(glu_gfix_ctx *c) {
  if (!c) return gfix_fail(-1, "null context");
  launch_reunit(c->U64, c->g.V, 4, c->stream);
  return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : gfix_fail(-2, "reunit failed");
}

// ---------------------------------------------------------------------------
// stage-level entry points (FP64)
// ---------------------------------------------------------------------------
static int fields_in(glu_gfix_ctx *c, const double *f, void *dst) {
  return hipMemcpyAsync(dst, f, (size_t)c->g.V * 4 * sizeof(double2), hipMemcpyHostToDevice, c->stream) ==
                 hipSuccess
             ? 0
             : gfix_fail(-2, "field upload failed");
}
static int fields_out(glu_gfix_ctx *c, const void *src, double *f) {
  if (hipMemcpyAsync(f, src, (size_t)c->g.V * 4 * sizeof(double2), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return gfix_fail(-2, "field download failed");
  return 0;
}
static int ctrl_live(glu_gfix_ctx *c, double alpha, double theta) {
  memset(c->hctrl, 0, sizeof(Ctrl));
  c->hctrl->alpha = alpha;
  c->hctrl->gf_alpha = alpha;
  c->hctrl->theta = theta;
  c->hctrl->cg_tol = kCgTol;
  return write_ctrl(c);
}

extern "C" int glu_gfix_stage_deriv(glu_gfix_ctx *c, double *fields, double *theta) {
  if (!c || !fields || !theta) return gfix_fail(-1, "null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "stage entry points are Landau-only");
  if (int rc = ctrl_live(c, 0.0, 1.0)) return rc;
  launch_deriv<double>(c->U64, c->F, c->part, c->g, c->ctrl, c->stream);
  launch_fin_theta(c->part, c->nblk, c->ctrl, landau_norm(c), c->stream);
  if (int rc = fields_out(c, c->F, fields)) return rc;
  if (int rc = read_ctrl(c)) return rc;
  *theta = c->hctrl->theta;
  return 0;
}

extern "C" int glu_gfix_stage_fa(glu_gfix_ctx *c, const double *fin, double *fout) {
  if (!c || !fin || !fout) return gfix_fail(-1, "null argument");
  if (int rc = ctrl_live(c, 0.0, 1.0)) return rc;
  if (int rc = fields_in(c, fin, c->F)) return rc;
  enq_fa<double>(c, c->F, c->ctrl, nullptr);
  return fields_out(c, c->F, fout);
}

extern "C" int glu_gfix_stage_line_search(glu_gfix_ctx *c, const double *fields, double probes[3], double *alpha) {
  if (!c || !fields || !probes || !alpha) return gfix_fail(-1, "null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "stage entry points are Landau-only");
  if (int rc = ctrl_live(c, 0.0, 1.0)) return rc;
  if (int rc = fields_in(c, fields, c->sn)) return rc;
  launch_line_search<double>(c->U64, c->sn, c->part, c->g, c->ctrl, 1, c->stream, c->ls_tile);
  launch_fin_ls(c->part, c->nblk, c->ctrl, nullptr, 1, 0.15, 0.30, c->stream);
  if (int rc = read_ctrl(c)) return rc;
  for (int k = 0; k < 3; k++) probes[k] = c->hctrl->probe[k];
  *alpha = c->hctrl->alpha;
  return 0;
}

extern "C" int glu_gfix_stage_egauge(glu_gfix_ctx *c, const double *fields, double alpha) {
  if (!c || !fields) return gfix_fail(-1, "null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "stage entry points are Landau-only");
  if (int rc = ctrl_live(c, alpha, 1.0)) return rc;
  if (int rc = fields_in(c, fields, c->sn)) return rc;
  launch_egauge<double>(c->U64, c->sn, nullptr, c->g, c->ctrl, 0, c->stream);
  return hipStreamSynchronize(c->stream) == hipSuccess ? 0 : gfix_fail(-2, "egauge failed");
}

// ---------------------------------------------------------------------------
// per-kernel timing
// ---------------------------------------------------------------------------
extern "C" int glu_gfix_profile(glu_gfix_ctx *c, int mask) {
  if (!c) return gfix_fail(-1, "null context");
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "profile: stream sync failed");
  prof_harvest(c);
  c->prof = mask;
  if (mask)
    for (auto &v : c->prof_ms) v.clear();
  return 0;
}

extern "C" int glu_gfix_profile_read(glu_gfix_ctx *c, int kernel, size_t *launches, double *mean_ms,
                                     double *total_ms) {
  if (!c || kernel < 0 || kernel >= 2 * PK_COUNT || !launches || !mean_ms || !total_ms)
    return gfix_fail(-1, "glu_gfix_profile_read: bad argument");
  if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "profile read: stream sync failed");
  prof_harvest(c);
  // every timed launch did its work: while timing, a batch holds one
  // iteration (nothing is enqueued past a stop) and the timings of line
  // searches skipped below CG_TOL are dropped when the batch ends
  const std::vector<float> &v = c->prof_ms[kernel];
  double tot = 0.0;
  for (float ms : v) tot += ms;
  const size_t n = v.size();
  *launches = n;
  *total_ms = tot;
  *mean_ms = n ? tot / n : 0.0;
  return 0;
}

extern "C" int glu_gfix_set_switch_theta(glu_gfix_ctx *c, double theta) {
  if (!c || !(theta >= 0.0)) return gfix_fail(-1, "glu_gfix_set_switch_theta: bad argument");
  c->switch_theta = theta;
  return 0;
}
