// tune.hip -- launch-shape autotuner of a context.
//
// The tunable shapes are those that change speed and nothing else:
//   - pencils per workgroup of each LDS FFT axis pass (three variants),
//   - the fused x-y plane pass against two separate axis passes,
//   - the x values per workgroup of the z-t column pass,
//   - the workgroup order of the fused update and of the line search
//     (plain, XCD-contiguous, XCD time slabs).
// glu_gfix_autotune times every candidate with HIP events on the context's
// own scratch fields (the FFT passes and the derivative write only the
// field planes and the partial sums) and keeps the fastest.  The choices
// are cached per (dims, gftype, precision) for the life of the process, so
// pooled or repeated contexts of one shape tune once.
#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

struct Tuning {
  int pv[4];
  int use_xy;
  int zt_pv;
  int ord_eg, ord_ls;
  int ls_tile;
  std::string report;
};

std::mutex g_tune_mu;
std::map<std::tuple<size_t, size_t, size_t, size_t, int, int>, Tuning> g_tune_cache;

// mean duration of `reps` launches after one untimed launch; a candidate
// whose launch fails (a shape the device refuses) times as 1e30
template <class Fn> float time_ms(glu_gfix_ctx *c, Fn &&launch, int reps) {
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1e30f;
  (void)hipGetLastError();  // clear an earlier, already reported error
  launch();
  if (hipGetLastError() != hipSuccess) {
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return 1e30f;
  }
  (void)hipEventRecord(a, c->stream);
  for (int r = 0; r < reps; r++) launch();
  (void)hipEventRecord(b, c->stream);
  float ms = 1e30f;
  if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) ms /= (float)reps;
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

template <typename T> Tuning tune(glu_gfix_ctx *c) {
  Tuning t{};
  const int D = c->ps.D;
  const long long V = c->g.V;
  const int ns = D == 4 ? 1 : c->g.L[3];
  const long long sl = D == 4 ? 0 : c->g.Lcu;
  void *F = sizeof(T) == 4 ? c->F32 : c->F;
  void *U = sizeof(T) == 4 ? c->U32 : c->U64;
  constexpr int kReps = 5;
  char line[160];
  // FFT axis passes: the last axis is the fused forward * p^2 * backward pass
  float best_axis[4] = {0, 0, 0, 0};
  for (int a = 0; a < D; a++) {
    const int mode = a == D - 1 ? 2 : 0;
    float best = 1e30f;
    for (int v = 0; v < 3; v++) {
      FftAxis ax = c->ax[a];
      ax.pv = v;
      const float ms = time_ms(
          c, [&] { launch_fft_axis<T>(F, V, 4, ax, mode, c->ps, nullptr, c->stream, ns, sl, nullptr); }, kReps);
      snprintf(line, sizeof(line), "fft axis %d (N=%d) variant %d: %.4f ms\n", a, ax.N, v, ms);
      t.report += line;
      if (ms < best) {
        best = ms;
        t.pv[a] = v;
      }
    }
    best_axis[a] = best;
  }
  // the x-y plane pass against the two axis passes it replaces
  t.use_xy = 0;
  const long long nplane = (D == 4 ? V : c->g.Lcu) / c->g.S[2];
  bool have_xy = false;
  const float xy = time_ms(
      c,
      [&] {
        have_xy = launch_fft_xy<T>(F, V, 4, c->ax[0], c->ax[1], 0, nullptr, c->stream, nplane, ns, sl, nullptr);
      },
      kReps);
  if (have_xy) {
    t.use_xy = xy < best_axis[0] + best_axis[1];
    snprintf(line, sizeof(line), "fft x-y plane pass: %.4f ms against %.4f ms for two passes -> %s\n", xy,
             best_axis[0] + best_axis[1], t.use_xy ? "plane" : "axes");
    t.report += line;
  }
  // the z-t column pass (Landau, after a plane pass): x values per workgroup
  t.zt_pv = 0;
  if (t.use_xy && D == 4 &&
      launch_fft_zt<T>(F, V, 4, c->ax[0], c->ax[1], c->ax[2], c->ax[3], c->ps, nullptr, c->stream, false, 0)) {
    float best = 1e30f;
    for (int v = 0; v < 3; v++) {
      // a variant whose x values do not divide Lx has no kernel: the pass
      // would fall back to single-axis passes, so it is not a candidate
      if (!launch_fft_zt<T>(F, V, 4, c->ax[0], c->ax[1], c->ax[2], c->ax[3], c->ps, nullptr, c->stream, false, v))
        continue;
      const float ms = time_ms(
          c,
          [&] {
            launch_fft_zt<T>(F, V, 4, c->ax[0], c->ax[1], c->ax[2], c->ax[3], c->ps, nullptr, c->stream, true, v);
          },
          kReps);
      snprintf(line, sizeof(line), "fft z-t column pass variant %d: %.4f ms\n", v, ms);
      t.report += line;
      if (ms < best) {
        best = ms;
        t.zt_pv = v;
      }
    }
  }
  // workgroup order of the two stencil kernels that dominate an iteration,
  // each timed on its own with a private control block (not converged,
  // fixed step): the fused update writes only scratch (the second link
  // buffer, the fields, the partials, and G, which a run resets)
  Ctrl h{};
  h.alpha = 0.1;
  h.theta = 1.0;
  h.cg_tol = 5e-11;
  Ctrl *dc = nullptr;
  if (hipMalloc(&dc, sizeof(Ctrl)) == hipSuccess &&
      hipMemcpyAsync(dc, &h, sizeof(Ctrl), hipMemcpyHostToDevice, c->stream) == hipSuccess) {
    void *Unext = sizeof(T) == 4 ? c->U32b : c->U64b;
    void *sn = sizeof(T) == 4 ? c->sn32 : c->sn;
    void *Fold = sizeof(T) == 4 ? c->old32 : c->old;
    void *G = sizeof(T) == 4 ? c->G64 : nullptr;
    const int norder = geo_slab_order_ok(c->g) ? 3 : 2;
    const char *names[3] = {"plain", "XCD-contiguous", "XCD t-slabs"};
    float eg[3] = {1e30f, 1e30f, 1e30f}, ls[3] = {1e30f, 1e30f, 1e30f};
    for (int o = 0; o < norder; o++) {
      const Geo g = geo_order(c, o);
      eg[o] = time_ms(c, [&] { launch_egauge_deriv<T>(U, Unext, sn, G, Fold, F, c->part, g, dc, c->stream); }, kReps);
      ls[o] = time_ms(c, [&] { launch_line_search<T>(U, sn, c->part, g, dc, 1, c->stream); }, kReps);
      snprintf(line, sizeof(line), "order %s: gauge update + derivative %.4f ms, line search %.4f ms\n", names[o],
               eg[o], ls[o]);
      t.report += line;
    }
    t.ord_eg = t.ord_ls = 0;
    for (int o = 1; o < norder; o++) {
      if (eg[o] < eg[t.ord_eg]) t.ord_eg = o;
      if (ls[o] < ls[t.ord_ls]) t.ord_ls = o;
    }
    // the tiled line search (in the XCD-contiguous tile order) against the
    // best one-site-per-thread order
    t.ls_tile = 0;
    float best_ls = ls[t.ord_ls];
    for (int k = 1; k <= kLsTiles; k++) {
      if (!ls_tile_fits(c->g, k)) continue;
      const Geo g = geo_order(c, 1);
      const float ms = time_ms(c, [&] { launch_line_search<T>(U, sn, c->part, g, dc, 1, c->stream, k); }, kReps);
      const int *s = kLsTileShape[k - 1];
      snprintf(line, sizeof(line), "line search tile %dx%dx%dx%d, links %s: %.4f ms\n", s[0], s[1], s[2], s[3],
               k <= kLsTiles / 2 ? "held" : "reloaded", ms);
      t.report += line;
      if (ms < best_ls) {
        best_ls = ms;
        t.ls_tile = k;
      }
    }
    if (t.ls_tile) t.ord_ls = 1;
    snprintf(line, sizeof(line), "stencil order: update %s, line search %s%s\n", names[t.ord_eg], names[t.ord_ls],
             t.ls_tile ? " (tiles)" : "");
    t.report += line;
  } else {
    t.ord_eg = t.ord_ls = 1;
    t.ls_tile = c->ls_tile;
  }
  (void)hipStreamSynchronize(c->stream);
  if (dc) (void)hipFree(dc);
  return t;
}

void apply(glu_gfix_ctx *c, const Tuning &t) {
  for (int a = 0; a < 4; a++) c->ax[a].pv = t.pv[a];
  c->use_xy = t.use_xy;
  c->zt_pv = t.zt_pv;
  c->ord_eg = t.ord_eg;
  c->ord_ls = t.ord_ls;
  c->ls_tile = t.ls_tile;
  for (hipGraphExec_t &gx : c->graph)  // captured launches used the old shapes
    if (gx) {
      (void)hipGraphExecDestroy(gx);
      gx = nullptr;
    }
}

}  // namespace

extern "C" int glu_gfix_autotune(glu_gfix_ctx *c, char *report, size_t cap) {
  if (!c) return gfix_fail(-1, "null context");
  if (c->run.inflight) return gfix_fail(-1, "a run is in flight on this context");
  const auto key = std::make_tuple(c->dims[0], c->dims[1], c->dims[2], c->dims[3], c->gftype, c->precision);
  Tuning t;
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      t = it->second;
      cached = true;
    }
  }
  if (!cached) {
    if (hipStreamSynchronize(c->stream) != hipSuccess) return gfix_fail(-2, "autotune: stream not idle");
    // the FP32 kernels carry a mixed Landau run; Coulomb runs in FP64
    t = (c->precision == GLU_GFIX_MIXED && c->gftype == GLU_GFIX_LANDAU) ? tune<float>(c) : tune<double>(c);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
      return gfix_fail(-2, "autotune failed: %s", hipGetErrorString(hipGetLastError()));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = t;
  }
  apply(c, t);
  c->deriv_ready[0] = c->deriv_ready[1] = 0;  // the fields and partials were used as scratch
  if (report && cap) {
    const std::string r = (cached ? std::string("(cached)\n") : std::string()) + t.report;
    snprintf(report, cap, "%s", r.c_str());
  }
  return 0;
}
