// glueprop.hip -- transverse and longitudinal gluon propagators of the
// device links: the GLUON_PROPS branch of GLU's cuts_struct
// (Cuts/cuts.c:114-184), downstream of the gauge fixing.
//
// GLU transforms all nine entries of the four Lie fields (36 complex FFTs).
// A_mu = Hermitian_proj(U_mu) is Hermitian and traceless, so four complex
// fields per direction carry it: the two real diagonals packed as
// a00 + i a11, and the upper entries a01, a02, a12.  Their transforms at p
// and -p give every entry at p (the lower ones and the split diagonals by
// the symmetry of the transform of a real / conjugated field), so one
// 16-field transform replaces GLU's 36.  The momentum list is GLU's
// (get_mom_veclist, Cuts/cut_routines.c:383-455) and is built on the host
// from the same formulas, like a transform plan; the projections run on the
// device, one thread per listed momentum.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

constexpr double kTwoPi = 6.283185307179586;   // TWOPI, Headers/GLU_definitions.h:139
constexpr double kPrecTol = 3 * 1.0e-14;       // PREC_TOL, Headers/GLU_definitions.h:362
constexpr double kDegToRad = 0.01745329251994330;  // ang(), Cuts/cut_routines.c:55

// the four packed Lie fields of every direction: Hermitian_proj
// (Matrix_Ops/exactQ.c:358-373) of each link, Z[(mu * 4 + f) * V + i]
__global__ __launch_bounds__(kBlock) void k_lie_pack(const double2 *__restrict__ U, double2 *__restrict__ Z,
                                                    long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  // all four directions' loads in flight at once
#pragma unroll
  for (int mu = 0; mu < kND; mu++) {
    const su3<double> u = load_link<double>(U, mu, i, V);
    const double c0 = u.e[0].im, c4 = u.e[4].im, c8 = u.e[8].im;
    // -i (a - conj(b)) / 2
    auto off = [](cx<double> a, cx<double> b) {
      return make_double2(0.5 * (a.im + b.im), -0.5 * (a.re - b.re));
    };
    Z[(long long)(mu * 4 + 0) * V + i] = make_double2((2 * c0 - c4 - c8) / 3., (2 * c4 - c0 - c8) / 3.);
    Z[(long long)(mu * 4 + 1) * V + i] = off(u.e[1], u.e[3]);
    Z[(long long)(mu * 4 + 2) * V + i] = off(u.e[2], u.e[6]);
    Z[(long long)(mu * 4 + 3) * V + i] = off(u.e[5], u.e[7]);
  }
}

__device__ __forceinline__ double2 cmulz(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 conjz(double2 a) { return make_double2(a.x, -a.y); }

// the listed momentum tables: axis mu of entry j -> (2 sin(k pi/L), phase)
struct MomTabs {
  const int *mom;
  const long long *site, *msite;
  const double *sin2, *cosk;
  int off[4], half[4];
};

// the nine entries of A_mu at listed entry j from the packed transforms at
// its site and the site of -p, times the Landau phase exp(i p_mu / 2) of
// correct_pspace_landau (Cuts/pspace_landau.c:28-86); p_mu -> *pmu
__device__ __forceinline__ void lie_at(const double2 *__restrict__ Z, long long V, const MomTabs &m, int j, int mu,
                                       double2 (&A)[9], double *pmu) {
  const int t = m.off[mu] + m.mom[4 * j + mu] + m.half[mu];
  *pmu = m.sin2[t];
  const double2 e = make_double2(m.cosk[t], 0.5 * m.sin2[t]);
  const long long s = m.site[j], ms = m.msite[j];
  const double2 *z = Z + (long long)mu * 4 * V;
  const double2 z0p = z[s], z0m = z[ms];
  // diagonal: a00 = (Z0(p) + conj Z0(-p)) / 2, a11 = (Z0(p) - conj Z0(-p)) / 2i
  const double2 a00 = make_double2(0.5 * (z0p.x + z0m.x), 0.5 * (z0p.y - z0m.y));
  const double2 a11 = make_double2(0.5 * (z0p.y + z0m.y), -0.5 * (z0p.x - z0m.x));
  A[0] = a00;
  A[4] = a11;
  A[8] = make_double2(-a00.x - a11.x, -a00.y - a11.y);
  // upper entries at p, lower ones from conj of the upper at -p
  A[1] = z[V + s];
  A[2] = z[2 * V + s];
  A[5] = z[3 * V + s];
  A[3] = conjz(z[V + ms]);
  A[6] = conjz(z[2 * V + ms]);
  A[7] = conjz(z[3 * V + ms]);
#pragma unroll
  for (int k = 0; k < 9; k++) A[k] = cmulz(A[k], e);
}

// One listed momentum per thread: the nine entries of every A_mu(p) from
// the packed transforms at p and -p, the exp(i p_mu / 2) factor of
// correct_pspace_landau (Cuts/pspace_landau.c:28-86), and
// project_trans_long (Cuts/glueprop.c:60-90) with its normalisation.
// sin2 / cosk hold 2 sin(k pi / L) and cos(k pi / L) per axis for
// k = -L/2 .. L - 1 - L/2 (offset L/2, axis mu at off[mu]), looked up with
// the listed momentum (GLU's psq_calc recovers it from the site, the same
// for even extents; for an odd extent it misplaces k = (L-1)/2).  err
// receives the momentum-space gauge condition term
// |sum_mu,j p_mu A_mu,j(p)| / NC.
__global__ __launch_bounds__(kBlock) void k_gluon_prop(const double2 *__restrict__ Z, long long V, int n, MomTabs m,
                                                      int zero_pos, double g2_norm, double g0_norm, double long_norm,
                                                      double *__restrict__ G, double *__restrict__ F,
                                                      double *__restrict__ err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double p[4];
  double2 A[4][9];
  double2 esum = make_double2(0.0, 0.0);
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    lie_at(Z, V, m, j, mu, A[mu], &p[mu]);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      esum.x += A[mu][k].x * p[mu];
      esum.y += A[mu][k].y * p[mu];
    }
  }
  double tspsq = 0.0;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) tspsq += p[mu] * p[mu];
  const double spsq = 1.0 / ((tspsq < kPrecTol) ? 1.0 : tspsq);
  double tr_ = 0.0, lo = 0.0;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
      double re = 0.0;  // Re tr(A_mu A_nu^dag), trace_ab_dag (Matrix_Ops/U_Nops.c:828-835)
#pragma unroll
      for (int k = 0; k < 9; k++) re += A[mu][k].x * A[nu][k].x + A[mu][k].y * A[nu][k].y;
      const double common = p[mu] * p[nu] * spsq * re;
      tr_ += (mu != nu) ? -common : re - common;
      lo += common;
    }
  G[j] = tr_ * (j == zero_pos ? g0_norm : g2_norm);
  F[j] = lo * long_norm;
  err[j] = sqrt(esum.x * esum.x + esum.y * esum.y) / 3.0;
}

__device__ __forceinline__ void mul33(double2 (&c)[9], const double2 (&a)[9], const double2 (&b)[9]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 3; s++) {
      double2 t = cmulz(a[r * 3], b[s]);
      const double2 u = cmulz(a[r * 3 + 1], b[3 + s]), v = cmulz(a[r * 3 + 2], b[6 + s]);
      t.x += u.x + v.x;
      t.y += u.y + v.y;
      c[r * 3 + s] = t;
    }
}
// tr(a b), trace_ab (Matrix_Ops/U_Nops.c:799-807)
__device__ __forceinline__ double2 tr_ab(const double2 (&a)[9], const double2 (&b)[9]) {
  double2 t = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 3; s++) {
      const double2 u = cmulz(a[r * 3 + s], b[s * 3 + r]);
      t.x += u.x;
      t.y += u.y;
    }
  return t;
}

// write_exceptional_g2g3_MOMgg (Cuts/MOMgg.c:141-215) with its default
// contraction_matrices (:94-121), one listed momentum p per thread; -p is
// the mirror entry of the list and 0 its middle entry:
//   g2 = sum_mu Re tr(A_mu(p) A_mu(-p)),
//   g3 = Re [tr(AA Arho) + tr(Arho AP AM)] / p^2, AA = sum_mu A_mu(p)
//   A_mu(-p), Arho = sum_mu p_mu A_mu(0), AP = sum_mu p_mu A_mu(p),
//   AM = -sum_mu p_mu A_mu(-p) / p^2.
__global__ __launch_bounds__(kBlock) void k_gluon_exceptional(const double2 *__restrict__ Z, long long V, int n,
                                                             MomTabs m, double g2_norm, double g0_norm,
                                                             double g3_norm, double *__restrict__ g2,
                                                             double *__restrict__ g3, double *__restrict__ err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int cj = n - 1 - j, zero = (n - 1) / 2;
  double p[4], spsq = 0.0;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    p[mu] = m.sin2[m.off[mu] + m.mom[4 * j + mu] + m.half[mu]];
    spsq += p[mu] * p[mu];
  }
  if (spsq < kPrecTol) spsq = 1.0;
  double2 Arho[9], AA[9], AP[9], AM[9];
#pragma unroll
  for (int k = 0; k < 9; k++) Arho[k] = AA[k] = AP[k] = AM[k] = make_double2(0.0, 0.0);
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    double2 a0[9];
    double q;
    lie_at(Z, V, m, zero, mu, a0, &q);
#pragma unroll
    for (int k = 0; k < 9; k++) {
      Arho[k].x += a0[k].x * p[mu];
      Arho[k].y += a0[k].y * p[mu];
    }
  }
  double s2 = 0.0;
  double2 esum = make_double2(0.0, 0.0);
#pragma unroll 1
  for (int mu = 0; mu < 4; mu++) {
    double2 ap[9], am[9], t[9];
    double q;
    lie_at(Z, V, m, j, mu, ap, &q);
    lie_at(Z, V, m, cj, mu, am, &q);
    s2 += tr_ab(ap, am).x;
    mul33(t, ap, am);
    const double wm = -p[mu] / spsq;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      AA[k].x += t[k].x;
      AA[k].y += t[k].y;
      AP[k].x += ap[k].x * p[mu];
      AP[k].y += ap[k].y * p[mu];
      AM[k].x += am[k].x * wm;
      AM[k].y += am[k].y * wm;
      esum.x += ap[k].x * p[mu];
      esum.y += ap[k].y * p[mu];
    }
  }
  g2[j] = s2 * (j != zero ? g2_norm : g0_norm);
  double2 t[9];
  mul33(t, Arho, AP);
  const double2 trans = tr_ab(AA, Arho), wiv = tr_ab(t, AM);
  g3[j] = (trans.x + wiv.x) / spsq * g3_norm;
  err[j] = sqrt(esum.x * esum.x + esum.y * esum.y) / 3.0;
}

// cylinder tests of Cuts/cut_routines.c:191-298 on the SIN_MOM momentum q
void diagonal(double x[4], int i) {
  int subvol = 1;
  for (int mu = 0; mu < 4; mu++) {
    x[mu] = (double)(((i - i % subvol) / subvol) % 2);
    if (x[mu] == 0) x[mu] = -1;
    subvol *= 2;
  }
}
bool cylinder_df(const double q[4], double width) {
  const double norm = 1. / (2. * std::sqrt(4.0));
  for (int d = 0; d < 16; d++) {
    double x[4], sp = 0.0, mod = 0.0;
    diagonal(x, d);
    for (int nu = 0; nu < 4; nu++) sp += q[nu] * x[nu];
    sp *= norm;
    for (int nu = 0; nu < 4; nu++) {
      const double t = q[nu] - sp * x[nu];
      mod += t * t;
    }
    if (std::sqrt(mod) <= width) return true;
  }
  return false;
}
bool cylinder_cjd(const double q[4], double width, double arc) {
  double qnorm = 0.0;
  for (int mu = 0; mu < 4; mu++) qnorm += q[mu] * q[mu];
  if (qnorm < kPrecTol) return true;
  qnorm = 1.0 / std::sqrt(qnorm);
  for (int d = 0; d < 16; d++) {
    double x[4], nnorm = 0.0, sp = 0.0;
    diagonal(x, d);
    for (int nu = 0; nu < 4; nu++) nnorm += x[nu] * x[nu];
    nnorm = qnorm / std::sqrt(nnorm);
    for (int nu = 0; nu < 4; nu++) sp += (q[nu] * x[nu] * nnorm);
    const double theta = std::acos(sp);
    const double delta_q = std::sin(theta) / qnorm;
    if ((delta_q <= width) && (theta <= arc * kDegToRad)) return true;
  }
  return false;
}

// get_mom_veclist with safe_momenta (Cuts/cut_routines.c:73-95, 300-455):
// momenta k_mu = c_mu - L_mu / 2 enumerated with x fastest, kept by the cut
int momentum_list(const size_t dims[4], int type, size_t max_mom, double cyl_width, double angle,
                  std::vector<int> &mom, std::vector<long long> &site) {
  if (type < GLU_GFIX_CUT_HYPERCUBIC || type > GLU_GFIX_CUT_CONICAL) return -1;  // check_psq
  size_t small = dims[0];
  for (int mu = 1; mu < 4; mu++) small = std::min(small, dims[mu]);
  double rats[4];
  for (int mu = 0; mu < 4; mu++) rats[mu] = (double)small / (double)dims[mu];
  const double latt_width = kTwoPi * cyl_width / small;
  const size_t V = dims[0] * dims[1] * dims[2] * dims[3];
  mom.clear();
  site.clear();
  for (size_t i = 0; i < V; i++) {
    int k[4];
    size_t subvol = 1;
    for (int mu = 0; mu < 4; mu++) {
      k[mu] = (int)(((i - i % subvol) / subvol) % dims[mu]) - (int)(dims[mu] >> 1);
      subvol *= dims[mu];
    }
    double psq = 0.0;
    for (int mu = 0; mu < 4; mu++) {
      const double t = rats[mu] * k[mu];
      psq += t * t;
    }
    bool keep;
    if (type == GLU_GFIX_CUT_HYPERCUBIC) {
      const double root = std::sqrt((double)max_mom);
      int flag = 0;
      for (int mu = 0; mu < 4; mu++)
        if (std::fabs((double)k[mu] * rats[mu]) <= root) flag++;
      keep = flag == 4;
    } else if (type == GLU_GFIX_CUT_PSQ) {
      keep = psq <= (double)max_mom;
    } else {
      keep = false;
      if (psq <= (double)max_mom) {
        double q[4];
        for (int mu = 0; mu < 4; mu++) q[mu] = 2. * std::sin((k[mu] * (kTwoPi / (double)dims[mu])) * 0.5);
        keep = type == GLU_GFIX_CUT_CYLINDER ? cylinder_df(q, latt_width) : cylinder_cjd(q, latt_width, angle);
      }
    }
    if (!keep) continue;
    long long s = 0, st = 1;
    for (int mu = 0; mu < 4; mu++) {
      mom.push_back(k[mu]);
      s += (long long)(k[mu] < 0 ? k[mu] + (int)dims[mu] : k[mu]) * st;
      st *= (long long)dims[mu];
    }
    site.push_back(s);
  }
  return 0;
}

// big-endian output of write_mom_veclist + write_g2g3_to_list
// (Cuts/cut_output.c:30-52, 311-336, 371-400)
int write_props(const char *path, const std::vector<int> &mom, const std::vector<double> &G,
                const std::vector<double> &F) {
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  auto put32 = [&](uint32_t v) {
    v = __builtin_bswap32(v);
    fwrite(&v, 4, 1, f);
  };
  auto put64 = [&](double d) {
    uint64_t v;
    memcpy(&v, &d, 8);
    v = __builtin_bswap64(v);
    fwrite(&v, 8, 1, f);
  };
  const size_t n = G.size();
  put32((uint32_t)n);
  for (size_t i = 0; i < n; i++) {
    put32(4);
    for (int mu = 0; mu < 4; mu++) put32((uint32_t)mom[4 * i + mu]);
  }
  put32((uint32_t)n);
  for (double g : G) put64(g);
  put32((uint32_t)n);
  for (double g : F) put64(g);
  const bool ok = !ferror(f);
  return (fclose(f) == 0 && ok) ? 0 : -1;
}

// momentum lists are built once per (dims, cut) for the process, as GLU
// keeps them in a file per shape (Cuts/cut_routines.c:500-600)
struct ListKey {
  size_t d[4];
  int cut;
  size_t max_mom;
  double width, angle;
  bool operator<(const ListKey &o) const {
    return std::tie(d[0], d[1], d[2], d[3], cut, max_mom, width, angle) <
           std::tie(o.d[0], o.d[1], o.d[2], o.d[3], o.cut, o.max_mom, o.width, o.angle);
  }
};
struct List {
  std::vector<int> mom;
  std::vector<long long> site;
};
std::mutex g_list_mu;
std::map<ListKey, List> g_lists;

int cached_list(const size_t dims[4], int cut, size_t max_mom, double width, double angle, List &out) {
  const ListKey k{{dims[0], dims[1], dims[2], dims[3]}, cut, max_mom, width, angle};
  {
    std::lock_guard<std::mutex> lk(g_list_mu);
    auto it = g_lists.find(k);
    if (it != g_lists.end()) {
      out = it->second;
      return 0;
    }
  }
  List l;
  if (momentum_list(dims, cut, max_mom, width, angle, l.mom, l.site)) return -1;
  std::lock_guard<std::mutex> lk(g_list_mu);
  g_lists[k] = l;
  out = std::move(l);
  return 0;
}

// the device copy of a momentum list with its -p sites and the per-axis
// tables of psq_calc / correct_pspace_landau, kept with the context for
// the last list it measured
struct GlueDev {
  ListKey key;
  int n = 0, ntab = 0, off[4] = {}, half[4] = {};
  int *mom = nullptr;
  long long *site = nullptr, *msite = nullptr;
  double *tab = nullptr, *out = nullptr;
};

void free_dev(GlueDev &d) {
  for (void *p : {(void *)d.mom, (void *)d.site, (void *)d.msite, (void *)d.tab, (void *)d.out})
    if (p) (void)hipFree(p);
  d = GlueDev{};
}

int glue_device(glu_gfix_ctx *c, const ListKey &key, const List &list) {
  GlueDev *d = (GlueDev *)c->glue;
  if (d && !(d->key < key) && !(key < d->key)) return 0;
  if (!d) c->glue = d = new GlueDev;
  free_dev(*d);
  d->key = key;
  const int n = (int)list.site.size();
  d->n = n;
  // sites of -p
  std::vector<long long> msite(n);
  for (int i = 0; i < n; i++) {
    long long s = 0, st = 1;
    for (int mu = 0; mu < 4; mu++) {
      const int L = (int)c->dims[mu], k = ((-list.mom[4 * i + mu]) % L + L) % L;
      s += (long long)k * st;
      st *= L;
    }
    msite[i] = s;
  }
  // 2 sin(k pi / L) and cos(k pi / L), k = -L/2 .. L - 1 - L/2: the momenta
  // of psq_calc (Cuts/glueprop.c:32-50) and the phase of
  // correct_pspace_landau (Cuts/pspace_landau.c:46-60)
  for (int mu = 0; mu < 4; mu++) {
    d->off[mu] = d->ntab;
    d->half[mu] = (int)(c->dims[mu] >> 1);
    d->ntab += (int)c->dims[mu];
  }
  std::vector<double> tab(2 * (size_t)d->ntab);
  for (int mu = 0; mu < 4; mu++) {
    const double tw = kTwoPi / (double)c->dims[mu];
    for (int t = 0; t < (int)c->dims[mu]; t++) {
      const int k = t - d->half[mu];
      tab[d->off[mu] + t] = 2.0 * std::sin(k * tw * 0.5);
      tab[d->ntab + d->off[mu] + t] = std::cos(0.5 * k * tw);
    }
  }
  hipStream_t s = c->stream;
  if (hipMalloc(&d->mom, (size_t)n * 4 * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->site, (size_t)n * sizeof(long long)) != hipSuccess ||
      hipMalloc(&d->msite, (size_t)n * sizeof(long long)) != hipSuccess ||
      hipMalloc(&d->tab, tab.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&d->out, (size_t)n * 3 * sizeof(double)) != hipSuccess ||
      hipMemcpyAsync(d->mom, list.mom.data(), (size_t)n * 4 * sizeof(int), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipMemcpyAsync(d->site, list.site.data(), (size_t)n * sizeof(long long), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipMemcpyAsync(d->msite, msite.data(), (size_t)n * sizeof(long long), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    free_dev(*d);
    d->key = ListKey{};
    return gfix_fail(-2, "gluon propagators: momentum list upload failed");
  }
  return 0;
}

}  // namespace

// the GLUON_PROPS (exceptional = 0) or EXCEPTIONAL (1) branch of
// cuts_struct (Cuts/cuts.c:114-184) on the device links: out1 / out2 are
// G / F or g2 / g3
static int gluon_cuts(glu_gfix_ctx *c, int exceptional, int cut, size_t max_mom, double cyl_width, double angle,
                      const char *path, size_t *nmom, int *mom, double *out1, double *out2, size_t cap,
                      double *pspace_err) {
  if (!c || !nmom) return gfix_fail(-1, "gluon cuts: null argument");
  if (c->gftype != GLU_GFIX_LANDAU) return gfix_fail(-1, "gluon cuts need a Landau context (4D transforms)");
  if (c->run.inflight) return gfix_fail(-1, "a run is in flight on this context");
  List list;
  if (cached_list(c->dims, cut, max_mom, cyl_width, angle, list))
    return gfix_fail(-1, "unknown momentum cut %d", cut);
  const std::vector<int> &km = list.mom;
  const std::vector<long long> &site = list.site;
  const int n = (int)site.size();
  *nmom = (size_t)n;
  if ((mom || out1 || out2) && cap < (size_t)n)
    return gfix_fail(-1, "momentum list has %d entries, the output arrays %zu", n, cap);
  if (n == 0) return gfix_fail(-1, "empty momentum list");
  if (!mom && !out1 && !out2 && !path && !pspace_err) return 0;  // the count only
  // GLU checks that list[i] = -list[n - 1 - i] and warns otherwise
  // (Cuts/cut_routines.c:415-437); its use of the mirror entry as -p is
  // then not exact
  for (int i = 0; i < n; i++) {
    bool sym = true;
    for (int mu = 0; mu < 4; mu++) sym = sym && km[4 * i + mu] + km[4 * (n - 1 - i) + mu] == 0;
    if (!sym) fprintf(stderr, "[CUTS] NON +/- Symmetric Momenta @ %d \n", i);
  }
  const long long V = c->g.V;
  if (int rc = glue_device(c, ListKey{{c->dims[0], c->dims[1], c->dims[2], c->dims[3]}, cut, max_mom, cyl_width,
                                      angle},
                           list))
    return rc;
  // the 16 packed fields stay with the context for later measurements
  if (!c->lie && hipMalloc(&c->lie, (size_t)V * 16 * sizeof(double2)) != hipSuccess) {
    c->lie = nullptr;
    return gfix_fail(-2, "gluon cuts: field allocation failed");
  }
  const GlueDev &d = *(const GlueDev *)c->glue;
  MomTabs tabs{d.mom, d.site, d.msite, d.tab, d.tab + d.ntab, {}, {}};
  for (int mu = 0; mu < 4; mu++) {
    tabs.off[mu] = d.off[mu];
    tabs.half[mu] = d.half[mu];
  }
  const double g2_norm = 2.0 / ((3 * 3 - 1) * (4 - 1) * (double)V);
  const double g0_norm = 2.0 / ((3 * 3 - 1) * (4) * (double)V);
  const double long_norm = 2.0 / ((3 * 3 - 1) * (double)V);
  const double g3_norm = 4.0 / (2.0 * 3 * (3 * 3 - 1) * (4 - 1) * (double)V);
  hipStream_t s = c->stream;
  void *Z = c->lie;
  GFIX_PROF(c, PK_LIE, hipLaunchKernelGGL(k_lie_pack, dim3((unsigned)c->nblk), dim3(kBlock), 0, s,
                                          (const double2 *)c->U64, (double2 *)Z, V));
  // unnormalised exp(+ip.x) transform of the 16 fields (Cuts/cuts.c:74-86)
  const long long nplane = V / c->g.S[2];
  bool xy = false;
  GFIX_PROF(c, PK_FFT,
            xy = c->use_xy &&
                 launch_fft_xy<double>(Z, V, 16, c->ax[0], c->ax[1], 1, nullptr, s, nplane, 1, 0, nullptr));
  for (int a = xy ? 2 : 0; a < 4; a++)
    GFIX_PROF(c, PK_FFT, launch_fft_axis<double>(Z, V, 16, c->ax[a], 1, c->ps, nullptr, s));
  double *out = d.out;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  if (exceptional)
    GFIX_PROF(c, PK_PROP,
              hipLaunchKernelGGL(k_gluon_exceptional, grid, dim3(kBlock), 0, s, (const double2 *)Z, V, n, tabs,
                                 g2_norm, g0_norm, g3_norm, out, out + n, out + 2 * n));
  else
    GFIX_PROF(c, PK_PROP,
              hipLaunchKernelGGL(k_gluon_prop, grid, dim3(kBlock), 0, s, (const double2 *)Z, V, n, tabs, n >> 1,
                                 g2_norm, g0_norm, long_norm, out, out + n, out + 2 * n));
  std::vector<double> h((size_t)n * 3);
  if (hipMemcpyAsync(h.data(), out, h.size() * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return gfix_fail(-2, "gluon cuts failed: %s", hipGetErrorString(hipGetLastError()));
  if (c->prof) prof_harvest(c);
  std::vector<double> A(h.begin(), h.begin() + n), B(h.begin() + n, h.begin() + 2 * n);
  if (!exceptional) {
    // compute_gluon_prop fills the second half from the first by p -> -p
    // (Cuts/glueprop.c:139-143)
    const int zero_pos = n >> 1;
    for (int i = zero_pos + 1; i < n; i++) {
      A[i] = A[n - i - 1];
      B[i] = B[n - i - 1];
    }
  }
  double e = 0.0;
  for (int i = 0; i < n; i++) e += h[2 * (size_t)n + i];
  if (pspace_err) *pspace_err = e / (double)n;
  if (c->verbose) fprintf(stdout, "\n[CUTS] P-Space gauge fixing error :: %1.5e \n\n", e / (double)n);
  if (mom) memcpy(mom, km.data(), (size_t)n * 4 * sizeof(int));
  if (out1) memcpy(out1, A.data(), (size_t)n * sizeof(double));
  if (out2) memcpy(out2, B.data(), (size_t)n * sizeof(double));
  if (path && write_props(path, km, A, B)) return gfix_fail(-3, "cannot write %s", path);
  return 0;
}

extern "C" int glu_gfix_gluon_props(glu_gfix_ctx *c, int cut, size_t max_mom, double cyl_width, double angle,
                                    const char *path, size_t *nmom, int *mom, double *transverse,
                                    double *longitudinal, size_t cap, double *pspace_err) {
  return gluon_cuts(c, 0, cut, max_mom, cyl_width, angle, path, nmom, mom, transverse, longitudinal, cap,
                    pspace_err);
}

extern "C" int glu_gfix_gluon_exceptional(glu_gfix_ctx *c, int cut, size_t max_mom, double cyl_width, double angle,
                                          const char *path, size_t *nmom, int *mom, double *g2, double *g3,
                                          size_t cap, double *pspace_err) {
  return gluon_cuts(c, 1, cut, max_mom, cyl_width, angle, path, nmom, mom, g2, g3, cap, pspace_err);
}

void glueprop_free(glu_gfix_ctx *c) {
  if (c->glue) {
    GlueDev *d = (GlueDev *)c->glue;
    free_dev(*d);
    delete d;
    c->glue = nullptr;
  }
}
