// landau_kernels.hip -- site-parallel streaming kernels of the Landau
// FACG/FASD iteration: derivative + theta, Polak-Ribiere sums, CG update,
// three-probe line search, gauge update, control kernels, and the
// link (re)packing / measurement kernels around them.
//
// One thread per lattice site, 256-thread blocks, SoA planes (coalesced),
// FP64 per-block partial sums reduced by a single-block kernel in a fixed
// order (bitwise reproducible, independent of what else runs on the GPU).
#include "gfix_kernels.h"

namespace gfix {

// ---------------------------------------------------------------------------
// derivative: Gfix/lin_derivs.c:75-193 + packing Gfix/FACG.c:38-73
// ---------------------------------------------------------------------------
// Delta(x) and its packing: DerivAcc (gfix_device.h)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_deriv(const vec2<T> *__restrict__ U, vec2<T> *__restrict__ F,
                                                 double *__restrict__ part, Geo g, const Ctrl *__restrict__ ctrl) {
  if (ctrl && ctrl->done) return;
  const long long V = g.V;
  const unsigned lb = logical_block(g);
  const long long i = (long long)lb * blockDim.x + threadIdx.x;
  double loc[1] = {0.0};
  if (i < V) {
    int x[4];
    coords(g, i, x);
    DerivAcc<T> d;
#pragma unroll
    for (int mu = 0; mu < kND; mu++)
      d.add(load_link<T>(U, mu, i, V), load_link<T>(U, mu, nb_bwd(g, i, x, mu), V));
    loc[0] = d.store(F, V, i);
  }
  block_sum_store<1>(loc, part + lb);
}

// ---------------------------------------------------------------------------
// trace-metric sums of the packed field: Gfix/FACG.c:99-178
//   s1 = <in, in>,  s2 = <in, in - old>
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pr_sums(const vec2<T> *__restrict__ F, const vec2<T> *__restrict__ O,
                                                   double *__restrict__ part, long long V,
                                                   const Ctrl *__restrict__ ctrl) {
  if (ctrl && ctrl->done) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (i < V) {
    double s1 = 0.0, s2 = 0.0;
    const cx<T> z0t = ld<T>(F + i), o0t = ld<T>(O + i);
    const double zr = z0t.re, zi = z0t.im, tr = zr - (double)o0t.re, ti = zi - (double)o0t.im;
    s1 += 2.0 * (zr * zr + zi * zi + zr * zi);
    s2 += 2.0 * (zr * tr + zi * ti);
    s2 += zr * ti + zi * tr;
#pragma unroll
    for (int f = 1; f < 4; f++) {
      const cx<T> z = ld<T>(F + f * V + i), o = ld<T>(O + f * V + i);
      const double ar = z.re, ai = z.im;
      s1 += 2.0 * (ar * ar + ai * ai);
      s2 += 2.0 * (ar * (ar - (double)o.re) + ai * (ai - (double)o.im));
    }
    v[0] = s1;
    v[1] = s2;
  }
  block_sum_store<2>(v, part + 2 * blockIdx.x);
}

// sn = in + beta sn (Gfix/FACG.c:315-323).  GLU's old = in needs no copy:
// the fused update writes the next derivative into the field buffer that
// held in_old, and the host swaps the two, so `in` becomes in_old in place.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(const vec2<T> *__restrict__ F, vec2<T> *__restrict__ sn,
                                                     long long n4, const Ctrl *__restrict__ ctrl) {
  if (ctrl->done) return;
  const T beta = (T)ctrl->beta;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n4; j += (long long)gridDim.x * blockDim.x) {
    const cx<T> a = ld<T>(F + j), b = ld<T>(sn + j);
    st<T>(sn + j, mk<T>(a.re + beta * b.re, a.im + beta * b.im));
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_copy(const vec2<T> *__restrict__ F, vec2<T> *__restrict__ a, long long n4,
                                                const Ctrl *__restrict__ ctrl) {
  if (ctrl->done) return;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n4; j += (long long)gridDim.x * blockDim.x)
    a[j] = F[j];
}

template <typename T>
__device__ __forceinline__ su3<T> gauge_at(const vec2<T> *__restrict__ F, long long V, long long i, T alpha) {
  return gauge_matrix<T>(ld<T>(F + i), ld<T>(F + V + i), ld<T>(F + 2 * V + i), ld<T>(F + 3 * V + i), alpha);
}

// ---------------------------------------------------------------------------
// line search probes: Gfix/line_search.c:274-341 (probes PL1 = 0.15,
// PL2 = 0.30 at :34-35).  Skipped (no partials) when theta <= CG_TOL unless
// forced (the SD start always searches).  The probes are evaluations of the
// functional, so they are computed in FP64 whatever the storage type: in
// FP32 the rounding of the reunitarised probe matrices biases the O(12 V)
// sums by more than the differences the parabola fit needs.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ su3<double> promote(const su3<T> &m) {
  su3<double> d;
#pragma unroll
  for (int k = 0; k < 9; k++) d.e[k] = mk<double>((double)m.e[k].re, (double)m.e[k].im);
  return d;
}
template <typename T>
__device__ __forceinline__ su3<double> gauge_at64(const vec2<T> *__restrict__ F, long long V, long long i,
                                                  double alpha) {
  auto f = [&](long long k) {
    const cx<T> z = ld<T>(F + k);
    return mk<double>((double)z.re, (double)z.im);
  };
  return gauge_matrix<double>(f(i), f(V + i), f(2 * V + i), f(3 * V + i), alpha);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_line_search(const vec2<T> *__restrict__ U, const vec2<T> *__restrict__ F,
                                                       double *__restrict__ part, Geo g,
                                                       const Ctrl *__restrict__ ctrl, int force) {
  if (ctrl->done) return;
  if (!force && !(ctrl->theta > ctrl->cg_tol)) return;
  const long long V = g.V;
  const unsigned lb = logical_block(g);
  const long long i = (long long)lb * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};
  if (i < V) {
    int x[4];
    coords(g, i, x);
    // one probe at a time keeps a single probe matrix per site live (the
    // link loads of the second probe hit the caches)
#pragma unroll 1
    for (int pr = 1; pr <= 2; pr++) {
      const double al = pr == 1 ? 0.15 : 0.30;
      const su3<double> A = gauge_at64<T>(F, V, i, al);
      double acc = 0.0;
#pragma unroll 1
      for (int mu = 0; mu < kND; mu++) {
        // rows 0, 1 of A u first: the link is dead while the neighbour's
        // matrix is built, which keeps the register peak lower
        cx<double> au[6];
        {
          const su3<double> u = promote<T>(load_link<T>(U, mu, i, V));
          if (pr == 1) v[0] += u.e[0].re + u.e[4].re + u.e[8].re;
          ab_rows01<double>(A, u, au);
        }
        acc += retrace_rows01_dag<double>(au, gauge_at64<T>(F, V, nb_fwd(g, i, x, mu), al));
      }
      v[pr] = acc;
    }
  }
  block_sum_store<3>(v, part + 3 * lb);
}

// ---------------------------------------------------------------------------
// The same probes with each probe matrix built once.  A workgroup owns a
// TX x TY x TZ x TT tile of 256 sites and builds, per probe, the matrices
// g(x; alpha) of its sites and of the forward faces of the tile (the
// neighbours x + mu outside it) into LDS: 2 + 256 / TX + ... builds per
// 256 sites instead of 5 per site.  Only columns 0 and 1 are kept: with
// M = A u (columns 0, 1 from u's stored columns) and B the neighbour's
// matrix, both in SU(3) so that their third columns are conj(c0 x c1),
//   Re tr(A u B^dag) = Re q00 + Re q11 + Re(q00 q11 - q01 q10),
//   q_jk = sum_r M_j[r] conj(B_k[r])
// (the Binet-Cauchy identity for the third columns), so neither u's nor
// B's third column is ever formed; the probe at 0 takes u's third diagonal
// entry from the same identity.  HOLD keeps a site's link columns in
// registers for both probes; without it they are loaded again for the
// second probe (from the caches), which frees registers for more waves.
// The arithmetic is FP64, as in k_line_search.
// ---------------------------------------------------------------------------
template <typename T, int TX, int TY, int TZ, int TT, bool HOLD>
__global__ __launch_bounds__(kBlock) void k_line_search_tile(const vec2<T> *__restrict__ U,
                                                            const vec2<T> *__restrict__ F, double *__restrict__ part,
                                                            Geo g, const Ctrl *__restrict__ ctrl, int force) {
  constexpr int NT = TX * TY * TZ * TT;
  static_assert(NT == kBlock, "one site per thread");
  constexpr int HX = TY * TZ * TT, HY = TX * TZ * TT, HZ = TX * TY * TT, HT = TX * TY * TZ;
  constexpr int OX = NT, OY = OX + HX, OZ = OY + HY, OT = OZ + HZ, NB = OT + HT;
  __shared__ double lds[12 * NB];  // entry e of slot j (re, im of 6 complex) at lds[e * NB + j]
  if (ctrl->done) return;
  if (!force && !(ctrl->theta > ctrl->cg_tol)) return;
  const long long V = g.V;
  const unsigned lb = logical_block(g);
  const int ntx = g.L[0] / TX, nty = g.L[1] / TY, ntz = g.L[2] / TZ;
  unsigned r = lb;
  const int x0 = (int)(r % (unsigned)ntx) * TX;
  r /= (unsigned)ntx;
  const int y0 = (int)(r % (unsigned)nty) * TY;
  r /= (unsigned)nty;
  const int z0 = (int)(r % (unsigned)ntz) * TZ;
  const int t0 = (int)(r / (unsigned)ntz) * TT;
  const int tid = threadIdx.x;
  const int lx = tid % TX, ly = (tid / TX) % TY, lz = (tid / (TX * TY)) % TZ, lt = tid / (TX * TY * TZ);
  auto site = [&](int x, int y, int z, int t) -> long long {
    return (long long)x + g.S[1] * y + g.S[2] * z + g.S[3] * t;
  };
  const long long me = site(x0 + lx, y0 + ly, z0 + lz, t0 + lt);
  // LDS slot of the forward neighbour in each direction
  int nbs[4];
  nbs[0] = lx < TX - 1 ? tid + 1 : OX + ly + TY * (lz + TZ * lt);
  nbs[1] = ly < TY - 1 ? tid + TX : OY + lx + TX * (lz + TZ * lt);
  nbs[2] = lz < TZ - 1 ? tid + TX * TY : OZ + lx + TX * (ly + TY * lt);
  nbs[3] = lt < TT - 1 ? tid + TX * TY * TZ : OT + lx + TX * (ly + TY * lz);
  // this site's link columns (col 0 rows 0..2, col 1 rows 0..2)
  cx<T> uh[HOLD ? kND : 1][6];
  auto load_u = [&](int mu, cx<T> (&w)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) w[k] = ld<T>(U + ((long long)mu * 6 + k) * V + me);
  };
  double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int mu = 0; mu < kND; mu++) {
    cx<T> w[6];
    load_u(mu, w);
    if constexpr (HOLD) {
#pragma unroll
      for (int k = 0; k < 6; k++) uh[mu][k] = w[k];
    }
    const double u00r = w[0].re, u00i = w[0].im, u10r = w[1].re, u10i = w[1].im;
    const double u01r = w[3].re, u01i = w[3].im, u11r = w[4].re, u11i = w[4].im;
    // Re U00 + Re U11 + Re U22, U22 = conj(U00 U11 - U10 U01)
    v[0] += u00r + u11r + (u00r * u11r - u00i * u11i) - (u10r * u01r - u10i * u01i);
  }
#pragma unroll 1
  for (int pr = 1; pr <= 2; pr++) {
    const double al = pr == 1 ? 0.15 : 0.30;
    if (pr == 2) __syncthreads();  // the first probe's matrices are no longer read
    for (int j = tid; j < NB; j += kBlock) {
      long long s = me;
      if (j >= NT) {
        int h = j - NT;
        if (h < HX) {
          s = site((x0 + TX) % g.L[0], y0 + h % TY, z0 + (h / TY) % TZ, t0 + h / (TY * TZ));
        } else if ((h -= HX) < HY) {
          s = site(x0 + h % TX, (y0 + TY) % g.L[1], z0 + (h / TX) % TZ, t0 + h / (TX * TZ));
        } else if ((h -= HY) < HZ) {
          s = site(x0 + h % TX, y0 + (h / TX) % TY, (z0 + TZ) % g.L[2], t0 + h / (TX * TY));
        } else {
          h -= HZ;
          s = site(x0 + h % TX, y0 + (h / TX) % TY, z0 + h / (TX * TY), (t0 + TT) % g.L[3]);
        }
      }
      auto f = [&](long long k) {
        const cx<T> z = ld<T>(F + k * V + s);
        return mk<double>((double)z.re, (double)z.im);
      };
      const su3<double> m = gauge_cols01<double>(f(0), f(1), f(2), f(3), al);
      const cx<double> c[6] = {m.e[0], m.e[3], m.e[6], m.e[1], m.e[4], m.e[7]};
#pragma unroll
      for (int e = 0; e < 6; e++) {
        lds[(2 * e) * NB + j] = c[e].re;
        lds[(2 * e + 1) * NB + j] = c[e].im;
      }
    }
    __syncthreads();
    // this site's matrix A, completed: a[r][k] = A_rk
    cx<double> a[3][3];
#pragma unroll
    for (int rr = 0; rr < 3; rr++)
#pragma unroll
      for (int k = 0; k < 2; k++)
        a[rr][k] = mk<double>(lds[(2 * (3 * k + rr)) * NB + tid], lds[(2 * (3 * k + rr) + 1) * NB + tid]);
    a[0][2] = cconj(csub(cmul(a[1][0], a[2][1]), cmul(a[2][0], a[1][1])));
    a[1][2] = cconj(csub(cmul(a[2][0], a[0][1]), cmul(a[0][0], a[2][1])));
    a[2][2] = cconj(csub(cmul(a[0][0], a[1][1]), cmul(a[1][0], a[0][1])));
    // one direction's trace: nb is the LDS slot of the forward neighbour
    auto trace_mu = [&](int mu, int nb) -> double {
      cx<T> w[6];
      if constexpr (HOLD) {
#pragma unroll
        for (int k = 0; k < 6; k++) w[k] = uh[mu][k];
      } else {
        load_u(mu, w);
      }
      // M = A u, columns 0 and 1
      cx<double> M[2][3];
#pragma unroll
      for (int jc = 0; jc < 2; jc++) {
        const cx<double> u0 = mk<double>((double)w[3 * jc].re, (double)w[3 * jc].im);
        const cx<double> u1 = mk<double>((double)w[3 * jc + 1].re, (double)w[3 * jc + 1].im);
        const cx<double> u2 = mk<double>((double)w[3 * jc + 2].re, (double)w[3 * jc + 2].im);
#pragma unroll
        for (int rr = 0; rr < 3; rr++)
          M[jc][rr] = cadd(cadd(cmul(a[rr][0], u0), cmul(a[rr][1], u1)), cmul(a[rr][2], u2));
      }
      cx<double> q[2][2];
#pragma unroll
      for (int k = 0; k < 2; k++) {
        cx<double> bk[3];
#pragma unroll
        for (int rr = 0; rr < 3; rr++)
          bk[rr] = mk<double>(lds[(2 * (3 * k + rr)) * NB + nb], lds[(2 * (3 * k + rr) + 1) * NB + nb]);
#pragma unroll
        for (int jc = 0; jc < 2; jc++)
          q[jc][k] = cadd(cadd(cmulc(M[jc][0], bk[0]), cmulc(M[jc][1], bk[1])), cmulc(M[jc][2], bk[2]));
      }
      const cx<double> d = csub(cmul(q[0][0], q[1][1]), cmul(q[0][1], q[1][0]));
      return q[0][0].re + q[1][1].re + d.re;
    };
    double acc = 0.0;
    if constexpr (HOLD) {
#pragma unroll
      for (int mu = 0; mu < kND; mu++) acc += trace_mu(mu, nbs[mu]);
    } else {
      // one direction at a time: fewer values live
#pragma unroll 1
      for (int mu = 0; mu < kND; mu++)
        acc += trace_mu(mu, mu == 0 ? nbs[0] : mu == 1 ? nbs[1] : mu == 2 ? nbs[2] : nbs[3]);
    }
    v[pr] = acc;
  }
  block_sum_store<3>(v, part + 3 * lb);
}

// ---------------------------------------------------------------------------
// U_mu(x) <- g(x) U_mu(x) g(x+mu)^dagger: Gfix/line_search.c:60-97.
// With G != null the site transform is also accumulated in FP64,
// G(x) <- g(x) G(x) with g rebuilt in FP64 from the same field values: the
// mixed schedule keeps the FP64 links untouched during the FP32 phase and
// replays G on them exactly at the switch.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_egauge(vec2<T> *__restrict__ U, const vec2<T> *__restrict__ F,
                                                  double2 *__restrict__ G, Geo g, const Ctrl *__restrict__ ctrl,
                                                  int fixed_step) {
  if (ctrl->done) return;
  const long long V = g.V;
  const long long i = (long long)logical_block(g) * blockDim.x + threadIdx.x;
  if (i >= V) return;
  const double alpha = fixed_step ? ctrl->gf_alpha : ctrl->alpha;
  int x[4];
  coords(g, i, x);
  const su3<T> A = gauge_at<T>(F, V, i, (T)alpha);
#pragma unroll
  for (int mu = 0; mu < kND; mu++) {
    const su3<T> B = gauge_at<T>(F, V, nb_fwd(g, i, x, mu), (T)alpha);
    const su3<T> u = load_link<T>(U, mu, i, V);
    su3<T> n = gtransform<T>(A, u, B);
    // FP32 links are kept on SU(3): unchecked, FP32 rounding of the columns
    // feeds back through the completed third column over hundreds of updates
    if (sizeof(T) == 4) gram_reunit<T>(n);
    store_link<T>(U, mu, i, V, n);
  }
  if (G) {
    const su3<double> A64 = gauge_at64<T>(F, V, i, alpha);
    store_link<double>(G, 0, i, V, mul_cols<double>(A64, load_link<double>(G, 0, i, V)));
  }
}

// ---------------------------------------------------------------------------
// The gauge update of one CG iteration fused with the derivative of the
// next (Gfix/FACG.c:325-334 followed by :276-280).  The update is out of
// place, Uin -> Uout: each site also forms its backward links
// U'_mu(x - mu) = g(x - mu) U_mu(x - mu) g(x)^dag from the old links, with
// the arithmetic of the site that stores them, and the old links of the
// neighbours must stay readable meanwhile.  The derivative of the updated
// links then costs no second pass over them.  D is the search direction;
// Fout (the buffer of in_old, which this iteration no longer needs)
// receives the packed derivative and part its theta partials; the host
// then swaps the link buffers and the field buffers Fin <-> Fout.  Past
// convergence the launch does nothing: the control block counts the
// updates that ran (Ctrl::upd), and the host undoes the swaps of those
// that did not once the batch has ended.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_egauge_deriv(const vec2<T> *__restrict__ Uin, vec2<T> *__restrict__ Uout,
                                                        const vec2<T> *__restrict__ D, double2 *__restrict__ G,
                                                        vec2<T> *__restrict__ Fin, vec2<T> *__restrict__ Fout,
                                                        double *__restrict__ part, Geo g, const Ctrl *__restrict__ ctrl) {
  if (ctrl->done) return;
  const long long V = g.V;
  const unsigned lb = logical_block(g);
  const long long i = (long long)lb * blockDim.x + threadIdx.x;
  double loc[1] = {0.0};
  if (i < V) {
    const double alpha = ctrl->alpha;
    int x[4];
    coords(g, i, x);
    const su3<T> A = gauge_at<T>(D, V, i, (T)alpha);
    DerivAcc<T> d;
#pragma unroll 1
    for (int mu = 0; mu < kND; mu++) {
      const su3<T> B = gauge_at<T>(D, V, nb_fwd(g, i, x, mu), (T)alpha);
      su3<T> n = gtransform<T>(A, load_link<T>(Uin, mu, i, V), B);
      if (sizeof(T) == 4) gram_reunit<T>(n);
      store_link<T>(Uout, mu, i, V, n);
      const long long xb = nb_bwd(g, i, x, mu);
      const su3<T> C = gauge_at<T>(D, V, xb, (T)alpha);
      su3<T> sb = gtransform<T>(C, load_link<T>(Uin, mu, xb, V), A);
      if (sizeof(T) == 4) gram_reunit<T>(sb);
      d.add(n, sb);
    }
    loc[0] = d.store(Fout, V, i);
    if (G) {
      const su3<double> A64 = gauge_at64<T>(D, V, i, alpha);
      store_link<double>(G, 0, i, V, mul_cols<double>(A64, load_link<double>(G, 0, i, V)));
    }
  }
  block_sum_store<1>(loc, part + lb);
}

// ---------------------------------------------------------------------------
// single-block finalisers (fixed summation order)
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void final_sum(const double *__restrict__ part, long long nb, double (&out)[NV]) {
  __shared__ double sh[NV][1024 / kWave];
  double v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = 0.0;
  for (long long b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] += part[b * NV + k];
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) sh[k][w] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double s = 0.0;
    for (int j = 0; j < (int)(blockDim.x / kWave); j++) s += sh[k][j];
    out[k] = s;
  }
}

// theta = sum / (NC V): GFNORM_LANDAU, Headers/GLU_definitions.h:427
__global__ void k_fin_theta(const double *__restrict__ part, long long nb, Ctrl *ctrl, double norm) {
  if (ctrl->done) return;
  double s[1];
  final_sum<1>(part, nb, s);
  if (threadIdx.x == 0) {
    double th = s[0] * norm;
    if (ctrl->sd_start) th += ctrl->stale_theta;  // GLU's stale red[5] slots
    ctrl->theta = th;
  }
}

// beta: Polak-Ribiere clipped at 0, Fletcher-Reeves below CG_TOL
// (Gfix/FACG.c:306-313)
__global__ void k_fin_pr(const double *__restrict__ part, long long nb, Ctrl *ctrl) {
  if (ctrl->done) return;
  double s[2];
  final_sum<2>(part, nb, s);
  if (threadIdx.x == 0) {
    const double insum = s[0], sc = s[1];
    double beta = sc / ctrl->inold;
    if (beta < 0.0) beta = 0.0;
    if (ctrl->theta < ctrl->cg_tol) beta = insum / ctrl->inold;
    if (ctrl->switch_theta > 0.0) {  // safeguards of the mixed FP32 phase
      if ((ctrl->sg_mode & 1) && ctrl->sg_uphill) beta = 0.0;
      if ((ctrl->sg_mode & 2) && beta > ctrl->beta_cap) beta = ctrl->beta_cap;
    }
    ctrl->sg_uphill = 0;
    ctrl->inold = insum;
    ctrl->insum = insum;
    ctrl->sum_conj = sc;
    ctrl->beta = beta;
  }
}

// end of the SD start: inold = sum_DER3 (+ GLU's stale red[3] slots)
__global__ void k_fin_sd(const double *__restrict__ part, long long nb, Ctrl *ctrl) {
  if (ctrl->done) return;
  double s[2];
  final_sum<2>(part, nb, s);
  if (threadIdx.x == 0) {
    ctrl->inold = s[0] + ctrl->stale_insum;
    ctrl->beta = 0.0;
    if (ctrl->trace_n < ctrl->trace_cap) ctrl->trace_n++;  // row already written
    ctrl->sd_start = 0;
    ctrl->stale_insum = 0.0;
    ctrl->stale_theta = 0.0;
    ctrl->stale_probe[0] = ctrl->stale_probe[1] = ctrl->stale_probe[2] = 0.0;
  }
}

__global__ void k_fin_ls(const double *__restrict__ part, long long nb, Ctrl *ctrl, double *trace, int force,
                         double pl1, double pl2) {
  if (ctrl->done) return;
  const int do_ls = force || (ctrl->theta > ctrl->cg_tol);
  double s[3] = {0.0, 0.0, 0.0};
  if (do_ls) final_sum<3>(part, nb, s);
  if (threadIdx.x == 0) {
    ctrl->upd++;  // the gauge update enqueued after this launch will run
    if (do_ls) {
      double val[3];
      for (int k = 0; k < 3; k++) {
        const double tot = s[k] + (ctrl->sd_start ? ctrl->stale_probe[k] : 0.0);
        ctrl->probe[k] = tot;
        val[k] = -tot;
      }
      const double al[3] = {0.0, pl1, pl2};
      ctrl->alpha = approx_minimum(al, val);
      ctrl->did_ls = 1;
      // mixed FP32 phase: the functional rises along the direction
      const int sg = ctrl->switch_theta > 0.0 && !force ? ctrl->sg_mode : 0;
      ctrl->sg_uphill = sg && val[1] > val[0];
      if ((sg & 4) && ctrl->sg_uphill) ctrl->alpha = 0.0;
    } else {
      ctrl->alpha = ctrl->gf_alpha;
      ctrl->did_ls = 0;
    }
    if (ctrl->sd_start && trace && ctrl->trace_n < ctrl->trace_cap) {
      double *r = trace + 3 * ctrl->trace_n;
      r[0] = ctrl->theta;
      r[1] = ctrl->alpha;
      r[2] = 0.0;
    }
  }
}

// end of a loop iteration: count, record, stop test (Gfix/FACG.c:331-334);
// returns the done flag
__device__ int iter_end_body(Ctrl *ctrl, double *trace, int fasd) {
  if (fasd) {
    ctrl->alpha = ctrl->gf_alpha;
    ctrl->beta = 0.0;
  }
  const double theta = ctrl->theta;
  const long long iters = ctrl->iters + 1;
  ctrl->iters = iters;
  ctrl->total++;
  if (trace && ctrl->trace_n < ctrl->trace_cap) {
    double *r = trace + 3 * ctrl->trace_n;
    r[0] = theta;
    r[1] = ctrl->alpha;
    r[2] = ctrl->beta;
    ctrl->trace_n++;
  }
  ctrl->stale_insum = ctrl->insum;
  ctrl->stale_theta = theta;
  for (int k = 0; k < 3; k++) ctrl->stale_probe[k] = ctrl->did_ls ? ctrl->probe[k] : 0.0;
  int done = !(theta > ctrl->acc && iters < ctrl->max_iters);
  if (ctrl->switch_theta > 0.0) {  // mixed FP32 phase: the switch to FP64
    const double tmin = ctrl->theta_min;
    if (theta < tmin) ctrl->theta_min = theta;
    const bool stalled = ctrl->stall_factor > 0.0 && tmin < ctrl->stall_theta && theta > ctrl->stall_factor * tmin;
    if (theta < ctrl->switch_theta || stalled) {
      ctrl->switched = stalled && !(theta < ctrl->switch_theta) ? 2 : 1;
      done = 1;
    }
  }
  ctrl->done = done;
  return done;
}

__global__ void k_iter_end(Ctrl *ctrl, double *trace, int fasd) {
  if (ctrl->done) return;
  iter_end_body(ctrl, trace, fasd);
}

// The end of a CG iteration and the theta of the next one in one launch:
// the stop test on theta_k, then, unless it stopped, theta_k+1 from the
// partials of the fused update + derivative (as k_fin_theta; the SD start
// is over, so no stale slots enter).  A stopped run keeps the theta it
// tested, and the derivative it left is taken again if a segment resumes.
__global__ void k_fin_iter(const double *__restrict__ part, long long nb, Ctrl *ctrl, double *trace, double norm) {
  if (ctrl->done) return;
  double s[1];
  final_sum<1>(part, nb, s);
  if (threadIdx.x == 0 && !iter_end_body(ctrl, trace, 0)) ctrl->theta = s[0] * norm;
}

// plain sum of nb partials into *out (fixed order)
__global__ void k_sum_part(const double *__restrict__ part, long long nb, double *out) {
  double s[1];
  final_sum<1>(part, nb, s);
  if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_begin_segment(Ctrl *ctrl, int sd) {
  ctrl->done = 0;
  ctrl->iters = 0;
  ctrl->sd_start = sd;
  ctrl->did_ls = 0;
}

// ---------------------------------------------------------------------------
// layout conversion: GLU AoS double[V][4][18] <-> device column planes
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack(const double *__restrict__ aos, vec2<T> *__restrict__ U,
                                                long long V, int nd, long long site0, long long nsite) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsite) return;
  const long long i = site0 + j;
  for (int mu = 0; mu < nd; mu++) {
    const double *a = aos + (j * nd + mu) * 18;
    su3<T> m;
#pragma unroll
    for (int k = 0; k < 9; k++) m.e[k] = mk<T>((T)a[2 * k], (T)a[2 * k + 1]);
    store_link<T>(U, mu, i, V, m);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_unpack(const vec2<T> *__restrict__ U, double *__restrict__ aos,
                                                  long long V, int nd, long long site0, long long nsite) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nsite) return;
  const long long i = site0 + j;
  for (int mu = 0; mu < nd; mu++) {
    const su3<T> m = load_link<T>(U, mu, i, V);
    double *a = aos + (j * nd + mu) * 18;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      a[2 * k] = (double)m.e[k].re;
      a[2 * k + 1] = (double)m.e[k].im;
    }
  }
}

// FP64 links -> FP32 copy
__global__ __launch_bounds__(kBlock) void k_demote(const double2 *__restrict__ U, float2 *__restrict__ W, long long n) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    const double2 v = U[j];
    W[j] = make_float2((float)v.x, (float)v.y);
  }
}

// identity site transform (column planes of one "direction")
template <typename T>
__global__ __launch_bounds__(kBlock) void k_identity(vec2<T> *__restrict__ G, long long V) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  su3<T> m;
#pragma unroll
  for (int k = 0; k < 9; k++) m.e[k] = mk<T>(T(0), T(0));
  m.e[0].re = m.e[4].re = m.e[8].re = T(1);
  store_link<T>(G, 0, i, V, m);
}

// U_mu(x) <- G(x) U_mu(x) G(x+mu)^dagger in FP64 with the accumulated site
// transform reunitarised first (Gfix/gtrans.c:169-189)
__global__ __launch_bounds__(kBlock) void k_apply_gauge(double2 *__restrict__ U, const double2 *__restrict__ G,
                                                       Geo g) {
  const long long V = g.V;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  int x[4];
  coords(g, i, x);
  auto fetch = [&](long long s) {
    su3<double> d = load_link<double>(G, 0, s, V);
    gram_reunit<double>(d);
    return d;
  };
  const su3<double> A = fetch(i);
#pragma unroll
  for (int mu = 0; mu < kND; mu++) {
    const su3<double> B = fetch(nb_fwd(g, i, x, mu));
    const su3<double> u = load_link<double>(U, mu, i, V);
    store_link<double>(U, mu, i, V, gtransform<double>(A, u, B));
  }
}

// packed fields FP32 -> FP64 (the CG state crosses the precision switch)
__global__ __launch_bounds__(kBlock) void k_promote(const float2 *__restrict__ a, double2 *__restrict__ b, long long n) {
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
    const float2 v = a[j];
    b[j] = make_double2(v.x, v.y);
  }
}

// reunitarise every link (Geometry/random_config.c:43-56)
__global__ __launch_bounds__(kBlock) void k_reunit(double2 *__restrict__ U, long long V, int nd) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= V) return;
  for (int mu = 0; mu < nd; mu++) {
    su3<double> m = load_link<double>(U, mu, i, V);
    gram_reunit<double>(m);
    store_link<double>(U, mu, i, V, m);
  }
}

// plaquette / link-trace partial sums: Field/plaqs_links.c:243-265, 399-448
__global__ __launch_bounds__(kBlock) void k_measure(const double2 *__restrict__ U, double *__restrict__ part, Geo g) {
  const long long V = g.V;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v[3] = {0.0, 0.0, 0.0};  // plaquette sum, link trace, temporal trace
  double mx = -1e300;             // largest single-link trace
  if (i < V) {
    int x[4];
    coords(g, i, x);
    su3<double> u[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) u[mu] = load_link<double>(U, mu, i, V);
    for (int mu = 0; mu < 4; mu++) {
      const double r = retrace<double>(u[mu]);
      v[1] += r;
      if (mu == 3) v[2] += r;
      mx = r > mx ? r : mx;
      const long long t = nb_fwd(g, i, x, mu);
      for (int nu = 0; nu < mu; nu++) {
        const long long s = nb_fwd(g, i, x, nu);
        const su3<double> b = load_link<double>(U, nu, t, V);
        const su3<double> c = load_link<double>(U, mu, s, V);
        // Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag)
        const su3<double> ab = mul_su3<double>(u[mu], b);
        const su3<double> dc = mul_su3<double>(u[nu], c);
        double t2 = 0.0;
        for (int k = 0; k < 9; k++) t2 += ab.e[k].re * dc.e[k].re + ab.e[k].im * dc.e[k].im;
        v[0] += t2;
      }
    }
  }
  __shared__ double mxs[kBlock];
  mxs[threadIdx.x] = mx;
  block_sum_store<3>(v, part + 4 * blockIdx.x);
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = -1e300;
    for (int k = 0; k < (int)blockDim.x; k++) m = mxs[k] > m ? mxs[k] : m;
    part[4 * blockIdx.x + 3] = m;
  }
}

// ---------------------------------------------------------------------------
// explicit instantiations + launch helpers
// ---------------------------------------------------------------------------
static inline unsigned nblk(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename T>
void launch_deriv(const void *U, void *F, double *part, const Geo &g, const Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_deriv<T>, dim3(nblk(g.V)), dim3(kBlock), 0, s, (const vec2<T> *)U, (vec2<T> *)F, part, g,
                     ctrl);
}
template <typename T>
void launch_pr_sums(const void *F, const void *O, double *part, long long V, const Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_pr_sums<T>, dim3(nblk(V)), dim3(kBlock), 0, s, (const vec2<T> *)F, (const vec2<T> *)O, part,
                     V, ctrl);
}
template <typename T>
void launch_cg_update(const void *F, void *sn, long long V, const Ctrl *ctrl, hipStream_t s) {
  const long long n4 = 4 * V;
  const unsigned b = (unsigned)std::min<long long>(nblk(n4), 8192);
  hipLaunchKernelGGL(k_cg_update<T>, dim3(b), dim3(kBlock), 0, s, (const vec2<T> *)F, (vec2<T> *)sn, n4, ctrl);
}
template <typename T> void launch_copy(const void *F, void *a, long long V, const Ctrl *ctrl, hipStream_t s) {
  const long long n4 = 4 * V;
  const unsigned b = (unsigned)std::min<long long>(nblk(n4), 8192);
  hipLaunchKernelGGL(k_copy<T>, dim3(b), dim3(kBlock), 0, s, (const vec2<T> *)F, (vec2<T> *)a, n4, ctrl);
}
template <typename T, int TX, int TY, int TZ, int TT, bool HOLD>
static void launch_ls_tile(const void *U, const void *F, double *part, Geo g, const Ctrl *ctrl, int force,
                           hipStream_t s) {
  g.xcd = g.xcd ? 1 : 0;  // tiles are numbered x fastest: plain or XCD-contiguous
  hipLaunchKernelGGL((k_line_search_tile<T, TX, TY, TZ, TT, HOLD>), dim3(nblk(g.V)), dim3(kBlock), 0, s,
                     (const vec2<T> *)U, (const vec2<T> *)F, part, g, ctrl, force);
}

bool ls_tile_fits(const Geo &g, int tile) {
  if (tile < 1 || tile > kLsTiles) return false;
  const int *t = kLsTileShape[tile - 1];
  for (int mu = 0; mu < 4; mu++)
    if (g.L[mu] % t[mu]) return false;
  return true;
}

template <typename T>
void launch_line_search(const void *U, const void *F, double *part, const Geo &g, const Ctrl *ctrl, int force,
                        hipStream_t s, int tile) {
  switch (ls_tile_fits(g, tile) ? tile : 0) {
    case 1: return launch_ls_tile<T, 8, 4, 4, 2, true>(U, F, part, g, ctrl, force, s);
    case 2: return launch_ls_tile<T, 4, 4, 4, 4, true>(U, F, part, g, ctrl, force, s);
    case 3: return launch_ls_tile<T, 8, 8, 2, 2, true>(U, F, part, g, ctrl, force, s);
    case 4: return launch_ls_tile<T, 16, 4, 2, 2, true>(U, F, part, g, ctrl, force, s);
    case 5: return launch_ls_tile<T, 8, 4, 4, 2, false>(U, F, part, g, ctrl, force, s);
    case 6: return launch_ls_tile<T, 4, 4, 4, 4, false>(U, F, part, g, ctrl, force, s);
    case 7: return launch_ls_tile<T, 8, 8, 2, 2, false>(U, F, part, g, ctrl, force, s);
    case 8: return launch_ls_tile<T, 16, 4, 2, 2, false>(U, F, part, g, ctrl, force, s);
    default:
      hipLaunchKernelGGL(k_line_search<T>, dim3(nblk(g.V)), dim3(kBlock), 0, s, (const vec2<T> *)U,
                         (const vec2<T> *)F, part, g, ctrl, force);
  }
}
template <typename T>
void launch_egauge(void *U, const void *F, void *G, const Geo &g, const Ctrl *ctrl, int fixed, hipStream_t s) {
  hipLaunchKernelGGL(k_egauge<T>, dim3(nblk(g.V)), dim3(kBlock), 0, s, (vec2<T> *)U, (const vec2<T> *)F,
                     (double2 *)G, g, ctrl, fixed);
}
template <typename T>
void launch_egauge_deriv(const void *Uin, void *Uout, const void *D, void *G, void *Fin, void *Fout, double *part,
                         const Geo &g, const Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_egauge_deriv<T>, dim3(nblk(g.V)), dim3(kBlock), 0, s, (const vec2<T> *)Uin, (vec2<T> *)Uout,
                     (const vec2<T> *)D, (double2 *)G, (vec2<T> *)Fin, (vec2<T> *)Fout, part, g, ctrl);
}
template <typename T>
void launch_pack(const double *aos, void *U, long long V, int nd, long long site0, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_pack<T>, dim3(nblk(n)), dim3(kBlock), 0, s, aos, (vec2<T> *)U, V, nd, site0, n);
}
template <typename T>
void launch_unpack(const void *U, double *aos, long long V, int nd, long long site0, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_unpack<T>, dim3(nblk(n)), dim3(kBlock), 0, s, (const vec2<T> *)U, aos, V, nd, site0, n);
}
template <typename T> void launch_identity(void *G, long long V, hipStream_t s) {
  hipLaunchKernelGGL(k_identity<T>, dim3(nblk(V)), dim3(kBlock), 0, s, (vec2<T> *)G, V);
}

#define GFIX_INST(T)                                                                                          \
  template void launch_deriv<T>(const void *, void *, double *, const Geo &, const Ctrl *, hipStream_t);      \
  template void launch_pr_sums<T>(const void *, const void *, double *, long long, const Ctrl *, hipStream_t); \
  template void launch_cg_update<T>(const void *, void *, long long, const Ctrl *, hipStream_t);               \
  template void launch_copy<T>(const void *, void *, long long, const Ctrl *, hipStream_t);                    \
  template void launch_line_search<T>(const void *, const void *, double *, const Geo &, const Ctrl *, int,   \
                                      hipStream_t, int);                                                      \
  template void launch_egauge<T>(void *, const void *, void *, const Geo &, const Ctrl *, int, hipStream_t);  \
  template void launch_egauge_deriv<T>(const void *, void *, const void *, void *, void *, void *, double *,     \
                                       const Geo &, const Ctrl *, hipStream_t);                                   \
  template void launch_pack<T>(const double *, void *, long long, int, long long, long long, hipStream_t);     \
  template void launch_unpack<T>(const void *, double *, long long, int, long long, long long, hipStream_t);   \
  template void launch_identity<T>(void *, long long, hipStream_t);
GFIX_INST(double)
GFIX_INST(float)

void launch_fin_theta(const double *part, long long nb, Ctrl *ctrl, double norm, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_theta, dim3(1), dim3(1024), 0, s, part, nb, ctrl, norm);
}
void launch_fin_pr(const double *part, long long nb, Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_pr, dim3(1), dim3(1024), 0, s, part, nb, ctrl);
}
void launch_fin_sd(const double *part, long long nb, Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_sd, dim3(1), dim3(1024), 0, s, part, nb, ctrl);
}
void launch_fin_ls(const double *part, long long nb, Ctrl *ctrl, double *trace, int force, double pl1, double pl2,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_fin_ls, dim3(1), dim3(1024), 0, s, part, nb, ctrl, trace, force, pl1, pl2);
}
void launch_fin_iter(const double *part, long long nb, Ctrl *ctrl, double *trace, double norm, hipStream_t s) {
  hipLaunchKernelGGL(k_fin_iter, dim3(1), dim3(1024), 0, s, part, nb, ctrl, trace, norm);
}
void launch_sum_part(const double *part, long long nb, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_part, dim3(1), dim3(1024), 0, s, part, nb, out);
}
void launch_iter_end(Ctrl *ctrl, double *trace, int fasd, hipStream_t s) {
  hipLaunchKernelGGL(k_iter_end, dim3(1), dim3(1), 0, s, ctrl, trace, fasd);
}
void launch_begin_segment(Ctrl *ctrl, int sd, hipStream_t s) {
  hipLaunchKernelGGL(k_begin_segment, dim3(1), dim3(1), 0, s, ctrl, sd);
}
void launch_demote(const void *U, void *W, long long n, hipStream_t s) {
  const unsigned b = (unsigned)std::min<long long>(nblk(n), 16384);
  hipLaunchKernelGGL(k_demote, dim3(b), dim3(kBlock), 0, s, (const double2 *)U, (float2 *)W, n);
}
void launch_apply_gauge(void *U, const void *G, const Geo &g, hipStream_t s) {
  hipLaunchKernelGGL(k_apply_gauge, dim3(nblk(g.V)), dim3(kBlock), 0, s, (double2 *)U, (const double2 *)G, g);
}
void launch_promote(const void *a, void *b, long long n, hipStream_t s) {
  const unsigned bl = (unsigned)std::min<long long>(nblk(n), 16384);
  hipLaunchKernelGGL(k_promote, dim3(bl), dim3(kBlock), 0, s, (const float2 *)a, (double2 *)b, n);
}
void launch_reunit(void *U, long long V, int nd, hipStream_t s) {
  hipLaunchKernelGGL(k_reunit, dim3(nblk(V)), dim3(kBlock), 0, s, (double2 *)U, V, nd);
}
void launch_measure(const void *U, double *part, const Geo &g, hipStream_t s) {
  hipLaunchKernelGGL(k_measure, dim3(nblk(g.V)), dim3(kBlock), 0, s, (const double2 *)U, part, g);
}

}  // namespace gfix
