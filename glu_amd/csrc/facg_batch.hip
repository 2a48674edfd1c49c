// facg_batch.hip -- FACG() (Gfix/FACG.c:238-392) for many small lattices of
// one shape in a single launch: one workgroup owns one lattice and runs its
// whole gauge fixing -- theta test, steepest-descent start, Polak-Ribiere
// CG with Fourier acceleration and the three-probe line search, GLU's
// continuation below 1e3 acc and its 123456789 sentinel -- with workgroup
// barriers only.  No workgroup waits for another, so the launch needs no
// grid-wide synchronisation and every workgroup ends when its own lattice
// stops.
//
// On a small lattice the iteration of the context path is bound by the
// latency of its dozen dependent launches (an 8^4 iteration takes about
// 75 us of which the bytes account for a few); here one launch carries
// every iteration of hundreds of lattices.
//
// Arithmetic is FP64 throughout.  The caller's links (GLU's layout, fixed
// in place) are copied at the start into device scratch as the first two
// columns of each link in structure-of-arrays planes, the layout the
// context path streams, and written back in GLU's layout at the end with
// the third column completed.  The fields in, sn and in_old (4 packed
// complex fields each, Gfix/FACG.c:57-60) live in the same scratch, sn
// twice: the CG update is formed inside the line-search probes, which
// write the new direction beside the old one.
// The Fourier acceleration transforms each axis (extents up to 16) on
// pencils in registers, by radix-2 FFT where the extent is a power of two
// and by direct sums otherwise: the x, y, z transforms of each time slice
// staged in LDS, the t transforms with the p^2 factor between them, so
// three passes over the fields.  Reductions are workgroup sums in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/glu_gfix.h"
#include "gfix_ctx.h"

using namespace gfix;

namespace {

constexpr int kMaxWgWaves = 8;  // workgroups of up to 512 threads
constexpr int kMaxExtent = 16;
constexpr double kCgTolB = 5e-11;     // CG_TOL, Headers/GLU_definitions.h:449-455
constexpr double kPl1 = 0.15, kPl2 = 0.30;  // Gfix/line_search.c:34-35
// scratch per site of each lattice, in complex numbers: the links'
// columns 0 and 1 (24), the fields in, in_old and two of sn (4 each)
constexpr long long kScratchPerSite = 40;
constexpr int kMaxSegments = 100;  // continuation segments; GLU does not bound them

// mL, mS: multipliers for division by the extents and strides (CDNA has
// no integer divide instruction; a division by a runtime value expands to
// a few dozen instructions, 64-bit ones to many more): for 2 <= d and
// x < 2^32 / d, x / d == umulhi(x, floor((2^32 - 1) / d) + 1).  Lattices
// of at most 16^4 sites keep every quotient here in that range.
struct BGeo {
  int L[4];
  long long S[4];
  long long V;
  unsigned mL[4], mS[4];  // mS[0] unused (S[0] = 1)
};

__device__ __forceinline__ unsigned bquot(unsigned x, unsigned m) { return __umulhi(x, m); }

__device__ __forceinline__ void bcoords(const BGeo &g, long long i, int x[4]) {
  unsigned r = (unsigned)i;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    const unsigned q = bquot(r, g.mL[mu]);
    x[mu] = (int)(r - q * (unsigned)g.L[mu]);
    r = q;
  }
}
__device__ __forceinline__ long long bfwd(const BGeo &g, long long i, const int x[4], int mu) {
  return x[mu] < g.L[mu] - 1 ? i + g.S[mu] : i - (long long)(g.L[mu] - 1) * g.S[mu];
}
__device__ __forceinline__ long long bbwd(const BGeo &g, long long i, const int x[4], int mu) {
  return x[mu] > 0 ? i - g.S[mu] : i + (long long)(g.L[mu] - 1) * g.S[mu];
}

// GLU's layout: link (i, mu) is 9 complex numbers, row-major
__device__ __forceinline__ const double2 *glu_link(const double2 *U, long long i, int mu) { return U + (i * 4 + mu) * 9; }

// The solve works on a copy of the links in scratch: columns 0 and 1 of
// link (i, mu) as 6 planes of V complex numbers, element k of direction mu
// at ((mu * 6 + k) V + i), so a wave's loads of one element are one
// contiguous run; column 2 is completed on every load, as the context
// path stores and completes its links
constexpr int kCol01[6] = {0, 1, 3, 4, 6, 7};
__device__ __forceinline__ su3<double> bload(const double2 *__restrict__ Uc, long long V, long long i, int mu) {
  su3<double> m;
#pragma unroll
  for (int k = 0; k < 6; k++) m.e[kCol01[k]] = ld<double>(Uc + (mu * 6 + k) * V + i);
  complete_col2(m);
  return m;
}
__device__ __forceinline__ void bstore(double2 *__restrict__ Uc, long long V, long long i, int mu, const su3<double> &m) {
#pragma unroll
  for (int k = 0; k < 6; k++) st<double>(Uc + (mu * 6 + k) * V + i, m.e[kCol01[k]]);
}

__device__ __forceinline__ su3<double> bgauge(const double2 *__restrict__ F, long long V, long long i, double al) {
  auto f = [&](long long k) { return ld<double>(F + k); };
  return gauge_matrix<double>(f(i), f(V + i), f(2 * V + i), f(3 * V + i), al);
}

// workgroup sum of NV values, fixed order, the result in every thread
template <int WG, int NV> __device__ void wg_sum(double (&v)[NV], double (*sh)[kMaxWgWaves]) {
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
    for (int j = 0; j < WG / kWave; j++) s += sh[k][j];
    v[k] = s;
  }
  __syncthreads();
}

// Delta and its packing into F (Gfix/FACG.c:38-73); returns sum tr Delta^2
template <int WG> __device__ __forceinline__ double bderiv(const double2 *__restrict__ U, double2 *__restrict__ F, const BGeo &g,
                         double (*sh)[kMaxWgWaves]) {
  double v[1] = {0.0};
  for (long long i = threadIdx.x; i < g.V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    DerivAcc<double> d;
    for (int mu = 0; mu < 4; mu++) d.add(bload(U, g.V, i, mu), bload(U, g.V, bbwd(g, i, x, mu), mu));
    v[0] += d.store(F, g.V, i);
  }
  wg_sum<WG, 1>(v, sh);
  return v[0];
}

// the site's own-link terms of Delta, left by begauge_own in F (the 8
// accumulators as 4 complex planes), read back by bderiv_rest
__device__ __forceinline__ void own_store(double2 *__restrict__ F, long long V, long long i, const DerivAcc<double> &d) {
  F[i] = make_double2(d.R0, d.R1);
  F[V + i] = make_double2(d.I1, d.R2);
  F[2 * V + i] = make_double2(d.I2, d.R3);
  F[3 * V + i] = make_double2(d.R4, d.I4);
}

// bderiv after begauge_own: the own-link terms from F, the backward
// neighbours' links added, packed into F in place (each site by its own
// thread).  The terms are summed in another order than bderiv's.
template <int WG> __device__ __forceinline__ double bderiv_rest(const double2 *__restrict__ U, double2 *__restrict__ F,
                                                                const BGeo &g, double (*sh)[kMaxWgWaves]) {
  double v[1] = {0.0};
  const long long V = g.V;
  for (long long i = threadIdx.x; i < V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    DerivAcc<double> d;
    const double2 p0 = F[i], p1 = F[V + i], p2 = F[2 * V + i], p3 = F[3 * V + i];
    d.R0 = p0.x, d.R1 = p0.y, d.I1 = p1.x, d.R2 = p1.y, d.I2 = p2.x, d.R3 = p2.y, d.R4 = p3.x, d.I4 = p3.y;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) d.add_bwd(bload(U, V, bbwd(g, i, x, mu), mu));
    v[0] += d.store(F, V, i);
  }
  wg_sum<WG, 1>(v, sh);
  return v[0];
}

// One direct DFT of a pencil held in registers: out_k = sum_j a_j w_jk,
// j ascending, w_jk = tw[j k mod N] forward (sign -1) and its conjugate
// backward; outputs k and k + 1 are summed together, two independent
// chains over the same inputs.  ld(j) reads input j, st(k, v) writes
// output k; every input is read before the first output is written, so a
// pencil may be transformed in place.
// bit reversal of i over log2 N bits
template <int N> __device__ __forceinline__ constexpr int brev(int i) {
  int r = 0;
  for (int n = N; n > 1; n >>= 1) {
    r = (r << 1) | (i & 1);
    i >>= 1;
  }
  return r;
}

// In place, the DFT of N = 2, 4, 8 or 16 numbers in registers by radix-2
// decimation in time (unnormalised, the sign convention of the direct
// sums): the inputs taken in bit-reversed order, then log2 N stages of
// butterflies with twiddles tw[j N / len] from the length-N table, the
// j = 0 twiddle (1) not multiplied.  N log2 N / 2 complex products instead
// of the direct sums' N^2; every index is a compile-time constant.
template <int N>
__device__ __forceinline__ void fft_pow2(cx<double> (&a)[kMaxExtent], int sign, const double2 *tw) {
  cx<double> b[N];
#pragma unroll
  for (int i = 0; i < N; i++) b[i] = a[brev<N>(i)];
#pragma unroll
  for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
    for (int i = 0; i < N; i += len) {
#pragma unroll
      for (int j = 0; j < len / 2; j++) {
        cx<double> t = b[i + j + len / 2];
        if (j > 0) {
          const double2 w = tw[j * (N / len)];
          t = cmul(t, mk<double>(w.x, sign < 0 ? w.y : -w.y));
        }
        const cx<double> u = b[i + j];
        b[i + j] = cadd(u, t);
        b[i + j + len / 2] = csub(u, t);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++) a[i] = b[i];
}

template <int N, typename Ld, typename St>
__device__ __forceinline__ void pencil_fft(int sign, const double2 *tw, Ld ld, St st) {
  cx<double> a[kMaxExtent];
#pragma unroll
  for (int j = 0; j < N; j++) a[j] = ld(j);
  fft_pow2<N>(a, sign, tw);
#pragma unroll
  for (int k = 0; k < N; k++) st(k, a[k]);
}

// pencil_dft: power-of-two extents by fft_pow2, the others by the direct
// sums
template <typename Ld, typename St>
__device__ __forceinline__ void pencil_dft(int N, int sign, const double2 *tw, Ld ld, St st) {
  switch (N) {
    case 2: pencil_fft<2>(sign, tw, ld, st); return;
    case 4: pencil_fft<4>(sign, tw, ld, st); return;
    case 8: pencil_fft<8>(sign, tw, ld, st); return;
    case 16: pencil_fft<16>(sign, tw, ld, st); return;
    default: break;
  }
  cx<double> a[kMaxExtent];
#pragma unroll
  for (int j = 0; j < kMaxExtent; j++)
    if (j < N) a[j] = ld(j);
  for (int k = 0; k < N; k += 2) {
    const int k2 = k + 1 < N ? k + 1 : k;
    cx<double> acc = mk<double>(0.0, 0.0), acc2 = acc;
    int jk = 0, jk2 = 0;  // j k mod N
#pragma unroll
    for (int j = 0; j < kMaxExtent; j++) {
      if (j < N) {
        const double2 w = tw[jk], w2 = tw[jk2];
        acc = cadd(acc, cmul(a[j], mk<double>(w.x, sign < 0 ? w.y : -w.y)));
        acc2 = cadd(acc2, cmul(a[j], mk<double>(w2.x, sign < 0 ? w2.y : -w2.y)));
        jk += k;
        if (jk >= N) jk -= N;
        jk2 += k2;
        if (jk2 >= N) jk2 -= N;
      }
    }
    st(k, acc);
    if (k2 != k) st(k2, acc2);
  }
}

// one DFT along axis mu of every pencil of the 4 fields (sign -1 forward,
// +1 backward, unnormalised; tw[k] = exp(-2 pi i k / N)) in place in F:
// the nine-pass form, for slabs beyond the LDS stage
template <int WG> __device__ __forceinline__ void bdft_axis(double2 *__restrict__ F, const BGeo &g, int mu, int sign, const double2 *tw) {
  const int N = g.L[mu];
  const unsigned S = (unsigned)g.S[mu], V = (unsigned)g.V;
  const unsigned np = V / N;  // pencils per field
  for (unsigned p = threadIdx.x; p < 4 * np; p += WG) {
    const unsigned f = p / np, q = p % np;
    double2 *b = F + (long long)f * V + (q % S) + (q / S) * S * N;
    pencil_dft(N, sign, tw, [&](int j) { return ld<double>(b + (long long)j * S); },
               [&](int k, cx<double> v) { st<double>(b + (long long)k * S, v); });
  }
  __syncthreads();
}

// The x, y and z transforms of the time slices of the fields (slabs of
// Lx Ly Lz sites, contiguous in F) run in LDS, as many slabs at a time as
// the stage holds: the slabs cross HBM once for the three axes instead of
// three times.  Each axis is transformed by pencil_dft on whole pencils
// in registers, in place.  Every x pencil is
// padded by one number in LDS (site (x, y, r) at x + (Lx + 1)(y + Ly r)),
// so that a wave's threads, one pencil each, read and write at a stride
// that spreads over the LDS banks.  The stage holds 8192 complex numbers
// (128 KiB) with 512 threads, whose workgroup holds a CU alone, and 4096
// with fewer, whose workgroups share a CU.
template <int WG> constexpr int stage_cx() { return WG >= 512 ? 8192 : 4096; }

// whether a slab, padded, fits the stage
template <int WG> __device__ __forceinline__ bool xyz_staged(const BGeo &g) {
  return (long long)(g.L[0] + 1) * g.L[1] * g.L[2] <= stage_cx<WG>();
}

// x, y, z transforms of the 4 T slabs of the 4 fields (requires
// xyz_staged); the z sums are written straight back to F
template <int WG>
__device__ __forceinline__ void bdft_xyz(double2 *__restrict__ F, const BGeo &g, int sign,
                                         const double2 (*tw)[kMaxExtent], double2 *__restrict__ lds) {
  const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2], s3 = (int)g.S[3], sxy = (int)g.S[2];
  const int px = Lx + 1, per = stage_cx<WG>() / (px * Ly * Lz), nslab = 4 * g.L[3];
  for (int s0 = 0; s0 < nslab; s0 += per) {
    const int npts = min(per, nslab - s0) * s3;
    double2 *src = F + (long long)s0 * s3;  // slab (f, t) starts at (f T + t) s3
    for (int p = threadIdx.x; p < npts; p += WG) lds[p + (int)bquot((unsigned)p, g.mL[0])] = src[p];
    __syncthreads();
    // x: pencil q at q px, stride 1
    for (int q = threadIdx.x; q < npts / Lx; q += WG) {
      double2 *b = lds + q * px;
      pencil_dft(Lx, sign, tw[0], [&](int j) { return ld<double>(b + j); },
                 [&](int k, cx<double> v) { st<double>(b + k, v); });
    }
    __syncthreads();
    // y: pencil q = x + Lx r (r = z + Lz s, s the slab of the round) at
    // x + px Ly r, stride px
    for (int q = threadIdx.x; q < npts / Ly; q += WG) {
      const int r = (int)bquot((unsigned)q, g.mL[0]);
      double2 *b = lds + (q - r * Lx) + px * Ly * r;
      pencil_dft(Ly, sign, tw[1], [&](int j) { return ld<double>(b + j * px); },
                 [&](int k, cx<double> v) { st<double>(b + k * px, v); });
    }
    __syncthreads();
    // z: pencil q = x + Lx y + Lx Ly s at x + px (y + Ly Lz s), stride
    // px Ly, into F at s s3 + x + Lx y, stride Lx Ly
    for (int q = threadIdx.x; q < npts / Lz; q += WG) {
      const int sl = (int)bquot((unsigned)q, g.mS[2]), xy = q - sl * sxy;
      const int y = (int)bquot((unsigned)xy, g.mL[0]), x = xy - y * Lx;
      const double2 *b = lds + x + px * (y + Ly * Lz * sl);
      double2 *o = src + (long long)sl * s3 + xy;
      pencil_dft(Lz, sign, tw[2], [&](int j) { return ld<double>(b + j * px * Ly); },
                 [&](int k, cx<double> v) { st<double>(o + (long long)k * sxy, v); });
    }
    __syncthreads();  // the stage is free for the next round, F complete
  }
}

// along t, on pencils in registers: forward sums, x MAX / p^2, backward
// sums -- the middle of the Fourier acceleration in one pass.  The scaled
// forward values go back to the pencil's own sites and are read again for
// the backward sums (the thread's own stores, still in cache), so that
// every register array is indexed by constants.  Outputs k and k + 1 are
// summed together, two independent chains over the same inputs.
template <int N, int WG>
__device__ __forceinline__ void bdft_t_scale_fft(double2 *__restrict__ F, const BGeo &g, const double2 (*tw)[kMaxExtent]) {
  const long long S = g.S[3], V = g.V;
  const double mx = 4.0 * 4 / (double)V;
  for (unsigned p = threadIdx.x; p < 4 * S; p += WG) {
    const unsigned f = bquot(p, g.mS[3]), q = p - f * (unsigned)S;
    const long long base = (long long)f * V + q;
    int x[4];
    bcoords(g, q, x);
    const double kxyz = tw[0][x[0]].x + tw[1][x[1]].x + tw[2][x[2]].x;
    cx<double> a[kMaxExtent];
#pragma unroll
    for (int j = 0; j < N; j++) a[j] = ld<double>(F + base + j * S);
    fft_pow2<N>(a, -1, tw[3]);
#pragma unroll
    for (int k = 0; k < N; k++) {
      const double kc = kxyz + tw[3][k].x;  // sum of cos(2 pi n_mu / L_mu), mu in order
      const double fac = mx / ((kc == 4.0) ? 1.0 : 2.0 * (4.0 - kc));
      a[k] = mk<double>(a[k].re * fac, a[k].im * fac);
    }
    fft_pow2<N>(a, +1, tw[3]);
#pragma unroll
    for (int k = 0; k < N; k++) st<double>(F + base + k * S, a[k]);
  }
  __syncthreads();
}

template <int WG>
__device__ __forceinline__ void bdft_t_scale(double2 *__restrict__ F, const BGeo &g, const double2 (*tw)[kMaxExtent]) {
  const int N = g.L[3];
  switch (N) {  // powers of two: the FFT, both directions in registers
    case 2: bdft_t_scale_fft<2, WG>(F, g, tw); return;
    case 4: bdft_t_scale_fft<4, WG>(F, g, tw); return;
    case 8: bdft_t_scale_fft<8, WG>(F, g, tw); return;
    case 16: bdft_t_scale_fft<16, WG>(F, g, tw); return;
    default: break;
  }
  const long long S = g.S[3], V = g.V;
  const double mx = 4.0 * 4 / (double)V;
  for (unsigned p = threadIdx.x; p < 4 * S; p += WG) {
    const unsigned f = bquot(p, g.mS[3]), q = p - f * (unsigned)S;
    const long long base = (long long)f * V + q;
    int x[4];
    bcoords(g, q, x);
    const double kxyz = tw[0][x[0]].x + tw[1][x[1]].x + tw[2][x[2]].x;
    cx<double> a[kMaxExtent];
#pragma unroll
    for (int j = 0; j < kMaxExtent; j++)
      if (j < N) a[j] = ld<double>(F + base + j * S);
    for (int k = 0; k < N; k += 2) {
      const int k2 = k + 1 < N ? k + 1 : k;
      cx<double> acc = mk<double>(0.0, 0.0), acc2 = acc;
      int jk = 0, jk2 = 0;  // j k mod N
#pragma unroll
      for (int j = 0; j < kMaxExtent; j++) {
        if (j < N) {
          const double2 w = tw[3][jk], w2 = tw[3][jk2];
          acc = cadd(acc, cmul(a[j], mk<double>(w.x, w.y)));
          acc2 = cadd(acc2, cmul(a[j], mk<double>(w2.x, w2.y)));
          jk += k;
          if (jk >= N) jk -= N;
          jk2 += k2;
          if (jk2 >= N) jk2 -= N;
        }
      }
      const double kc = kxyz + tw[3][k].x;  // sum of cos(2 pi n_mu / L_mu), mu in order
      const double fac = mx / ((kc == 4.0) ? 1.0 : 2.0 * (4.0 - kc));
      st<double>(F + base + k * S, mk<double>(acc.re * fac, acc.im * fac));
      if (k2 != k) {
        const double kc2 = kxyz + tw[3][k2].x;
        const double fac2 = mx / ((kc2 == 4.0) ? 1.0 : 2.0 * (4.0 - kc2));
        st<double>(F + base + k2 * S, mk<double>(acc2.re * fac2, acc2.im * fac2));
      }
    }
#pragma unroll
    for (int j = 0; j < kMaxExtent; j++)
      if (j < N) a[j] = ld<double>(F + base + j * S);
    for (int k = 0; k < N; k += 2) {
      const int k2 = k + 1 < N ? k + 1 : k;
      cx<double> acc = mk<double>(0.0, 0.0), acc2 = acc;
      int jk = 0, jk2 = 0;  // j k mod N
#pragma unroll
      for (int j = 0; j < kMaxExtent; j++) {
        if (j < N) {
          const double2 w = tw[3][jk], w2 = tw[3][jk2];
          acc = cadd(acc, cmul(a[j], mk<double>(w.x, -w.y)));
          acc2 = cadd(acc2, cmul(a[j], mk<double>(w2.x, -w2.y)));
          jk += k;
          if (jk >= N) jk -= N;
          jk2 += k2;
          if (jk2 >= N) jk2 -= N;
        }
      }
      st<double>(F + base + k * S, acc);
      if (k2 != k) st<double>(F + base + k2 * S, acc2);
    }
  }
  __syncthreads();
}

// forward transform, x MAX / p^2 (p^2(0) := 1, MAX = 4 ND / V), backward
// transform: Gfix/FACG.c:77-95 with Gfix/Landau.c:275-284
template <int WG>
__device__ __forceinline__ void bfourier_accelerate(double2 *__restrict__ F, const BGeo &g,
                                                    const double2 (*tw)[kMaxExtent], double2 *__restrict__ lds) {
  if (xyz_staged<WG>(g)) {  // three passes over the fields
    bdft_xyz<WG>(F, g, -1, tw, lds);
    bdft_t_scale<WG>(F, g, tw);
    bdft_xyz<WG>(F, g, +1, tw, lds);
    return;
  }
  for (int mu = 0; mu < 4; mu++) bdft_axis<WG>(F, g, mu, -1, tw[mu]);
  const double mx = 4.0 * 4 / (double)g.V;
  for (long long i = threadIdx.x; i < g.V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    double kc = 0.0;
    for (int mu = 0; mu < 4; mu++) kc += tw[mu][x[mu]].x;  // cos(2 pi n / L)
    const double p2 = (kc == 4.0) ? 1.0 : 2.0 * (4.0 - kc);
    const double fac = mx / p2;
    for (int f = 0; f < 4; f++) {
      const cx<double> z = ld<double>(F + f * g.V + i);
      st<double>(F + f * g.V + i, mk<double>(z.re * fac, z.im * fac));
    }
  }
  __syncthreads();
  for (int mu = 0; mu < 4; mu++) bdft_axis<WG>(F, g, mu, +1, tw[mu]);
}

// <in, in> and <in, in - old> in GLU's trace metric (Gfix/FACG.c:99-178)
template <int WG> __device__ __forceinline__ void bpr_sums(const double2 *__restrict__ F, const double2 *__restrict__ O, const BGeo &g, double out[2],
                         double (*sh)[kMaxWgWaves]) {
  double v[2] = {0.0, 0.0};
  const long long V = g.V;
  for (long long i = threadIdx.x; i < V; i += WG) {
    const cx<double> z0 = ld<double>(F + i), o0 = ld<double>(O + i);
    const double zr = z0.re, zi = z0.im, tr = zr - o0.re, ti = zi - o0.im;
    double s1 = 2.0 * (zr * zr + zi * zi + zr * zi);
    double s2 = 2.0 * (zr * tr + zi * ti) + zr * ti + zi * tr;
    for (int f = 1; f < 4; f++) {
      const cx<double> z = ld<double>(F + f * V + i), o = ld<double>(O + f * V + i);
      s1 += 2.0 * (z.re * z.re + z.im * z.im);
      s2 += 2.0 * (z.re * (z.re - o.re) + z.im * (z.im - o.im));
    }
    v[0] += s1;
    v[1] += s2;
  }
  wg_sum<WG, 2>(v, sh);
  out[0] = v[0];
  out[1] = v[1];
}

// functional probes at 0, PL1, PL2 along the packed direction D
// (Gfix/line_search.c:274-341): each link is loaded once for the three
// probes.  The direction loop stays rolled: unrolled, the compiler hoists
// the loads of all four links and their neighbours' gauge matrices and
// spills.
template <int WG> __device__ __forceinline__ void bprobes(const double2 *__restrict__ U, const double2 *__restrict__ D, const BGeo &g, double out[3],
                        double (*sh)[kMaxWgWaves]) {
  double v[3] = {0.0, 0.0, 0.0};
  const long long V = g.V;
  for (long long i = threadIdx.x; i < V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    // the site's gauge matrices are rebuilt from its four field values in
    // every direction rather than held across the loop: fewer registers
    // live (34 -> 24 spilled at 512 threads), the same values
    const cx<double> f0 = ld<double>(D + i), f1 = ld<double>(D + V + i), f2 = ld<double>(D + 2 * V + i),
                     f3 = ld<double>(D + 3 * V + i);
    double acc1 = 0.0, acc2 = 0.0;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      const su3<double> u = bload(U, V, i, mu);
      v[0] += u.e[0].re + u.e[4].re + u.e[8].re;
      cx<double> au1[6], au2[6];
      ab_rows01<double>(gauge_matrix<double>(f0, f1, f2, f3, kPl1), u, au1);
      ab_rows01<double>(gauge_matrix<double>(f0, f1, f2, f3, kPl2), u, au2);
      const long long f = bfwd(g, i, x, mu);
      acc1 += retrace_rows01_dag<double>(au1, bgauge(D, V, f, kPl1));
      acc2 += retrace_rows01_dag<double>(au2, bgauge(D, V, f, kPl2));
    }
    v[1] += acc1;
    v[2] += acc2;
  }
  wg_sum<WG, 3>(v, sh);
  for (int k = 0; k < 3; k++) out[k] = v[k];
}

// bprobes along the new direction sn + beta... : D = I + beta S formed
// where it is read (at each site and its forward neighbours, with the
// arithmetic of the CG update), and each site's own value written to N,
// the next direction -- the CG update folded into the probes.  S is only
// read here, so no thread waits for another's update.
template <int WG>
__device__ __forceinline__ void bprobes_cg(const double2 *__restrict__ U, const double2 *__restrict__ I,
                                           const double2 *__restrict__ S, double2 *__restrict__ N, double beta,
                                           const BGeo &g, double out[3], double (*sh)[kMaxWgWaves]) {
  double v[3] = {0.0, 0.0, 0.0};
  const long long V = g.V;
  auto dir = [&](long long k) {
    const double2 a = I[k], s = S[k];
    return make_double2(a.x + beta * s.x, a.y + beta * s.y);
  };
  for (long long i = threadIdx.x; i < V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    const double2 d0 = dir(i), d1 = dir(V + i), d2 = dir(2 * V + i), d3 = dir(3 * V + i);
    N[i] = d0;
    N[V + i] = d1;
    N[2 * V + i] = d2;
    N[3 * V + i] = d3;
    const cx<double> f0 = mk<double>(d0.x, d0.y), f1 = mk<double>(d1.x, d1.y), f2 = mk<double>(d2.x, d2.y),
                     f3 = mk<double>(d3.x, d3.y);
    double acc1 = 0.0, acc2 = 0.0;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      const su3<double> u = bload(U, V, i, mu);
      v[0] += u.e[0].re + u.e[4].re + u.e[8].re;
      cx<double> au1[6], au2[6];
      ab_rows01<double>(gauge_matrix<double>(f0, f1, f2, f3, kPl1), u, au1);
      ab_rows01<double>(gauge_matrix<double>(f0, f1, f2, f3, kPl2), u, au2);
      const long long f = bfwd(g, i, x, mu);
      const double2 n0 = dir(f), n1 = dir(V + f), n2 = dir(2 * V + f), n3 = dir(3 * V + f);
      const cx<double> g0 = mk<double>(n0.x, n0.y), g1 = mk<double>(n1.x, n1.y), g2 = mk<double>(n2.x, n2.y),
                       g3 = mk<double>(n3.x, n3.y);
      acc1 += retrace_rows01_dag<double>(au1, gauge_matrix<double>(g0, g1, g2, g3, kPl1));
      acc2 += retrace_rows01_dag<double>(au2, gauge_matrix<double>(g0, g1, g2, g3, kPl2));
    }
    v[1] += acc1;
    v[2] += acc2;
  }
  wg_sum<WG, 3>(v, sh);
  for (int k = 0; k < 3; k++) out[k] = v[k];
}

// U_mu(x) <- g(x) U_mu(x) g(x + mu)^dag with g from D at step alpha, in
// place: g depends on D only, and each link is written by its own site
// (Gfix/line_search.c:60-97)
template <int WG> __device__ __forceinline__ void begauge(double2 *__restrict__ U, const double2 *__restrict__ D, const BGeo &g, double alpha) {
  for (long long i = threadIdx.x; i < g.V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    const su3<double> A = bgauge(D, g.V, i, alpha);
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++)
      bstore(U, g.V, i, mu, gtransform<double>(A, bload(U, g.V, i, mu), bgauge(D, g.V, bfwd(g, i, x, mu), alpha)));
  }
  __syncthreads();
}

__device__ __forceinline__ void swap_fields(double2 *&a, double2 *&b) {
  double2 *t = a;
  a = b;
  b = t;
}

// begauge, and the updated links' own-site terms of the next Delta into
// F (own_store): the next derivative then reads only the backward
// neighbours' links (bderiv_rest)
template <int WG>
__device__ __forceinline__ void begauge_own(double2 *__restrict__ U, const double2 *__restrict__ D, double2 *__restrict__ F,
                                            const BGeo &g, double alpha) {
  const long long V = g.V;
  for (long long i = threadIdx.x; i < V; i += WG) {
    int x[4];
    bcoords(g, i, x);
    const su3<double> A = bgauge(D, V, i, alpha);
    DerivAcc<double> d;
#pragma unroll 1
    for (int mu = 0; mu < 4; mu++) {
      const su3<double> m = gtransform<double>(A, bload(U, V, i, mu), bgauge(D, V, bfwd(g, i, x, mu), alpha));
      bstore(U, V, i, mu, m);
      d.add_own(m);
    }
    own_store(F, V, i, d);
  }
  __syncthreads();
}

struct BTrace {
  double *rows;  // 3 per row
  long long cap, n;
  __device__ void rec(double th, double al, double be) {
    if (rows && n < cap && threadIdx.x == 0) {
      rows[3 * n] = th;
      rows[3 * n + 1] = al;
      rows[3 * n + 2] = be;
    }
    if (n < cap) n++;
  }
};

// FACG of lattice blockIdx.x: links U (GLU layout, V sites), scratch
// fields in, in_old and two of sn (4 V complex each).  Writes its iteration count (GLU's
// return value) and final theta.
template <int WG>
__global__ __launch_bounds__(WG, WG >= 256 ? 2 : 1) void k_facg_batch(double2 *const *__restrict__ lat, double2 *__restrict__ scratch,
                                                          BGeo g, int nlat, double gf, double acc,
                                                          long long max_iters, long long *iters_out,
                                                          double *theta_out, double *trace, long long trace_cap,
                                                          int keep_failed) {
  __shared__ double sh[3][kMaxWgWaves];
  __shared__ double2 tw[4][kMaxExtent];
  __shared__ double2 stage[stage_cx<WG>()];
  const int b = blockIdx.x;
  if (b >= nlat) return;
  const long long V = g.V;
  double2 *glu = lat[b];
  double2 *U = scratch + (long long)b * V * kScratchPerSite;  // the links, columns 0 and 1
  double2 *in = U + 24 * V, *sn = in + 4 * V, *old = in + 8 * V, *sn2 = in + 12 * V;
  for (long long j = threadIdx.x; j < 4 * V; j += WG) {
    const long long i = j >> 2;
    const int mu = (int)(j & 3);
    const double2 *p = glu_link(glu, i, mu);
#pragma unroll
    for (int k = 0; k < 6; k++) U[(mu * 6 + k) * V + i] = p[kCol01[k]];
  }
  for (int t = threadIdx.x; t < 4 * kMaxExtent; t += WG) {
    const int mu = t / kMaxExtent, k = t % kMaxExtent;
    const double a = 6.283185307179586 * (double)k / (double)g.L[mu];
    tw[mu][k] = make_double2(cos(a), -sin(a));
  }
  __syncthreads();
  BTrace tr{trace ? trace + (long long)b * trace_cap * 3 : nullptr, trace_cap, 0};
  const double norm = 1.0 / (3.0 * (double)V);  // GFNORM_LANDAU
  // theta_test_lin (FACG.c:349): 2 x the FA theta of the input links
  double theta = 2.0 * bderiv<WG>(U, in, g, sh) * norm;
  long long iters = 0;
  if (theta < acc) {
    if (threadIdx.x == 0) {
      iters_out[b] = 0;
      theta_out[b] = theta;
    }
    return;
  }
  // GLU's red[] slots carried into a continuation's SD start (FACG.c:249-265)
  double stale_theta = 0.0, stale_insum = 0.0, stale_probe[3] = {0.0, 0.0, 0.0};
  for (int seg = 0; seg < kMaxSegments; seg++) {
    // steepest-descent start
    theta = bderiv<WG>(U, in, g, sh) * norm + stale_theta;
    bfourier_accelerate<WG>(in, g, tw, stage);
    double pr[3];
    bprobes<WG>(U, in, g, pr, sh);
    double val[3];
    for (int k = 0; k < 3; k++) val[k] = -(pr[k] + stale_probe[k]);
    const double als[3] = {0.0, kPl1, kPl2};
    double alpha = approx_minimum(als, val);
    begauge<WG>(U, in, g, alpha);
    tr.rec(theta, alpha, 0.0);
    for (long long j = threadIdx.x; j < 4 * V; j += WG) sn[j] = in[j];
    // in_old <- in: the two buffers trade roles instead of copying, and
    // the next derivative is written into the other
    swap_fields(in, old);
    double s2[2];
    bpr_sums<WG>(old, old, g, s2, sh);  // <in, in> (sum_DER3)
    double inold = s2[0] + stale_insum;
    long long loc = 0;
    double insum = 0.0, probe[3] = {0.0, 0.0, 0.0};
    int did_ls = 0;
    int own = 0;  // whether begauge_own left the own-link terms in `in`
    for (;;) {
      theta = (own ? bderiv_rest<WG>(U, in, g, sh) : bderiv<WG>(U, in, g, sh)) * norm;
      bfourier_accelerate<WG>(in, g, tw, stage);
      bpr_sums<WG>(in, old, g, s2, sh);
      insum = s2[0];
      double beta = s2[1] / inold;
      if (beta < 0.0) beta = 0.0;
      if (theta < kCgTolB) beta = insum / inold;  // Fletcher-Reeves below CG_TOL
      inold = insum;
      if (theta > kCgTolB) {
        // sn <- in + beta sn inside the probes, into the other buffer
        bprobes_cg<WG>(U, in, sn, sn2, beta, g, probe, sh);
        swap_fields(sn, sn2);
        for (int k = 0; k < 3; k++) val[k] = -probe[k];
        alpha = approx_minimum(als, val);
        did_ls = 1;
      } else {
        for (long long j = threadIdx.x; j < 4 * V; j += WG) {
          const double2 a = in[j], s = sn[j];
          sn[j] = make_double2(a.x + beta * s.x, a.y + beta * s.y);
        }
        __syncthreads();
        alpha = gf;
        did_ls = 0;
      }
      swap_fields(in, old);  // in_old <- in
      begauge_own<WG>(U, sn, in, g, alpha);  // `in` is free until the next derivative
      own = 1;
      tr.rec(theta, alpha, beta);
      loc++;
      if (!(theta > acc && loc < max_iters)) break;
    }
    stale_theta = theta;
    stale_insum = insum;
    for (int k = 0; k < 3; k++) stale_probe[k] = did_ls ? probe[k] : 0.0;
    iters = loc;
    if (loc >= max_iters) {
      if (theta < 1e3 * acc) continue;  // continuation (FACG.c:372-378)
      iters = 123456789;
    }
    break;
  }
  if (iters == max_iters && !(theta < acc)) iters = 123456789;  // the segment bound ran out
  __syncthreads();
  // the fixed links back in GLU's layout, column 2 completed (a failed
  // lattice keeps its input when the caller restarts from it)
  for (long long j = threadIdx.x; j < 4 * V && !(keep_failed && iters == 123456789); j += WG) {
    const long long i = j >> 2;
    const int mu = (int)(j & 3);
    const su3<double> m = bload(U, V, i, mu);
    double2 *p = glu + (i * 4 + mu) * 9;
#pragma unroll
    for (int k = 0; k < 9; k++) st<double>(p + k, m.e[k]);
  }
  if (threadIdx.x == 0) {
    iters_out[b] = iters;
    theta_out[b] = theta;
  }
}

}  // namespace

// the shape and sizes one launch takes, checked before any device work
static int batch_geometry(const size_t dims[4], size_t nlat, size_t max_iters, BGeo *g) {
  if (!dims) return gfix_fail(-1, "glu_gfix_facg_batch: null argument");
  if (nlat > 65535) return gfix_fail(-1, "glu_gfix_facg_batch: %zu lattices exceed the 65535 of one launch", nlat);
  if (max_iters == 0) return gfix_fail(-1, "max_iters must be positive");
  long long V = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (dims[mu] < 2 || dims[mu] > (size_t)kMaxExtent)
      return gfix_fail(-1, "glu_gfix_facg_batch: extent %zu outside 2..%d", dims[mu], kMaxExtent);
    g->L[mu] = (int)dims[mu];
    g->S[mu] = V;
    V *= (long long)dims[mu];
  }
  g->V = V;
  // division multipliers (extents are at least 2, strides past the first too)
  auto mul = [](long long d) { return (unsigned)(0xFFFFFFFFull / (unsigned long long)d) + 1u; };
  for (int mu = 0; mu < 4; mu++) {
    g->mL[mu] = mul(g->L[mu]);
    g->mS[mu] = mu ? mul(g->S[mu]) : 0u;
  }
  return 0;
}

// FACG of nlat lattices of one shape whose links (GLU's layout, device
// memory) are at lat[k], a host array of device pointers, on a stream of
// its own; with keep_failed, a lattice that returns 123456789 keeps its
// input links
int facg_batch_run(const size_t dims[4], size_t nlat, double *const *lat, double gf_tune, double accuracy,
                   size_t max_iters, size_t *iters, double *theta, double *trace, size_t trace_cap,
                   int keep_failed) {
  if (!lat || !iters || !theta) return gfix_fail(-1, "glu_gfix_facg_batch: null argument");
  BGeo g{};
  if (int rc = batch_geometry(dims, nlat, max_iters, &g)) return rc;
  if (nlat == 0) return 0;
  const long long V = g.V;
  for (size_t k = 0; k < nlat; k++)
    if (!lat[k]) return gfix_fail(-1, "glu_gfix_facg_batch: null links of lattice %zu", k);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return gfix_fail(-3, "no HIP device: the gauge-fixing library has no CPU path");
  double2 *scratch = nullptr, **d_lat = nullptr;
  long long *d_it = nullptr;
  double *d_th = nullptr, *d_tr = nullptr;
  hipStream_t st = nullptr;
  int rc = 0;
  const size_t tcap = trace ? trace_cap : 0;
  if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    st = nullptr;
    rc = gfix_fail(-2, "glu_gfix_facg_batch: stream creation failed");
  } else if (hipMalloc(&scratch, nlat * (size_t)V * kScratchPerSite * sizeof(double2)) != hipSuccess ||
      hipMalloc(&d_lat, nlat * sizeof(double2 *)) != hipSuccess ||
      hipMalloc(&d_it, nlat * sizeof(long long)) != hipSuccess || hipMalloc(&d_th, nlat * sizeof(double)) != hipSuccess ||
      (tcap && hipMalloc(&d_tr, nlat * tcap * 3 * sizeof(double)) != hipSuccess)) {
    rc = gfix_fail(-2, "glu_gfix_facg_batch: scratch allocation failed");
  } else if (hipMemcpyAsync(d_lat, lat, nlat * sizeof(double2 *), hipMemcpyHostToDevice, st) != hipSuccess) {
    rc = gfix_fail(-2, "glu_gfix_facg_batch: pointer upload failed");
  } else {
    // threads per lattice: 512 unless GLU_GFIX_BATCH_WG picks 128 or 256
    // (512 8^4 lattices: 9350, 8940 and 7080 configs/s at 512, 256 and 128)
    int wg = 512;
    if (const char *e = getenv("GLU_GFIX_BATCH_WG")) wg = atoi(e);
    auto launch = [&](auto kernel, int threads) {
      hipLaunchKernelGGL(kernel, dim3((unsigned)nlat), dim3(threads), 0, st, (double2 *const *)d_lat, scratch, g,
                         (int)nlat, gf_tune, accuracy, (long long)max_iters, d_it, d_th, d_tr, (long long)tcap,
                         keep_failed);
    };
    if (wg == 128)
      launch(k_facg_batch<128>, 128);
    else if (wg == 256)
      launch(k_facg_batch<256>, 256);
    else
      launch(k_facg_batch<512>, 512);
    std::vector<long long> it(nlat);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(it.data(), d_it, nlat * sizeof(long long), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(theta, d_th, nlat * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        (tcap && hipMemcpyAsync(trace, d_tr, nlat * tcap * 3 * sizeof(double), hipMemcpyDeviceToHost, st) !=
                     hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = gfix_fail(-2, "glu_gfix_facg_batch failed: %s", hipGetErrorString(hipGetLastError()));
    } else {
      for (size_t k = 0; k < nlat; k++) iters[k] = (size_t)it[k];
    }
  }
  // teardown: nothing to undo if a free fails
  if (st) (void)hipStreamSynchronize(st);
  if (scratch) (void)hipFree(scratch);
  if (d_lat) (void)hipFree(d_lat);
  if (d_it) (void)hipFree(d_it);
  if (d_th) (void)hipFree(d_th);
  if (d_tr) (void)hipFree(d_tr);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

extern "C" int glu_gfix_facg_batch(const size_t dims[4], size_t nlat, double *links_dev, double gf_tune,
                                   double accuracy, size_t max_iters, size_t *iters, double *theta, double *trace,
                                   size_t trace_cap) {
  if (!links_dev || !iters || !theta) return gfix_fail(-1, "glu_gfix_facg_batch: null argument");
  BGeo g{};
  if (int rc = batch_geometry(dims, nlat, max_iters, &g)) return rc;
  if (nlat == 0) return 0;
  const size_t V = (size_t)g.V;
  std::vector<double *> lat(nlat);
  for (size_t k = 0; k < nlat; k++) lat[k] = links_dev + k * V * 72;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return gfix_fail(-3, "no HIP device: the gauge-fixing library has no CPU path");
  // the links may still be in flight on any of the caller's streams
  if (hipDeviceSynchronize() != hipSuccess)
    return gfix_fail(-2, "glu_gfix_facg_batch: device synchronisation failed: %s", hipGetErrorString(hipGetLastError()));
  return facg_batch_run(dims, nlat, lat.data(), gf_tune, accuracy, max_iters, iters, theta, trace, trace_cap, 0);
}
