// fft_kernels.hip -- LDS-staged mixed-radix (2, 3, 4, 5, 7) complex FFT along
// one lattice axis, batched over every pencil of the 4 packed fields, with
// the p^2 preconditioner fused into the last forward axis.
//
// This replaces the FFTW transforms of Gfix/FACG.c:77-95 (Landau, 4D) and
// Gfix/CFACG.c:70-89 (Coulomb, 3D per time slice): forward transform,
// multiply by psq = MAX / p^2 (Gfix/Landau.c:278-284), backward transform,
// both unnormalised.
//
// A workgroup stages P pencils of length N in LDS (N*P <= 2048 complex),
// runs a Stockham auto-sort radix-4/2/3 sequence on them, and writes them
// back.  Pencils along x are contiguous rows; along y, z, t consecutive
// pencils are consecutive addresses, so loads and stores are coalesced in
// both cases.  Twiddles come from a per-axis table computed on the host.
#include "gfix_kernels.h"

namespace gfix {

template <typename T> __device__ __forceinline__ cx<T> tw_at(const double2 *tw, int idx, int sign) {
  const double2 w = tw[idx];
  return mk<T>((T)w.x, sign < 0 ? (T)w.y : (T)-w.y);
}

template <typename T, int R> __device__ __forceinline__ void dft_small(cx<T> (&v)[R], int sign) {
  if constexpr (R == 2) {
    const cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const cx<T> a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const cx<T> c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // forward: y1 = b - i d, y3 = b + i d ; backward swaps
    const cx<T> id = sign < 0 ? mk<T>(d.im, -d.re) : mk<T>(-d.im, d.re);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, id);
    v[3] = csub(b, id);
  } else if constexpr (R == 8) {
    // two 4-point DFTs of the even and odd inputs, combined with the
    // eighth roots of unity e^{sign i pi k / 4}
    cx<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<T, 4>(e, sign);
    dft_small<T, 4>(o, sign);
    const T h = T(0.70710678118654752440);
    const T sg = sign < 0 ? T(-1) : T(1);
    cx<T> w[4];
    w[0] = o[0];
    w[1] = mk<T>(h * (o[1].re - sg * o[1].im), h * (o[1].im + sg * o[1].re));  // * h (1 + i sg)
    w[2] = mk<T>(-sg * o[2].im, sg * o[2].re);                                 // * i sg
    w[3] = mk<T>(-h * (o[3].re + sg * o[3].im), h * (sg * o[3].re - o[3].im)); // * h (-1 + i sg)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = cadd(e[k], w[k]);
      v[k + 4] = csub(e[k], w[k]);
    }
  } else if constexpr (R == 5 || R == 7) {
    // plain O(R^2) DFT for the odd primes GLU's extents may contain
    constexpr double c5[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                              0.30901699437494745};
    constexpr double s5[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                              -0.9510565162951535};
    constexpr double c7[7] = {1.0, 0.6234898018587336, -0.22252093395631434, -0.9009688679024191,
                              -0.9009688679024191, -0.22252093395631434, 0.6234898018587336};
    constexpr double s7[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.43388373911755823,
                              -0.43388373911755823, -0.9749279121818236, -0.7818314824680298};
    cx<T> y[R];
#pragma unroll
    for (int m = 0; m < R; m++) {
      cx<T> acc = v[0];
#pragma unroll
      for (int q = 1; q < R; q++) {
        const int k = (q * m) % R;
        const T c = (T)(R == 5 ? c5[k] : c7[k]), sn = (T)(sign * (R == 5 ? s5[k] : s7[k]));
        acc = cadd(acc, cmul(v[q], mk<T>(c, sn)));
      }
      y[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; m++) v[m] = y[m];
  } else if constexpr (R == 3) {
    const T c1 = T(-0.5), s1 = T(sign) * T(0.86602540378443864676);
    const cx<T> a = v[0], b = v[1], c = v[2];
    const cx<T> sbc = cadd(b, c), dbc = csub(b, c);
    v[0] = cadd(a, sbc);
    const cx<T> m = mk<T>(a.re + c1 * sbc.re, a.im + c1 * sbc.im);
    // i s1 (b - c)
    const cx<T> r = mk<T>(-s1 * dbc.im, s1 * dbc.re);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
}

constexpr int kFftThreads = 256;
constexpr int kFftMaxItems = 4;

// one Stockham stage of radix R on P pencils of length N held in LDS
template <typename T, int R>
__device__ __forceinline__ void fft_stage(vec2<T> *lds, int ldP, int N, int P, int Ns, const double2 *tw,
                                          int sign) {
  const int nbf = N / R;
  const int total = nbf * P;
  const int tstep = N / (Ns * R);
  cx<T> v[kFftMaxItems][R];
#pragma unroll
  for (int it = 0; it < kFftMaxItems; it++) {
    const int w = threadIdx.x + it * kFftThreads;
    if (w < total) {
      const int j = w % P, b = w / P, k = b % Ns;
#pragma unroll
      for (int q = 0; q < R; q++) v[it][q] = ld<T>(lds + (b + q * nbf) * ldP + j);
      if (Ns > 1) {
#pragma unroll
        for (int q = 1; q < R; q++) v[it][q] = cmul(v[it][q], tw_at<T>(tw, (q * k * tstep) % N, sign));
      }
      dft_small<T, R>(v[it], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kFftMaxItems; it++) {
    const int w = threadIdx.x + it * kFftThreads;
    if (w < total) {
      const int j = w % P, b = w / P, k = b % Ns;
      const int out = (b / Ns) * Ns * R + k;
#pragma unroll
      for (int q = 0; q < R; q++) st<T>(lds + (out + q * Ns) * ldP + j, v[it][q]);
    }
  }
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ void fft_run(vec2<T> *lds, int ldP, const FftAxis &ax, int sign) {
  int Ns = 1;
  for (int s = 0; s < ax.nst; s++) {
    const int R = ax.rad[s];
    if (R == 4)
      fft_stage<T, 4>(lds, ldP, ax.N, ax.P, Ns, ax.tw, sign);
    else if (R == 2)
      fft_stage<T, 2>(lds, ldP, ax.N, ax.P, Ns, ax.tw, sign);
    else if (R == 3)
      fft_stage<T, 3>(lds, ldP, ax.N, ax.P, Ns, ax.tw, sign);
    else if (R == 5)
      fft_stage<T, 5>(lds, ldP, ax.N, ax.P, Ns, ax.tw, sign);
    else
      fft_stage<T, 7>(lds, ldP, ax.N, ax.P, Ns, ax.tw, sign);
    Ns *= R;
  }
}

// mode: 0 forward, 1 backward, 2 forward * psq * backward
// blockIdx.z selects a batch slice of slice_len elements (Coulomb: one time
// slice); slices with slice_done[z] set are skipped
template <typename T>
__global__ __launch_bounds__(kFftThreads) void k_fft_axis(vec2<T> *__restrict__ F, long long fstride, FftAxis ax,
                                                         int mode, PsqInfo ps, const Ctrl *__restrict__ ctrl,
                                                         long long slice_len, const int *__restrict__ slice_done,
    const int *__restrict__ slice_map) {
  if (ctrl && ctrl->done) return;
  // batch row blockIdx.z is time slice zs (the rows of running slices only)
  const int zs = slice_map ? slice_map[blockIdx.z] : (int)blockIdx.z;
  if (slice_done && slice_done[zs]) return;
  extern __shared__ unsigned char fft_smem[];
  vec2<T> *lds = reinterpret_cast<vec2<T> *>(fft_smem);
  vec2<T> *base = F + (long long)blockIdx.y * fstride + (long long)zs * slice_len;
  const long long p0 = (long long)blockIdx.x * ax.P;
  const int ldP = ax.P + 1;
  const int tot = ax.N * ax.P;
  for (int idx = threadIdx.x; idx < tot; idx += kFftThreads) {
    int n, j;
    if (ax.S == 1) {
      j = idx / ax.N;
      n = idx % ax.N;
    } else {
      j = idx % ax.P;
      n = idx / ax.P;
    }
    const long long p = p0 + j;
    if (p < ax.npencil) {
      const long long gi = (p % ax.S) + (p / ax.S) * ax.S * ax.N + (long long)n * ax.S;
      lds[n * ldP + j] = base[gi];
    }
  }
  __syncthreads();
  fft_run<T>(lds, ldP, ax, mode == 1 ? +1 : -1);
  if (mode == 2) {
    for (int idx = threadIdx.x; idx < tot; idx += kFftThreads) {
      const int j = idx % ax.P, n = idx / ax.P;
      const long long p = p0 + j;
      if (p < ax.npencil) {
        long long gi = (p % ax.S) + (p / ax.S) * ax.S * ax.N + (long long)n * ax.S;
        // p^2 with the zero mode set to 1: Geometry/geometry.c:170-186
        double kcos = 0.0;
        for (int mu = 0; mu < ps.D; mu++) {
          const int xm = (int)(gi % ps.L[mu]);
          gi /= ps.L[mu];
          kcos += ps.cosv[ps.off[mu] + xm];
        }
        const double p2 = (kcos == (double)ps.D) ? 1.0 : 2.0 * ((double)ps.D - kcos);
        const T f = (T)(ps.max / p2);
        cx<T> z = ld<T>(lds + n * ldP + j);
        st<T>(lds + n * ldP + j, mk<T>(z.re * f, z.im * f));
      }
    }
    __syncthreads();
    fft_run<T>(lds, ldP, ax, +1);
  }
  for (int idx = threadIdx.x; idx < tot; idx += kFftThreads) {
    int n, j;
    if (ax.S == 1) {
      j = idx / ax.N;
      n = idx % ax.N;
    } else {
      j = idx % ax.P;
      n = idx / ax.P;
    }
    const long long p = p0 + j;
    if (p < ax.npencil) {
      const long long gi = (p % ax.S) + (p / ax.S) * ax.S * ax.N + (long long)n * ax.S;
      base[gi] = lds[n * ldP + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Compile-time specialisation for the common extents: N, P and the radix
// plan are constants, so every index, modulo and loop bound folds away.
// Each thread keeps one pencil for the whole pass (P divides the block), the
// twiddles sit in LDS, and the p^2 factor of the fused pass is the pencil's
// cos sum plus one table entry per element.
// ---------------------------------------------------------------------------
struct RadPlan {
  int n;
  int r[8];
};
// radix 8 first: two stages where radix 4/2 needs three (N = 32, 64), so
// one LDS round trip and two barriers fewer per pass
constexpr RadPlan make_plan(int N) {
  RadPlan p{0, {0, 0, 0, 0, 0, 0, 0, 0}};
  while (N % 8 == 0) {
    p.r[p.n++] = 8;
    N /= 8;
  }
  while (N % 4 == 0) {
    p.r[p.n++] = 4;
    N /= 4;
  }
  if (N % 2 == 0) {
    p.r[p.n++] = 2;
    N /= 2;
  }
  while (N % 3 == 0) {
    p.r[p.n++] = 3;
    N /= 3;
  }
  return p;
}
constexpr int fast_P(int N) {
  int p = 1;
  while (p * 2 <= 2048 / N && p * 2 <= 256) p *= 2;
  return p;
}

// element n of pencil j at lds[n * SN + (j % PJ) * SJ + (j / PJ) * PS]:
// pencils come in groups of PJ (one group per plane of a multi-plane tile)
template <typename T, int N, int P, int R, int Ns, int SN, int SJ, int PJ = P, int PS = 0>
__device__ __forceinline__ void fast_stage(vec2<T> *lds, const cx<T> *twl, int sign) {
  constexpr int nbf = N / R, total = nbf * P, tstep = N / (Ns * R);
  constexpr int items = (total + kFftThreads - 1) / kFftThreads;
  cx<T> v[items][R];
#pragma unroll
  for (int it = 0; it < items; it++) {
    const int w = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads == 0 || w < total) {
      const int j = w % P, b = w / P, k = b % Ns;
      const int jo = (j % PJ) * SJ + (j / PJ) * PS;
#pragma unroll
      for (int q = 0; q < R; q++) v[it][q] = ld<T>(lds + (b + q * nbf) * SN + jo);
      if constexpr (Ns > 1) {
#pragma unroll
        for (int q = 1; q < R; q++) {
          cx<T> t = twl[(q * k * tstep) % N];
          if (sign > 0) t.im = -t.im;
          v[it][q] = cmul(v[it][q], t);
        }
      }
      dft_small<T, R>(v[it], sign);
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < items; it++) {
    const int w = threadIdx.x + it * kFftThreads;
    if (total % kFftThreads == 0 || w < total) {
      const int j = w % P, b = w / P, k = b % Ns;
      const int jo = (j % PJ) * SJ + (j / PJ) * PS;
      const int out = (b / Ns) * Ns * R + k;
#pragma unroll
      for (int q = 0; q < R; q++) st<T>(lds + (out + q * Ns) * SN + jo, v[it][q]);
    }
  }
  __syncthreads();
}

template <typename T, int N, int P, int S_, int Ns, int SN = P + 1, int SJ = 1, int PJ = P, int PS = 0>
__device__ __forceinline__ void fast_run(vec2<T> *lds, const cx<T> *twl, int sign) {
  constexpr RadPlan pl = make_plan(N);
  if constexpr (S_ < pl.n) {
    constexpr int R = pl.r[S_];
    fast_stage<T, N, P, R, Ns, SN, SJ, PJ, PS>(lds, twl, sign);
    fast_run<T, N, P, S_ + 1, Ns * R, SN, SJ, PJ, PS>(lds, twl, sign);
  }
}

// both in-plane axes of NPL consecutive x-y planes in one pass: the planes
// (contiguous in memory) are staged as lds[x * (NPL NY + 1) + p NY + y]; the
// x transform runs on the NPL NY pencils (p, y) with stride NPL NY + 1
// between points, the y transform on the NPL NX pencils (p, x) with stride
// 1; the row length NPL NY + 1 is odd, so both access patterns are
// bank-conflict free.  With two 32 x 32 planes per workgroup every thread
// is busy in every stage (one plane leaves half of them idle in the
// radix-8 stages).
template <typename T, int NX, int NY, int NPL>
__global__ __launch_bounds__(kFftThreads) void k_fft_xy(vec2<T> *__restrict__ F, long long fstride,
                                                       const double2 *__restrict__ twx,
                                                       const double2 *__restrict__ twy, int mode,
                                                       const Ctrl *__restrict__ ctrl, long long slice_len,
                                                       const int *__restrict__ slice_done,
    const int *__restrict__ slice_map) {
  constexpr int LD = NPL * NY + 1, NE = NX * NY, NT = NPL * NE;
  if (ctrl && ctrl->done) return;
  // batch row blockIdx.z is time slice zs (the rows of running slices only)
  const int zs = slice_map ? slice_map[blockIdx.z] : (int)blockIdx.z;
  if (slice_done && slice_done[zs]) return;
  __shared__ vec2<T> lds[NX * LD];
  __shared__ cx<T> tx[NX], ty[NY];
  for (int k = threadIdx.x; k < NX; k += kFftThreads) tx[k] = mk<T>((T)twx[k].x, (T)twx[k].y);
  for (int k = threadIdx.x; k < NY; k += kFftThreads) ty[k] = mk<T>((T)twy[k].x, (T)twy[k].y);
  vec2<T> *base =
      F + (long long)blockIdx.y * fstride + (long long)zs * slice_len + (long long)blockIdx.x * NT;
  // element e of the tile: plane e / NE, row (e % NE) / NX, column e % NX
#pragma unroll 4
  for (int e = threadIdx.x; e < NT; e += kFftThreads) lds[(e % NX) * LD + e / NX] = base[e];
  __syncthreads();
  const int sign = mode == 1 ? +1 : -1;
  fast_run<T, NX, NPL * NY, 0, 1, LD, 1>(lds, tx, sign);
  fast_run<T, NY, NPL * NX, 0, 1, 1, LD, NX, NY>(lds, ty, sign);
#pragma unroll 4
  for (int e = threadIdx.x; e < NT; e += kFftThreads) base[e] = lds[(e % NX) * LD + e / NX];
}

template <typename T, int N>
static bool try_xy(void *F, long long fstride, int nfield, const FftAxis &ax, const FftAxis &ay, int mode,
                   const Ctrl *ctrl, hipStream_t s, long long nplane, int nslice, long long slice_len,
                   const int *slice_done, const int *slice_map) {
  if (ax.N != N || ay.N != N) return false;
  if (sizeof(T) * 2 * N * (N + 1) > 48 * 1024) return false;
  // two planes per tile in FP32 where a plane has at most 1024 points (the
  // tile's row e / NX runs over (plane, y)); in FP64 the doubled tile
  // halves the workgroups an LDS holds and gained nothing at 32 x 32
  constexpr int NPL = (sizeof(T) == 4 && N * N <= 1024) ? 2 : 1;
  if (NPL == 2 && nplane % 2 == 0)
    hipLaunchKernelGGL((k_fft_xy<T, N, N, NPL>), dim3((unsigned)(nplane / 2), nfield, nslice), dim3(kFftThreads),
                       0, s, (vec2<T> *)F, fstride, ax.tw, ay.tw, mode, ctrl, slice_len, slice_done, slice_map);
  else
    hipLaunchKernelGGL((k_fft_xy<T, N, N, 1>), dim3((unsigned)nplane, nfield, nslice), dim3(kFftThreads), 0, s,
                       (vec2<T> *)F, fstride, ax.tw, ay.tw, mode, ctrl, slice_len, slice_done, slice_map);
  return true;
}

template <typename T>
bool launch_fft_xy(void *F, long long fstride, int nfield, const FftAxis &ax, const FftAxis &ay, int mode,
                   const Ctrl *ctrl, hipStream_t s, long long nplane, int nslice, long long slice_len,
                   const int *slice_done, const int *slice_map) {
  return try_xy<T, 8>(F, fstride, nfield, ax, ay, mode, ctrl, s, nplane, nslice, slice_len, slice_done, slice_map) ||
         try_xy<T, 16>(F, fstride, nfield, ax, ay, mode, ctrl, s, nplane, nslice, slice_len, slice_done, slice_map) ||
         try_xy<T, 24>(F, fstride, nfield, ax, ay, mode, ctrl, s, nplane, nslice, slice_len, slice_done, slice_map) ||
         try_xy<T, 32>(F, fstride, nfield, ax, ay, mode, ctrl, s, nplane, nslice, slice_len, slice_done, slice_map) ||
         try_xy<T, 48>(F, fstride, nfield, ax, ay, mode, ctrl, s, nplane, nslice, slice_len, slice_done, slice_map);
}
template bool launch_fft_xy<double>(void *, long long, int, const FftAxis &, const FftAxis &, int, const Ctrl *,
                                    hipStream_t, long long, int, long long, const int *, const int *);
template bool launch_fft_xy<float>(void *, long long, int, const FftAxis &, const FftAxis &, int, const Ctrl *,
                                   hipStream_t, long long, int, long long, const int *, const int *);

// The z and t axes of the 4D transform in one LDS residency: forward z,
// forward t, x p^2, backward t, backward z, for PX consecutive x values of
// one y (a "column" of z-t planes), after the forward x-y plane pass and
// before the backward one.  Three passes over the fields per Fourier
// acceleration instead of five.  The pencils' arithmetic is that of the
// single-axis passes (same radix plan, twiddles and p^2 summation order).
// Element (j, z, t) of the column sits at lds[t RZ + z PX + j]; a t row is
// padded to RZ = NZ PX + 1 entries.  Workgroups are dealt to the 8 XCDs
// round-robin; logical block l = (b % 8) nb/8 + b / 8 gives each XCD a
// contiguous run of columns, so the x groups sharing a cache line are
// fetched into one L2.
template <typename T, int NZ, int NT, int PX>
__global__ __launch_bounds__(kFftThreads) void k_fft_zt(vec2<T> *__restrict__ F, long long V, int Lx, int Ly,
                                                       const double2 *__restrict__ twz,
                                                       const double2 *__restrict__ twt, PsqInfo ps,
                                                       const Ctrl *__restrict__ ctrl) {
  constexpr int RZ = NZ * PX + 1, NE = NZ * NT * PX;
  if (ctrl && ctrl->done) return;
  __shared__ vec2<T> lds[NT * RZ];
  __shared__ cx<T> tz[NZ], tt[NT];
  for (int k = threadIdx.x; k < NZ; k += kFftThreads) tz[k] = mk<T>((T)twz[k].x, (T)twz[k].y);
  for (int k = threadIdx.x; k < NT; k += kFftThreads) tt[k] = mk<T>((T)twt[k].x, (T)twt[k].y);
  // logical block -> (field, y, x group); gridDim.x is a multiple of 8
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned l = (b & 7u) * (nb >> 3) + (b >> 3);
  const int ngx = Lx / PX;
  const unsigned cols = (unsigned)(ngx * Ly);
  const int field = (int)(l / cols), col = (int)(l % cols);
  const int xg = col % ngx, y = col / ngx;
  const long long Sz = (long long)Lx * Ly, St = Sz * NZ;
  vec2<T> *base = F + (long long)field * V + (long long)xg * PX + (long long)y * Lx;
  // the column's loads are all issued before any is stored to LDS: with
  // runs of only PX elements per (z, t) row the pass is bound by the bytes
  // each thread keeps in flight
  if constexpr (NE % kFftThreads == 0) {
    constexpr int K = NE / kFftThreads;
    vec2<T> r[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = threadIdx.x + k * kFftThreads;
      const int j = e % PX, z = (e / PX) % NZ, t = e / (PX * NZ);
      r[k] = base[j + z * Sz + t * St];
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int e = threadIdx.x + k * kFftThreads;
      const int j = e % PX, z = (e / PX) % NZ, t = e / (PX * NZ);
      lds[t * RZ + z * PX + j] = r[k];
    }
  } else {
    for (int e = threadIdx.x; e < NE; e += kFftThreads) {
      const int j = e % PX, z = (e / PX) % NZ, t = e / (PX * NZ);
      lds[t * RZ + z * PX + j] = base[j + z * Sz + t * St];
    }
  }
  __syncthreads();
  // forward z: pencil (j, t) at j + t RZ, points PX apart
  fast_run<T, NZ, PX * NT, 0, 1, PX, 1, PX, RZ>(lds, tz, -1);
  // forward t: pencil (j, z) at z PX + j, points RZ apart
  fast_run<T, NT, PX * NZ, 0, 1, RZ, 1>(lds, tt, -1);
  // p^2 = 2 (4 - sum cos) with p^2(0) := 1 (Geometry/geometry.c:170-186),
  // the cos of x, y, z, t added in that order
  const double *cx_ = ps.cosv + ps.off[0], *cy_ = ps.cosv + ps.off[1];
  const double *cz_ = ps.cosv + ps.off[2], *ct_ = ps.cosv + ps.off[3];
#pragma unroll 4
  for (int e = threadIdx.x; e < NE; e += kFftThreads) {
    const int j = e % PX, z = (e / PX) % NZ, t = e / (PX * NZ);
    double kc = 0.0;
    kc += cx_[xg * PX + j];
    kc += cy_[y];
    kc += cz_[z];
    const double kcos = kc + ct_[t];
    const double p2 = (kcos == 4.0) ? 1.0 : 2.0 * (4.0 - kcos);
    const T f = (T)(ps.max / p2);
    const int a = t * RZ + z * PX + j;
    const cx<T> v = ld<T>(lds + a);
    st<T>(lds + a, mk<T>(v.re * f, v.im * f));
  }
  __syncthreads();
  fast_run<T, NT, PX * NZ, 0, 1, RZ, 1>(lds, tt, +1);
  fast_run<T, NZ, PX * NT, 0, 1, PX, 1, PX, RZ>(lds, tz, +1);
#pragma unroll 8
  for (int e = threadIdx.x; e < NE; e += kFftThreads) {
    const int j = e % PX, z = (e / PX) % NZ, t = e / (PX * NZ);
    base[j + z * Sz + t * St] = lds[t * RZ + z * PX + j];
  }
}

// x values (pencils) of a z-t column per workgroup: variant 0 the largest
// power of two whose tile fits 40 KiB of LDS (2 in FP32, 1 in FP64 at
// 32 x 64), variant 1 half of it, variant 2 twice it where the tile stays
// within 96 KiB (gfx950 gives a workgroup up to 160 KiB)
template <typename T, int NZ, int NT> constexpr int zt_px(int v) {
  constexpr int bytes = NZ * NT * (int)sizeof(vec2<T>);
  int p = 1;
  while (p * 2 * bytes <= 40 * 1024 && p < 8) p *= 2;
  if (v == 1) return p > 1 ? p / 2 : p;
  if (v == 2) return (2 * p * NZ + 1) * NT * (int)sizeof(vec2<T>) <= 98304 ? 2 * p : p;
  return p;
}

template <typename T, int NZ, int NT, int PX>
static bool zt_go(void *F, long long V, int nfield, const FftAxis &ax, const FftAxis &ay, const FftAxis &az,
                  const FftAxis &at, const PsqInfo &ps, const Ctrl *ctrl, hipStream_t s, bool launch) {
  if (ax.N % PX) return false;
  const long long nb = (long long)(ax.N / PX) * ay.N * nfield;
  if (nb % 8 || nb > (1ll << 30)) return false;
  if (!launch) return true;
  hipLaunchKernelGGL((k_fft_zt<T, NZ, NT, PX>), dim3((unsigned)nb), dim3(kFftThreads), 0, s, (vec2<T> *)F, V,
                     ax.N, ay.N, az.tw, at.tw, ps, ctrl);
  return true;
}

template <typename T, int NZ, int NT>
static bool try_zt(void *F, long long V, int nfield, const FftAxis &ax, const FftAxis &ay, const FftAxis &az,
                   const FftAxis &at, const PsqInfo &ps, const Ctrl *ctrl, hipStream_t s, bool launch, int v) {
  if (az.N != NZ || at.N != NT) return false;
  if (v == 1) return zt_go<T, NZ, NT, zt_px<T, NZ, NT>(1)>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch);
  if (v == 2) return zt_go<T, NZ, NT, zt_px<T, NZ, NT>(2)>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch);
  return zt_go<T, NZ, NT, zt_px<T, NZ, NT>(0)>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch);
}

template <typename T>
bool launch_fft_zt(void *F, long long V, int nfield, const FftAxis &ax, const FftAxis &ay, const FftAxis &az,
                   const FftAxis &at, const PsqInfo &ps, const Ctrl *ctrl, hipStream_t s, bool launch, int v) {
  if (ps.D != 4) return false;
  return try_zt<T, 8, 8>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch, v) ||
         try_zt<T, 16, 16>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch, v) ||
         try_zt<T, 16, 32>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch, v) ||
         try_zt<T, 24, 48>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch, v) ||
         try_zt<T, 32, 64>(F, V, nfield, ax, ay, az, at, ps, ctrl, s, launch, v);
}
template bool launch_fft_zt<double>(void *, long long, int, const FftAxis &, const FftAxis &, const FftAxis &,
                                    const FftAxis &, const PsqInfo &, const Ctrl *, hipStream_t, bool, int);
template bool launch_fft_zt<float>(void *, long long, int, const FftAxis &, const FftAxis &, const FftAxis &,
                                   const FftAxis &, const PsqInfo &, const Ctrl *, hipStream_t, bool, int);

template <typename T, int N, int P>
__global__ __launch_bounds__(kFftThreads) void k_fft_fast(vec2<T> *__restrict__ F, long long fstride, long long S,
                                                         long long npencil, const double2 *__restrict__ tw, int mode,
                                                         PsqInfo ps, const Ctrl *__restrict__ ctrl,
                                                         long long slice_len, const int *__restrict__ slice_done,
    const int *__restrict__ slice_map) {
  constexpr int LD = P + 1, TOT = N * P;
  static_assert(LD > 0, "");
  static_assert(kFftThreads % P == 0, "a pencil per thread needs P | threads");
  if (ctrl && ctrl->done) return;
  // batch row blockIdx.z is time slice zs (the rows of running slices only)
  const int zs = slice_map ? slice_map[blockIdx.z] : (int)blockIdx.z;
  if (slice_done && slice_done[zs]) return;
  __shared__ vec2<T> lds[N * LD];
  __shared__ cx<T> twl[N];
  for (int k = threadIdx.x; k < N; k += kFftThreads) twl[k] = mk<T>((T)tw[k].x, (T)tw[k].y);
  vec2<T> *base = F + (long long)blockIdx.y * fstride + (long long)zs * slice_len;
  const long long p0 = (long long)blockIdx.x * P;
  // this thread's pencil and first element (strided axes)
  const int jt = threadIdx.x % P, n0 = threadIdx.x / P;
  constexpr int nstep = kFftThreads / P;
  const long long pt = p0 + jt;
  const bool valid = pt < npencil;
  long long a0 = 0;
  if (S != 1 && valid) a0 = (pt % S) + (pt / S) * S * N;
  if (S == 1) {
    const long long lim = (npencil - p0) * N;  // elements left in this chunk
    const vec2<T> *src = base + p0 * N;
    for (int idx = threadIdx.x; idx < TOT; idx += kFftThreads)
      if (idx < lim) lds[(idx % N) * LD + idx / N] = src[idx];
  } else if (valid) {
#pragma unroll 4
    for (int n = n0; n < N; n += nstep) lds[n * LD + jt] = base[a0 + (long long)n * S];
  }
  __syncthreads();
  if (mode == 1)
    fast_run<T, N, P, 0, 1>(lds, twl, +1);
  else
    fast_run<T, N, P, 0, 1>(lds, twl, -1);
  if (mode == 2) {
    // p^2 = 2 (D - sum cos) with p^2(0) := 1 (Geometry/geometry.c:170-186);
    // the fused axis is the last of the D, so its cos enters the sum last
    double kc = 0.0;
    long long r = a0;
    for (int mu = 0; mu < ps.D - 1; mu++) {
      kc += ps.cosv[ps.off[mu] + (int)(r % ps.L[mu])];
      r /= ps.L[mu];
    }
    const double *cl = ps.cosv + ps.off[ps.D - 1];
    if (valid && S != 1) {
#pragma unroll 4
      for (int n = n0; n < N; n += nstep) {
        const double kcos = kc + cl[n];
        const double p2 = (kcos == (double)ps.D) ? 1.0 : 2.0 * ((double)ps.D - kcos);
        const T f = (T)(ps.max / p2);
        const cx<T> z = ld<T>(lds + n * LD + jt);
        st<T>(lds + n * LD + jt, mk<T>(z.re * f, z.im * f));
      }
    }
    __syncthreads();
    fast_run<T, N, P, 0, 1>(lds, twl, +1);
  }
  if (S == 1) {
    const long long lim = (npencil - p0) * N;
    vec2<T> *dst = base + p0 * N;
    for (int idx = threadIdx.x; idx < TOT; idx += kFftThreads)
      if (idx < lim) dst[idx] = lds[(idx % N) * LD + idx / N];
  } else if (valid) {
#pragma unroll 4
    for (int n = n0; n < N; n += nstep) base[a0 + (long long)n * S] = lds[n * LD + jt];
  }
}

// pencils per workgroup of variant v: the default fast_P(N), half of it,
// or twice it where the LDS tile stays within 64 KB
template <typename T, int N> constexpr int variant_P(int v) {
  constexpr int P0 = fast_P(N);
  if (v == 1) return P0 > 1 ? P0 / 2 : P0;
  if (v == 2) return (2 * P0 <= kFftThreads && N * (2 * P0 + 1) * (int)sizeof(vec2<T>) <= 65536) ? 2 * P0 : P0;
  return P0;
}

template <typename T, int N, int P>
static void run_fast(void *F, long long fstride, int nfield, const FftAxis &ax, int mode, const PsqInfo &ps,
                     const Ctrl *ctrl, hipStream_t s, int nslice, long long slice_len, const int *slice_done, const int *slice_map) {
  const unsigned nb = (unsigned)((ax.npencil + P - 1) / P);
  hipLaunchKernelGGL((k_fft_fast<T, N, P>), dim3(nb, nfield, nslice), dim3(kFftThreads), 0, s, (vec2<T> *)F,
                     fstride, ax.S, ax.npencil, ax.tw, mode, ps, ctrl, slice_len, slice_done, slice_map);
}

template <typename T, int N>
static bool try_fast(void *F, long long fstride, int nfield, const FftAxis &ax, int mode, const PsqInfo &ps,
                     const Ctrl *ctrl, hipStream_t s, int nslice, long long slice_len, const int *slice_done, const int *slice_map) {
  if (ax.N != N) return false;
  // the fused p^2 pass runs on a strided axis (it is always the last one)
  if (mode == 2 && ax.S == 1) return false;
  if (ax.pv == 1)
    run_fast<T, N, variant_P<T, N>(1)>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map);
  else if (ax.pv == 2)
    run_fast<T, N, variant_P<T, N>(2)>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map);
  else
    run_fast<T, N, variant_P<T, N>(0)>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map);
  return true;
}

template <typename T>
void launch_fft_axis(void *F, long long fstride, int nfield, const FftAxis &ax, int mode, const PsqInfo &ps,
                     const Ctrl *ctrl, hipStream_t s, int nslice, long long slice_len, const int *slice_done, const int *slice_map) {
  if (try_fast<T, 4>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 6>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 8>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 12>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 16>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 24>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 32>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 48>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 64>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map) ||
      try_fast<T, 96>(F, fstride, nfield, ax, mode, ps, ctrl, s, nslice, slice_len, slice_done, slice_map))
    return;
  const unsigned nb = (unsigned)((ax.npencil + ax.P - 1) / ax.P);
  const size_t sm = (size_t)ax.N * (ax.P + 1) * sizeof(vec2<T>);
  hipLaunchKernelGGL(k_fft_axis<T>, dim3(nb, nfield, nslice), dim3(kFftThreads), sm, s, (vec2<T> *)F, fstride, ax,
                     mode, ps, ctrl, slice_len, slice_done, slice_map);
}
template void launch_fft_axis<double>(void *, long long, int, const FftAxis &, int, const PsqInfo &, const Ctrl *,
                                      hipStream_t, int, long long, const int *, const int *);
template void launch_fft_axis<float>(void *, long long, int, const FftAxis &, int, const PsqInfo &, const Ctrl *,
                                     hipStream_t, int, long long, const int *, const int *);

}  // namespace gfix
