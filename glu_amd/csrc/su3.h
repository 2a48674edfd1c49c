// su3.h -- register-resident SU(3) algebra for the gauge-fixing kernels.
//
// A link on the device is stored as its first two COLUMNS (6 complex
// numbers); the third column is rebuilt as conj(col0 x col1).  This is the
// same completion GLU applies after every gauge transformation
// (Gfix/gtrans.c:122-136) and every reunitarisation
// (Matrix_Ops/gramschmidt.c:73-165), so nothing is lost for SU(3) links and
// a third of the link traffic disappears.
#pragma once

#include <hip/hip_runtime.h>

namespace gfix {

template <typename T> struct vec2_t;
template <> struct vec2_t<double> { using type = double2; };
template <> struct vec2_t<float> { using type = float2; };
template <typename T> using vec2 = typename vec2_t<T>::type;

template <typename T> struct cx {
  T re, im;
};

template <typename T> __device__ __forceinline__ cx<T> mk(T r, T i) { return {r, i}; }
template <typename T> __device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T> __device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T> __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// a * conj(b)
template <typename T> __device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> b) {
  return {a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
template <typename T> __device__ __forceinline__ cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
template <typename T> __device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T> __device__ __forceinline__ T cnorm2(cx<T> a) { return a.re * a.re + a.im * a.im; }

template <typename T> __device__ __forceinline__ cx<T> ld(const vec2<T> *p) {
  const vec2<T> v = *p;
  return {v.x, v.y};
}
template <typename T> __device__ __forceinline__ void st(vec2<T> *p, cx<T> a) {
  vec2<T> v;
  v.x = a.re;
  v.y = a.im;
  *p = v;
}

// 3x3 complex matrix, row-major e[a*3+b] = M_ab
template <typename T> struct su3 {
  cx<T> e[9];
};

// column 2 from columns 0 and 1: conj(col0 x col1)
template <typename T> __device__ __forceinline__ void complete_col2(su3<T> &m) {
  m.e[2] = cconj(csub(cmul(m.e[3], m.e[7]), cmul(m.e[6], m.e[4])));
  m.e[5] = cconj(csub(cmul(m.e[6], m.e[1]), cmul(m.e[0], m.e[7])));
  m.e[8] = cconj(csub(cmul(m.e[0], m.e[4]), cmul(m.e[3], m.e[1])));
}

// Device link layout: for direction mu the six column entries
// (U00, U10, U20, U01, U11, U21) are six planes of V complex numbers:
// entry k of site i at U[(mu*6 + k)*V + i].
template <typename T>
__device__ __forceinline__ su3<T> load_link(const vec2<T> *__restrict__ U, int mu, size_t i, size_t V) {
  const vec2<T> *b = U + (size_t)mu * 6 * V + i;
  su3<T> m;
  m.e[0] = ld<T>(b);
  m.e[3] = ld<T>(b + V);
  m.e[6] = ld<T>(b + 2 * V);
  m.e[1] = ld<T>(b + 3 * V);
  m.e[4] = ld<T>(b + 4 * V);
  m.e[7] = ld<T>(b + 5 * V);
  complete_col2(m);
  return m;
}

template <typename T>
__device__ __forceinline__ void store_link(vec2<T> *__restrict__ U, int mu, size_t i, size_t V, const su3<T> &m) {
  vec2<T> *b = U + (size_t)mu * 6 * V + i;
  st<T>(b, m.e[0]);
  st<T>(b + V, m.e[3]);
  st<T>(b + 2 * V, m.e[6]);
  st<T>(b + 3 * V, m.e[1]);
  st<T>(b + 4 * V, m.e[4]);
  st<T>(b + 5 * V, m.e[7]);
}

// 1/sqrt(x): the hardware estimate refined by Newton steps (two in FP64,
// one in FP32) to within an ulp of 1/sqrt(x), without the division
__device__ __forceinline__ double rsqrt_nr(double x) {
  double r = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  r = r * (1.5 - h * r * r);
  r = r * (1.5 - h * r * r);
  return r;
}
__device__ __forceinline__ float rsqrt_nr(float x) {
  float r = __builtin_amdgcn_rsqf(x);
  r = r * (1.5f - 0.5f * x * r * r);
  return r;
}

// Gram-Schmidt on columns 0 and 1 (column 2 untouched)
template <typename T> __device__ __forceinline__ void gram_cols01(su3<T> &m) {
  T n = cnorm2(m.e[0]) + cnorm2(m.e[3]) + cnorm2(m.e[6]);
  n = rsqrt_nr(n);
  m.e[0] = cscale(m.e[0], n);
  m.e[3] = cscale(m.e[3], n);
  m.e[6] = cscale(m.e[6], n);
  // p = sum_k conj(col0_k) col1_k
  cx<T> p = cadd(cadd(cmulc(m.e[1], m.e[0]), cmulc(m.e[4], m.e[3])), cmulc(m.e[7], m.e[6]));
  m.e[1] = csub(m.e[1], cmul(p, m.e[0]));
  m.e[4] = csub(m.e[4], cmul(p, m.e[3]));
  m.e[7] = csub(m.e[7], cmul(p, m.e[6]));
  T q = cnorm2(m.e[1]) + cnorm2(m.e[4]) + cnorm2(m.e[7]);
  q = rsqrt_nr(q);
  m.e[1] = cscale(m.e[1], q);
  m.e[4] = cscale(m.e[4], q);
  m.e[7] = cscale(m.e[7], q);
}

// Gram-Schmidt on columns 0 and 1, column 2 completed:
// Matrix_Ops/gramschmidt.c:73-165
template <typename T> __device__ __forceinline__ void gram_reunit(su3<T> &m) {
  gram_cols01(m);
  complete_col2(m);
}

// g = P_SU(3)(1 + alpha X) from the packed field values f0..f3 at one site:
// the exp_approx branch of set_gauge_matrix, Gfix/CG.c:47-100
template <typename T>
__device__ __forceinline__ su3<T> gauge_matrix(cx<T> f0, cx<T> f1, cx<T> f2, cx<T> f3, T alpha) {
  su3<T> g;
  g.e[0] = mk<T>(T(1), alpha * f0.im);
  g.e[1] = cscale(f1, alpha);
  g.e[3] = mk<T>(-g.e[1].re, g.e[1].im);
  g.e[4] = mk<T>(T(1), -alpha * f0.re);
  g.e[6] = mk<T>(-alpha * f2.re, alpha * f2.im);
  g.e[7] = mk<T>(-alpha * f3.re, alpha * f3.im);
  gram_reunit(g);
  return g;
}

// columns 0 and 1 of gauge_matrix (the two the third is completed from)
template <typename T>
__device__ __forceinline__ su3<T> gauge_cols01(cx<T> f0, cx<T> f1, cx<T> f2, cx<T> f3, T alpha) {
  su3<T> g;
  g.e[0] = mk<T>(T(1), alpha * f0.im);
  g.e[1] = cscale(f1, alpha);
  g.e[3] = mk<T>(-g.e[1].re, g.e[1].im);
  g.e[4] = mk<T>(T(1), -alpha * f0.re);
  g.e[6] = mk<T>(-alpha * f2.re, alpha * f2.im);
  g.e[7] = mk<T>(-alpha * f3.re, alpha * f3.im);
  gram_cols01(g);
  return g;
}

// b <- a b c^dagger with columns 0,1 of the product and column 2 completed:
// Gfix/gtrans.c:32-141
template <typename T>
__device__ __forceinline__ su3<T> gtransform(const su3<T> &a, const su3<T> &b, const su3<T> &c) {
  cx<T> t[6];  // (b c^dagger) columns 0 and 1, rows 0..2: t[r*2+s]
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 2; s++)
      t[r * 2 + s] = cadd(cadd(cmulc(b.e[r * 3], c.e[s * 3]), cmulc(b.e[r * 3 + 1], c.e[s * 3 + 1])),
                          cmulc(b.e[r * 3 + 2], c.e[s * 3 + 2]));
  su3<T> o;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 2; s++)
      o.e[r * 3 + s] = cadd(cadd(cmul(a.e[r * 3], t[s]), cmul(a.e[r * 3 + 1], t[2 + s])),
                            cmul(a.e[r * 3 + 2], t[4 + s]));
  complete_col2(o);
  return o;
}

// Re tr(a b c^dagger) with rows 0,1 of a.b and row 2 from the cofactors:
// Matrix_Ops/trace_abc_SSE.c:34-120
template <typename T>
__device__ __forceinline__ T retrace_abc_dag(const su3<T> &a, const su3<T> &b, const su3<T> &c) {
  cx<T> ab[9];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < 3; s++)
      ab[r * 3 + s] = cadd(cadd(cmul(a.e[r * 3], b.e[s]), cmul(a.e[r * 3 + 1], b.e[3 + s])),
                           cmul(a.e[r * 3 + 2], b.e[6 + s]));
  ab[6] = cconj(csub(cmul(ab[1], ab[5]), cmul(ab[2], ab[4])));
  ab[7] = cconj(csub(cmul(ab[2], ab[3]), cmul(ab[0], ab[5])));
  ab[8] = cconj(csub(cmul(ab[0], ab[4]), cmul(ab[1], ab[3])));
  T re = T(0), im = T(0);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    re += ab[k].re * c.e[k].re;
    im += ab[k].im * c.e[k].im;
  }
  return re + im;
}

// retrace_abc_dag in two steps, with the same arithmetic: rows 0, 1 of
// a b, then Re tr((a b) c^dagger) with row 2 of a b from the cofactors
template <typename T> __device__ __forceinline__ void ab_rows01(const su3<T> &a, const su3<T> &b, cx<T> (&ab)[6]) {
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < 3; s++)
      ab[r * 3 + s] = cadd(cadd(cmul(a.e[r * 3], b.e[s]), cmul(a.e[r * 3 + 1], b.e[3 + s])),
                           cmul(a.e[r * 3 + 2], b.e[6 + s]));
}
template <typename T> __device__ __forceinline__ T retrace_rows01_dag(const cx<T> (&ab01)[6], const su3<T> &c) {
  cx<T> ab[9];
#pragma unroll
  for (int k = 0; k < 6; k++) ab[k] = ab01[k];
  ab[6] = cconj(csub(cmul(ab[1], ab[5]), cmul(ab[2], ab[4])));
  ab[7] = cconj(csub(cmul(ab[2], ab[3]), cmul(ab[0], ab[5])));
  ab[8] = cconj(csub(cmul(ab[0], ab[4]), cmul(ab[1], ab[3])));
  T re = T(0), im = T(0);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    re += ab[k].re * c.e[k].re;
    im += ab[k].im * c.e[k].im;
  }
  return re + im;
}

// product a b with columns 0,1 computed and column 2 completed: the same
// completion as the storage format, so repeated products stay consistent
template <typename T> __device__ __forceinline__ su3<T> mul_cols(const su3<T> &a, const su3<T> &b) {
  su3<T> o;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int s = 0; s < 2; s++)
      o.e[r * 3 + s] = cadd(cadd(cmul(a.e[r * 3], b.e[s]), cmul(a.e[r * 3 + 1], b.e[3 + s])),
                            cmul(a.e[r * 3 + 2], b.e[6 + s]));
  complete_col2(o);
  return o;
}

// general product a b (rows 0,1; row 2 completed): MMUL/MMUL_SUNC_SSE.c
template <typename T> __device__ __forceinline__ su3<T> mul_su3(const su3<T> &b, const su3<T> &c) {
  su3<T> a;
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int s = 0; s < 3; s++)
      a.e[r * 3 + s] = cadd(cadd(cmul(b.e[r * 3], c.e[s]), cmul(b.e[r * 3 + 1], c.e[3 + s])),
                            cmul(b.e[r * 3 + 2], c.e[6 + s]));
  a.e[6] = cconj(csub(cmul(a.e[1], a.e[5]), cmul(a.e[2], a.e[4])));
  a.e[7] = cconj(csub(cmul(a.e[2], a.e[3]), cmul(a.e[0], a.e[5])));
  a.e[8] = cconj(csub(cmul(a.e[0], a.e[4]), cmul(a.e[1], a.e[3])));
  return a;
}

template <typename T> __device__ __forceinline__ T retrace(const su3<T> &m) {
  return m.e[0].re + m.e[4].re + m.e[8].re;
}

}  // namespace gfix
