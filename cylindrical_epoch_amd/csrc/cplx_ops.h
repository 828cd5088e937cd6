// Complex helpers in the operand order of the reference's Fortran
// expressions.  Used by every kernel; with -ffp-contract=off the FP64 path
// evaluates exactly these operations.
#pragma once

#include "common.h"

namespace cylpic {

template <class T>
__device__ __forceinline__ cplx<T> mk(T re, T im) {
  cplx<T> z;
  z.re = re;
  z.im = im;
  return z;
}
template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk(a.re + b.re, a.im + b.im); }
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return mk(a.re - b.re, a.im - b.im); }
template <class T>
__device__ __forceinline__ cplx<T> cneg(cplx<T> a) { return mk(-a.re, -a.im); }
// real * complex
template <class T>
__device__ __forceinline__ cplx<T> rmul(T s, cplx<T> a) { return mk(s * a.re, s * a.im); }
// complex * real
template <class T>
__device__ __forceinline__ cplx<T> cmulr(cplx<T> a, T s) { return mk(a.re * s, a.im * s); }
template <class T>
__device__ __forceinline__ cplx<T> cdivr(cplx<T> a, T s) { return mk(a.re / s, a.im / s); }
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// (0, s) * a
template <class T>
__device__ __forceinline__ cplx<T> imul(T s, cplx<T> a) { return mk(-(s * a.im), s * a.re); }
template <class T>
__device__ __forceinline__ T re_mul(cplx<T> a, cplx<T> b) { return a.re * b.re - a.im * b.im; }

template <class D, class S>
__device__ __forceinline__ cplx<D> conv(cplx<S> a) { return mk((D)a.re, (D)a.im); }

template <class T>
__device__ __forceinline__ cplx<T> czero() { return mk((T)0, (T)0); }

}  // namespace cylpic
