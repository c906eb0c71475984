// Shared device/host helpers for the 6D Fourier mass-operator library.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace fame6d {

// Interleaved complex number; layout-compatible with C99 `double complex` /
// `float complex`, numpy complex128/complex64 and torch complex tensors.
template <class T>
struct alignas(2 * sizeof(T)) cpx {
  T re, im;
};

template <class T>
__host__ __device__ __forceinline__ cpx<T> mk(T re, T im) {
  cpx<T> c;
  c.re = re;
  c.im = im;
  return c;
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) {
  return mk(a.re + b.re, a.im + b.im);
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) {
  return mk(a.re - b.re, a.im - b.im);
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a) {
  return mk(-a.re, -a.im);
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) {
  return mk(s * a.re, s * a.im);
}
// conj(a) * b
template <class T>
__host__ __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
template <class T>
__host__ __device__ __forceinline__ cpx<T> conj(cpx<T> a) {
  return mk(a.re, -a.im);
}
template <class T>
__host__ __device__ __forceinline__ T norm2(cpx<T> a) {
  return a.re * a.re + a.im * a.im;
}
// multiply by sigma*i (sigma = +1 or -1)
template <int SIGMA, class T>
__host__ __device__ __forceinline__ cpx<T> mul_i(cpx<T> a) {
  return SIGMA > 0 ? mk(-a.im, a.re) : mk(a.im, -a.re);
}

template <class To, class Ti>
__host__ __device__ __forceinline__ cpx<To> cvt(cpx<Ti> a) {
  return mk(static_cast<To>(a.re), static_cast<To>(a.im));
}

}  // namespace fame6d
