// Register-resident small DFTs for the 6D grid kernels.
//
// dft<T, N, SIGMA>(v) overwrites v[0..N) with X_k = sum_j v_j exp(SIGMA*2*pi*i*j*k/N).
// SIGMA = -1 is MATLAB's fft (unnormalised), SIGMA = +1 the unnormalised inverse;
// the 1/NF of ifftn is folded into the permittivity samples by the callers.
// Prime lengths 2, 3, 5 and length 4 are written out from the DFT definition;
// composite lengths 6, 8, 10, 12 use one Cooley-Tukey split N = R*M with
// compile-time twiddles cos/sin(2*pi*m/120) (120 = lcm of all lengths).
#pragma once

#include "common.hpp"
#include "twiddles.hpp"

namespace fame6d {

template <int N>
struct DftTag {};

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<1>) {}

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<2>) {
  cpx<T> a = v[0], b = v[1];
  v[0] = a + b;
  v[1] = a - b;
}

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<3>) {
  // X1 = v0 - (v1+v2)/2 + SIGMA*i*(sqrt(3)/2)*(v1-v2); X2 = conjugate partner.
  const T s3 = static_cast<T>(kSin120[40]);  // sin(2*pi/3)
  cpx<T> t1 = v[1] + v[2];
  cpx<T> t2 = v[0] - static_cast<T>(0.5) * t1;
  cpx<T> d = v[1] - v[2];
  cpx<T> t3 = mul_i<SIGMA>(s3 * d);
  v[0] = v[0] + t1;
  v[1] = t2 + t3;
  v[2] = t2 - t3;
}

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<4>) {
  cpx<T> a = v[0] + v[2], b = v[0] - v[2];
  cpx<T> c = v[1] + v[3], d = mul_i<SIGMA>(v[1] - v[3]);
  v[0] = a + c;
  v[2] = a - c;
  v[1] = b + d;
  v[3] = b - d;
}

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<5>) {
  // With c1=cos(2pi/5), c2=cos(4pi/5), s1=sin(2pi/5), s2=sin(4pi/5):
  //   X1,X4 = v0 + c1 a1 + c2 a2 +- SIGMA*i*(s1 b1 + s2 b2)
  //   X2,X3 = v0 + c2 a1 + c1 a2 +- SIGMA*i*(s2 b1 - s1 b2)
  // where a1=v1+v4, a2=v2+v3, b1=v1-v4, b2=v2-v3.
  const T c1 = static_cast<T>(kCos120[24]), c2 = static_cast<T>(kCos120[48]);
  const T s1 = static_cast<T>(kSin120[24]), s2 = static_cast<T>(kSin120[48]);
  cpx<T> a1 = v[1] + v[4], a2 = v[2] + v[3];
  cpx<T> b1 = v[1] - v[4], b2 = v[2] - v[3];
  cpx<T> v0 = v[0];
  cpx<T> p1 = v0 + c1 * a1 + c2 * a2;
  cpx<T> p2 = v0 + c2 * a1 + c1 * a2;
  cpx<T> q1 = mul_i<SIGMA>(s1 * b1 + s2 * b2);
  cpx<T> q2 = mul_i<SIGMA>(s2 * b1 - s1 * b2);
  v[0] = v0 + a1 + a2;
  v[1] = p1 + q1;
  v[4] = p1 - q1;
  v[2] = p2 + q2;
  v[3] = p2 - q2;
}

template <class T, int N, int SIGMA>
__device__ __forceinline__ void dft(cpx<T>* v);

// One Cooley-Tukey step: j = M*j1 + j2, k = k1 + R*k2.
template <class T, int N, int R, int SIGMA>
__device__ __forceinline__ void dft_ct(cpx<T>* v) {
  constexpr int M = N / R;
  cpx<T> a[R][M];
#pragma unroll
  for (int j2 = 0; j2 < M; ++j2) {
    cpx<T> t[R];
#pragma unroll
    for (int j1 = 0; j1 < R; ++j1) t[j1] = v[M * j1 + j2];
    dft<T, R, SIGMA>(t);
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) {
      constexpr int step = 120 / N;
      const int m = (j2 * k1 * step) % 120;
      if (m == 0) {
        a[k1][j2] = t[k1];
      } else {
        const T c = static_cast<T>(kCos120[m]);
        const T s = static_cast<T>(SIGMA) * static_cast<T>(kSin120[m]);
        a[k1][j2] = t[k1] * mk(c, s);
      }
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) {
    dft<T, M, SIGMA>(a[k1]);
#pragma unroll
    for (int k2 = 0; k2 < M; ++k2) v[k1 + R * k2] = a[k1][k2];
  }
}

template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<6>) { dft_ct<T, 6, 2, SIGMA>(v); }
template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<8>) { dft_ct<T, 8, 2, SIGMA>(v); }
template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<10>) { dft_ct<T, 10, 2, SIGMA>(v); }
template <class T, int SIGMA>
__device__ __forceinline__ void dft_impl(cpx<T>* v, DftTag<12>) { dft_ct<T, 12, 4, SIGMA>(v); }

template <class T, int N, int SIGMA>
__device__ __forceinline__ void dft(cpx<T>* v) {
  dft_impl<T, SIGMA>(v, DftTag<N>{});
}

}  // namespace fame6d
