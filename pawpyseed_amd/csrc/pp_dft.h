// Register-level backward DFTs (e^{+2 pi i n k / R}) of the mixed-radix
// Stockham FFT passes: radices 2, 3, 4, 5 directly, and the composite
// radices 6 = 2x3, 8 = 2x4, 9 = 3x3, 12 = 4x3, 16 = 4x4 as two nested
// small DFTs with constant twiddles (one LDS round trip per composite
// stage instead of two).  Host and device code, so the arithmetic can be
// checked on the CPU.
#pragma once
#include "pp_common.h"

namespace pp {

template <typename T> struct Consts;
template <> struct Consts<double> {
  static constexpr double s3 = 0.86602540378443864676;
  static constexpr double c51 = 0.30901699437494742410, c52 = -0.80901699437494742410;
  static constexpr double s51 = 0.95105651629515357212, s52 = 0.58778525229247312917;
};
template <> struct Consts<float> {
  static constexpr float s3 = 0.86602540378443864676f;
  static constexpr float c51 = 0.30901699437494742410f, c52 = -0.80901699437494742410f;
  static constexpr float s51 = 0.95105651629515357212f, s52 = 0.58778525229247312917f;
};

// Roots e^{+2 pi i m / R}, m < R, of the composite radices (cos, sin).
static constexpr double kRoot6[6][2] = {{1.0, 0.0}, {0.5, 0.8660254037844386467637}, {-0.5, 0.8660254037844386467637}, {-1.0, 0.0}, {-0.5, -0.8660254037844386467637}, {0.5, -0.8660254037844386467637}};
static constexpr double kRoot8[8][2] = {{1.0, 0.0}, {0.7071067811865475244008, 0.7071067811865475244008}, {0.0, 1.0}, {-0.7071067811865475244008, 0.7071067811865475244008}, {-1.0, 0.0}, {-0.7071067811865475244008, -0.7071067811865475244008}, {0.0, -1.0}, {0.7071067811865475244008, -0.7071067811865475244008}};
static constexpr double kRoot9[9][2] = {{1.0, 0.0}, {0.7660444431189780352024, 0.6427876096865393263226}, {0.1736481776669303488517, 0.9848077530122080593667}, {-0.5, 0.8660254037844386467637}, {-0.9396926207859083840541, 0.3420201433256687330441}, {-0.9396926207859083840541, -0.3420201433256687330441}, {-0.5, -0.8660254037844386467637}, {0.1736481776669303488517, -0.9848077530122080593667}, {0.7660444431189780352024, -0.6427876096865393263226}};
static constexpr double kRoot12[12][2] = {{1.0, 0.0}, {0.8660254037844386467637, 0.5}, {0.5, 0.8660254037844386467637}, {0.0, 1.0}, {-0.5, 0.8660254037844386467637}, {-0.8660254037844386467637, 0.5}, {-1.0, 0.0}, {-0.8660254037844386467637, -0.5}, {-0.5, -0.8660254037844386467637}, {0.0, -1.0}, {0.5, -0.8660254037844386467637}, {0.8660254037844386467637, -0.5}};
static constexpr double kRoot16[16][2] = {{1.0, 0.0}, {0.9238795325112867561282, 0.3826834323650897717285}, {0.7071067811865475244008, 0.7071067811865475244008}, {0.3826834323650897717285, 0.9238795325112867561282}, {0.0, 1.0}, {-0.3826834323650897717285, 0.9238795325112867561282}, {-0.7071067811865475244008, 0.7071067811865475244008}, {-0.9238795325112867561282, 0.3826834323650897717285}, {-1.0, 0.0}, {-0.9238795325112867561282, -0.3826834323650897717285}, {-0.7071067811865475244008, -0.7071067811865475244008}, {-0.3826834323650897717285, -0.9238795325112867561282}, {0.0, -1.0}, {0.3826834323650897717285, -0.9238795325112867561282}, {0.7071067811865475244008, -0.7071067811865475244008}, {0.9238795325112867561282, -0.3826834323650897717285}};

template <int R> __host__ __device__ constexpr double root_cos(int m);
template <int R> __host__ __device__ constexpr double root_sin(int m);
#define PP_ROOT(R)                                                                        \
  template <> __host__ __device__ constexpr double root_cos<R>(int m) { return kRoot##R[m % R][0]; } \
  template <> __host__ __device__ constexpr double root_sin<R>(int m) { return kRoot##R[m % R][1]; }
PP_ROOT(6)
PP_ROOT(8)
PP_ROOT(9)
PP_ROOT(12)
PP_ROOT(16)
#undef PP_ROOT

template <typename T, int R> struct Dft;

template <typename T> struct Dft<T, 2> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) {
    const cplx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};
template <typename T> struct Dft<T, 3> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) {
    const cplx<T> s = v[1] + v[2], d = v[1] - v[2];
    const cplx<T> m = v[0] - scale(s, (T)0.5);
    const cplx<T> t = mul_i(scale(d, Consts<T>::s3));
    v[0] = v[0] + s;
    v[1] = m + t;
    v[2] = m - t;
  }
};
template <typename T> struct Dft<T, 4> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) {
    const cplx<T> a0 = v[0] + v[2], a1 = v[0] - v[2], b0 = v[1] + v[3], b1 = mul_i(v[1] - v[3]);
    v[0] = a0 + b0;
    v[2] = a0 - b0;
    v[1] = a1 + b1;
    v[3] = a1 - b1;
  }
};
template <typename T> struct Dft<T, 5> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) {
    const cplx<T> a1 = v[1] + v[4], b1 = v[1] - v[4], a2 = v[2] + v[3], b2 = v[2] - v[3];
    const cplx<T> m1 = v[0] + scale(a1, Consts<T>::c51) + scale(a2, Consts<T>::c52);
    const cplx<T> m2 = v[0] + scale(a1, Consts<T>::c52) + scale(a2, Consts<T>::c51);
    const cplx<T> t1 = mul_i(scale(b1, Consts<T>::s51) + scale(b2, Consts<T>::s52));
    const cplx<T> t2 = mul_i(scale(b1, Consts<T>::s52) - scale(b2, Consts<T>::s51));
    v[0] = v[0] + a1 + a2;
    v[1] = m1 + t1;
    v[4] = m1 - t1;
    v[2] = m2 + t2;
    v[3] = m2 - t2;
  }
};

// R = R1 R2, input index n = R2 n1 + n2, output index k = k1 + R1 k2:
// X[k1 + R1 k2] = sum_n2 W_R2^(n2 k2) [W_R^(n2 k1) sum_n1 W_R1^(n1 k1) x[R2 n1 + n2]].
template <typename T, int R1, int R2>
__host__ __device__ __forceinline__ void dft_composite(cplx<T> *v) {
  constexpr int R = R1 * R2;
  cplx<T> y[R];
#pragma unroll
  for (int n2 = 0; n2 < R2; n2++) {
    cplx<T> t[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; n1++)
      t[n1] = v[R2 * n1 + n2];
    Dft<T, R1>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < R1; k1++) {
      const int m = (n2 * k1) % R;
      y[k1 * R2 + n2] = m == 0 ? t[k1] : t[k1] * mk<T>((T)root_cos<R>(m), (T)root_sin<R>(m));
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < R1; k1++) {
    cplx<T> u[R2];
#pragma unroll
    for (int n2 = 0; n2 < R2; n2++)
      u[n2] = y[k1 * R2 + n2];
    Dft<T, R2>::run(u);
#pragma unroll
    for (int k2 = 0; k2 < R2; k2++)
      v[k1 + R1 * k2] = u[k2];
  }
}

template <typename T> struct Dft<T, 6> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) { dft_composite<T, 2, 3>(v); }
};
template <typename T> struct Dft<T, 8> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) { dft_composite<T, 2, 4>(v); }
};
template <typename T> struct Dft<T, 9> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) { dft_composite<T, 3, 3>(v); }
};
template <typename T> struct Dft<T, 12> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) { dft_composite<T, 4, 3>(v); }
};
template <typename T> struct Dft<T, 16> {
  static __host__ __device__ __forceinline__ void run(cplx<T> *v) { dft_composite<T, 4, 4>(v); }
};

// radices with a register-level DFT above
__host__ __device__ constexpr bool dft_supported(int R) {
  return R == 2 || R == 3 || R == 4 || R == 5 || R == 6 || R == 8 || R == 9 || R == 12 || R == 16;
}
__host__ __device__ constexpr bool dft_composite_radix(int R) { return R == 6 || R == 8 || R == 9 || R == 12 || R == 16; }

}  // namespace pp
