// Shared device/host helpers of the pawpyseed_amd HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/pawpy_abi.h"

// Interleaved complex of real type T; layout-compatible with C99
// `double complex` / `float complex` and std::complex<T>.
template <typename T> struct alignas(2 * sizeof(T)) cplx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T a, T b) {
  cplx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> scale(cplx<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}
// acc += a * b
template <typename T>
__host__ __device__ __forceinline__ void cmac(cplx<T> &acc, cplx<T> a, cplx<T> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ void cmacc(cplx<T> &acc, cplx<T> a, cplx<T> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
// multiply by +i
template <typename T> __host__ __device__ __forceinline__ cplx<T> mul_i(cplx<T> a) {
  return mk<T>(-a.y, a.x);
}
template <typename To, typename Ti>
__host__ __device__ __forceinline__ cplx<To> cvt(cplx<Ti> a) {
  return mk<To>((To)a.x, (To)a.y);
}

// Non-temporal (streaming) access to one complex value in device memory.
template <typename T> __device__ __forceinline__ cplx<T> load_nt(const cplx<T> *p) {
  using V = T __attribute__((ext_vector_type(2)));
  const V v = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
  return mk<T>(v.x, v.y);
}
template <typename T> __device__ __forceinline__ void store_nt(cplx<T> *p, cplx<T> v) {
  using V = T __attribute__((ext_vector_type(2)));
  V w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V *>(p));
}

// Two adjacent complex floats as one 16-byte access (p 16-byte aligned).
__device__ __forceinline__ void load2(const cplx<float> *p, bool nt, cplx<float> &a, cplx<float> &b) {
  using V = float __attribute__((ext_vector_type(4)));
  const V v = nt ? __builtin_nontemporal_load(reinterpret_cast<const V *>(p)) : *reinterpret_cast<const V *>(p);
  a = mk<float>(v.x, v.y);
  b = mk<float>(v.z, v.w);
}
__device__ __forceinline__ void store2(cplx<float> *p, bool nt, cplx<float> a, cplx<float> b) {
  using V = float __attribute__((ext_vector_type(4)));
  V w;
  w.x = a.x;
  w.y = a.y;
  w.z = b.x;
  w.w = b.y;
  if (nt)
    __builtin_nontemporal_store(w, reinterpret_cast<V *>(p));
  else
    *reinterpret_cast<V *>(p) = w;
}

// ----------------------------------------------------------------- errors
namespace pp {
void set_error(int code, const std::string &msg);
int last_code();
}  // namespace pp

#define PP_CHECK_HIP(expr)                                                       \
  do {                                                                           \
    hipError_t _e = (expr);                                                      \
    if (_e != hipSuccess) {                                                      \
      pp::set_error(PP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return PP_ERR_HIP;                                                         \
    }                                                                            \
  } while (0)

#define PP_REQUIRE(cond, code, msg)                                              \
  do {                                                                           \
    if (!(cond)) {                                                               \
      pp::set_error((code), (msg));                                              \
      return (code);                                                             \
    }                                                                            \
  } while (0)

#define PP_TRY(expr)                                                             \
  do {                                                                           \
    int _s = (expr);                                                             \
    if (_s != PP_OK)                                                             \
      return _s;                                                                 \
  } while (0)
