// Shared helpers for the sawfem HIP library (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct c64 { float x, y; };
struct c128 { double x, y; };

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };

template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  V r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { V r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { V r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// a - b*c
template <typename V> __device__ __forceinline__ V cfms(V a, V b, V c) {
  V r; r.x = a.x - (b.x * c.x - b.y * c.y); r.y = a.y - (b.x * c.y + b.y * c.x); return r;
}
// a + b*c
template <typename V> __device__ __forceinline__ V cfma(V a, V b, V c) {
  V r; r.x = a.x + (b.x * c.x - b.y * c.y); r.y = a.y + (b.x * c.y + b.y * c.x); return r;
}
template <typename V> __device__ __forceinline__ V cdiv(V a, V b) {
  // Smith-free plain division; pivots are far from overflow in this application.
  auto d = b.x * b.x + b.y * b.y;
  V r; r.x = (a.x * b.x + a.y * b.y) / d; r.y = (a.y * b.x - a.x * b.y) / d; return r;
}
template <typename V> __device__ __forceinline__ V czero() { V r; r.x = 0; r.y = 0; return r; }

__device__ __forceinline__ double2 to_d2(float2 a) { return make_double2((double)a.x, (double)a.y); }
__device__ __forceinline__ double2 to_d2(double2 a) { return a; }
template <typename V> __device__ __forceinline__ V from_d2(double2 a);
template <> __device__ __forceinline__ float2 from_d2<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
template <> __device__ __forceinline__ double2 from_d2<double2>(double2 a) { return a; }

// thread-local last error for the C ABI
const char* sawfem_set_error(const char* msg);

// status 4 for a failed device allocation, 3 for any other HIP runtime error
#define SAW_HIP(call)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      sawfem_set_error(hipGetErrorString(e_));                                                \
      return e_ == hipErrorOutOfMemory ? 4 : 3;                                               \
    }                                                                                         \
  } while (0)
