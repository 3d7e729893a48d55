// cx.h — complex arithmetic on double2/float2 with num-complex 0.4.6's
// formulas (Mul: (ac - bd, ad + bc); Complex/f64 divides each part). The
// library is compiled with -ffp-contract=off, so each operation rounds
// separately, as in the reference's Rust; kernels that want fused
// multiply-adds call fma explicitly.
#pragma once
#include <hip/hip_runtime.h>

namespace fpmhip {

template <class R>
struct cx;

template <>
struct cx<double> {
  using V = double2;
  __host__ __device__ static inline V make(double r, double i) { return make_double2(r, i); }
  __host__ __device__ static inline V add(V a, V b) { return make_double2(a.x + b.x, a.y + b.y); }
  __host__ __device__ static inline V sub(V a, V b) { return make_double2(a.x - b.x, a.y - b.y); }
  __host__ __device__ static inline V mul(V a, V b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __host__ __device__ static inline V mulr(V a, double s) { return make_double2(a.x * s, a.y * s); }
  __host__ __device__ static inline V scale(double s, V a) { return make_double2(s * a.x, s * a.y); }
  __host__ __device__ static inline V divr(V a, double s) { return make_double2(a.x / s, a.y / s); }
  __host__ __device__ static inline V conj(V a) { return make_double2(a.x, -a.y); }
  __host__ __device__ static inline double norm2(V a) { return a.x * a.x + a.y * a.y; }
};

template <>
struct cx<float> {
  using V = float2;
  __host__ __device__ static inline V make(float r, float i) { return make_float2(r, i); }
  __host__ __device__ static inline V add(V a, V b) { return make_float2(a.x + b.x, a.y + b.y); }
  __host__ __device__ static inline V sub(V a, V b) { return make_float2(a.x - b.x, a.y - b.y); }
  __host__ __device__ static inline V mul(V a, V b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
  }
  __host__ __device__ static inline V mulr(V a, float s) { return make_float2(a.x * s, a.y * s); }
  __host__ __device__ static inline V scale(float s, V a) { return make_float2(s * a.x, s * a.y); }
  __host__ __device__ static inline V divr(V a, float s) { return make_float2(a.x / s, a.y / s); }
  __host__ __device__ static inline V conj(V a) { return make_float2(a.x, -a.y); }
  __host__ __device__ static inline float norm2(V a) { return a.x * a.x + a.y * a.y; }
};

// conversions between the session precision and the FP64 object spectrum
__host__ __device__ inline double2 to_d2(double2 v) { return v; }
__host__ __device__ inline double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
template <class V>
__host__ __device__ inline V from_d2(double2 v);
template <>
__host__ __device__ inline double2 from_d2<double2>(double2 v) {
  return v;
}
template <>
__host__ __device__ inline float2 from_d2<float2>(double2 v) {
  return make_float2((float)v.x, (float)v.y);
}

}  // namespace fpmhip
