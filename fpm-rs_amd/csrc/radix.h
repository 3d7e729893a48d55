// radix.h — complex64 helpers and the in-register radix-2^k DFT of the
// fused sweep's transform stages (also used by the DFT trial program).
#pragma once

#include <hip/hip_runtime.h>

namespace fpmhip {
namespace {

__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float2 fadd(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 fsub(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 fmul(float2 a, float2 b) {
  return f2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 fmulc(float2 a, float2 b) {  // conj(a) * b
  return f2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.x, b.y, -a.y * b.x));
}

// 16th roots of unity: cos(pi m / 8), sin(pi m / 8), m = 0..15.
__device__ constexpr float kCos16[16] = {1.0f, 0.92387953251128674f, 0.70710678118654757f, 0.38268343236508978f,
                                         0.0f, -0.38268343236508978f, -0.70710678118654757f, -0.92387953251128674f,
                                         -1.0f, -0.92387953251128674f, -0.70710678118654757f, -0.38268343236508978f,
                                         0.0f, 0.38268343236508978f, 0.70710678118654757f, 0.92387953251128674f};
__device__ constexpr float kSin16[16] = {0.0f, 0.38268343236508978f, 0.70710678118654757f, 0.92387953251128674f,
                                         1.0f, 0.92387953251128674f, 0.70710678118654757f, 0.38268343236508978f,
                                         0.0f, -0.38268343236508978f, -0.70710678118654757f, -0.92387953251128674f,
                                         -1.0f, -0.92387953251128674f, -0.70710678118654757f, -0.38268343236508978f};

// v * W_16^M (forward: exp(-2 pi i M/16), inverse: conjugate); quarter turns exact.
template <int M, bool INV>
__device__ __forceinline__ float2 mul_w16(float2 v) {
  constexpr int m = M & 15;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (m == 8) {
    return f2(-v.x, -v.y);
  } else if constexpr (m == 4) {
    return INV ? f2(-v.y, v.x) : f2(v.y, -v.x);
  } else if constexpr (m == 12) {
    return INV ? f2(v.y, -v.x) : f2(-v.y, v.x);
  } else {
    const float c = kCos16[m], s = INV ? kSin16[m] : -kSin16[m];
    return f2(fmaf(v.x, c, -v.y * s), fmaf(v.x, s, v.y * c));
  }
}

template <int B, int V>
struct BitRev {
  static constexpr int value = (BitRev<B - 1, (V >> 1)>::value) | ((V & 1) << (B - 1));
};
template <int V>
struct BitRev<0, V> {
  static constexpr int value = 0;
};
template <int P>
struct Log2 {
  static constexpr int value = 1 + Log2<P / 2>::value;
};
template <>
struct Log2<1> {
  static constexpr int value = 0;
};

template <int P, bool INV, int I = 0>
__device__ __forceinline__ void bitrev_permute(float2* x) {
  if constexpr (I < P) {
    constexpr int J = BitRev<Log2<P>::value, I>::value;
    if constexpr (J > I) {
      const float2 t = x[I];
      x[I] = x[J];
      x[J] = t;
    }
    bitrev_permute<P, INV, I + 1>(x);
  }
}

template <int P, bool INV, int LEN, int I, int K>
__device__ __forceinline__ void dit_level(float2* x) {
  if constexpr (I < P) {
    if constexpr (K < LEN / 2) {
      const float2 t = mul_w16<K * (16 / LEN), INV>(x[I + K + LEN / 2]);
      const float2 u = x[I + K];
      x[I + K] = fadd(u, t);
      x[I + K + LEN / 2] = fsub(u, t);
      dit_level<P, INV, LEN, I, K + 1>(x);
    } else {
      dit_level<P, INV, LEN, I + LEN, 0>(x);
    }
  }
}

template <int P, bool INV, int LEN = 2>
__device__ __forceinline__ void dit_levels(float2* x) {
  if constexpr (LEN <= P) {
    dit_level<P, INV, LEN, 0, 0>(x);
    dit_levels<P, INV, LEN * 2>(x);
  }
}

// In-place natural-order DFT of P <= 16 register values (iterative radix-2
// decimation in time, compile-time twiddles).
template <int P, bool INV>
__device__ __forceinline__ void reg_dft(float2* x) {
  static_assert(P <= 16 && (P & (P - 1)) == 0, "radix");
  bitrev_permute<P, INV>(x);
  dit_levels<P, INV>(x);
}

}  // namespace
}  // namespace fpmhip
