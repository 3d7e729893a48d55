// fused.hip — one persistent launch per sweep for a batch of independent
// tiles: each 1024-thread workgroup owns one tile and runs every LED update of
// the sweep (projection_update, src/algorithms/common.rs:57-306, single
// source per frame) with the tile's low-resolution field resident in LDS:
//
//   gather 4-tap bilinear patch from the FP64 object spectrum (crop.rs:129-167)
//   x pupil -> ifftshift -> LDS
//   inverse 2-D FFT in LDS (Stockham radix-16/8 stages, FP32)
//   |field|^2, objective terms (FP64), amplitude replacement (common.rs:123-228)
//   forward 2-D FFT in LDS (x 1/N), fftshift
//   U = conj(P) (psi' - patch P) / denominator (common.rs:240-261)
//   adjoint insert as a gather over the (n+1)^2 window, FP64 RMW (crop.rs:188-256)
//   every `interval` counted frames: mPIE momentum over the tile's R^2 grid
//   (common.rs:308-387)
//
// Arithmetic: FFTs, fields and updates in FP32 (explicit FMAs), object
// spectrum, momentum, denominators and objective reductions in FP64.
// Workgroup-internal reductions use a fixed order, so results are
// deterministic run to run. No atomics anywhere.
#include "fused.h"

namespace fpmhip {

namespace {

constexpr int kThreads = 1024;

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

// One Stockham stage (radix P, stride S) over all N lines of an N x N array
// in LDS. Rows: consecutive threads walk a row; columns: consecutive threads
// walk across columns. Loads complete (barrier) before any store, so the
// stage runs in place in a single LDS buffer.
template <int N, int P, int S, bool ROWS, bool INV>
__device__ __forceinline__ void stage(float2* lds, const float2* tw) {
  constexpr int M = N / (P * S);
  constexpr int BPL = N / P;
  constexpr int TOTAL = N * BPL;
  constexpr int PER = (TOTAL + kThreads - 1) / kThreads;
  float2 a[PER][P];
  int line[PER], pp[PER], q[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = threadIdx.x + u * kThreads;
    line[u] = ROWS ? b / BPL : b % N;
    const int t = ROWS ? b % BPL : b / N;
    pp[u] = t / S;
    q[u] = t % S;
    if (b < TOTAL) {
#pragma unroll
      for (int r = 0; r < P; ++r) {
        const int e = q[u] + S * (pp[u] + r * M);
        a[u][r] = ROWS ? lds[line[u] * N + e] : lds[e * N + line[u]];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = threadIdx.x + u * kThreads;
    if (b < TOTAL) {
      reg_dft<P, INV>(a[u]);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int e = q[u] + S * (P * pp[u] + k);
        float2 v = a[u][k];
        if (k > 0 && M > 1) v = fmul(v, tw[(pp[u] * k * S) % N]);
        if (ROWS)
          lds[line[u] * N + e] = v;
        else
          lds[e * N + line[u]] = v;
      }
    }
  }
  __syncthreads();
}

template <int N>
struct Plan;
template <>
struct Plan<32> {
  static constexpr int P0 = 8, P1 = 4;
};
template <>
struct Plan<64> {
  static constexpr int P0 = 8, P1 = 8;
};
template <>
struct Plan<128> {
  static constexpr int P0 = 16, P1 = 8;
};

// Unscaled 2-D DFT in LDS with the direction given by `tw`.
template <int N, bool INV>
__device__ __forceinline__ void fft2_lds(float2* lds, const float2* tw) {
  constexpr int P0 = Plan<N>::P0, P1 = Plan<N>::P1;
  static_assert(P0 * P1 == N, "plan");
  stage<N, P0, 1, true, INV>(lds, tw);
  stage<N, P1, P0, true, INV>(lds, tw);
  stage<N, P0, 1, false, INV>(lds, tw);
  stage<N, P1, P0, false, INV>(lds, tw);
}

// Fixed-order workgroup sum of one double per thread.
__device__ __forceinline__ double wg_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) s += red[w];
  return s;  // valid in thread 0
}

__device__ __forceinline__ double point_loss(double predicted, double measured, int loss) {
  if (loss == FPMHIP_LOSS_INTENSITY_MSE) {
    const double r = predicted - measured;
    return r * r;
  }
  if (loss == FPMHIP_LOSS_POISSON_NLL) {
    const double mean = fmax(predicted, 1e-12);
    return mean - fmax(measured, 0.0) * log(mean);
  }
  const double r = sqrt(fmax(predicted, 0.0)) - sqrt(fmax(measured, 0.0));
  if (loss == FPMHIP_LOSS_HUBER_AMPLITUDE) {
    const double a = fabs(r);
    return a <= 1.0 ? 0.5 * r * r : a - 0.5;
  }
  return r * r;
}

// Bilinear sample of the FP64 spectrum for centred patch pixel (r, c)
// (32-bit offsets: a tile spectrum has far fewer than 2^32 elements).
__device__ __forceinline__ double2 sample(const double2* __restrict__ spec, uint32_t rw, const SrcParam& sp,
                                          int r, int c) {
  if (sp.zero) return spec[((uint32_t)sp.sr + r) * rw + (uint32_t)sp.sc + c];
  const uint32_t l0 = (sp.ar.lower + r) * rw + sp.ac.lower + c;
  const uint32_t du = (sp.ar.upper - sp.ar.lower) * rw, dc = sp.ac.upper - sp.ac.lower;
  const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
               wuu = sp.ar.wu * sp.ac.wu;
  const double2 a = spec[l0], b = spec[l0 + dc], cc = spec[l0 + du], d = spec[l0 + du + dc];
  return make_double2(fma(d.x, wuu, fma(cc.x, wul, fma(b.x, wlu, a.x * wll))),
                      fma(d.y, wuu, fma(cc.y, wul, fma(b.y, wlu, a.y * wll))));
}

// Object-update denominator of common.rs:247-258 for pupil value P.
__device__ __forceinline__ float inv_denominator(const TileDesc& T, float2 P, double maxp) {
  const double local = (double)P.x * P.x + (double)P.y * P.y;
  double den;
  if (T.denom == FPMHIP_DENOMINATOR_LOCAL)
    den = local + T.eps;
  else if (T.denom == FPMHIP_DENOMINATOR_GLOBAL)
    den = maxp + T.eps;
  else
    den = (1.0 - T.alpha) * local + T.alpha * maxp + T.eps;
  return (float)(1.0 / den);
}

// One sweep of one tile. The centred pixel p = (r, c) and its swap partner
// ((r + N/2) mod N, (c + N/2) mod N) are owned by the same thread: for even
// N, fftshift and ifftshift are both that swap, so the shifted LDS reads and
// writes of phases 1 and 5 stay inside the owning thread.
template <int N>
__device__ void sweep_tile(const TileDesc& T, float2* lds, const float2* twf, const float2* twi, double* red) {
  constexpr int NN = N * N;
  constexpr int HALF = N / 2;
  constexpr int NPAIR = NN / 2;  // pairs: pixels with r < N/2 and their partners
  constexpr int PR = (NPAIR + kThreads - 1) / kThreads;
  const int tid = threadIdx.x;
  const float inv_n = 1.0f / (float)NN;
  double2* __restrict__ spec = T.spec;
  const float2* __restrict__ pupil = T.pupil;
  const double eps = T.eps;
  const double maxp = fmax(T.max_p, eps);
  uint64_t counter = *T.counter;
  double obj_sum = 0.0, w_sum = 0.0;
  for (uint32_t pos = 0; pos < T.F; ++pos) {
    const uint32_t f = T.order ? T.order[pos] : pos;
    const double fw = T.weights[f];
    if (fw == 0.0) {
      if (tid == 0) T.frame_loss[pos] = 0.0;
      continue;
    }
    const SrcParam sp = T.src[f];
    const double gain = T.gains[f];
    // 1. patch x pupil at p and its partner, written swapped (ifftshift)
    float2 ex0[PR], ex1[PR];
#pragma unroll
    for (int j = 0; j < PR; ++j) {
      const int i = tid + j * kThreads;
      if (i < NPAIR) {
        const int r = i / N, c = i % N;
        const int r1 = r + HALF, c1 = (c + HALF) % N;
        const int p0 = r * N + c, p1 = r1 * N + c1;
        const double2 v0 = sample(spec, T.rw, sp, r, c);
        const double2 v1 = sample(spec, T.rw, sp, r1, c1);
        ex0[j] = fmul(f2((float)v0.x, (float)v0.y), pupil[p0]);
        ex1[j] = fmul(f2((float)v1.x, (float)v1.y), pupil[p1]);
        lds[p1] = ex0[j];
        lds[p0] = ex1[j];
      }
    }
    __syncthreads();
    // 2. inverse FFT (unscaled)
    fft2_lds<N, true>(lds, twi);
    // 3. amplitude replacement and objective terms (pointwise, natural order)
    double lsum = 0.0, lcnt = 0.0;
    const uint64_t fbase = (uint64_t)f * NN;
    for (int p = tid; p < NN; p += kThreads) {
      const float2 fv = lds[p];
      const double pred = (double)fv.x * fv.x + (double)fv.y * fv.y;
      const bool masked = T.masks && T.masks[fbase + p] == 0;
      float2 d = fv;
      if (!masked) {
        const double b = T.bg_len == 0 ? 0.0 : T.bg[T.bg_len == (uint64_t)NN ? p : fbase + p];
        const double y = T.is_u16 ? (double)static_cast<const uint16_t*>(T.frames)[fbase + p]
                                  : static_cast<const double*>(T.frames)[fbase + p];
        lsum += point_loss(gain * pred + b, y, T.loss);
        lcnt += 1.0;
        const double target = fmax((y - b) / gain, 0.0);
        if (pred > eps) {
          const float ratio = (float)sqrt(target / pred);
          d = f2(fv.x * ratio, fv.y * ratio);
        } else {
          d = f2((float)sqrt(target), 0.0f);
        }
      }
      lds[p] = d;
    }
    const double ls = wg_sum(lsum, red);
    const double lc = wg_sum(lcnt, red);
    if (tid == 0) {
      const double L = lc > 0.0 ? ls / lc : 0.0;
      T.frame_loss[pos] = L;
      obj_sum += fw * L;
      w_sum += fw;
    }
    __syncthreads();
    // 4. forward FFT (x 1/N applied in phase 5)
    fft2_lds<N, false>(lds, twf);
    // 5. U = conj(P) (fftshift(FFT)/N - patch P) / den at p and its partner
#pragma unroll
    for (int j = 0; j < PR; ++j) {
      const int i = tid + j * kThreads;
      if (i < NPAIR) {
        const int r = i / N, c = i % N;
        const int p0 = r * N + c, p1 = (r + HALF) * N + (c + HALF) % N;
        const float2 s0 = lds[p1], s1 = lds[p0];
        const float2 P0 = pupil[p0], P1 = pupil[p1];
        const float2 d0 = fsub(f2(s0.x * inv_n, s0.y * inv_n), ex0[j]);
        const float2 d1 = fsub(f2(s1.x * inv_n, s1.y * inv_n), ex1[j]);
        const float2 u0 = fmulc(P0, d0), u1 = fmulc(P1, d1);
        const float k0 = inv_denominator(T, P0, maxp), k1 = inv_denominator(T, P1, maxp);
        lds[p0] = f2(u0.x * k0, u0.y * k0);
        lds[p1] = f2(u1.x * k1, u1.y * k1);
      }
    }
    __syncthreads();
    // 6. adjoint insert: gather over the destination window, FP64 RMW
    const double scale = fw * T.object_step;
    if (sp.zero) {
      for (int p = tid; p < NN; p += kThreads) {
        const int r = p / N, c = p % N;
        double2& d = spec[(sp.sr + r) * T.rw + sp.sc + c];
        const float2 u = lds[p];
        d = make_double2(fma(scale, (double)u.x, d.x), fma(scale, (double)u.y, d.y));
      }
    } else {
      const bool fr = sp.ar.upper != sp.ar.lower, fc = sp.ac.upper != sp.ac.lower;
      const int wh = N + (fr ? 1 : 0), ww = N + (fc ? 1 : 0);
      const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
                   wuu = sp.ar.wu * sp.ac.wu;
      for (int qd = tid; qd < wh * ww; qd += kThreads) {
        const int a = qd / ww, b = qd % ww;
        double2& dst = spec[(sp.ar.lower + a) * (uint64_t)T.rw + sp.ac.lower + b];
        double ax = 0.0, ay = 0.0;
        auto tap = [&](int sr_, int sc_, double wt) {
          if (sr_ >= 0 && sr_ < N && sc_ >= 0 && sc_ < N) {
            const float2 u = lds[sr_ * N + sc_];
            ax = fma((double)u.x, wt, ax);
            ay = fma((double)u.y, wt, ay);
          }
        };
        if (fr && fc) {
          tap(a - 1, b - 1, wuu);
          tap(a - 1, b, wul);
          tap(a, b - 1, wlu);
          tap(a, b, wll);
        } else if (fr) {
          tap(a - 1, b, wul);
          tap(a, b, wll);
        } else if (fc) {
          tap(a, b - 1, wlu);
          tap(a, b, wll);
        } else {
          tap(a, b, wll);
        }
        const double2 o = dst;
        dst = make_double2(fma(scale, ax, o.x), fma(scale, ay, o.y));
      }
    }
    __syncthreads();
    // 7. mPIE momentum over the whole tile spectrum (common.rs:308-387)
    if (T.momentum) {
      counter += 1;
      if (counter >= T.interval) {
        const uint64_t M = (uint64_t)T.rh * T.rw;
        bool bad = false;
        for (uint64_t i = tid; i < M; i += kThreads) {
          const double2 o = spec[i], v = T.vel[i], an = T.anc[i];
          const double2 nv = make_double2(fma(T.friction, v.x, o.x - an.x), fma(T.friction, v.y, o.y - an.y));
          const double2 no =
              T.feedback == 0.0 ? o : make_double2(fma(T.feedback, nv.x, o.x), fma(T.feedback, nv.y, o.y));
          bad |= !isfinite(nv.x) || !isfinite(nv.y) || !isfinite(no.x) || !isfinite(no.y);
          T.vel[i] = nv;
          spec[i] = no;
          T.anc[i] = no;
        }
        if (bad) T.status[0] = 1;
        counter = 0;
        __syncthreads();
      }
    }
  }
  if (tid == 0) {
    const uint32_t it = *T.iteration;
    if (it < T.trace_len) T.trace[it] = w_sum > 0.0 ? obj_sum / w_sum : __builtin_nan("");
    *T.iteration = it + 1;
    *T.counter = counter;
  }
}

__global__ void __launch_bounds__(kThreads) k_sweep(const TileDesc* __restrict__ tiles, const float2* __restrict__ tw_all) {
  extern __shared__ unsigned char smem[];
  float2* lds = reinterpret_cast<float2*>(smem);
  const TileDesc T = tiles[blockIdx.x];
  const int N = (int)T.n;
  // twiddles W_N^j, forward then inverse, after the N x N field
  float2* twf = lds + N * N;
  float2* twi = twf + N;
  double* red = reinterpret_cast<double*>(twi + N);
  const float2* src = tw_all + tw_offset(N);
  for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) twf[j] = src[j];
  __syncthreads();
  switch (N) {
    case 32:
      sweep_tile<32>(T, lds, twf, twi, red);
      break;
    case 64:
      sweep_tile<64>(T, lds, twf, twi, red);
      break;
    case 128:
      sweep_tile<128>(T, lds, twf, twi, red);
      break;
    default:
      break;
  }
}

}  // namespace

size_t sweep_lds_bytes(uint32_t nmax) {
  return (size_t)nmax * nmax * sizeof(float2) + 2 * (size_t)nmax * sizeof(float2) + 16 * sizeof(double);
}

void prepare_sweep(uint32_t nmax) {
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sweep_lds_bytes(nmax)));
}

void launch_sweep(const TileDesc* tiles_dev, uint32_t count, const float2* tw_all, uint32_t nmax, hipStream_t stream) {
  const size_t smem = sweep_lds_bytes(nmax);
  hipLaunchKernelGGL(k_sweep, dim3(count), dim3(kThreads), smem, stream, tiles_dev, tw_all);
  HIP_CHECK(hipGetLastError());
}

}  // namespace fpmhip
