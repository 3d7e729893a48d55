// fused.hip — one persistent launch per sweep for a batch of independent
// tiles: each 1024-thread workgroup owns one tile and runs every LED update of
// the sweep (projection_update, src/algorithms/common.rs:57-306, single
// source per frame) with the tile's low-resolution field resident in LDS:
//
//   gather 4-tap bilinear patch from the FP64 object spectrum (crop.rs:129-167)
//   x pupil -> ifftshift -> LDS
//   inverse 2-D FFT in LDS (in-place radix-16/8 stages, decimation in
//   frequency, FP32; 256^2 and 512^2 tiles: Stockham stages over chunks)
//   |field|^2, objective terms (FP64), amplitude replacement (common.rs:123-228),
//   kept as the field change psi' - psi, in the inverse's digit-reversed order
//   forward 2-D FFT in LDS (decimation in time, back to natural order; x 1/N),
//   fftshift: the exit-wave difference
//   U = conj(P) difference / denominator (common.rs:240-261)
//   adjoint insert as a gather over the (n+1)^2 window, FP64 RMW (crop.rs:188-256)
//   every `interval` counted frames: mPIE momentum over the tile's R^2 grid
//   (common.rs:308-387)
//
// Arithmetic: FFTs, fields and updates in FP32 (explicit FMAs), object
// spectrum, momentum, denominators and objective reductions in FP64.
// Workgroup-internal reductions use a fixed order, so results are
// deterministic run to run. No atomics anywhere.
#include <algorithm>

#include "fused.h"
#include "radix.h"

namespace fpmhip {

namespace {

constexpr int kThreads = 1024;


// Row pitch of the N x N field in LDS: one padding element per row makes
// both row-wise and column-wise accesses of consecutive lanes hit distinct
// banks (a stride of N+1 complex64 values is an odd number of 8-byte words).
template <int N>
__host__ __device__ constexpr int pitch() {
  return N + 1;
}

// The lines of a stage that carry data: nl lines starting at line l0,
// cyclically (all lines when LIVE is off).
struct LiveRows {
  int nl;     // number of live lines
  int l0;     // first live line
  float rcp;  // 1 / nl
};

// One Stockham stage (radix P, stride S) over LINES lines of length LEN in
// LDS; sample e of line l sits at l * LS + e * ES. Consecutive lanes take
// consecutive lines, so with LS = 1 or LS = an odd number of 8-byte words
// every LDS access of a wave is conflict free. Loads complete (barrier)
// before any store, so the stage runs in place in one buffer. With LIVE,
// only the live lines are transformed: the work items are packed onto the
// first threads, so whole waves skip the stage.
template <int LEN, int LINES, int P, int S, int LS, int ES, bool INV, bool LIVE = false>
__device__ __forceinline__ void stage(float2* lds, const float2* tw, int tid, LiveRows lr = {LINES, 0, 0.0f}) {
  asm volatile("" : "+v"(tid));
  constexpr int M = LEN / (P * S);
  constexpr int BPL = LEN / P;
  constexpr int TOTAL = LINES * BPL;
  constexpr int PER = (TOTAL + kThreads - 1) / kThreads;
  const int total = LIVE ? lr.nl * BPL : TOTAL;
  float2 a[PER][P];
  int line[PER], pp[PER], q[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = tid + u * kThreads;
    int t;
    if (LIVE) {
      // b = t * nl + li, the quotient from the FP32 reciprocal, corrected
      t = (int)((float)b * lr.rcp);
      int li = b - t * lr.nl;
      if (li < 0) {
        --t;
        li += lr.nl;
      } else if (li >= lr.nl) {
        ++t;
        li -= lr.nl;
      }
      line[u] = (lr.l0 + li) & (LINES - 1);
    } else {
      line[u] = b % LINES;
      t = b / LINES;
    }
    pp[u] = t / S;
    q[u] = t % S;
    if (b < total) {
#pragma unroll
      for (int r = 0; r < P; ++r) {
        const int e = q[u] + S * (pp[u] + r * M);
        a[u][r] = lds[line[u] * LS + e * ES];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = tid + u * kThreads;
    if (b < total) {
      reg_dft<P, INV>(a[u]);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        const int e = q[u] + S * (P * pp[u] + k);
        float2 v = a[u][k];
        if (k > 0 && M > 1) v = fmul(v, tw[(pp[u] * k * S) % LEN]);
        lds[line[u] * LS + e * ES] = v;
      }
    }
  }
  __syncthreads();
}

// Two-stage radix plans of an N-point line transform of a large tile: first
// stage radix P0, second N / P0 (each at most 16).
template <int N>
struct DefaultPlan;
template <>
struct DefaultPlan<256> {
  static constexpr int P0 = 16;
};

// Line of work item b of a stage over the live lines (b = t * nl + li,
// the quotient from the FP32 reciprocal, corrected): returns the line and
// sets t.
__device__ __forceinline__ int live_line(int b, int lines, LiveRows lr, int& t) {
  t = (int)((float)b * lr.rcp);
  int li = b - t * lr.nl;
  if (li < 0) {
    --t;
    li += lr.nl;
  } else if (li >= lr.nl) {
    ++t;
    li -= lr.nl;
  }
  return (lr.l0 + li) & (lines - 1);
}

// In-place two-stage transforms of the LDS-resident tile (N = P0 P1, each
// radix at most 16). The inverse is decimation in frequency: from natural
// order, stage 1 (radix P0) takes the samples g + P1 k of butterfly g < P1
// and multiplies its output k by W^(g k); stage 2 (radix P1) transforms the
// contiguous groups g P1 + m, g < P0. Position j1 + P1 j0 then holds output
// j0 + P0 j1 (digit-reversed order). The forward transform is decimation in
// time, the inverse's steps reversed: stage 1 (radix P1) on the contiguous
// groups, output m of group g multiplied by W^(-g m), then stage 2 (radix P0)
// on the samples g + P1 k, which leaves natural order. A butterfly writes
// back to the slots it read and no other thread touches them during the
// stage, so a stage needs one barrier (after its writes), not one between
// its reads and writes as well.
template <int P0, int P1>
__device__ __forceinline__ int digit_rev(int p) {
  return p / P1 + P0 * (p % P1);
}

// One in-place stage over LINES lines (sample e of line l at l * LS + e * ES):
// radix P on the strided samples g + k N/P (STRIDED) or on the contiguous
// groups g P + k; with TW, output k is multiplied by tw[g k]. Consecutive
// lanes take consecutive lines, as in `stage`. With LIVE only the live lines
// are transformed (work items packed onto the first threads).
template <int N, int LINES, int P, bool STRIDED, bool TW, int LS, int ES, bool INV, bool LIVE = false>
__device__ __forceinline__ void ip_stage(float2* lds, const float2* tw, int tid, LiveRows lr = {LINES, 0, 0.0f}) {
  asm volatile("" : "+v"(tid));
  constexpr int BPL = N / P;
  constexpr int TOTAL = LINES * BPL;
  constexpr int PER = (TOTAL + kThreads - 1) / kThreads;
  constexpr int STEP = (STRIDED ? BPL : 1) * ES;
  const int total = LIVE ? lr.nl * BPL : TOTAL;
  float2 a[PER][P];
  int base[PER], g[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int b = tid + u * kThreads;
    int line;
    if (LIVE) {
      line = live_line(b, LINES, lr, g[u]);
    } else {
      line = b % LINES;
      g[u] = b / LINES;
    }
    base[u] = line * LS + (STRIDED ? g[u] : g[u] * P) * ES;
    if (b < total) {
#pragma unroll
      for (int k = 0; k < P; ++k) a[u][k] = lds[base[u] + k * STEP];
    }
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    if (tid + u * kThreads < total) {
      reg_dft<P, INV>(a[u]);
#pragma unroll
      for (int k = 0; k < P; ++k) {
        float2 v = a[u][k];
        if (TW && k > 0) v = fmul(v, tw[g[u] * k]);
        lds[base[u] + k * STEP] = v;
      }
    }
  }
  __syncthreads();
}

// The thread index, hidden from the optimiser so that per-thread addresses
// are recomputed inside the LED loop instead of being hoisted out of it and
// spilled for the whole sweep.
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// Reduction slots in LDS, one double per wave in each of five sets: two
// sums (frame objective), and three maxima of EPRY (|patch|^2 over the
// support pixels, over the whole patch, and |P|^2 of the updated pupil), each
// written before a barrier and read after it.
constexpr int kWaves = kThreads / 64;
constexpr int kSlotLower = 2 * kWaves, kSlotPatch = 3 * kWaves, kSlotPupil = 4 * kWaves;
constexpr int kRedSlots = 5 * kWaves;

// Fixed-order workgroup sums of two doubles per thread (one pair of
// barriers for both); red holds kRedSlots doubles.
__device__ __forceinline__ double2 wg_sum2(double a, double b, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  const int wave = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[wave] = a;
    red[kThreads / 64 + wave] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kThreads / 64; ++w) {
      sa += red[w];
      sb += red[kThreads / 64 + w];
    }
  return make_double2(sa, sb);  // valid in thread 0
}

// The first half of wg_sum2 without its barriers: each wave's sums of a and
// b into red[wave] and red[kThreads / 64 + wave]. The caller's next barrier
// publishes them (and a barrier must separate this from the last read of
// red); thread 0 then adds the slots in wave order.
__device__ __forceinline__ void wave_sums2_to(double* red, double a, double b) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off, 64);
    b += __shfl_xor(b, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = a;
    red[kThreads / 64 + (threadIdx.x >> 6)] = b;
  }
}

// A value that is the same in every lane, as a scalar register value.
__device__ __forceinline__ double uniform(double v) {
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Fixed-order workgroup maximum of one double per thread; valid in every
// thread (red holds kThreads / 64 doubles).
__device__ __forceinline__ double wg_max(double a, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  double r = red[0];
  for (int w = 1; w < kThreads / 64; ++w) r = fmax(r, red[w]);
  __syncthreads();
  return uniform(r);
}

// Object-update denominator of common.rs:247-258 as a + b |P|^2 with FP64
// constants a, b rounded once: Local: (eps, 1); Global: (max + eps, 0);
// Rpie: (alpha max + eps, 1 - alpha).
struct Denominator {
  float a, b;
};
__device__ __forceinline__ Denominator make_denominator(const TileDesc& T, double maxp) {
  if (T.denom == FPMHIP_DENOMINATOR_LOCAL) return {(float)T.eps, 1.0f};
  if (T.denom == FPMHIP_DENOMINATOR_GLOBAL) return {(float)(maxp + T.eps), 0.0f};
  return {(float)(T.alpha * maxp + T.eps), (float)(1.0 - T.alpha)};
}

// Hardware square root and reciprocal (1 ulp): the FP32 per-pixel terms
// feed FP64 accumulations, far inside the 1e-5 error budget.
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ float point_loss_f(float predicted, float measured, int loss) {
  if (loss == FPMHIP_LOSS_INTENSITY_MSE) {
    const float r = predicted - measured;
    return r * r;
  }
  if (loss == FPMHIP_LOSS_POISSON_NLL) {
    const float mean = fmaxf(predicted, 1e-12f);
    return mean - fmaxf(measured, 0.0f) * logf(mean);
  }
  const float r = fsqrt(fmaxf(predicted, 0.0f)) - fsqrt(fmaxf(measured, 0.0f));
  if (loss == FPMHIP_LOSS_HUBER_AMPLITUDE) {
    const float a = fabsf(r);
    return a <= 1.0f ? 0.5f * r * r : a - 0.5f;
  }
  return r * r;
}

// Global-memory access policies. Plain: ordinary cached loads and stores
// (a tile owned by one workgroup). Coherent: relaxed device-scope atomic
// loads and stores (64-bit each), which the hardware serves from the
// device-coherent level, so that workgroups of one cooperative tile group on
// different CUs / XCDs see each other's writes after a group barrier without
// invalidating or writing back whole caches.
// A pointer read from a tile descriptor, as a global-memory pointer: its
// accesses become global ones, not flat ones (a flat access also counts
// against the LDS counter, so waiting for LDS data would wait for it too).
template <class P>
__device__ __forceinline__ const __attribute__((address_space(1))) P* gptr(const P* p) {
  return (const __attribute__((address_space(1))) P*)p;
}
template <class P>
__device__ __forceinline__ __attribute__((address_space(1))) P* gptr_mut(P* p) {
  return (__attribute__((address_space(1))) P*)p;
}

struct PlainMem {
  static constexpr bool kCoherent = false;
  // component-wise through global pointers (the pair is one wide access)
  __device__ static __forceinline__ double2 ld(const double2* p) {
    const auto q = gptr(reinterpret_cast<const double*>(p));
    return make_double2(q[0], q[1]);
  }
  __device__ static __forceinline__ void st(double2* p, double2 v) {
    const auto q = gptr_mut(reinterpret_cast<double*>(p));
    q[0] = v.x;
    q[1] = v.y;
  }
  __device__ static __forceinline__ float2 ld(const float2* p) {
    const auto q = gptr(reinterpret_cast<const float*>(p));
    return make_float2(q[0], q[1]);
  }
  __device__ static __forceinline__ void st(float2* p, float2 v) {
    const auto q = gptr_mut(reinterpret_cast<float*>(p));
    q[0] = v.x;
    q[1] = v.y;
  }
};

struct CoherentMem {
  static constexpr bool kCoherent = true;
  __device__ static __forceinline__ uint64_t ld64(const void* p) {
    return __hip_atomic_load(reinterpret_cast<uint64_t*>(const_cast<void*>(p)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static __forceinline__ void st64(void* p, uint64_t v) {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static __forceinline__ double2 ld(const double2* p) {
    return make_double2(__longlong_as_double((long long)ld64(p)),
                        __longlong_as_double((long long)ld64(reinterpret_cast<const double*>(p) + 1)));
  }
  __device__ static __forceinline__ void st(double2* p, double2 v) {
    st64(p, (uint64_t)__double_as_longlong(v.x));
    st64(reinterpret_cast<double*>(p) + 1, (uint64_t)__double_as_longlong(v.y));
  }
  __device__ static __forceinline__ float2 ld(const float2* p) {
    const uint64_t u = ld64(p);
    return make_float2(__uint_as_float((uint32_t)u), __uint_as_float((uint32_t)(u >> 32)));
  }
  __device__ static __forceinline__ void st(float2* p, float2 v) {
    st64(p, (uint64_t)__float_as_uint(v.x) | ((uint64_t)__float_as_uint(v.y) << 32));
  }
};

// Raw buffer accesses of one array (ranged by its byte size): a lane whose
// byte offset is kSkip (beyond every range) loads zero without touching
// memory and its store is dropped. Conditional accesses written this way
// need no branch, so all of a thread's loads can be in flight at once
// instead of one basic block (and one memory round trip) per condition.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t kSkip = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle

__device__ __forceinline__ Rsrc make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}
template <int AUX = 0>
__device__ __forceinline__ double2 bld_d2(Rsrc r, uint32_t byte_off) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  const v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, AUX);
  return make_double2(__hiloint2double(v.y, v.x), __hiloint2double(v.w, v.z));
}
template <int AUX = 0>
__device__ __forceinline__ void bst_d2(Rsrc r, uint32_t byte_off, double2 d) {
  typedef int v4i __attribute__((ext_vector_type(4)));
  v4i v;
  v.x = __double2loint(d.x);
  v.y = __double2hiint(d.x);
  v.z = __double2loint(d.y);
  v.w = __double2hiint(d.y);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, byte_off, 0, AUX);
}
__device__ __forceinline__ float bld_u16f(Rsrc r, uint32_t byte_off) {
  return (float)(unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, byte_off, 0, 0);
}
__device__ __forceinline__ float bld_f64f(Rsrc r, uint32_t byte_off) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  const v2i v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, 0);
  return (float)__hiloint2double(v.y, v.x);
}
template <int AUX = 0>
__device__ __forceinline__ void bst_f2(Rsrc r, uint32_t byte_off, float2 f) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  v2i v;
  v.x = __float_as_int(f.x);
  v.y = __float_as_int(f.y);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, byte_off, 0, AUX);
}
template <int AUX = 0>
__device__ __forceinline__ float2 bld_f2(Rsrc r, uint32_t byte_off) {
  typedef int v2i __attribute__((ext_vector_type(2)));
  const v2i v = __builtin_amdgcn_raw_buffer_load_b64(r, byte_off, 0, AUX);
  return make_float2(__int_as_float(v.x), __int_as_float(v.y));
}

// Cache policy bits of a buffer access under a memory policy: the coherent
// policy's accesses carry sc1, as its device-scope atomics do.
template <class MEM>
constexpr int kAux = MEM::kCoherent ? 16 : 0;

// Measured value of one pixel (uint16 counts or f64).
__device__ __forceinline__ float measured(const TileDesc& T, uint64_t i) {
  return T.is_u16 ? (float)gptr(static_cast<const uint16_t*>(T.frames))[i]
                  : (float)gptr(static_cast<const double*>(T.frames))[i];
}

// Amplitude replacement of pixel p of frame f (fbase = f * NN) kept as the
// field change psi' - psi (common.rs:172-228); adds the pixel's objective
// term unless it is masked (a masked pixel keeps psi' = psi).
__device__ __forceinline__ float2 replace_amplitude(const TileDesc& T, uint64_t fbase, uint64_t NN, int p, float2 fv,
                                                    float y, float gain_f, float inv_gain, float eps_f, float& lsum,
                                                    int& lcnt) {
  if (T.masks && gptr(T.masks)[fbase + p] == 0) return f2(0.0f, 0.0f);
  const float pred = fmaf(fv.x, fv.x, fv.y * fv.y);
  const float b = T.bg_len == 0 ? 0.0f : (float)gptr(T.bg)[T.bg_len == NN ? (uint64_t)p : fbase + p];
  lsum += point_loss_f(fmaf(gain_f, pred, b), y, T.loss);
  lcnt += 1;
  const float target = fmaxf((y - b) * inv_gain, 0.0f);
  if (pred > eps_f) {
    const float rm1 = fsqrt(target * frcp(pred)) - 1.0f;
    return f2(fv.x * rm1, fv.y * rm1);
  }
  return f2(fsqrt(target) - fv.x, -fv.y);
}

// mPIE momentum event over the tile's whole R^2 spectrum (common.rs:308-387):
// V <- friction V + (O - A), O <- O + feedback V, A <- O, evaluated in FP64
// with V and A stored as complex64; workgroup `member` of `members` takes
// every members-th block of elements.
__device__ __forceinline__ void momentum_element(const TileDesc& T, double2& o, float2& v, float2& an, bool& bad) {
  const double2 nv = make_double2(fma(T.friction, (double)v.x, o.x - (double)an.x),
                                  fma(T.friction, (double)v.y, o.y - (double)an.y));
  const double2 no = T.feedback == 0.0 ? o : make_double2(fma(T.feedback, nv.x, o.x), fma(T.feedback, nv.y, o.y));
  bad |= !isfinite(nv.x) || !isfinite(nv.y) || !isfinite(no.x) || !isfinite(no.y);
  o = no;
  v = f2((float)nv.x, (float)nv.y);
  an = f2((float)no.x, (float)no.y);
}

template <class MEM = PlainMem>
__device__ __forceinline__ void momentum_event(const TileDesc& T, int tid, uint32_t member = 0,
                                               uint32_t members = 1) {
  constexpr int AUX = kAux<MEM>;
  const uint32_t M = T.rh * T.rw;
  bool bad = false;
  constexpr int U = 4;  // chunks in flight per wave before their use
  // branch-free: pixels past a chunk load zero and store nothing
  const Rsrc rs = make_rsrc(T.spec, M * 16u), rv = make_rsrc(T.vel, M * 8u), ra = make_rsrc(T.anc, M * 8u);
  // Only the pixels an LED update can change, as chunks of up to 64 pixels
  // of one row (TileDesc::mspan): one chunk per wave and step, U chunks in
  // flight. Elsewhere the velocity stays zero and the object equals the
  // anchor, so the event would leave them unchanged (common.rs:355-387).
  const auto span = gptr(T.mspan);
  const uint32_t nspan = __builtin_amdgcn_readfirstlane(span[0]);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = members * (kThreads / 64);
  const uint32_t wave = __builtin_amdgcn_readfirstlane(member * (kThreads / 64) + (threadIdx.x >> 6));
  for (uint32_t c0 = wave * U; c0 < nspan; c0 += waves * U) {
    double2 o[U];
    float2 v[U], an[U];
    uint32_t idx[U];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t c = c0 + u;
      const uint32_t e = span[1 + (c < nspan ? c : nspan - 1)];
      idx[u] = (e & 0x1FFFFFFu) + lane;
      live[u] = c < nspan && lane <= (e >> 25);
      o[u] = bld_d2<AUX>(rs, live[u] ? idx[u] * 16u : kSkip);
      v[u] = bld_f2<AUX>(rv, live[u] ? idx[u] * 8u : kSkip);
      an[u] = bld_f2<AUX>(ra, live[u] ? idx[u] * 8u : kSkip);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      momentum_element(T, o[u], v[u], an[u], bad);
      bst_f2<AUX>(rv, live[u] ? idx[u] * 8u : kSkip, v[u]);
      bst_d2<AUX>(rs, live[u] ? idx[u] * 16u : kSkip, o[u]);
      bst_f2<AUX>(ra, live[u] ? idx[u] * 8u : kSkip, an[u]);
    }
  }
  if (bad) T.status[0] = 1;
  __syncthreads();
}

// End of a tile's sweep: iteration objective into the trace, counters back.
__device__ __forceinline__ void finish_sweep(const TileDesc& T, double obj_sum, double w_sum, uint64_t counter) {
  if (threadIdx.x == 0) {
    const uint32_t it = *T.iteration;
    if (it < T.trace_len) T.trace[it] = w_sum > 0.0 ? obj_sum / w_sum : __builtin_nan("");
    *T.iteration = it + 1;
    *T.counter = counter;
  }
}

// Whether column c lies in a row's support extent (lo | hi << 16).
__device__ __forceinline__ bool in_extent(uint32_t e, int c) {
  return c >= (int)(e & 0xFFFFu) && c < (int)(e >> 16);
}

// Upper bound of |O|^2 over the spectrum rows r0 .. r1 and columns c0 .. c1
// (inclusive, clamped to the spectrum) from the block bounds (T.bounds), read
// as relaxed workgroup-scope atomics (the maximum with 0 compiles to a plain
// load). The raises of this workgroup's earlier inserts (atomic maxima of the
// same scope) are seen: the workgroup barriers between an insert and these
// reads order them (one CU; its vector L1 is coherent for the workgroup).
__device__ __forceinline__ double block_bound(const TileDesc& T, uint32_t r0, uint32_t r1, uint32_t c0, uint32_t c1) {
  r1 = r1 < T.rh ? r1 : T.rh - 1;  // a tap row or column past the spectrum is never read
  c1 = c1 < T.rw ? c1 : T.rw - 1;
  uint64_t m = 0;
  for (uint32_t br = r0 / kBoundBlock; br <= r1 / kBoundBlock; ++br)
    for (uint32_t bc = c0 / kBoundBlock; bc <= c1 / kBoundBlock; ++bc) {
      const uint64_t v =
          __hip_atomic_fetch_max(T.bounds + br * T.bcols + bc, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      m = v > m ? v : m;
    }
  return __longlong_as_double((long long)m);
}

// max |patch|^2 over the whole n x n bilinear patch of a source (the EPRY
// pupil-step denominator, common.rs:271-277), in FP64: thread t takes column
// t % N and N^2 / kThreads consecutive rows from row (t / N) N^2 / kThreads,
// tap rows shared between consecutive pixels, in groups of G rows. With
// block bounds (T.bounds), a group whose taps' blocks cannot exceed `lower`
// (|patch|^2 of a patch pixel, so a lower bound of the maximum) is skipped:
// a bilinear value's magnitude is at most its largest tap's (non-negative
// weights summing to one), and the 1e-14 relative margin covers the FP64
// rounding of both sides, so the maximum is the full pass's, bit for bit.
template <int N>
__device__ __forceinline__ double patch_max(const TileDesc& T, Rsrc rspec, const SrcParam& sp, uint32_t rw, int tid,
                                            double lower) {
  constexpr int PP = N * N / kThreads, G = PP < 8 ? PP : 8;
  static_assert(PP >= 1 && PP % G == 0, "patch rows per thread");
  const int c = tid % N, r0 = (tid / N) * PP;
  const bool fc = sp.ac.upper != sp.ac.lower, fr = sp.ar.upper != sp.ar.lower;
  const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
               wuu = sp.ar.wu * sp.ac.wu;
  const double2 zero = make_double2(0.0, 0.0);
  double m = 0.0;
#pragma unroll
  for (int g = 0; g < PP; g += G) {
    const uint32_t o = (sp.ar.lower + (uint32_t)(r0 + g)) * rw + sp.ac.lower + (uint32_t)c;
    bool skip = false;
    if (T.bounds) {
      const uint32_t rr = sp.ar.lower + (uint32_t)(r0 + g), cc = sp.ac.lower + (uint32_t)c;
      skip = block_bound(T, rr, rr + G, cc, cc + 1) * (1.0 + 1e-14) <= lower;
    }
    double2 t0[G + 1], t1[G + 1];
#pragma unroll
    for (int j = 0; j <= G; ++j) {
      const bool need = !skip && (j < G || fr);
      t0[j] = bld_d2(rspec, need ? (o + j * rw) * 16u : kSkip);
      t1[j] = bld_d2(rspec, need && fc ? (o + j * rw + 1) * 16u : kSkip);
    }
    if (!skip) {
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const double2 ul = fr ? t0[j + 1] : zero, uu = fr ? t1[j + 1] : zero;
        const double2 v = make_double2(fma(uu.x, wuu, fma(ul.x, wul, fma(t1[j].x, wlu, t0[j].x * wll))),
                                       fma(uu.y, wuu, fma(ul.y, wul, fma(t1[j].y, wlu, t0[j].y * wll))));
        m = fmax(m, v.x * v.x + v.y * v.y);
      }
    }
  }
  return m;
}

// Exact block bounds of one tile: max |O|^2 over each kBoundBlock x
// kBoundBlock block of the spectrum (partial blocks at the edges), as FP64
// bit patterns. Wave w takes block rows w, w + waves, ...; lane l the columns
// c0 + l of a 64-column strip, the 8 rows of the block row, then a maximum
// over the 8 lanes of each block.
__global__ void __launch_bounds__(kThreads) k_patch_bounds(const TileDesc* __restrict__ tiles) {
  const TileDesc& T = tiles[blockIdx.x];
  if (!T.bounds || T.member != 0) return;
  const uint32_t rh = T.rh, rw = T.rw, brows = (rh + kBoundBlock - 1) / kBoundBlock;
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, waves = kThreads / 64;
  const Rsrc rs = make_rsrc(T.spec, rh * rw * 16u);
  for (uint32_t br = wave; br < brows; br += waves)
    for (uint32_t c0 = 0; c0 < rw; c0 += 64) {
      const uint32_t c = c0 + lane;
      double m = 0.0;
#pragma unroll
      for (uint32_t j = 0; j < kBoundBlock; ++j) {
        const uint32_t r = br * kBoundBlock + j;
        const double2 v = bld_d2(rs, (r < rh && c < rw) ? (r * rw + c) * 16u : kSkip);
        m = fmax(m, v.x * v.x + v.y * v.y);
      }
#pragma unroll
      for (int off = 1; off < (int)kBoundBlock; off <<= 1) m = fmax(m, __shfl_xor(m, off, 64));
      if (lane % kBoundBlock == 0 && c < rw)
        T.bounds[br * T.bcols + c / kBoundBlock] = (uint64_t)__double_as_longlong(m);
    }
}

// Bilinear patch x pupil (crop.rs:129-167) over the tile's list of support
// strips `sl` (count ns; kGatherStrip vertically consecutive pixels of one
// column, centred patch coordinates, with a mask of the pixels in the
// support; `en` is entry tid, loaded by the caller), written to the N x N
// LDS field at the ifftshifted position; the rest of the field is zero
// (written first, then a barrier). Consecutive
// pixels of a strip share a row of taps: S + 1 rows of two spectrum taps and
// S pupil values per strip, all independent loads issued together. An axis
// without a fraction has an upper weight of exactly zero and its upper taps
// are neither loaded nor used (they contribute zero).
template <int N, bool EPRY = false>
__device__ __forceinline__ double gather_strips(Rsrc rspec, Rsrc rpup, const SrcParam& sp, uint32_t rw, int tid,
                                                const uint32_t* sl, uint32_t ns, uint32_t en, float2* lds,
                                                Rsrc rpatch) {
  constexpr int LD = pitch<N>(), HALF = N / 2, NN = N * N, S = (int)kGatherStrip;
  static_assert(NN % kThreads == 0, "pixels per thread");
  const Rsrc rl = make_rsrc(sl, ns * 4u);
  {
    const int row0 = tid / N, col = tid % N;
#pragma unroll
    for (int j = 0; j < NN / kThreads; ++j) lds[(row0 + j * (kThreads / N)) * LD + col] = f2(0.0f, 0.0f);
  }
  __syncthreads();
  const bool fc = sp.ac.upper != sp.ac.lower, fr = sp.ar.upper != sp.ar.lower;
  const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
               wuu = sp.ar.wu * sp.ac.wu;
  const double2 zero = make_double2(0.0, 0.0);
  double vmax = 0.0;  // EPRY: max |patch|^2 over this thread's support pixels
  for (uint32_t i = (uint32_t)tid; i < ns; i += kThreads) {
    const uint32_t e = i == (uint32_t)tid ? en : (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rl, i * 4u, 0, 0);
    const uint32_t r0 = e & 0x3FFu, c = (e >> 10) & 0x3FFu, mask = e >> 20;
    const uint32_t o = (sp.ar.lower + r0) * rw + sp.ac.lower + c;
    double2 t0[S + 1], t1[S + 1];
    float2 P[S];
#pragma unroll
    for (int j = 0; j <= S; ++j) {
      // tap row j serves pixel j (lower row) and, with a row fraction, pixel j - 1
      const bool need = (j < S && (mask >> j & 1u)) || (fr && j > 0 && (mask >> (j - 1) & 1u));
      t0[j] = bld_d2(rspec, need ? (o + j * rw) * 16u : kSkip);
      t1[j] = bld_d2(rspec, need && fc ? (o + j * rw + 1) * 16u : kSkip);
    }
#pragma unroll
    for (int j = 0; j < S; ++j) P[j] = bld_f2(rpup, (mask >> j & 1u) ? ((r0 + j) * N + c) * 8u : kSkip);
    const int cc = (c + HALF) % N;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const double2 ul = fr ? t0[j + 1] : zero, uu = fr ? t1[j + 1] : zero;
      const double2 v = make_double2(fma(uu.x, wuu, fma(ul.x, wul, fma(t1[j].x, wlu, t0[j].x * wll))),
                                     fma(uu.y, wuu, fma(ul.y, wul, fma(t1[j].y, wlu, t0[j].y * wll))));
      const float2 pv = f2((float)v.x, (float)v.y);
      if (mask >> j & 1u) lds[((r0 + j + HALF) % N) * LD + cc] = fmul(pv, P[j]);
      // EPRY keeps the patch of the support pixels for the pupil gradient
      if (EPRY) {
        bst_f2(rpatch, (mask >> j & 1u) ? ((r0 + j) * N + c) * 8u : kSkip, pv);
        if (mask >> j & 1u) vmax = fmax(vmax, v.x * v.x + v.y * v.y);
      }
    }
  }
  return vmax;
}

// gather_strips for one row chunk of a large tile: the strips of list
// entries [s0, s1) (the bands covering centred rows R0 .. R0 + nk - 1),
// restricted to those rows, written to an LDS row chunk (pitch LD; chunk row
// r - R0, ifftshifted column); the chunk's rows are zeroed first (then a
// barrier). AUX: cache policy of the spectrum reads.
template <int N, int LD, int AUX>
__device__ __forceinline__ void gather_strips_rows(Rsrc rspec, Rsrc rpup, const SrcParam& sp, uint32_t rw, int tid,
                                                   const uint32_t* sl, uint32_t s0, uint32_t s1, int R0, int nk,
                                                   float2* lds) {
  constexpr int HALF = N / 2, S = (int)kGatherStrip;
  for (int i = tid; i < nk * N; i += kThreads) lds[(i / N) * LD + i % N] = f2(0.0f, 0.0f);
  __syncthreads();
  const Rsrc rl = make_rsrc(sl, s1 * 4u);
  const bool fc = sp.ac.upper != sp.ac.lower, fr = sp.ar.upper != sp.ar.lower;
  const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
               wuu = sp.ar.wu * sp.ac.wu;
  const double2 zero = make_double2(0.0, 0.0);
  for (uint32_t i = s0 + (uint32_t)tid; i < s1; i += kThreads) {
    const uint32_t e = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rl, i * 4u, 0, 0);
    const int r0 = (int)(e & 0x3FFu), c = (int)((e >> 10) & 0x3FFu);
    uint32_t mask = e >> 20;
#pragma unroll
    for (int j = 0; j < S; ++j)  // rows of other chunks
      if (r0 + j < R0 || r0 + j >= R0 + nk) mask &= ~(1u << j);
    const uint32_t o = (sp.ar.lower + r0) * rw + sp.ac.lower + c;
    double2 t0[S + 1], t1[S + 1];
    float2 P[S];
#pragma unroll
    for (int j = 0; j <= S; ++j) {
      const bool need = (j < S && (mask >> j & 1u)) || (fr && j > 0 && (mask >> (j - 1) & 1u));
      t0[j] = bld_d2<AUX>(rspec, need ? (o + j * rw) * 16u : kSkip);
      t1[j] = bld_d2<AUX>(rspec, need && fc ? (o + j * rw + 1) * 16u : kSkip);
    }
#pragma unroll
    for (int j = 0; j < S; ++j) P[j] = bld_f2(rpup, (mask >> j & 1u) ? ((r0 + j) * N + c) * 8u : kSkip);
    const int cs = c ^ HALF;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const double2 ul = fr ? t0[j + 1] : zero, uu = fr ? t1[j + 1] : zero;
      const double2 v = make_double2(fma(uu.x, wuu, fma(ul.x, wul, fma(t1[j].x, wlu, t0[j].x * wll))),
                                     fma(uu.y, wuu, fma(ul.y, wul, fma(t1[j].y, wlu, t0[j].y * wll))));
      if (mask >> j & 1u) lds[(r0 + j - R0) * LD + cs] = fmul(f2((float)v.x, (float)v.y), P[j]);
    }
  }
}

// Adjoint insert of one patch update U (crop.rs:188-256) over the tile's
// list of live destinations `dl` (count nd; a | b << 16 in row-major order,
// (a, b) in the (n+1)^2 window): each destination sums its <= 4 weighted taps
// of U (u(row, col) in centred patch coordinates) and is read-modified-
// written in FP64 when that sum is non-zero. The list holds every window
// pixel that a tap of the support can reach, so no live destination is
// missed and no thread iterates over dead ones.
constexpr int kInsertGroup = 5;  // destinations per thread in flight

// The first group of a thread's destination entries (issued ahead of the
// insert, so that its read-modify-writes need not wait for them).
__device__ __forceinline__ void insert_list_first(const uint32_t* dl, uint32_t nd, int tid,
                                                  uint32_t (&en)[kInsertGroup]) {
  const Rsrc rd = make_rsrc(dl, nd * 4u);
#pragma unroll
  for (int k = 0; k < kInsertGroup; ++k) {
    const uint32_t i = (uint32_t)tid + k * kThreads;
    en[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rd, i < nd ? i * 4u : kSkip, 0, 0);
  }
}

// With `bounds` (EPRY's block bounds of |O|^2, bcols blocks per spectrum
// row), every written value raises its block's bound when it exceeds it.
template <int N, class MEM = PlainMem, class GetU>
__device__ __forceinline__ void insert_list(double2* __restrict__ spec, uint32_t spec_elems, const SrcParam& sp,
                                            uint32_t rw, double scale, int tid, const uint32_t* dl, uint32_t nd,
                                            uint32_t (&en)[kInsertGroup], GetU u, uint64_t* bounds = nullptr,
                                            uint32_t bcols = 0) {
  constexpr int AUX = kAux<MEM>;
  const bool fr = sp.ar.upper != sp.ar.lower, fc = sp.ac.upper != sp.ac.lower;
  const float wll = (float)(sp.ar.wl * sp.ac.wl), wlu = (float)(sp.ar.wl * sp.ac.wu),
              wul = (float)(sp.ar.wu * sp.ac.wl), wuu = (float)(sp.ar.wu * sp.ac.wu);
  const Rsrc rs = make_rsrc(spec, spec_elems * 16u), rd = make_rsrc(dl, nd * 4u);
  constexpr int UNROLL = kInsertGroup;
  // list entries of the next group are loaded with the current group's reads
  for (uint32_t base = (uint32_t)tid; base < nd; base += UNROLL * kThreads) {
    uint32_t e[UNROLL];
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
      e[k] = en[k];
      const uint32_t i = base + (UNROLL + k) * kThreads;
      en[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rd, i < nd ? i * 4u : kSkip, 0, 0);
    }
    float2 add[UNROLL];
    uint32_t off[UNROLL];
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
      const int a = (int)(e[k] & 0xFFFFu), b = (int)(e[k] >> 16);
      float ax = 0.0f, ay = 0.0f;
      auto tap = [&](int sr_, int sc_, float wt) {
        if (sr_ >= 0 && sr_ < N && sc_ >= 0 && sc_ < N) {
          const float2 v = u(sr_, sc_);
          ax = fmaf(v.x, wt, ax);
          ay = fmaf(v.y, wt, ay);
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
      add[k] = f2(ax, ay);
      const bool hit = base + k * kThreads < nd && (ax != 0.0f || ay != 0.0f);
      off[k] = hit ? ((sp.ar.lower + (uint32_t)a) * rw + sp.ac.lower + (uint32_t)b) * 16u : kSkip;
    }
    double2 old[UNROLL];
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) old[k] = bld_d2<AUX>(rs, off[k]);
#pragma unroll
    for (int k = 0; k < UNROLL; ++k) {
      const double2 nv = make_double2(fma(scale, (double)add[k].x, old[k].x), fma(scale, (double)add[k].y, old[k].y));
      bst_d2<AUX>(rs, off[k], nv);
      if (bounds && off[k] != kSkip) {
        const uint32_t px = off[k] / 16u;  // spectrum pixel
        uint64_t* bp = bounds + (px / rw / kBoundBlock) * bcols + (px % rw) / kBoundBlock;
        const uint64_t n2 = (uint64_t)__double_as_longlong(nv.x * nv.x + nv.y * nv.y);
        // a stale (lower) cached bound only costs a redundant atomic
        if (n2 > *bp) __hip_atomic_fetch_max(bp, n2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
}

// Per-LED record staged in LDS, as 32-bit words: the source's crop
// parameters (SrcParam), the frame weight and gain, and the frame index.
// Each LED's record is fetched from global memory during the previous LED
// (double buffer), so no LED starts with a wait on global loads.
constexpr int kRecSrc = (int)(sizeof(SrcParam) / 4);
constexpr int kRecWords = kRecSrc + 5;  // + weight (2), gain (2), frame (1)
constexpr int kRecPitch = 24;           // words per slot
static_assert(sizeof(SrcParam) % 4 == 0 && kRecWords <= kRecPitch, "LED record");

__device__ __forceinline__ uint32_t frame_at(const TileDesc& T, uint32_t pos) {
  return T.order ? gptr(T.order)[pos] : pos;
}
// Word t (< kRecWords) of the record of frame f.
__device__ __forceinline__ uint32_t rec_word(const TileDesc& T, uint32_t f, int t) {
  if (t < kRecSrc) return gptr(reinterpret_cast<const uint32_t*>(T.src + f))[t];
  if (t < kRecSrc + 2) return gptr(reinterpret_cast<const uint32_t*>(T.weights + f))[t - kRecSrc];
  if (t < kRecSrc + 4) return gptr(reinterpret_cast<const uint32_t*>(T.gains + f))[t - kRecSrc - 2];
  return f;
}
__device__ __forceinline__ double rec_double(const uint32_t* r, int t) {
  return __hiloint2double((int)__builtin_amdgcn_readfirstlane(r[t + 1]), (int)__builtin_amdgcn_readfirstlane(r[t]));
}
__device__ __forceinline__ SrcParam rec_src(const uint32_t* r) {
  SrcParam sp;
  uint32_t* w = reinterpret_cast<uint32_t*>(&sp);
#pragma unroll
  for (int t = 0; t < kRecSrc; ++t) w[t] = r[t];
  return sp;
}

// Amplitude replacement and loss term of one pixel of an unmasked,
// background-free amplitude-MSE frame (common.rs:123-228), as the field
// change psi' - psi, with two transcendental operations per pixel:
// sy = sqrt(max(y, 0)) and rp = 1/sqrt(pred), so that sqrt(g pred) =
// sqrt(g) pred rp and sqrt(max(y/g, 0) / pred) = sy sqrt(1/g) rp
// (sg = sqrt(g), sig = sqrt(1/g)).
__device__ __forceinline__ float2 replace_plain(float2 fv, float y, float sg, float sig, float eps_f, float& lsum) {
  const float pred = fmaf(fv.x, fv.x, fv.y * fv.y);
  const float sy = fsqrt(fmaxf(y, 0.0f));
  const float rp = __builtin_amdgcn_rsqf(fmaxf(pred, 1e-30f));
  const float r = sg * (pred * rp) - sy;  // point_loss with b = 0
  lsum += r * r;
  const float a = sy * sig;
  const float qm1 = a * rp - 1.0f;
  return pred > eps_f ? f2(fv.x * qm1, fv.y * qm1) : f2(a - fv.x, -fv.y);
}

// One sweep of one tile. The support pixels are handled in strips of
// kGatherStrip vertically consecutive pixels (the tile's strip list), one
// strip per thread: the gather writes each pixel's patch x pupil at its
// ifftshifted (natural) slot and the update replaces the transform there by
// U; for even N, fftshift and ifftshift are both the swap
// ((r + N/2) mod N, (c + N/2) mod N).
template <int N, int RADIX, bool EPRY = false>
__device__ __forceinline__ void sweep_tile(const TileDesc& T, float2* lds, const float2* twf, const float2* twi,
                                           double* red, const uint32_t* ext, uint32_t* rec) {
  constexpr int NN = N * N;
  constexpr int HALF = N / 2;
  constexpr int LD = pitch<N>();
  const float inv_n = 1.0f / (float)NN;
  double2* __restrict__ spec = T.spec;
  float2* pupil = T.pupil;
  const double eps = T.eps;
  const float eps_f = (float)eps;
  // EPRY pupil recovery: the pupil changes every LED, and so do max |P|^2
  // and the Global denominator made from it (common.rs:198-205)
  constexpr bool epry = EPRY;
  const Rsrc rpatch = make_rsrc(T.patch, EPRY ? NN * 8u : 0u);
  const Rsrc rp64 = make_rsrc(T.pupil64, EPRY ? NN * 16u : 0u);
  double maxp = T.max_p;
  Denominator den = make_denominator(T, fmax(maxp, eps));
  const Rsrc rspec = make_rsrc(spec, T.rh * T.rw * 16u), rpup = make_rsrc(pupil, NN * 8u);
  // the common case: every pixel counts, no background, amplitude MSE
  const bool plain_frames = T.masks == nullptr && T.bg_len == 0 && T.loss == FPMHIP_LOSS_AMPLITUDE_MSE;
  const uint32_t ndest = __builtin_amdgcn_readfirstlane(gptr(T.dest)[0]);
  const uint32_t nstrip = __builtin_amdgcn_readfirstlane(gptr(T.dest)[1]);
  const uint32_t* slist = T.dest + 2 + (N + 1) * (N + 1);
  // this thread's support strip (the same for every LED): gather and update
  const uint32_t my_strip = threadIdx.x < nstrip ? gptr(slist)[threadIdx.x] : 0u;
  uint64_t counter = *T.counter;
  // Rows of the transforms that carry data: the centred rows rlo..rhi with
  // pupil support, i.e. the field rows (rlo + N/2 + i) mod N, i < nl (the
  // patch x pupil is zero elsewhere and U is read only there).
  LiveRows live;
  {
    int rlo = N, rhi = -1;
    for (int r = 0; r < N; ++r) {
      const uint32_t e = ext[r];
      if ((e & 0xFFFFu) < (e >> 16)) {
        rlo = rlo < r ? rlo : r;
        rhi = r;
      }
    }
    const int nl = rhi >= rlo ? rhi - rlo + 1 : 0;
    live.nl = __builtin_amdgcn_readfirstlane(nl);
    live.l0 = __builtin_amdgcn_readfirstlane((rlo + HALF) & (N - 1));
    live.rcp = live.nl ? 1.0f / (float)live.nl : 0.0f;
  }
  if (threadIdx.x < kRecWords) rec[threadIdx.x] = rec_word(T, frame_at(T, 0), threadIdx.x);
  __syncthreads();
  uint32_t slot = 0;
  for (uint32_t sweep = 0; sweep < T.sweeps; ++sweep) {
    double obj_sum = 0.0, w_sum = 0.0;
    if (epry) {  // max |P|^2 of the pupil this sweep starts from (the gauge rescales it)
      double m = 0.0;
      for (int p = opaque_tid(); p < NN; p += kThreads) {
        const float2 P = bld_f2(rpup, (uint32_t)p * 8u);
        m = fmax(m, (double)P.x * (double)P.x + (double)P.y * (double)P.y);
      }
      maxp = wg_max(m, red);
    }
    for (uint32_t pos = 0; pos < T.F; ++pos, slot ^= 1) {
      const uint32_t* r = rec + slot * kRecPitch;
      uint32_t* rn = rec + (slot ^ 1) * kRecPitch;
      const uint32_t f = __builtin_amdgcn_readfirstlane(r[kRecWords - 1]);
      const double fw = rec_double(r, kRecSrc);
      // the next LED: the first of the next sweep after the last
      const bool more = pos + 1 < T.F || sweep + 1 < T.sweeps;
      const uint32_t npos = pos + 1 < T.F ? pos + 1 : 0;
      const bool fetch = more && threadIdx.x < kRecWords;
      const int tid = opaque_tid();
      if (fw == 0.0) {
        if (tid == 0) T.frame_loss[pos] = 0.0;
        if (fetch) rn[threadIdx.x] = rec_word(T, frame_at(T, npos), threadIdx.x);
        __syncthreads();
        continue;
      }
      const SrcParam sp = rec_src(r);
      const double gain = rec_double(r, kRecSrc + 2);
      const float gain_f = (float)gain;
      const float inv_gain = (float)(1.0 / gain);
      const float sg = (float)sqrt(gain), sig = (float)sqrt(1.0 / gain);
      // EPRY: the frame's step denominators, max |P|^2 of the current pupil
      // (Global) and max |patch|^2 over the whole patch (pupil step)
      float pupil_k = 0.0f;
      if (epry) {
        den = make_denominator(T, fmax(maxp, eps));
        den.a = uniform(den.a);
        den.b = uniform(den.b);
      }
      // 1. patch x pupil over the support pixels, written ifftshifted (EPRY:
      //    with each wave's maximum of |patch|^2 over the support pixels in the
      //    third set of reduction slots, read after the barrier)
      {
        double vm = gather_strips<N, EPRY>(rspec, rpup, sp, T.rw, tid, slist, nstrip, my_strip, lds, rpatch);
        if constexpr (EPRY) {
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) vm = fmax(vm, __shfl_xor(vm, off, 64));
          if ((threadIdx.x & 63) == 0) red[kSlotLower + (threadIdx.x >> 6)] = vm;
        }
      }
      __syncthreads();
      double lower = 0.0;
      if constexpr (EPRY) {
        // max |patch|^2 over the whole patch (common.rs:271-277); the support
        // pixels' maximum prunes the pass. Each wave's maximum goes to its
        // slot, read after the barrier of the fused column stage (the pupil
        // step needs it only at the update).
        lower = red[kSlotLower];
        for (int w = 1; w < kWaves; ++w) lower = fmax(lower, red[kSlotLower + w]);
        double mo = patch_max<N>(T, rspec, sp, T.rw, tid, lower);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mo = fmax(mo, __shfl_xor(mo, off, 64));
        if ((threadIdx.x & 63) == 0) red[kSlotPatch + (threadIdx.x >> 6)] = mo;
      }
      // the next LED's frame index, then its record words (written to the
      // other slot after the forward transform)
      uint32_t next_f = 0, next_w = 0;
      if (fetch) next_f = frame_at(T, npos);
      // 2.-4. inverse 2-D FFT, amplitude replacement, forward 2-D FFT, in place:
      //    inverse rows and columns by decimation in frequency (natural in,
      //    digit-reversed out), then the inverse's last column stage, the
      //    replacement and the forward's first column stage on the same
      //    registers, then the forward columns and rows by decimation in time
      //    (back to natural order; x 1/N applied in phase 5). Row stages take
      //    only the live rows.
      constexpr int P0 = RADIX, P1 = N / RADIX;
      static_assert(P0 * P1 == N && P0 <= 16 && P1 <= 16, "plan");
      ip_stage<N, N, P0, true, true, LD, 1, true, true>(lds, twi, tid, live);
      ip_stage<N, N, P1, false, false, LD, 1, true, true>(lds, twi, tid, live);
      if (fetch) next_w = rec_word(T, next_f, threadIdx.x);
      // 3. amplitude replacement and objective terms, pointwise at the
      //    natural pixel (digit_rev(r'), digit_rev(c')) of LDS slot (r', c').
      //    The field becomes psi' - psi: the forward transform is linear, so
      //    FFT(psi' - psi)/N equals FFT(psi')/N - patch P (common.rs:240-245)
      //    without cancelling two nearly equal spectra in FP32. Per-pixel
      //    terms are FP32; the frame objective is accumulated in FP64.
      //    Work item b: column b % N, rows (b / N) P1 .. + P1 - 1.
      constexpr int GROUPS = N * P0;
      constexpr int PG = (GROUPS + kThreads - 1) / kThreads;
      const uint64_t fbase = (uint64_t)f * NN;
      ip_stage<N, N, P0, true, true, 1, LD, true>(lds, twi, tid);
      {
        float lsum_f = 0.0f;
        int lcnt_i = 0;
        if (plain_frames) {
          // the inverse's last column stage, the replacement (branch-free: both
          // cases of the ratio, common.rs:180-190, evaluated and selected) and
          // the forward's first column stage on the same registers
          const int t = opaque_tid();
          const Rsrc rm = T.is_u16 ? make_rsrc(static_cast<const uint16_t*>(T.frames) + fbase, NN * 2u)
                                   : make_rsrc(static_cast<const double*>(T.frames) + fbase, NN * 8u);
#pragma unroll
          for (int u = 0; u < PG; ++u) {
            const int b = t + u * kThreads, c = b % N, j0 = b / N, nc = digit_rev<P0, P1>(c);
            if (b < GROUPS) {
              float y[P1];
              if (T.is_u16) {
#pragma unroll
                for (int m = 0; m < P1; ++m) y[m] = bld_u16f(rm, (uint32_t)((j0 + P0 * m) * N + nc) * 2u);
              } else {
#pragma unroll
                for (int m = 0; m < P1; ++m) y[m] = bld_f64f(rm, (uint32_t)((j0 + P0 * m) * N + nc) * 8u);
              }
              float2 a[P1];
#pragma unroll
              for (int m = 0; m < P1; ++m) a[m] = lds[(j0 * P1 + m) * LD + c];
              reg_dft<P1, true>(a);
#pragma unroll
              for (int m = 0; m < P1; ++m) a[m] = replace_plain(a[m], y[m], sg, sig, eps_f, lsum_f);
              reg_dft<P1, false>(a);
#pragma unroll
              for (int m = 0; m < P1; ++m) lds[(j0 * P1 + m) * LD + c] = m > 0 ? fmul(a[m], twf[j0 * m]) : a[m];
              lcnt_i += P1;
            }
          }
        } else {
          // masks, backgrounds or other losses: the stages and the
          // replacement as separate passes over LDS
          ip_stage<N, N, P1, false, false, 1, LD, true>(lds, twi, tid);
          constexpr int PP = NN / kThreads;  // pixels per thread
          static_assert(NN % kThreads == 0, "pixels per thread");
          const int t = opaque_tid();
#pragma unroll
          for (int g0 = 0; g0 < PP; g0 += 4) {
            float y[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int i = t + (g0 + j) * kThreads;
              y[j] = measured(T, fbase + (uint64_t)(digit_rev<P0, P1>(i / N) * N + digit_rev<P0, P1>(i % N)));
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int i = t + (g0 + j) * kThreads, lp = (i / N) * LD + i % N;
              lds[lp] = replace_amplitude(T, fbase, (uint64_t)NN,
                                          digit_rev<P0, P1>(i / N) * N + digit_rev<P0, P1>(i % N), lds[lp], y[j],
                                          gain_f, inv_gain, eps_f, lsum_f, lcnt_i);
            }
          }
          __syncthreads();
          ip_stage<N, N, P1, false, true, 1, LD, false>(lds, twf, tid);
        }
        // per-thread partials (at most 16 terms) in FP32; wave sums in FP64
        // into the reduction slots, read by thread 0 after the barrier
        wave_sums2_to(red, (double)lsum_f, (double)lcnt_i);
      }
      __syncthreads();
      if constexpr (EPRY) {
        double maxo = lower;
        for (int w = 0; w < kWaves; ++w) maxo = fmax(maxo, red[kSlotPatch + w]);
        pupil_k = uniform((float)(1.0 / (fmax(maxo, eps) + eps)) * inv_n);
      }
      if (tid == 0) {
        double ls = 0.0, lc = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) {
          ls += red[w];
          lc += red[kThreads / 64 + w];
        }
        const double L = lc > 0.0 ? ls / lc : 0.0;
        T.frame_loss[pos] = L;
        obj_sum += fw * L;
        w_sum += fw;
      }
      // 4. the rest of the forward transform
      ip_stage<N, N, P0, true, false, 1, LD, false>(lds, twf, tid);
      ip_stage<N, N, P1, false, true, LD, 1, false, true>(lds, twf, tid, live);
      // the pupil at this thread's first support strip and the first insert
      // destinations, loaded during the last stage (the update and the
      // insert then start without a memory round trip)
      constexpr int S = (int)kGatherStrip;
      float2 Pfirst[S];
      {
        const uint32_t mask = my_strip >> 20, r0 = my_strip & 0x3FFu, c = (my_strip >> 10) & 0x3FFu;
#pragma unroll
        for (int j = 0; j < S; ++j) Pfirst[j] = bld_f2(rpup, (mask >> j & 1u) ? ((r0 + j) * N + c) * 8u : kSkip);
      }
      uint32_t dfirst[kInsertGroup];
      insert_list_first(T.dest + 2, ndest, tid, dfirst);
      ip_stage<N, N, P0, true, false, LD, 1, false, true>(lds, twf, tid, live);
      if (fetch) rn[threadIdx.x] = next_w;
      // 5. U = conj(P) fftshift(FFT(psi' - psi))/N / den at the support pixels,
      //    in place: U of centred pixel (r, c) replaces the transform value at
      //    its natural slot ((r + N/2) mod N, (c + N/2) mod N). Outside the
      //    support the field keeps values that the insert does not read.
      {
        const Rsrc rl = make_rsrc(slist, nstrip * 4u);
        const float pk = uniform((float)(fw * T.pupil_step) * pupil_k);
        double pmax = 0.0;
        for (uint32_t i = (uint32_t)tid; i < nstrip; i += kThreads) {
          const uint32_t e = i == (uint32_t)tid ? my_strip : (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rl, i * 4u, 0, 0);
          const int r0 = (int)(e & 0x3FFu), c = (int)((e >> 10) & 0x3FFu);
          const uint32_t mask = e >> 20;
          const int cn = (c + HALF) % N;
          float2 P[S];
#pragma unroll
          for (int j = 0; j < S; ++j)
            P[j] = i == (uint32_t)tid ? Pfirst[j] : bld_f2(rpup, (mask >> j & 1u) ? ((r0 + j) * N + c) * 8u : kSkip);
          if constexpr (EPRY) {
            // EPRY pupil step at the strip's support pixels (common.rs:271-300):
            // g = conj(patch) diff / (max |patch|^2 + eps), P += w beta g, with
            // the old P in U; the support projection leaves the rest at zero
            // the strip's patch values and FP64 master pupil values, all loads
            // issued before the first store
            float2 pa[S];
            double2 m64[S];
#pragma unroll
            for (int j = 0; j < S; ++j) {
              const bool in = mask >> j & 1u;
              const uint32_t pix = (uint32_t)((r0 + j) * N + c);
              pa[j] = bld_f2(rpatch, in ? pix * 8u : kSkip);
              m64[j] = bld_d2(rp64, in ? pix * 16u : kSkip);
            }
#pragma unroll
            for (int j = 0; j < S; ++j) {
              if (mask >> j & 1u) {
                const int l = ((r0 + j + HALF) % N) * LD + cn;
                const float2 d = lds[l];
                const float k = inv_n / fmaf(den.b, fmaf(P[j].x, P[j].x, P[j].y * P[j].y), den.a);
                const float2 u = fmulc(P[j], d);
                lds[l] = f2(u.x * k, u.y * k);
                const float2 g = fmulc(pa[j], d);
                // the step accumulates into the FP64 master; P is its complex64 copy
                const uint32_t pix = (uint32_t)((r0 + j) * N + c);
                const double2 n64 =
                    make_double2(fma((double)pk, (double)g.x, m64[j].x), fma((double)pk, (double)g.y, m64[j].y));
                bst_d2(rp64, pix * 16u, n64);
                const float2 np = f2((float)n64.x, (float)n64.y);
                bst_f2(rpup, pix * 8u, np);
                pmax = fmax(pmax, (double)np.x * (double)np.x + (double)np.y * (double)np.y);
              }
            }
          } else {
#pragma unroll
            for (int j = 0; j < S; ++j) {
              if (mask >> j & 1u) {
                const int l = ((r0 + j + HALF) % N) * LD + cn;
                const float k = inv_n / fmaf(den.b, fmaf(P[j].x, P[j].x, P[j].y * P[j].y), den.a);
                const float2 u = fmulc(P[j], lds[l]);
                lds[l] = f2(u.x * k, u.y * k);
              }
            }
          }
        }
        if constexpr (EPRY) {  // the next frame's max |P|^2 (zero off the support), read after the barrier
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) pmax = fmax(pmax, __shfl_xor(pmax, off, 64));
          if ((threadIdx.x & 63) == 0) red[kSlotPupil + (threadIdx.x >> 6)] = pmax;
        }
      }
      __syncthreads();
      if constexpr (EPRY) {
        double m = red[kSlotPupil];
        for (int w = 1; w < kWaves; ++w) m = fmax(m, red[kSlotPupil + w]);
        maxp = uniform(m);
      }
      // 6. adjoint insert over the tile's live destinations: FP64 RMW of the
      //    pixels that receive a non-zero update (U is zero outside the support
      //    extents, where it is not read)
      const double scale = fw * T.object_step;
      insert_list<N>(
          spec, T.rh * T.rw, sp, T.rw, scale, tid, T.dest + 2, ndest, dfirst,
          [&](int sr_, int sc_) {
            return in_extent(ext[sr_], sc_) ? lds[((sr_ + HALF) % N) * LD + (sc_ + HALF) % N] : f2(0.0f, 0.0f);
          },
          EPRY ? T.bounds : nullptr, T.bcols);
      __syncthreads();
      // 7. mPIE momentum over the tile spectrum's changeable pixels (common.rs:308-387)
      if (T.momentum) {
        counter += 1;
        if (counter >= T.interval) {
          momentum_event(T, tid);
          counter = 0;
        }
      }
    }
    finish_sweep(T, obj_sum, w_sum, counter);
  }
}

// Rows 0..nk-1 of an LDS row chunk (pitch LD, rows of N) to / from rows
// row_of(j) of a row-major global array with N columns: every load of the
// thread is issued before its first store.
template <int CH, int N, int LD, class MEM, class RowOf>
__device__ __forceinline__ void rows_out(const float2* lds, float2* __restrict__ g, int nk, int tid, RowOf row_of) {
  constexpr int PER = CH * N / kThreads;
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    v[j] = i < nk * N ? lds[(i / N) * LD + i % N] : f2(0.0f, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    if (i < nk * N) MEM::st(g + row_of(i / N) * N + i % N, v[j]);
  }
}

template <int CH, int N, int LD, class MEM, class RowOf>
__device__ __forceinline__ void rows_in(float2* lds, const float2* __restrict__ g, int nk, int tid, RowOf row_of) {
  constexpr int PER = (CH * N + kThreads - 1) / kThreads;
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    v[j] = i < nk * N ? MEM::ld(g + row_of(i / N) * N + i % N) : f2(0.0f, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    if (i < nk * N) lds[(i / N) * LD + i % N] = v[j];
  }
}

// copy_in of a column block of width w <= W whose rows with live(row)
// false read as zero (without loading them); LDS columns w..W-1 are not
// written.
template <int TOTAL, int W, int LD, class MEM, class Live>
__device__ __forceinline__ void copy_in_live(float2* lds, const float2* __restrict__ g, int gstride, int w, int tid,
                                             Live live) {
  constexpr int PER = TOTAL / kThreads;
  static_assert(TOTAL % kThreads == 0, "copy shape");
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    v[j] = (i % W < w && live(i / W)) ? MEM::ld(g + (i / W) * gstride + i % W) : f2(0.0f, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    if (i % W < w) lds[(i / W) * LD + i % W] = v[j];
  }
}

// copy_out of the first w <= W columns of a column block.
template <int TOTAL, int W, int LD, class MEM>
__device__ __forceinline__ void copy_out_cols(const float2* lds, float2* __restrict__ g, int gstride, int w,
                                              int tid) {
  constexpr int PER = TOTAL / kThreads;
  static_assert(TOTAL % kThreads == 0, "copy shape");
  float2 v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    v[j] = lds[(i / W) * LD + i % W];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = tid + j * kThreads;
    if (i % W < w) MEM::st(g + (i / W) * gstride + i % W, v[j]);
  }
}

// Group barrier of a cooperative tile (members > 1). Every access of data
// shared between the members (the field, partial sums, the
// spectrum and momentum arrays of the tile) is device-coherent: relaxed
// agent-scope atomics or buffer accesses with the sc1 policy bit, which the
// hardware serves from the device-coherent level. Each wave first waits for
// its own stores to complete (s_waitcnt vmcnt(0): a completed device-scope
// store is visible to device-scope loads on every CU), then the workgroup
// synchronises (a workgroup-scope fence: the compiler keeps every access of
// the workgroup on its side of it), one thread arrives on the group's
// counter and spins, bounded, until every member has arrived, and a second
// workgroup barrier holds the other threads' accesses until then. The
// arrival and the spin are relaxed, not release / acquire: at agent scope
// those compile to a write-back and an invalidation of the whole L2 of the
// XCD, which the coherent accesses do not need and which cost the uneven
// batch (16 cooperative groups beside 80 single-workgroup tiles) 28 % of
// its sweep time (8.53 vs 6.66 ms per sweep). `passed` counts the group's
// barriers since its state was last reset. Returns false when the group
// stalled (the spin limit expired here or in another member): the caller
// abandons the sweep, so every workgroup still exits.
constexpr int kWaitVmcnt0 = 0x0F70;  // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15)
constexpr uint32_t kGroupSpinLimit = 1u << 20;

__device__ __forceinline__ bool group_barrier(const TileDesc& T, uint64_t& passed, int* bflag) {
  __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
  __syncthreads();
  passed += 1;
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&T.group[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t target = passed * T.members;
    uint32_t spins = 0;
    int ok = 1;
    while (__hip_atomic_load(&T.group[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(&T.group[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
          ++spins > kGroupSpinLimit) {
        __hip_atomic_store(&T.group[2], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        T.status[0] = 2;
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    *bflag = ok;
  }
  __syncthreads();
  return *bflag != 0;
}

// Tiles whose N x N complex64 field does not fit in LDS (N = 256: 512 KiB).
// The field lives in the tile's global scratch buffer T.field (it stays in
// L2 / the last-level cache) and passes through LDS in chunks of at most CH
// lines:
//
//   A  rows with pupil support, split evenly over the chunks: patch x pupil
//      (support strips of the tile's list) written ifftshifted, inverse row
//      transforms, out to their natural rows of T.field
//   B  columns (the rows without support read as zero): inverse column
//      transforms, amplitude replacement kept as the field change psi' - psi,
//      forward column transforms, back to T.field
//   C  the rows of A, in centred order: forward row transforms, fftshift by
//      index, U = conj(P) diff / N / den in LDS, then the adjoint insert of
//      the destination rows this chunk completes (the row above the chunk is
//      the previous chunk's last row of U: kept from that chunk when one
//      workgroup runs them all, transformed once more by a group member)
//
// One workgroup runs the chunks in turn (MEM = PlainMem), or a cooperative
// group of N/CH workgroups takes one chunk each (MEM = CoherentMem): group
// barriers separate A, B and C, and one LED from the next (three per LED,
// one more after a momentum event); per-member objective sums are
// combined by member 0 in member order. The LED loop, objective, momentum
// and arithmetic are those of sweep_tile, and the spectra do not depend on
// the group size.
template <int N>
struct BigPlan {
  static constexpr int CH = (int)big_chunk(N);  // lines per chunk
};

// The line transforms of a large tile: two stages (radix P0, then N / P0)
// for 256, three radix-8 stages for 512 (radices of at most 16 per stage).
template <int N, int LINES, int LS, int ES, bool INV, bool LIVE = false>
__device__ __forceinline__ void big_lines(float2* lds, const float2* tw, int tid, LiveRows lr = {LINES, 0, 0.0f}) {
  if constexpr (N == 512) {
    stage<N, LINES, 8, 1, LS, ES, INV, LIVE>(lds, tw, tid, lr);
    stage<N, LINES, 8, 8, LS, ES, INV, LIVE>(lds, tw, tid, lr);
    stage<N, LINES, 8, 64, LS, ES, INV, LIVE>(lds, tw, tid, lr);
  } else {
    constexpr int P0 = DefaultPlan<N>::P0, P1 = N / P0;
    stage<N, LINES, P0, 1, LS, ES, INV, LIVE>(lds, tw, tid, lr);
    stage<N, LINES, P1, P0, LS, ES, INV, LIVE>(lds, tw, tid, lr);
  }
}

// float2 elements of a large tile's chunk region: CH + 1 rows of N (+1 pad)
// in pass C of a group member, or N rows of CH (+1 pad) in pass B.
template <int N>
__host__ __device__ constexpr int big_field_elems() {
  return (BigPlan<N>::CH + 1) * (N + 1) > N * (BigPlan<N>::CH + 1) ? (BigPlan<N>::CH + 1) * (N + 1)
                                                                     : N * (BigPlan<N>::CH + 1);
}

// One chunk region for every large tile size (a single constant keeps the
// LDS offsets of all sizes one select apart).
constexpr uint32_t kBigFieldElems =
    big_field_elems<256>() > big_field_elems<512>() ? big_field_elems<256>() : big_field_elems<512>();

// float2 elements of the LDS field region for an N x N tile.
__host__ __device__ constexpr uint32_t sweep_field_elems(uint32_t n) {
  return n <= kFusedLdsMaxN ? n * (n + 1) : kBigFieldElems;
}

template <int N, class MEM>
__device__ __forceinline__ void sweep_tile_big(const TileDesc& T, float2* lds, float2* prev, const float2* twf,
                                               const float2* twi, double* red, int* bflag, const uint32_t* ext,
                                               uint32_t* rec) {
  constexpr int NN = N * N;
  constexpr int HALF = N / 2;
  constexpr int CH = BigPlan<N>::CH;
  constexpr int NCH = N / CH;
  constexpr int LDR = N + 1;   // row chunk: CH rows of N (+1 pad)
  constexpr int LDC = CH + 1;  // column chunk: N rows of CH (+1 pad)
  constexpr bool COOP = MEM::kCoherent;
  static_assert((N & (N - 1)) == 0 && HALF % CH == 0, "big tile shape");
  // a cooperative group has one member per chunk; a lone workgroup runs them all
  // a cooperative group has one member per chunk (8..32 members at 256^2,
  // 32 at 512^2); a lone workgroup runs the N / CH chunks in turn
  const int nch = COOP ? (int)__builtin_amdgcn_readfirstlane(T.members) : NCH;
  const int first = COOP ? (int)T.member : 0, step = COOP ? nch : 1;
  const bool lead = !COOP || T.member == 0;
  uint64_t passed = COOP ? T.group[1] : 0;
  const float inv_n = 1.0f / (float)NN;
  double2* __restrict__ spec = T.spec;
  float2* __restrict__ fld = T.field;
  float2* pupil = T.pupil;
  const float eps_f = (float)T.eps;
  const Denominator den = make_denominator(T, fmax(T.max_p, T.eps));
  constexpr int AUX = kAux<MEM>;
  const Rsrc rspec = make_rsrc(spec, T.rh * T.rw * 16u), rpup = make_rsrc(pupil, NN * 8u);
  const bool plain_frames = T.masks == nullptr && T.bg_len == 0 && T.loss == FPMHIP_LOSS_AMPLITUDE_MSE;
  uint64_t counter = *T.counter;
  // The centred rows rlo..rhi carry pupil support: the patch x pupil is zero
  // on the other rows and U is read only on these. Passes A and C split
  // them evenly over the chunks (chunk k: centred rows rlo + k crows, at most
  // crows <= CH of them), so no member spends its pass on rows without data;
  // pass B keeps whole column blocks and reads the other rows as zero.
  int rlo = N, rhi = -1;
  for (int r = 0; r < N; ++r) {
    const uint32_t e = ext[r];
    if ((e & 0xFFFFu) < (e >> 16)) {
      rlo = rlo < r ? rlo : r;
      rhi = r;
    }
  }
  rlo = __builtin_amdgcn_readfirstlane(rlo);
  rhi = __builtin_amdgcn_readfirstlane(rhi);
  const int nl = rhi >= rlo ? rhi - rlo + 1 : 0;
  const int crows = (nl + nch - 1) / nch;  // <= CH
  const int clast = nl ? (nl - 1) / crows : -1;  // last chunk with rows
  auto chunk_rows = [&](int k) { return k <= clast ? min(crows, nl - k * crows) : 0; };
  auto live_rows = [&](int nk) { return LiveRows{nk, 0, nk ? 1.0f / (float)nk : 0.0f}; };
  if (threadIdx.x < kRecWords) rec[threadIdx.x] = rec_word(T, frame_at(T, 0), threadIdx.x);
  __syncthreads();
  uint32_t slot = 0;
  for (uint32_t sweep = 0; sweep < T.sweeps; ++sweep) {
    double obj_sum = 0.0, w_sum = 0.0;
    for (uint32_t pos = 0; pos < T.F; ++pos, slot ^= 1) {
      const uint32_t* r = rec + slot * kRecPitch;
      uint32_t* rn = rec + (slot ^ 1) * kRecPitch;
      const uint32_t f = __builtin_amdgcn_readfirstlane(r[kRecWords - 1]);
      const double fw = rec_double(r, kRecSrc);
      const bool more = pos + 1 < T.F || sweep + 1 < T.sweeps;
      const uint32_t npos = pos + 1 < T.F ? pos + 1 : 0;
      const bool fetch = more && threadIdx.x < kRecWords;
      int tid = opaque_tid();
      if (fw == 0.0) {
        if (lead && tid == 0) T.frame_loss[pos] = 0.0;
        if (fetch) rn[threadIdx.x] = rec_word(T, frame_at(T, npos), threadIdx.x);
        __syncthreads();
        continue;
      }
      const SrcParam sp = rec_src(r);
      const double gain = rec_double(r, kRecSrc + 2);
      const float gain_f = (float)gain;
      const float inv_gain = (float)(1.0 / gain);
      const float sg = (float)sqrt(gain), sig = (float)sqrt(1.0 / gain);
      // the next LED's frame index (during A), then its record words (during
      // B), written to the other slot after B
      uint32_t next_f = 0, next_w = 0;
      if (fetch) next_f = frame_at(T, npos);
      // A. chunk k: centred rows R0 .. R0 + nk - 1 in LDS rows 0 .. nk - 1 (with
      //    the ifftshifted columns), inverse row transforms, out to the natural
      //    rows (R0 + j) ^ HALF of T.field (ifftshift for even N)
      for (int k = first; k < nch; k += step) {
        tid = opaque_tid();
        const int nk = chunk_rows(k), R0 = rlo + k * crows;
        {
          // the support strips of the bands that cover the chunk's rows
          const uint32_t* bands = T.dest + dest_band_starts_offset(N);
          const uint32_t st0 = nk > 0 ? __builtin_amdgcn_readfirstlane(gptr(bands)[R0 / (int)kGatherStrip]) : 0u;
          const uint32_t st1 =
              nk > 0 ? __builtin_amdgcn_readfirstlane(gptr(bands)[(R0 + nk - 1) / (int)kGatherStrip + 1]) : 0u;
          gather_strips_rows<N, LDR, AUX>(rspec, rpup, sp, T.rw, tid, T.dest + dest_strips_offset(N), st0, st1, R0,
                                          nk, lds);
        }
        __syncthreads();
        big_lines<N, CH, LDR, 1, true, true>(lds, twi, tid, live_rows(nk));
        rows_out<CH, N, LDR, MEM>(lds, fld, nk, opaque_tid(), [&](int j) { return (R0 + j) ^ HALF; });
        __syncthreads();
      }
      if (COOP && !group_barrier(T, passed, bflag)) return;
      // B. columns: natural image order after the inverse column transforms
      const uint64_t fbase = (uint64_t)f * NN;
      double lsum = 0.0;
      int lcnt = 0;
      if (fetch) next_w = rec_word(T, next_f, threadIdx.x);
      for (int jc = first; jc < nch; jc += step) {
        tid = opaque_tid();
        // the block's columns c0 .. c0 + w - 1 (w = CH unless the group has more
        // members than N / CH chunks); only its w columns are transformed
        const int c0 = (jc * N) / nch, w = ((jc + 1) * N) / nch - c0;
        copy_in_live<N * CH, CH, LDC, MEM>(lds, fld + c0, N, w, opaque_tid(), [&](int nr) {
          const int cr = nr ^ HALF;  // centred row of natural row nr
          return cr >= rlo && cr <= rhi;
        });
        __syncthreads();
        big_lines<N, CH, 1, LDC, true, true>(lds, twi, tid, live_rows(w));
        if (plain_frames) {
          // branch-free (as in sweep_tile): pixel (row0 + j RSTEP, c0 + col)
          constexpr int PP = N * CH / kThreads, RSTEP = kThreads / CH;
          const int tid = opaque_tid();
          const int row0 = tid / CH, col = tid % CH;
          const bool in = col < w;
          float y[PP];
          const uint32_t p0 = (uint32_t)(row0 * N + c0 + col);
          if (T.is_u16) {
            const Rsrc rm = make_rsrc(static_cast<const uint16_t*>(T.frames) + fbase, NN * 2u);
#pragma unroll
            for (int j = 0; j < PP; ++j) y[j] = bld_u16f(rm, in ? (p0 + j * RSTEP * N) * 2u : kSkip);
          } else {
            const Rsrc rm = make_rsrc(static_cast<const double*>(T.frames) + fbase, NN * 8u);
#pragma unroll
            for (int j = 0; j < PP; ++j) y[j] = bld_f64f(rm, in ? (p0 + j * RSTEP * N) * 8u : kSkip);
          }
          float lsum_f = 0.0f;
          if (in) {
#pragma unroll
            for (int j = 0; j < PP; ++j) {
              const int lp = (row0 + j * RSTEP) * LDC + col;
              const float2 fv = lds[lp];
              lds[lp] = replace_plain(fv, y[j], sg, sig, eps_f, lsum_f);
            }
            lcnt += PP;
          }
          lsum += (double)lsum_f;
        } else {
          // measurement loads in groups of G issued before their use
          constexpr int PP = N * CH / kThreads;
          constexpr int G = 8;
          const int tid = opaque_tid();
          float lsum_f = 0.0f;
#pragma unroll
          for (int g = 0; g < PP; g += G) {
            float y[G];
#pragma unroll
            for (int j = 0; j < G; ++j) {
              const int i = tid + (g + j) * kThreads;
              y[j] = i % CH < w ? measured(T, fbase + (uint64_t)(i / CH) * N + c0 + i % CH) : 0.0f;
            }
#pragma unroll
            for (int j = 0; j < G; ++j) {
              const int i = tid + (g + j) * kThreads;
              const int lp = (i / CH) * LDC + i % CH;
              if (i % CH < w)
                lds[lp] = replace_amplitude(T, fbase, (uint64_t)NN, (i / CH) * N + c0 + i % CH, lds[lp], y[j],
                                            gain_f, inv_gain, eps_f, lsum_f, lcnt);
            }
          }
          lsum += (double)lsum_f;
        }
        __syncthreads();
        big_lines<N, CH, 1, LDC, false, true>(lds, twf, tid, live_rows(w));
        copy_out_cols<N * CH, CH, LDC, MEM>(lds, fld + c0, N, w, opaque_tid());
        __syncthreads();
      }
      if (fetch) rn[threadIdx.x] = next_w;
      {
        const double2 lsc = wg_sum2(lsum, (double)lcnt, red);
        double ls = lsc.x, lc = lsc.y;
        if (COOP) {
          if (tid == 0) MEM::st(T.partial + T.member, make_double2(ls, lc));
          if (!group_barrier(T, passed, bflag)) return;
          if (lead && tid == 0) {
            ls = 0.0;
            lc = 0.0;
            for (uint32_t m = 0; m < T.members; ++m) {
              const double2 pm = MEM::ld(T.partial + m);
              ls += pm.x;
              lc += pm.y;
            }
          }
        }
        if (lead && tid == 0) {
          const double L = lc > 0.0 ? ls / lc : 0.0;
          T.frame_loss[pos] = L;
          obj_sum += fw * L;
          w_sum += fw;
        }
        __syncthreads();
      }
      // C. chunk q: centred rows R0 .. R0 + nk - 1 (natural rows (R0 + j) ^ HALF),
      //    forward row transforms, U, then the insert of destination rows
      //    R0 .. R0 + nk - 1 (the last chunk with rows also takes the row below
      //    its last, for a row fraction). The insert also reads the row of U
      //    above the chunk (zero above the first chunk): a group member
      //    transforms that row (the previous chunk's last) as well, as LDS row
      //    0, so that no row of U passes between members; one workgroup keeps
      //    it from the previous chunk in `prev`.
      const double scale = fw * T.object_step;
      constexpr int CROWS = COOP ? CH + 1 : CH;  // rows in LDS
      constexpr int CLINES = COOP ? 2 * CH : CH;  // power-of-two line count of the live-row stages
      for (int q = first; q < nch; q += step) {
        tid = opaque_tid();
        const int nk = chunk_rows(q), R0 = rlo + q * crows;
        const int ex = (COOP && q > 0 && nk > 0) ? 1 : 0;  // the row above, transformed here
        const int RB = R0 - ex, nb = nk + ex;              // centred rows RB .. RB + nb - 1 in LDS rows 0 .. nb - 1
        rows_in<CROWS, N, LDR, MEM>(lds, fld, nb, opaque_tid(), [&](int j) { return (RB + j) ^ HALF; });
        __syncthreads();
        big_lines<N, CLINES, LDR, 1, false, true>(lds, twf, tid, live_rows(nb));
        for (int i = tid; i < nb * N; i += kThreads) {
          const int lr = i / N, kc = i % N;
          const bool in = in_extent(ext[RB + lr], kc ^ HALF);  // U is zero outside the support
          const float2 P = bld_f2(rpup, in ? (uint32_t)((RB + lr) * N + (kc ^ HALF)) * 8u : kSkip);
          const float kk = inv_n / fmaf(den.b, fmaf(P.x, P.x, P.y * P.y), den.a);
          const float2 u = fmulc(P, lds[lr * LDR + kc]);
          lds[lr * LDR + kc] = in ? f2(u.x * kk, u.y * kk) : f2(0.0f, 0.0f);
        }
        __syncthreads();
        const int a0 = R0;
        // destination rows of this chunk
        if (sp.zero) {
          for (int i = opaque_tid(); i < nk * N; i += kThreads) {
            const int lr = i / N, c = i % N;
            const float2 u = lds[(lr + ex) * LDR + (c ^ HALF)];
            if (u.x != 0.0f || u.y != 0.0f) {
              double2* d = spec + (sp.sr + a0 + lr) * T.rw + sp.sc + c;
              const double2 o = MEM::ld(d);
              MEM::st(d, make_double2(fma(scale, (double)u.x, o.x), fma(scale, (double)u.y, o.y)));
            }
          }
        } else {
          // the chunk's destination rows a0 .. a1 - 1 of the tile's list
          const int a1 = a0 + nk + ((q == clast && sp.ar.upper != sp.ar.lower) ? 1 : 0);
          const uint32_t* rows = T.dest + dest_row_starts_offset(N);
          const uint32_t d0 = nk > 0 ? __builtin_amdgcn_readfirstlane(gptr(rows)[a0]) : 0u;
          const uint32_t d1 = nk > 0 ? __builtin_amdgcn_readfirstlane(gptr(rows)[a1]) : 0u;
          uint32_t dfirst[kInsertGroup];
          const int itid = opaque_tid();
          insert_list_first(T.dest + 2 + d0, d1 - d0, itid, dfirst);
          insert_list<N, MEM>(spec, T.rh * T.rw, sp, T.rw, scale, itid, T.dest + 2 + d0, d1 - d0, dfirst,
                              [&](int sr_, int sc_) {
                                if (sr_ >= a0 + nk) return f2(0.0f, 0.0f);  // below the support rows
                                if (sr_ >= RB) return lds[(sr_ - RB) * LDR + (sc_ ^ HALF)];
                                if (q == 0) return f2(0.0f, 0.0f);  // above the support rows
                                return prev[sc_ ^ HALF];            // one workgroup: the previous chunk's row
                              });
        }
        __syncthreads();
        if (!COOP && q + 1 < nch) {
          for (int i = tid; i < N; i += kThreads) prev[i] = nk > 0 ? lds[(nk - 1) * LDR + i] : f2(0.0f, 0.0f);
          __syncthreads();
        }
      }
      if (COOP && !group_barrier(T, passed, bflag)) return;
      if (T.momentum) {
        counter += 1;
        if (counter >= T.interval) {
          momentum_event<MEM>(T, tid, COOP ? T.member : 0, COOP ? T.members : 1);
          counter = 0;
          if (COOP && !group_barrier(T, passed, bflag)) return;
        }
      }
    }
    if (lead) finish_sweep(T, obj_sum, w_sum, counter);
  }
  if (COOP && T.member == 0 && threadIdx.x == 0) T.group[1] = passed;
}

// The fused sweep: k_sweep<false> takes tiles up to 256^2, k_sweep<true>
// every size up to 512^2 (two kernels, so that the 512^2 path's registers
// do not cost the batches without such tiles).
template <bool LARGE>
__global__ void __launch_bounds__(kThreads)
    k_sweep(const TileDesc* __restrict__ tiles, const float2* __restrict__ tw_all) {
  extern __shared__ unsigned char smem[];
  float2* lds = reinterpret_cast<float2*>(smem);
  const TileDesc& T = tiles[blockIdx.x];  // fields are read (scalar loads) where used
  const int N = (int)T.n;
  // LDS: field (or chunk) region, [big tiles: previous row of U], twiddles
  // W_N^j forward then inverse, reduction slots, barrier flag, pupil extents
  float2* prev = lds + sweep_field_elems(N);
  float2* twf = prev + (N > (int)kFusedLdsMaxN ? N : 0);
  float2* twi = twf + N;
  double* red = reinterpret_cast<double*>(twi + N);
  int* bflag = reinterpret_cast<int*>(red + kRedSlots);
  uint32_t* ext = reinterpret_cast<uint32_t*>(bflag + 4);
  uint32_t* rec = ext + N;  // two LED records
  const float2* src = tw_all + tw_offset(N);
  for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) twf[j] = src[j];
  for (int j = threadIdx.x; j < N; j += blockDim.x) ext[j] = T.extent[j];
  __syncthreads();
  switch (N) {
    case 32:
      if (T.plan == 4)
        sweep_tile<32, 4>(T, lds, twf, twi, red, ext, rec);
      else
        sweep_tile<32, 8>(T, lds, twf, twi, red, ext, rec);
      break;
    case 64:
      if (T.plan == 4)
        sweep_tile<64, 4>(T, lds, twf, twi, red, ext, rec);
      else if (T.plan == 16)
        sweep_tile<64, 16>(T, lds, twf, twi, red, ext, rec);
      else
        sweep_tile<64, 8>(T, lds, twf, twi, red, ext, rec);
      break;
    case 128:
      if (T.plan == 8)
        sweep_tile<128, 8>(T, lds, twf, twi, red, ext, rec);
      else
        sweep_tile<128, 16>(T, lds, twf, twi, red, ext, rec);
      break;
    case 256:
      if (T.members > 1)
        sweep_tile_big<256, CoherentMem>(T, lds, prev, twf, twi, red, bflag, ext, rec);
      else
        sweep_tile_big<256, PlainMem>(T, lds, prev, twf, twi, red, bflag, ext, rec);
      break;
    case 512:
      if constexpr (LARGE) {
        if (T.members > 1)
          sweep_tile_big<512, CoherentMem>(T, lds, prev, twf, twi, red, bflag, ext, rec);
        else
          sweep_tile_big<512, PlainMem>(T, lds, prev, twf, twi, red, bflag, ext, rec);
      }
      break;
    default:
      break;
  }
}

template __global__ void k_sweep<false>(const TileDesc*, const float2*);
template __global__ void k_sweep<true>(const TileDesc*, const float2*);

}  // namespace

// The fused sweep with EPRY pupil recovery (tiles up to kFusedLdsMaxN), a
// kernel of its own so that its extra live state does not cost the other
// algorithms' kernel registers. Same LDS layout as k_sweep.
__global__ void __launch_bounds__(kThreads)
    k_sweep_epry(const TileDesc* __restrict__ tiles, const float2* __restrict__ tw_all) {
  extern __shared__ unsigned char smem[];
  float2* lds = reinterpret_cast<float2*>(smem);
  const TileDesc& T = tiles[blockIdx.x];
  const int N = (int)T.n;
  float2* prev = lds + sweep_field_elems(N);
  float2* twf = prev + (N > (int)kFusedLdsMaxN ? N : 0);
  float2* twi = twf + N;
  double* red = reinterpret_cast<double*>(twi + N);
  int* bflag = reinterpret_cast<int*>(red + kRedSlots);
  uint32_t* ext = reinterpret_cast<uint32_t*>(bflag + 4);
  uint32_t* rec = ext + N;
  const float2* src = tw_all + tw_offset(N);
  for (int j = threadIdx.x; j < 2 * N; j += blockDim.x) twf[j] = src[j];
  for (int j = threadIdx.x; j < N; j += blockDim.x) ext[j] = T.extent[j];
  __syncthreads();
  switch (N) {
    case 32:
      sweep_tile<32, 8, true>(T, lds, twf, twi, red, ext, rec);
      break;
    case 64:
      sweep_tile<64, 8, true>(T, lds, twf, twi, red, ext, rec);
      break;
    case 128:
      sweep_tile<128, 16, true>(T, lds, twf, twi, red, ext, rec);
      break;
    default:
      break;
  }
}

// LDS of one tile of n x n pixels (the layout of k_sweep).
static size_t sweep_lds_bytes_one(uint32_t n) {
  return ((size_t)sweep_field_elems(n) + (n > kFusedLdsMaxN ? n : 0) + 2 * (size_t)n) * sizeof(float2) +
         kRedSlots * sizeof(double) + 16 + (size_t)n * sizeof(uint32_t) + 2 * kRecPitch * sizeof(uint32_t);
}

// A launch's LDS: the largest layout of any tile size up to nmax (a 256^2
// tile's chunk can need less than a 128^2 tile's whole field).
size_t sweep_lds_bytes(uint32_t nmax) {
  size_t bytes = 0;
  for (uint32_t n = 32; n <= nmax && n <= kFusedMaxN; n *= 2) bytes = std::max(bytes, sweep_lds_bytes_one(n));
  return bytes;
}

uint32_t sweep_radix_choices(uint32_t n, uint32_t out[3]) {
  switch (n) {
    case 32:
      out[0] = 8, out[1] = 4;
      return 2;
    case 64:
      out[0] = 8, out[1] = 4, out[2] = 16;
      return 3;
    case 128:
      out[0] = 16, out[1] = 8;
      return 2;
    case 256:
      out[0] = 16;
      return 1;
    case 512:
      out[0] = 8;  // three radix-8 stages
      return 1;
    default:
      return 0;
  }
}

size_t group_state_bytes() {
  // [0, 24) counters, [64, 64 + 16 kMaxGroupMembers) partial sums
  return 64 + 16 * (size_t)kMaxGroupMembers;
}

// The sweep kernel for tiles up to nmax.
static const void* sweep_kernel_for(uint32_t nmax) {
  return nmax > 256 ? reinterpret_cast<const void*>(&k_sweep<true>) : reinterpret_cast<const void*>(&k_sweep<false>);
}

void prepare_sweep(uint32_t nmax) {
  HIP_CHECK(hipFuncSetAttribute(sweep_kernel_for(nmax), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sweep_lds_bytes(nmax)));
  HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sweep_epry),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)sweep_lds_bytes(std::min(nmax, kFusedLdsMaxN))));
}

uint32_t sweep_coresident_limit(uint32_t nmax) {
  int dev = 0, per_cu = 0, cus = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sweep_kernel_for(nmax), kThreads,
                                                         sweep_lds_bytes(nmax)));
  return (uint32_t)(std::max(per_cu, 0) * std::max(cus, 0));
}

void launch_patch_bounds(const TileDesc* tiles_dev, uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(k_patch_bounds, dim3(count), dim3(kThreads), 0, stream, tiles_dev);
  HIP_CHECK(hipGetLastError());
}

void launch_sweep(const TileDesc* tiles_dev, uint32_t count, const float2* tw_all, uint32_t nmax, bool cooperative,
                  hipStream_t stream, bool epry) {
  const size_t smem = sweep_lds_bytes(nmax);
  if (epry) {  // no cooperative groups: EPRY tiles are at most kFusedLdsMaxN
    hipLaunchKernelGGL(k_sweep_epry, dim3(count), dim3(kThreads), smem, stream, tiles_dev, tw_all);
    HIP_CHECK(hipGetLastError());
    return;
  }
  if (cooperative) {
    // every workgroup of every cooperative group must be resident at once
    void* args[] = {(void*)&tiles_dev, (void*)&tw_all};
    HIP_CHECK(hipLaunchCooperativeKernel(sweep_kernel_for(nmax), dim3(count), dim3(kThreads), args, (unsigned)smem,
                                         stream));
  } else if (nmax > 256) {
    hipLaunchKernelGGL(k_sweep<true>, dim3(count), dim3(kThreads), smem, stream, tiles_dev, tw_all);
    HIP_CHECK(hipGetLastError());
  } else {
    hipLaunchKernelGGL(k_sweep<false>, dim3(count), dim3(kThreads), smem, stream, tiles_dev, tw_all);
    HIP_CHECK(hipGetLastError());
  }
}

}  // namespace fpmhip
