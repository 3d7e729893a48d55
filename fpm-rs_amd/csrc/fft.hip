// fft.hip — LDS-resident batched 1-D complex FFTs along rows or columns of
// row-major 2-D arrays, any length (Stockham autosort, decimation in
// frequency; radix-4/2 butterflies with exact quarter-turn rotations, direct
// DFT butterflies for 3, 5, 7 and other primes up to 64).
//
// Semantics of fpm-rs's Backend::fft2 (src/backend/cpu.rs:130-174): row
// transforms, then column transforms, forward divided by h*w, inverse
// unscaled. Used for the generic (per-stage) path, for init/result transforms
// on the smooth high-resolution grids (375 = 3*5^3, 729 = 3^6, ...) and by
// fpmhip_fft2_host. The FP64 instantiation is compiled without FMA
// contraction so that each butterfly rounds like the scalar FP64 definition.
#include <algorithm>

#include "cx.h"
#include "internal.h"

namespace fpmhip {

void unit_root(uint64_t j, uint64_t n, int sign, double* re, double* im) {
  j %= n;
  if ((4 * j) % n == 0) {
    static const double qr[4] = {1.0, 0.0, -1.0, 0.0};
    static const double qi[4] = {0.0, 1.0, 0.0, -1.0};
    uint64_t q = (4 * j) / n;
    *re = qr[q];
    *im = sign * qi[q];
    return;
  }
  long double theta = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
  *re = (double)cosl(theta);
  *im = (double)(sign * sinl(theta));
}

std::vector<uint32_t> factorize(uint32_t n) {
  std::vector<uint32_t> r;
  while (n % 4 == 0) {
    r.push_back(4);
    n /= 4;
  }
  while (n % 2 == 0) {
    r.push_back(2);
    n /= 2;
  }
  for (uint32_t p = 3; p * p <= n; p += 2)
    while (n % p == 0) {
      r.push_back(p);
      n /= p;
    }
  if (n > 1) r.push_back(n);
  return r;
}

static void fill_fft1(Fft1& f, uint32_t n, int sign) {
  std::vector<uint32_t> radices = factorize(n);
  if (radices.size() > (size_t)kMaxStages) fail(FPMHIP_UNSUPPORTED, "FFT length has too many factors");
  for (uint32_t p : radices)
    if (p > 64) fail(FPMHIP_UNSUPPORTED, "FFT length has a prime factor above 64");
  f.plan.n = n;
  f.plan.nstages = (uint32_t)radices.size();
  for (size_t i = 0; i < radices.size(); ++i) f.plan.radix[i] = radices[i];
  std::vector<double> t64(2 * (size_t)n);
  std::vector<float> t32(2 * (size_t)n);
  for (uint32_t j = 0; j < n; ++j) {
    double re, im;
    unit_root(j, n, sign, &re, &im);
    t64[2 * j] = re;
    t64[2 * j + 1] = im;
    long double theta = 2.0L * 3.14159265358979323846264338327950288L * (long double)(j % n) / (long double)n;
    // single rounding from extended precision for the FP32 table
    if ((4ull * j) % n == 0) {
      t32[2 * j] = (float)re;
      t32[2 * j + 1] = (float)im;
    } else {
      t32[2 * j] = (float)cosl(theta);
      t32[2 * j + 1] = (float)(sign * sinl(theta));
    }
  }
  f.tw_f64.allocate(t64.size() * sizeof(double));
  f.tw_f32.allocate(t32.size() * sizeof(float));
  HIP_CHECK(hipMemcpy(f.tw_f64.ptr, t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(f.tw_f32.ptr, t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice));
}

std::unique_ptr<Fft1Pair> make_fft1(uint32_t n) {
  auto p = std::make_unique<Fft1Pair>();
  fill_fft1(p->fwd, n, -1);
  fill_fft1(p->inv, n, +1);
  return p;
}

std::unique_ptr<Fft2> make_fft2(uint32_t h, uint32_t w) {
  auto f = std::make_unique<Fft2>();
  f->h = h;
  f->w = w;
  f->rows = make_fft1(w);
  f->cols = h == w ? nullptr : make_fft1(h);
  return f;
}

template <class R>
struct LineKernelArgs {
  using V = typename cx<R>::V;
  V* data;
  const V* tw;
  LinePlan plan;
  uint64_t nlines;       // lines in the batch
  uint64_t lines_per_array;
  uint64_t array_stride;  // elements between consecutive arrays
  uint64_t line_stride;   // elements between consecutive lines of one array
  uint64_t elem_stride;   // elements between consecutive samples of a line
  uint32_t lines_per_block;
  int inverse;
  double divide;  // 0: none; else every output is divided by this
};

template <class R>
__global__ void __launch_bounds__(256) k_fft_lines(LineKernelArgs<R> a) {
  using C = cx<R>;
  using V = typename C::V;
  extern __shared__ unsigned char smem_raw[];
  V* buf0 = reinterpret_cast<V*>(smem_raw);
  const uint32_t n = a.plan.n;
  const uint32_t lpb = a.lines_per_block;
  V* buf1 = buf0 + (size_t)n * lpb;
  const uint64_t line0 = (uint64_t)blockIdx.x * lpb;
  const uint32_t total = n * lpb;
  const bool rowwise = a.elem_stride == 1;
  // load: coalesced along rows (row transforms) or across columns (column transforms)
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    uint32_t line = rowwise ? i / n : i % lpb;
    uint32_t e = rowwise ? i % n : i / lpb;
    uint64_t gl = line0 + line;
    if (gl < a.nlines) {
      uint64_t arr = gl / a.lines_per_array, li = gl % a.lines_per_array;
      buf0[(size_t)line * n + e] = a.data[arr * a.array_stride + li * a.line_stride + (uint64_t)e * a.elem_stride];
    }
  }
  __syncthreads();
  V* src = buf0;
  V* dst = buf1;
  uint32_t ncur = n, s = 1;
  for (uint32_t st = 0; st < a.plan.nstages; ++st) {
    const uint32_t p = a.plan.radix[st];
    const uint32_t m = ncur / p;
    const uint32_t nb = m * s;
    for (uint32_t i = threadIdx.x; i < nb * lpb; i += blockDim.x) {
      const uint32_t line = i / nb, t = i % nb;
      const uint32_t pp = t / s, q = t % s;
      const V* in = src + (size_t)line * n;
      V* out = dst + (size_t)line * n;
      if (p == 4) {
        const V a0 = in[q + s * pp], a1 = in[q + s * (pp + m)], a2 = in[q + s * (pp + 2 * m)],
                a3 = in[q + s * (pp + 3 * m)];
        const V t0 = C::add(a0, a2), t1 = C::sub(a0, a2), t2 = C::add(a1, a3), d = C::sub(a1, a3);
        const V t3 = a.inverse ? C::make(-d.y, d.x) : C::make(d.y, -d.x);
        out[q + s * (p * pp)] = C::add(t0, t2);
        out[q + s * (p * pp + 1)] = C::mul(C::add(t1, t3), a.tw[(pp * 1 * s) % n]);
        out[q + s * (p * pp + 2)] = C::mul(C::sub(t0, t2), a.tw[(pp * 2 * s) % n]);
        out[q + s * (p * pp + 3)] = C::mul(C::sub(t1, t3), a.tw[(pp * 3 * s) % n]);
      } else if (p == 2) {
        const V a0 = in[q + s * pp], a1 = in[q + s * (pp + m)];
        out[q + s * (p * pp)] = C::add(a0, a1);
        out[q + s * (p * pp + 1)] = C::mul(C::sub(a0, a1), a.tw[(pp * s) % n]);
      } else {
        // direct DFT butterfly, each output written as soon as it is summed
        const uint32_t step = n / p;
        for (uint32_t k = 0; k < p; ++k) {
          V acc = in[q + s * pp];
          for (uint32_t r = 1; r < p; ++r)
            acc = C::add(acc, C::mul(in[q + s * (pp + r * m)], a.tw[((r * k) % p) * step]));
          out[q + s * (p * pp + k)] = k == 0 ? acc : C::mul(acc, a.tw[(pp * k * s) % n]);
        }
      }
    }
    __syncthreads();
    V* tmp = src;
    src = dst;
    dst = tmp;
    ncur = m;
    s *= p;
  }
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    uint32_t line = rowwise ? i / n : i % lpb;
    uint32_t e = rowwise ? i % n : i / lpb;
    uint64_t gl = line0 + line;
    if (gl < a.nlines) {
      uint64_t arr = gl / a.lines_per_array, li = gl % a.lines_per_array;
      V v = src[(size_t)line * n + e];
      if (a.divide != 0.0) v = C::divr(v, (R)a.divide);
      a.data[arr * a.array_stride + li * a.line_stride + (uint64_t)e * a.elem_stride] = v;
    }
  }
}

template <class R>
static void launch_lines(const Fft1& f, typename cx<R>::V* data, uint64_t narrays, uint64_t lines_per_array,
                         uint64_t array_stride, uint64_t line_stride, uint64_t elem_stride, bool inverse,
                         double divide, hipStream_t stream) {
  using V = typename cx<R>::V;
  const uint32_t n = f.plan.n;
  const size_t per_line = 2 * (size_t)n * sizeof(V);
  uint32_t lpb = (uint32_t)std::max<size_t>(1, std::min<size_t>(16, 32768 / per_line));
  if (elem_stride != 1) lpb = std::max<uint32_t>(lpb, 1);
  const uint64_t nlines = narrays * lines_per_array;
  lpb = (uint32_t)std::min<uint64_t>(lpb, nlines);
  LineKernelArgs<R> a;
  a.data = data;
  a.tw = reinterpret_cast<const V*>(std::is_same<R, double>::value ? f.tw_f64.ptr : f.tw_f32.ptr);
  a.plan = f.plan;
  a.nlines = nlines;
  a.lines_per_array = lines_per_array;
  a.array_stride = array_stride;
  a.line_stride = line_stride;
  a.elem_stride = elem_stride;
  a.lines_per_block = lpb;
  a.inverse = inverse ? 1 : 0;
  a.divide = divide;
  const size_t smem = per_line * lpb;
  const uint32_t blocks = (uint32_t)((nlines + lpb - 1) / lpb);
  if (smem > 65536)
    HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_lines<R>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(k_fft_lines<R>, dim3(blocks), dim3(256), smem, stream, a);
  HIP_CHECK(hipGetLastError());
}

template <class R>
static void fft2_impl(const Fft2& f, typename cx<R>::V* data, uint32_t count, bool inverse, hipStream_t stream) {
  const uint64_t h = f.h, w = f.w;
  const Fft1Pair& rp = *f.rows;
  const Fft1Pair& cp = f.cols ? *f.cols : *f.rows;
  const Fft1& r1 = inverse ? rp.inv : rp.fwd;
  const Fft1& c1 = inverse ? cp.inv : cp.fwd;
  // rows: `count` arrays of h lines of length w
  launch_lines<R>(r1, data, count, h, h * w, w, 1, inverse, 0.0, stream);
  // columns: w lines of length h, samples w apart; the forward divides by h*w
  launch_lines<R>(c1, data, count, w, h * w, 1, w, inverse, inverse ? 0.0 : (double)(h * w), stream);
}

void fft2_f64(const Fft2& f, double2* data, uint32_t count, bool inverse, hipStream_t stream) {
  fft2_impl<double>(f, data, count, inverse, stream);
}
void fft2_f32(const Fft2& f, float2* data, uint32_t count, bool inverse, hipStream_t stream) {
  fft2_impl<float>(f, data, count, inverse, stream);
}

}  // namespace fpmhip
