// dft16_trial.hip — the DFT-as-GEMM trial for the fused sweep's radix-16
// stages: batched 16-point complex64 DFTs computed
//   (a) as the sweep computes them: reg_dft<16> on VALU registers
//       (k_dft16_valu, one DFT per lane), and
//   (b) as matrix products on the matrix cores: Y = F X for 16 vectors at a
//       time with v_mfma_f32_16x16x4_f32 (k_dft16_mfma, one wave per 16
//       vectors; Yr = Fr Xr - Fi Xi and Yi = Fr Xi + Fi Xr, 16 MFMAs).
// Both apply the unitary DFT (x 1/4) R times to register-resident data read
// once, so the loops are bound by arithmetic, not memory. The program checks
// one application of each kernel against a double-precision DFT on the host
// and that R applications (R a multiple of 4, U^4 = I) return the input, then
// prints the time per 16-point DFT of each. Built by `make -C fpm-rs_amd/csrc
// trial`; not part of the library.
//
// MFMA operand layout (16x16x4, FP32): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[4 (l / 16) + r][l % 16] in its four
// accumulator registers r. With the DFT's columns taken in the order
// k = 4 (l / 16) + s at step s, the output registers of a lane are exactly
// the inputs it supplies at the next application, so no data moves between
// lanes.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../radix.h"

#define CHECK(x)                                                                  \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess) {                                                       \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                \
      std::exit(2);                                                               \
    }                                                                             \
  } while (0)

using fpmhip::reg_dft;
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;

// x: [vectors][16] complex64; one vector per thread
__global__ void __launch_bounds__(kBlock) k_dft16_valu(float2* __restrict__ x, int vectors, int reps) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= vectors) return;
  float2 a[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = x[(size_t)v * 16 + k];
  for (int r = 0; r < reps; ++r) {
    reg_dft<16, false>(a);
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = make_float2(a[k].x * 0.25f, a[k].y * 0.25f);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) x[(size_t)v * 16 + k] = a[k];
}

// x: [groups][16 vectors][16] complex64; one wave per group of 16 vectors
__global__ void __launch_bounds__(kBlock) k_dft16_mfma(float2* __restrict__ x, int groups, int reps) {
  const int lane = threadIdx.x & 63;
  const int group = (blockIdx.x * kBlock + threadIdx.x) >> 6;
  if (group >= groups) return;  // whole waves
  const int g = lane >> 4, j = lane & 15;  // lane group, vector (column)
  float2* base = x + (size_t)group * 256;
  // A operand per step s: F[j][4 g + s] (row j of the DFT, column 4 g + s), unitary
  float fr[4], fi[4];
  float xr[4], xi[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * g + s;
    const double th = -2.0 * M_PI * (double)((j * k) % 16) / 16.0;
    fr[s] = (float)(0.25 * cos(th));
    fi[s] = (float)(0.25 * sin(th));
    const float2 v = base[j * 16 + k];  // vector j, point k
    xr[s] = v.x;
    xi[s] = v.y;
  }
  for (int r = 0; r < reps; ++r) {
    v4f yr = {0.0f, 0.0f, 0.0f, 0.0f}, yi = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      yr = __builtin_amdgcn_mfma_f32_16x16x4f32(fr[s], xr[s], yr, 0, 0, 0);
      yr = __builtin_amdgcn_mfma_f32_16x16x4f32(-fi[s], xi[s], yr, 0, 0, 0);
      yi = __builtin_amdgcn_mfma_f32_16x16x4f32(fr[s], xi[s], yi, 0, 0, 0);
      yi = __builtin_amdgcn_mfma_f32_16x16x4f32(fi[s], xr[s], yi, 0, 0, 0);
    }
    // D[4 g + s][j]: point 4 g + s of vector j, the B input of step s next time
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      xr[s] = yr[s];
      xi[s] = yi[s];
    }
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) base[j * 16 + 4 * g + s] = make_float2(xr[s], xi[s]);
}

// unitary 16-point DFT of every vector, in double
static void host_dft(const std::vector<float2>& in, std::vector<std::complex<double>>& out, size_t vectors) {
  out.assign(vectors * 16, 0.0);
  for (size_t v = 0; v < vectors; ++v)
    for (int i = 0; i < 16; ++i) {
      std::complex<double> acc = 0.0;
      for (int k = 0; k < 16; ++k)
        acc += std::complex<double>(in[v * 16 + k].x, in[v * 16 + k].y) *
               std::polar(1.0, -2.0 * M_PI * (double)((i * k) % 16) / 16.0);
      out[v * 16 + i] = 0.25 * acc;
    }
}

static double rel_err(const std::vector<float2>& got, const std::vector<std::complex<double>>& ref) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < got.size(); ++i) {
    num += std::norm(std::complex<double>(got[i].x, got[i].y) - ref[i]);
    den += std::norm(ref[i]);
  }
  return std::sqrt(num / den);
}

int main() {
  int cus = 0;
  CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
  const int blocks = cus * 8;              // 8 x 256 threads per CU
  const int vectors = blocks * kBlock;     // VALU: one vector per thread
  const int groups = blocks * kBlock / 64;  // MFMA: one group of 16 vectors per wave
  const size_t n_valu = (size_t)vectors * 16, n_mfma = (size_t)groups * 256;
  std::vector<float2> h(std::max(n_valu, n_mfma));
  srand(7);
  for (auto& v : h) v = make_float2((float)rand() / RAND_MAX - 0.5f, (float)rand() / RAND_MAX - 0.5f);
  float2* d = nullptr;
  CHECK(hipMalloc(&d, h.size() * sizeof(float2)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  int ok = 1;
  // correctness: one application, and R = 64 applications return the input
  for (int kind = 0; kind < 2; ++kind) {
    const size_t n = kind ? n_mfma : n_valu, nv = n / 16;
    std::vector<float2> in(h.begin(), h.begin() + n), out(n);
    std::vector<std::complex<double>> ref;
    host_dft(in, ref, nv);
    for (int reps : {1, 64}) {
      CHECK(hipMemcpy(d, in.data(), n * sizeof(float2), hipMemcpyHostToDevice));
      if (kind)
        hipLaunchKernelGGL(k_dft16_mfma, dim3(blocks), dim3(kBlock), 0, 0, d, groups, reps);
      else
        hipLaunchKernelGGL(k_dft16_valu, dim3(blocks), dim3(kBlock), 0, 0, d, vectors, reps);
      CHECK(hipGetLastError());
      CHECK(hipMemcpy(out.data(), d, n * sizeof(float2), hipMemcpyDeviceToHost));
      std::vector<std::complex<double>> want(n);
      for (size_t i = 0; i < n; ++i) want[i] = reps == 1 ? ref[i] : std::complex<double>(in[i].x, in[i].y);
      const double err = rel_err(out, want);
      std::printf("%s reps %d: relative L2 error %.3e\n", kind ? "mfma" : "valu", reps, err);
      ok &= err < (reps == 1 ? 1e-6 : 1e-5);
    }
  }
  // timing: R applications per kernel, best of 5
  const int reps = 512;
  double ns[2];
  for (int kind = 0; kind < 2; ++kind) {
    float best = 1e30f;
    for (int t = 0; t < 5; ++t) {
      CHECK(hipEventRecord(e0, 0));
      if (kind)
        hipLaunchKernelGGL(k_dft16_mfma, dim3(blocks), dim3(kBlock), 0, 0, d, groups, reps);
      else
        hipLaunchKernelGGL(k_dft16_valu, dim3(blocks), dim3(kBlock), 0, 0, d, vectors, reps);
      CHECK(hipGetLastError());
      CHECK(hipEventRecord(e1, 0));
      CHECK(hipEventSynchronize(e1));
      float ms = 0.0f;
      CHECK(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    const double dfts = (double)(kind ? (size_t)groups * 16 : (size_t)vectors) * reps;
    ns[kind] = best * 1e6 / dfts;
    std::printf("%s: %d reps, %.3f ms, %.4f ns per 16-point DFT (whole GPU), %.1f GFLOP/s at 5 N log2 N\n",
                kind ? "mfma" : "valu", reps, best, ns[kind], dfts * 5.0 * 16 * 4 / (best * 1e6));
  }
  std::printf("mfma / valu time per DFT: %.2f\n", ns[1] / ns[0]);
  CHECK(hipFree(d));
  return ok ? 0 : 1;
}
