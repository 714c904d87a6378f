// VALU issue-rate probe for the roofline of the forward kernels: sustained throughput of independent
// non-packed FP32 FMA / MUL / ADD and FP64 FMA chains, and cycles per wave64 instruction by operand kind
// (VGPR, 32-bit literal, inline constant, SGPR, separate multiply and add).  Built with the library's
// flags (no packed FP32).  Prints one line per measurement; cycles assume a 2.4 GHz clock.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

// Throughput: kChains independent chains per lane, `iters` steps each, every operand in a VGPR (b and c
// differ per lane; an SGPR operand slows the instruction, see the operand kinds below).
constexpr int kChains = 16;
template <int OP, class T>
__global__ __launch_bounds__(256) void chain_kernel(T* out, int iters, T b0, T c0) {
  const T b = b0 + (T)1e-7 * (T)threadIdx.x, c = c0 * (T)(1 + threadIdx.x % 7);
  T a[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) a[i] = (T)(threadIdx.x + i) * (T)1e-3;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
      if (OP == 0) a[i] = fma(a[i], b, c);
      else if (OP == 1) a[i] = a[i] * b;
      else a[i] = a[i] + c;
    }
  }
  T s = 0;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;  // grid <= the buffer's blocks (checked by the caller)
}

namespace nc {
#pragma clang fp contract(off)
__device__ __forceinline__ float muladd(float a, float b, float c) { return a * b + c; }
}  // namespace nc

// Operand kinds.  KIND 0: b, c in VGPRs; 1: b a 32-bit literal; 2: b an inline constant (0.5); 3: b an SGPR
// (kernel argument); 4: separate multiply and add (no contraction), VGPRs.
constexpr int kKindChains = 16;
template <int KIND>
__global__ __launch_bounds__(256) void kind_kernel(float* out, int iters, float bs) {
  float b = 0.999f + 1e-7f * threadIdx.x, c = 1e-6f * (1 + threadIdx.x % 7);
  if (KIND == 1) b = 0.99912345f;
  if (KIND == 2) b = 0.5f;
  if (KIND == 3) b = bs;
  float a[kKindChains];
#pragma unroll
  for (int i = 0; i < kKindChains; ++i) a[i] = (float)(threadIdx.x + i) * 1e-3f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll 4
    for (int u = 0; u < 4; ++u) {
#pragma unroll
      for (int i = 0; i < kKindChains; ++i) {
        if (KIND == 4) a[i] = nc::muladd(a[i], b, c);
        else a[i] = fmaf(a[i], b, c);
      }
    }
  }
  float s = 0;
#pragma unroll
  for (int i = 0; i < kKindChains; ++i) s += a[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

float timed_ms(hipEvent_t e0, hipEvent_t e1) {
  float ms = 0.f;
  (void)hipEventSynchronize(e1);
  (void)hipEventElapsedTime(&ms, e0, e1);
  return ms;
}

template <int OP, class T>
double tflops(int blocks, int iters, T* out, hipEvent_t e0, hipEvent_t e1) {
  chain_kernel<OP, T><<<blocks, 256>>>(out, iters, (T)0.999, (T)1e-6);  // untimed: clocks settle
  (void)hipEventRecord(e0);
  chain_kernel<OP, T><<<blocks, 256>>>(out, iters, (T)0.999, (T)1e-6);
  (void)hipEventRecord(e1);
  const double ms = timed_ms(e0, e1);
  return (double)blocks * 256 * iters * kChains * (OP == 0 ? 2.0 : 1.0) / (ms * 1e-3) / 1e12;
}

template <int KIND>
double cycles(int cus, int wpsimd, int iters, float* out, hipEvent_t e0, hipEvent_t e1) {
  const int blocks = cus * wpsimd;  // 256 lanes = one wavefront per SIMD per workgroup
  kind_kernel<KIND><<<blocks, 256>>>(out, iters, 0.999f);
  (void)hipEventRecord(e0);
  kind_kernel<KIND><<<blocks, 256>>>(out, iters, 0.999f);
  (void)hipEventRecord(e1);
  const double ms = timed_ms(e0, e1);
  const double per_simd = (double)blocks * 4 * iters * 4 * kKindChains * (KIND == 4 ? 2 : 1) / (cus * 4);
  return ms * 1e-3 * 2.4e9 / per_simd;
}

}  // namespace

int main() {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess || cus <= 0) {
    std::fprintf(stderr, "no device\n");
    return 1;
  }
  constexpr int kMaxWaves = 8;  // waves per SIMD: grids of at most cus * kMaxWaves workgroups of 256 lanes
  const size_t lanes = (size_t)cus * kMaxWaves * 256;
  float* of = nullptr;
  double* od = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc(&of, lanes * sizeof(float)) != hipSuccess || hipMalloc(&od, lanes * sizeof(double)) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    std::fprintf(stderr, "allocation failed\n");
    return 1;
  }
  std::printf("CUs %d\n", cus);
  // ~0.5 s of work first: the clocks ramp up over the first tens of milliseconds
  for (int i = 0; i < 20; ++i) kind_kernel<0><<<cus * kMaxWaves, 256>>>(of, 16384, 0.999f);
  (void)hipDeviceSynchronize();
  for (int w : {1, 2, 4, 8}) {
    const int blocks = cus * w, it = 32768 / w;
    std::printf("throughput, %d waves/SIMD: fp32 fma %.1f TFLOP/s, fp32 mul %.1f, fp32 add %.1f, fp64 fma %.1f\n", w,
                tflops<0, float>(blocks, it, of, e0, e1), tflops<1, float>(blocks, it, of, e0, e1),
                tflops<2, float>(blocks, it, of, e0, e1), tflops<0, double>(blocks, it / 2, od, e0, e1));
  }
  for (int w : {2, 4, 8})
    std::printf("cycles per wave64 instruction, %d waves/SIMD: fma vgpr %.2f, fma literal %.2f, fma inline %.2f, "
                "fma sgpr %.2f, mul+add vgpr %.2f\n",
                w, cycles<0>(cus, w, 4096, of, e0, e1), cycles<1>(cus, w, 4096, of, e0, e1),
                cycles<2>(cus, w, 4096, of, e0, e1), cycles<3>(cus, w, 4096, of, e0, e1),
                cycles<4>(cus, w, 4096, of, e0, e1));
  const bool ok = hipDeviceSynchronize() == hipSuccess;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(of);
  (void)hipFree(od);
  return ok ? 0 : 1;
}
