// valu_probe.hip -- vector-ALU issue rates of this GPU, for the roofline of
// the backprojection (which is bound by VALU issue).
//
// Each thread runs 8 independent chains of one instruction for kIters
// iterations (no dependency stalls at 8 waves per SIMD) and stores one value,
// so that the chains are not removed; out holds one float per thread.  The
// rate printed is op instructions per SIMD and cycle at the device's peak
// clock: about 0.25 for a wave64 FP32 instruction on a 16-lane SIMD.
// rocprofv3 --pmc SQ_INSTS_VALU SQ_ACTIVE_INST_VALU GRBM_GUI_ACTIVE over this
// program calibrates those counters (the FMA chain saturates the VALU).
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kIters = 4096;
constexpr int kChains = 8;

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

template <int OP>
__global__ __launch_bounds__(256) void chains(float* out, float a, float b) {
  float x[kChains];
  int xi[kChains];
#pragma unroll
  for (int i = 0; i < kChains; ++i) {
    x[i] = threadIdx.x * 1e-3f + i;
    xi[i] = threadIdx.x + i;
  }
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int i = 0; i < kChains; ++i) {
      if (OP == 0) x[i] = __builtin_fmaf(x[i], a, b);   // v_fma_f32
      if (OP == 1) x[i] = x[i] * a;                      // v_mul_f32
      if (OP == 2)                                       // v_add_u32
        asm volatile("v_add_u32 %0, %0, %1" : "+v"(xi[i]) : "v"(xi[(i + 1) % kChains]));
      if (OP == 3)                                       // v_dot2_f32_f16
        x[i] = __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, xi[i]),
                                      __builtin_bit_cast(half2_t, xi[i]), x[i], false);
      if (OP == 4) x[i] = __builtin_amdgcn_sqrtf(x[i]);  // v_sqrt_f32
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < kChains; ++i) s += x[i] + (float)xi[i];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

#define CHECK(x)                                                             \
  do {                                                                       \
    hipError_t e_ = (x);                                                     \
    if (e_ != hipSuccess) {                                                  \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
      return false;                                                          \
    }                                                                        \
  } while (0)

template <int OP>
bool run(const char* name, float* d, int blocks, int ncu, double ghz, bool first) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(chains<OP>, dim3(blocks), dim3(256), 0, 0, d, 0.999f, 1e-3f);
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  const int reps = 5;
  CHECK(hipEventRecord(e0));
  for (int r = 0; r < reps; ++r)
    hipLaunchKernelGGL(chains<OP>, dim3(blocks), dim3(256), 0, 0, d, 0.999f, 1e-3f);
  CHECK(hipGetLastError());
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  ms /= reps;
  const double waves = blocks * 4.0;
  const double insts = waves * kIters * kChains;
  const double simd_cycles = ncu * 4.0 * ms * 1e-3 * ghz * 1e9;
  std::printf("%s\"%s\": {\"ms\": %.4f, \"per_simd_cycle\": %.4f}", first ? "" : ", ", name, ms,
              insts / simd_cycles);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return true;
}

}  // namespace

int main() {
  int ncu = 0, khz = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess ||
      hipDeviceGetAttribute(&khz, hipDeviceAttributeClockRate, 0) != hipSuccess || ncu <= 0) {
    std::fprintf(stderr, "no GPU\n");
    return 1;
  }
  const double ghz = khz * 1e-6;
  const int blocks = ncu * 8 * 4;   // 8 waves per SIMD, 4 rounds
  float* d = nullptr;
  if (hipMalloc(&d, (size_t)blocks * 256 * sizeof(float)) != hipSuccess) return 1;
  std::printf("{\"cus\": %d, \"clock_ghz\": %.3f, \"waves_per_simd\": 8, \"rates\": {", ncu, ghz);
  const bool ok = run<0>("v_fma_f32", d, blocks, ncu, ghz, true) &&
                  run<1>("v_mul_f32", d, blocks, ncu, ghz, false) &&
                  run<2>("v_add_u32", d, blocks, ncu, ghz, false) &&
                  run<3>("v_dot2_f32_f16", d, blocks, ncu, ghz, false) &&
                  run<4>("v_sqrt_f32", d, blocks, ncu, ghz, false);
  std::printf("}}\n");
  (void)hipFree(d);
  return ok ? 0 : 1;
}
