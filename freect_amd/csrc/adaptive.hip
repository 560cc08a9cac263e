// adaptive.hip -- adaptive filtration of the raw scan before the rebinning.
//
// FreeCT parses adaptive_filtration_s (parse_config.cpp:57, recon_structs.h:49)
// and lists adaptive filtration as missing (README.md:14); no implementation
// exists in its tree.  This one follows the idea of multi-dimensional
// adaptive filtering (Kachelriess, Watzke, Kalender, Med. Phys. 28 (2001)
// 475): the noise of a log-transformed reading grows like exp(p / 2), so only
// the most attenuated rays of a view are smoothed.  Per raw view, with vmax
// the largest finite |p| and s in (0, 1] the strength, a sample p is blended
// with the mean m of its 3 x 3 (row, channel) neighbourhood (the neighbours
// that exist at the detector's edges):
//   lambda = clamp((p - (1 - s) vmax) / (s vmax), 0, 1),  p' = p + lambda (m - p).
// Samples below (1 - s) vmax are unchanged; p' is a convex combination of
// raw values, so |p'| <= max |p| (the FP16 storage bound still holds).  A
// non-finite value counts as 0, as in the rebinning gather.
#include "wfbp_internal.h"

namespace fct {
namespace {

__device__ __forceinline__ float fin0(float v) { return __builtin_isfinite(v) ? v : 0.0f; }

__global__ __launch_bounds__(256) void view_absmax_kernel(const float* __restrict__ in,
                                                          float* __restrict__ vmax, int v0,
                                                          int per_view) {
  __shared__ float part[4];
  const int v = v0 + blockIdx.x;
  const float* x = in + (size_t)v * per_view;
  float m = 0.0f;
  for (int i = threadIdx.x; i < per_view; i += 256) m = fmaxf(m, fabsf(fin0(x[i])));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) vmax[v] = fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
}

__global__ __launch_bounds__(256) void adaptive_filter_kernel(const float* __restrict__ in,
                                                              float* __restrict__ out,
                                                              const float* __restrict__ vmax,
                                                              int v0, int nrows, int ncols,
                                                              float s) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int r = blockIdx.y;
  const int v = v0 + blockIdx.z;
  if (c >= ncols) return;
  const size_t base = (size_t)v * nrows * ncols;
  const float* x = in + base;
  const float p = fin0(x[(size_t)r * ncols + c]);
  const float vm = vmax[v];
  const float den = s * vm;
  const float lam = den > 0.0f ? fminf(fmaxf((p - (vm - den)) / den, 0.0f), 1.0f) : 0.0f;
  float res = p;
  if (lam > 0.0f) {
    float sum = 0.0f;
    int n = 0;
    for (int rr = max(r - 1, 0); rr <= min(r + 1, nrows - 1); ++rr)
      for (int cc = max(c - 1, 0); cc <= min(c + 1, ncols - 1); ++cc) {
        sum += fin0(x[(size_t)rr * ncols + cc]);
        ++n;
      }
    res = fmaf(lam, sum / (float)n - p, p);
  }
  out[base + (size_t)r * ncols + c] = res;
}

}  // namespace

hipError_t launch_adaptive_filter(const float* in, float* out, float* vmax, int v0, int v1,
                                  int nrows, int ncols, float s, hipStream_t st) {
  if (v1 <= v0) return hipSuccess;
  if (in == out) return hipErrorInvalidValue;
  hipLaunchKernelGGL(view_absmax_kernel, dim3(v1 - v0), dim3(256), 0, st, in, vmax, v0,
                     nrows * ncols);
  hipLaunchKernelGGL(adaptive_filter_kernel, dim3((ncols + 255) / 256, nrows, v1 - v0),
                     dim3(256), 0, st, in, out, vmax, v0, nrows, ncols, s);
  return hipGetLastError();
}

}  // namespace fct
