// rebin_filter.hip -- fused fan->parallel rebinning + ramp filtering.
//
// Replaces, per detector row sheet, the reference's rebin_kernel (texture
// bilinear, cuda_kernels.cuh:11-39), cuFFT R2C, multiply_filter
// (cuda_kernels.cuh:41-58), cuFFT C2R and reshape_rebin_into_final_array
// (cuda_kernels.cuh:60-77), launched from ReconstructionMachine.cu:241-292.
//
// One workgroup owns two filter lines (line = (parallel view j, row r)).  Both
// are rebinned straight into an LDS complex buffer as real and imaginary
// parts, transformed with one LDS-resident Stockham FFT, multiplied by the
// real filter spectrum H/N and transformed back through conj(FFT(conj(.))).
// Because H is real and even (the spatial kernel is symmetric), the real and
// imaginary parts of the result are the two filtered lines.  The lines occupy
// ncols <= N/2 samples, so the circular product equals the linear
// convolution y[i] = sum_j h[i-j] x[j], exactly what the reference computes.
//
// Storage: FP32 rows, or FP16/BF16 quad entries (wfbp_internal.h) divided by
// one power-of-two scale per scan.  The scale comes from a bound known before
// any line is filtered: |y| <= max|raw| * sum|h_k| (the rebinning is a convex
// combination of raw samples), with max|raw| from an abs-max pass over the
// raw scan.
#include <hip/hip_bf16.h>
#include <algorithm>
#include <hip/hip_fp16.h>

#include "fft_lds.h"
#include "wfbp_internal.h"

namespace fct {
namespace {

constexpr int kRfThreads = 256;

// A non-finite raw measurement (a dead or saturated detector reading) enters
// the rebinning as 0, so the filter cannot spread it over its line.
__device__ __forceinline__ float finite_or_zero(float v) { return __builtin_isfinite(v) ? v : 0.0f; }

__device__ __forceinline__ float rebin_sample(const float* __restrict__ raw, int j, int q,
                                              const RebinDev& e, const RebinFilterArgs& a) {
  // Bilinear at (projection j + aoff + fa of the entry's focal state, fan
  // channel b0 + fb) with the projection clamped to the scan, as the
  // reference's clamp-addressed texture; without FFS the state is 0 of 1 and
  // the projection is the fan view.  Every address stays inside the scan
  // also for an entry without a fan sample (state 0: channels 0, offset 0,
  // the state's mmax >= 0), read as focal state 0 -- the caller discards
  // that value.
  const int rho = max(e.state - 1, 0);
  const int ma = max(min(j + e.aoff, e.mmax), 0);
  const int mb = max(min(j + e.aoff + 1, e.mmax), 0);
  const int row = q >> a.zshift;
  const float* ra = raw + ((size_t)(ma * a.states + rho) * a.raw_nrows + row) * a.raw_ncols;
  const float* rb = raw + ((size_t)(mb * a.states + rho) * a.raw_nrows + row) * a.raw_ncols;
  const float a0 = finite_or_zero(ra[e.b0]), a1 = finite_or_zero(ra[e.b1]);
  const float c0 = finite_or_zero(rb[e.b0]), c1 = finite_or_zero(rb[e.b1]);
  const float lo = fmaf(e.fa, c0 - a0, a0);
  const float hi = fmaf(e.fa, c1 - a1, a1);
  return fmaf(e.fb, hi - lo, lo);
}

// Table entry of channel c for effective row q (one table per row parity
// with z FFS).
__device__ __forceinline__ const RebinDev& rebin_entry_of(const RebinFilterArgs& a, int q, int c) {
  return a.tab[(size_t)(q & a.zshift) * a.ncols + c];
}

// The raw scan as seen by one line pair (views j0 <= j1 <= j0 + 1): a
// buffer resource based at the first view group any of the pair's taps can
// reach, m_lo = max(0, min(j0 + aoff_min, mmax_min)) -- every clamped view
// index max(min(j + aoff, mmax), 0) of the table is >= m_lo -- so that the
// taps are 32-bit offsets from it for any scan size (the planner checks that
// the fan's span of view groups fits 4 GiB).
constexpr int kRawRsrcWord3 = 0x00020000;   // raw buffer, 32-bit data format (gfx9)

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

struct RawWindow {
  __amdgpu_buffer_rsrc_t rsrc;
  int m_lo;
};

__device__ __forceinline__ RawWindow raw_window(const RebinFilterArgs& a, int j0) {
  RawWindow w;
  w.m_lo = max(0, min(j0 + a.aoff_min, a.mmax_min));
  const size_t base = (size_t)w.m_lo * a.group_bytes;
  const size_t n = (size_t)a.raw_elems * sizeof(float) - base;
  w.rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<char*>(reinterpret_cast<const char*>(a.raw) + base), 0,
      (int)(uint32_t)min(n, (size_t)0xFFFFFFF0u), kRawRsrcWord3);
  return w;
}

// The four taps of rebin_sample: views ma / mb at fan channels b0 / b1.
struct RebinTaps {
  float a0, a1, c0, c1;
};

// What the gather needs of a table entry.  Without z FFS a thread's four
// entries are the same for every line pair: they are packed once into 16
// bytes each in LDS (boff | last | valid << 1, aoff in 12 bits | mmax << 12,
// fa, fb; the planner checks |aoff| < 2048 and mmax < 2^20) and read back
// with one LDS load per pair instead of two global loads.
struct GatherEntry {
  uint32_t boff;
  int aoff, mmax;
  float fa, fb;
  bool last, valid;   // b1 == b0 (last fan channel); a fan sample exists
};

__device__ __forceinline__ GatherEntry gather_entry(const RebinDev& e) {
  return GatherEntry{e.boff, e.aoff, e.mmax, e.fa, e.fb, e.b1 == e.b0, e.state != 0};
}

__device__ __forceinline__ uint4 pack_entry(const RebinDev& e) {
  return make_uint4(e.boff | (e.b1 == e.b0 ? 1u : 0u) | (e.state ? 2u : 0u),
                    ((uint32_t)e.aoff & 0xfffu) | ((uint32_t)e.mmax << 12), __float_as_uint(e.fa),
                    __float_as_uint(e.fb));
}

__device__ __forceinline__ GatherEntry unpack_entry(uint4 p) {
  return GatherEntry{p.x & ~3u, (int)(p.y << 20) >> 20, (int)(p.y >> 12), __uint_as_float(p.z),
                     __uint_as_float(p.w), (p.x & 1u) != 0, (p.x & 2u) != 0};
}

// rebin_sample's taps through the window (jr = j - m_lo; rowB = byte offset
// of the raw row): the same clamped views.  The two fan channels of a tap
// row are adjacent (b1 = b0 + 1) except at the last fan channel (b1 = b0):
// one 8-byte load per view reads both, starting at b0, or at b0 - 1 for the
// last channel (inside the row: plans have at least two channels), whose
// second element is then the tap for both -- the values of four single loads.
__device__ __forceinline__ RebinTaps rebin_taps_w(const RawWindow& w, int jr, uint32_t rowB,
                                                  const GatherEntry& e, uint32_t group_bytes) {
  const int x = jr + e.aoff;
  const int top = e.mmax - w.m_lo;
  const uint32_t ma = (uint32_t)max(min(x, top), -w.m_lo);
  const uint32_t mb = (uint32_t)max(min(x + 1, top), -w.m_lo);
  const bool last = e.last;
  const uint32_t o = e.boff + rowB - (last ? 4u : 0u);
  const uint32_t oa = __umul24(ma, group_bytes) + o, ob = __umul24(mb, group_bytes) + o;
  const u32x2 ta = __builtin_amdgcn_raw_buffer_load_b64(w.rsrc, (int)oa, 0, 0);
  const u32x2 tb = __builtin_amdgcn_raw_buffer_load_b64(w.rsrc, (int)ob, 0, 0);
  return RebinTaps{__uint_as_float(last ? ta.y : ta.x), __uint_as_float(ta.y),
                   __uint_as_float(last ? tb.y : tb.x), __uint_as_float(tb.y)};
}

// The bilinear value of rebin_sample from the taps, as they are (fine when
// every tap is finite) or with finite_or_zero per tap.
__device__ __forceinline__ float rebin_combine(const RebinTaps& t, const GatherEntry& e) {
  const float lo = fmaf(e.fa, t.c0 - t.a0, t.a0);
  const float hi = fmaf(e.fa, t.c1 - t.a1, t.a1);
  return fmaf(e.fb, hi - lo, lo);
}
__device__ __forceinline__ float rebin_combine_finite(const RebinTaps& t, const GatherEntry& e) {
  return rebin_combine(RebinTaps{finite_or_zero(t.a0), finite_or_zero(t.a1),
                                 finite_or_zero(t.c0), finite_or_zero(t.c1)},
                       e);
}

// Plain combinations v[m] of the taps tp[m], recombined with finite_or_zero
// per tap in lanes where some value is not finite (bitwise the same result:
// finite_or_zero leaves finite taps alone, and a non-finite tap makes the
// plain combination non-finite); 0 past ncols and without a fan sample
// (cuda_kernels.cuh:32-34).
template <int K>
__device__ __forceinline__ void rebin_values(const RebinTaps* tp, const GatherEntry* const* e,
                                             const bool* in, float* v) {
  bool finite = true;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = rebin_combine(tp[k], *e[k]);
    finite = finite && __builtin_isfinite(v[k]);
  }
  if (!finite) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = rebin_combine_finite(tp[k], *e[k]);
  }
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = in[k] && e[k]->valid ? v[k] : 0.0f;
}

// Rebinned samples of the pair's lines (j0, q0) and (j1, q1) at channels
// t + m T, m = 0..3 (the register kernel's ownership); vb = 0 without a
// second line.  Channels past ncols read the last channel's entry.  Without
// z FFS (ZFFS false) both lines use the same table entries, read once, and
// the taps of all eight samples are in flight together; with z FFS the row
// parities pick different tables and the lines are gathered one after the
// other.
template <bool ZFFS>
__device__ __forceinline__ void gather_pair(const RebinFilterArgs& a, int t, int T, int j0,
                                            int q0, int j1, int q1, bool has1, float va[4],
                                            float vb[4], const uint4* ents) {
  const RawWindow w = raw_window(a, j0);
  const uint32_t rstride = (uint32_t)a.raw_ncols * (uint32_t)sizeof(float);
  const uint32_t rowB0 = (uint32_t)(q0 >> a.zshift) * rstride;
  const uint32_t rowB1 = (uint32_t)(q1 >> a.zshift) * rstride;
  const int jr0 = j0 - w.m_lo, jr1 = (has1 ? j1 : j0) - w.m_lo;
  GatherEntry e[4];
  bool in[8];
  const GatherEntry* ep[8];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    e[m] = ZFFS ? gather_entry(rebin_entry_of(a, q0, min(t + m * T, a.ncols - 1)))
                : unpack_entry(ents[m * T + t]);
    in[m] = t + m * T < a.ncols;
    in[m + 4] = has1 && in[m];
    ep[m] = ep[m + 4] = &e[m];
  }
  if (!ZFFS) {
    RebinTaps tp[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      tp[m] = rebin_taps_w(w, jr0, rowB0, e[m], a.group_bytes);
      tp[m + 4] = rebin_taps_w(w, jr1, rowB1, e[m], a.group_bytes);
    }
    float v[8];
    rebin_values<8>(tp, ep, in, v);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      va[m] = v[m];
      vb[m] = v[m + 4];
    }
    return;
  }
  RebinTaps tp[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) tp[m] = rebin_taps_w(w, jr0, rowB0, e[m], a.group_bytes);
  rebin_values<4>(tp, ep, in, va);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    e[m] = gather_entry(rebin_entry_of(a, q1, min(t + m * T, a.ncols - 1)));
    tp[m] = rebin_taps_w(w, jr1, rowB1, e[m], a.group_bytes);
  }
  rebin_values<4>(tp, ep + 4, in + 4, vb);
}

__device__ FctCplx* fft_forward(FctCplx* a, FctCplx* b, const FctCplx* __restrict__ tw, int N,
                                int log2n) {
  int Ns = 1, rem = log2n;
  while (rem > 0) {
    if (rem >= 2) {
      for (int j = threadIdx.x; j < (N >> 2); j += kRfThreads) fct_fft_r4(a, b, tw, N, Ns, j);
      Ns <<= 2;
      rem -= 2;
    } else {
      for (int j = threadIdx.x; j < (N >> 1); j += kRfThreads) fct_fft_r2(a, b, tw, N, Ns, j);
      Ns <<= 1;
      rem -= 1;
    }
    __syncthreads();
    FctCplx* t = a;
    a = b;
    b = t;
  }
  return a;
}

template <class T>
__device__ __forceinline__ T to_store(float v);
template <>
__device__ __forceinline__ __half to_store<__half>(float v) { return __float2half_rn(v); }
template <>
__device__ __forceinline__ __hip_bfloat16 to_store<__hip_bfloat16>(float v) {
  return __float2bfloat16(v);
}

__device__ __forceinline__ float from_store(__half v) { return __half2float(v); }
__device__ __forceinline__ float from_store(__hip_bfloat16 v) { return __bfloat162float(v); }

// Scale of FP16 / BF16 storage from the abs-max pass: the power of two s with
// bound / s < 2^15, bound = max|raw| * sum|h| * (1 + 2^-7) (the margin covers
// the FP32 transform's round-off).  1 for an all-zero (or non-finite) scan.
__device__ __forceinline__ float scale_of_bound(const float* scale, float hsum) {
  const float m = __uint_as_float(reinterpret_cast<const uint32_t*>(scale)[1]);
  const float bound = m * hsum * 1.0078125f;
  if (!(bound > 0.0f) || !(bound <= 3.0e38f)) return 1.0f;
  int e = 0;
  frexpf(bound, &e);   // bound < 2^e
  return ldexpf(1.0f, e - 15);
}

__device__ __forceinline__ float storage_scale(const RebinFilterArgs& a) {
  return scale_of_bound(a.scale, a.hsum);
}

// Streaming FP16 / BF16 storage: the bound in scale[1] grows as views
// arrive.  scale[0] becomes the scale of the bound so far and scale[2] the
// factor s_old / s_new (a power of two, 1 while the scale is unchanged) by
// which the quads stored under the previous scale are multiplied.
__global__ void scale_update_kernel(float* scale, float hsum) {
  if (threadIdx.x != 0) return;
  const float s_new = scale_of_bound(scale, hsum);
  const float s_old = scale[0];
  scale[2] = s_old > 0.0f ? s_old / s_new : 1.0f;
  scale[0] = s_new;
}

// scale[3] = scale[0]: the scale the streaming backprojections read, changed
// only while none of them runs.
__global__ void scale_commit_kernel(float* scale) {
  if (threadIdx.x == 0) scale[3] = scale[0];
}

// Multiplies n 16-bit values (four per 8-byte entry) by scale[2] when that
// is not 1: exact for results in the format's normal range.
template <class T>
__global__ __launch_bounds__(256) void quad_rescale_kernel(uint2* q, size_t n_entries,
                                                           const float* scale) {
  const float r = scale[2];
  if (r == 1.0f) return;
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n_entries; i += stride) {
    uint2 v = q[i];
    T h[4];
    __builtin_memcpy(h, &v, sizeof v);
#pragma unroll
    for (int k = 0; k < 4; ++k) h[k] = to_store<T>(from_store(h[k]) * r);
    __builtin_memcpy(&v, h, sizeof v);
    q[i] = v;
  }
}

// Value h of line (view, row r) at channel c into its four quad slots:
// (r, c).0, (r, c-1).1, (r-1, c).2 and (r-1, c-1).3 of the same view; the
// slots that would hold channel cpitch or row nrows are zeroed.  e is the
// entry index of (view, r, c).  Every slot of the buffer is written by exactly
// one line, so concurrent lines never race.
template <class T>
__device__ __forceinline__ void quad_store(T* __restrict__ q, size_t e, int r, int c, int nrows,
                                           int cpitch, T h) {
  const T z = to_store<T>(0.0f);
  const bool last_c = c == cpitch - 1;
  q[e * 4 + 0] = h;
  if (c > 0) q[(e - 1) * 4 + 1] = h;
  if (last_c) q[e * 4 + 1] = z;
  if (r > 0) {
    const size_t u = e - (size_t)cpitch;
    q[u * 4 + 2] = h;
    if (c > 0) q[(u - 1) * 4 + 3] = h;
    if (last_c) q[u * 4 + 3] = z;
  }
  if (r == nrows - 1) {
    q[e * 4 + 2] = z;
    q[e * 4 + 3] = z;
  }
}

template <class T>
__device__ __forceinline__ void quad_store_lines(const RebinFilterArgs& a, size_t o0, size_t o1,
                                                 int r0, int r1, bool has1, int c, float y0,
                                                 float y1) {
  T* q = static_cast<T*>(a.out);
  quad_store<T>(q, o0 + c, r0, c, a.nrows, a.cpitch, to_store<T>(y0));
  if (has1) quad_store<T>(q, o1 + c, r1, c, a.nrows, a.cpitch, to_store<T>(y1));
}

template <class T>
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
  const T l = to_store<T>(lo), h = to_store<T>(hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, l) |
         ((uint32_t)__builtin_bit_cast(uint16_t, h) << 16);
}

// The filtered output is stored non-temporally (streaming, not kept in L2):
// it is read back only by the backprojection, from HBM, while the raw fan rows
// the gather reads are each used by the ~150 parallel views that follow, and
// output lines kept in L2 would evict them first.
typedef uint32_t nt_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(uint32_t* p, uint32_t v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_nt(uint32_t* p, uint32_t lo, uint32_t hi) {
  nt_u32x2 v;
  v.x = lo;
  v.y = hi;
  __builtin_nontemporal_store(v, reinterpret_cast<nt_u32x2*>(p));
}

// Rows r (even) and r + 1 of one view at channel c, entry index e of (r, c),
// with the values at c and c + 1 of both rows: the whole quad (r, c), the
// first half of (r + 1, c) (with a zero second half on the last row) and the
// second half of (r - 1, c).  Together with the other pairs of the view every
// slot is written exactly once.
template <class T>
__device__ __forceinline__ void quad_store_pair(const RebinFilterArgs& a, size_t e, int r,
                                                float y00, float y01, float y10, float y11) {
  uint32_t* q = static_cast<uint32_t*>(a.out);   // two 32-bit halves per entry
  const uint32_t lo = pack2<T>(y00, y01), hi = pack2<T>(y10, y11);
  st_nt(q + e * 2, lo, hi);
  const size_t e1 = e + (size_t)a.cpitch;
  if (r + 1 == a.nrows - 1)
    st_nt(q + e1 * 2, hi, 0u);
  else
    st_nt(q + e1 * 2, hi);
  if (r > 0) st_nt(q + (e - (size_t)a.cpitch) * 2 + 1, lo);
}

// max|x| over the raw scan as FP32 bits in scale[1] (non-negative floats
// order like their bit patterns).
__global__ void scale_reset_kernel(float* scale) {
  if (threadIdx.x == 0) reinterpret_cast<uint32_t*>(scale)[1] = 0u;
}

// |v| as ordered bits, 0 for a non-finite v (the rebinning reads those as 0).
__device__ __forceinline__ uint32_t abs_bits(float v) {
  const uint32_t u = __float_as_uint(fabsf(v));
  return u < 0x7f800000u ? u : 0u;
}

// The FP16 / BF16 bound from a caller's max|raw| (|m| as ordered bits).
__global__ void scale_set_kernel(float* scale, float m) {
  if (threadIdx.x == 0) reinterpret_cast<uint32_t*>(scale)[1] = abs_bits(m);
}

template <bool kVec>
__global__ __launch_bounds__(256) void absmax_kernel(const float* __restrict__ x, int64_t n,
                                                     float* scale) {
  uint32_t m = 0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (kVec) {
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const int64_t n4 = n >> 2;
    auto absmax4 = [](const float4 v) {
      return max(max(abs_bits(v.x), abs_bits(v.y)), max(abs_bits(v.z), abs_bits(v.w)));
    };
    for (; i + 3 * stride < n4; i += 4 * stride) {   // four loads in flight per thread
      const float4 v0 = x4[i], v1 = x4[i + stride], v2 = x4[i + 2 * stride],
                   v3 = x4[i + 3 * stride];
      m = max(max(m, max(absmax4(v0), absmax4(v1))), max(absmax4(v2), absmax4(v3)));
    }
    for (; i < n4; i += stride) m = max(m, absmax4(x4[i]));
    i = 4 * n4 + (int64_t)blockIdx.x * 256 + threadIdx.x;
  }
  for (; i < n; i += stride) m = max(m, abs_bits(x[i]));
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
  if ((threadIdx.x & 63) == 0 && m) atomicMax(reinterpret_cast<uint32_t*>(scale) + 1, m);
}

__global__ __launch_bounds__(kRfThreads) void rebin_filter_kernel(RebinFilterArgs a) {
  extern __shared__ FctCplx lds[];
  const int N = a.nfft;
  FctCplx* buf0 = lds;
  FctCplx* buf1 = lds + N;
  const int l0 = a.line0 + 2 * blockIdx.x;
  const int l1 = l0 + 1;
  const bool has1 = l1 < a.line0 + a.nlines;
  const int j0 = l0 / a.nrows, r0 = l0 - j0 * a.nrows;
  const int j1 = has1 ? l1 / a.nrows : 0, r1 = has1 ? l1 - j1 * a.nrows : 0;

  for (int c = threadIdx.x; c < N; c += kRfThreads) {
    float va = 0.0f, vb = 0.0f;
    if (a.lines_in) {   // filter-only: the lines are already parallel rows
      if (c < a.ncols) {
        va = a.lines_in[(size_t)l0 * a.cpitch + c];
        if (has1) vb = a.lines_in[(size_t)l1 * a.cpitch + c];
      }
    } else if (c < a.ncols) {
      // No fan sample (fan channel outside [0, ncols)): 0 (cuda_kernels.cuh:32-34).
      const RebinDev e0 = rebin_entry_of(a, r0, c);
      if (e0.state) va = rebin_sample(a.raw, j0, r0, e0, a);
      if (has1) {
        const RebinDev e1 = rebin_entry_of(a, r1, c);
        if (e1.state) vb = rebin_sample(a.raw, j1, r1, e1, a);
      }
    }
    buf0[c] = FctCplx{va, vb};
    if (a.rebin_only && c < a.cpitch) {   // rebinned rows, FP32, no filter
      float* out = static_cast<float*>(a.out);
      out[(size_t)l0 * a.cpitch + c] = va;
      if (has1) out[(size_t)l1 * a.cpitch + c] = vb;
    }
  }
  if (a.rebin_only) return;
  __syncthreads();
  const FctCplx* tw = reinterpret_cast<const FctCplx*>(a.tw);
  FctCplx* X = fft_forward(buf0, buf1, tw, N, a.log2n);
  FctCplx* other = X == buf0 ? buf1 : buf0;
  for (int c = threadIdx.x; c < N; c += kRfThreads) {
    const float h = a.Hn[c];
    const FctCplx z = X[c];
    X[c] = FctCplx{z.x * h, -z.y * h};   // conj(H/N * Z)
  }
  __syncthreads();
  FctCplx* Y = fft_forward(X, other, tw, N, a.log2n);   // y = conj(Y)

  const size_t o0 = (size_t)l0 * a.cpitch, o1 = (size_t)l1 * a.cpitch;
  if (a.precision == FCT_PRECISION_FP32) {
    float* out = static_cast<float*>(a.out);
    for (int c = threadIdx.x; c < a.cpitch; c += kRfThreads) {
      const bool in = c < a.ncols;
      out[o0 + c] = in ? Y[c].x : 0.0f;
      if (has1) out[o1 + c] = in ? -Y[c].y : 0.0f;
    }
    return;
  }
  const float s = storage_scale(a);
  const float is = 1.0f / s;   // exact: a power of two
  if (blockIdx.x == 0 && threadIdx.x == 0) a.scale[0] = s;
  for (int c = threadIdx.x; c < a.cpitch; c += kRfThreads) {
    const bool in = c < a.ncols;
    const float y0 = in ? Y[c].x * is : 0.0f, y1 = in ? -Y[c].y * is : 0.0f;
    if (a.precision == FCT_PRECISION_FP16)
      quad_store_lines<__half>(a, o0, o1, r0, r1, has1, c, y0, y1);
    else
      quad_store_lines<__hip_bfloat16>(a, o0, o1, r0, r1, has1, c, y0, y1);
  }
}

// Register-resident variant for N in {512, 1024, 2048, 4096}: N / 8
// threads, 8 values each (fft_lds.h).  The block is persistent over line
// pairs, so the filter spectrum at its frequencies and its twiddles are
// loaded once into registers.  Thread t owns
// channels and frequencies t + m * N / 8: the rebinned values enter the first
// pass without an exchange, the spectrum product sits between the two
// transforms without an exchange, and the last inverse pass writes the
// filtered channels directly (coalesced).  Lines have ncols <= N / 2
// channels, so registers m >= 4 start as zero.
template <int N, int P>
__device__ __forceinline__ void load_twiddles(FctCplx* out, const FctCplx* __restrict__ tw, int t) {
  constexpr int R = FctFftShape<N>::radix(P), Ns = FctFftShape<N>::span(P);
#pragma unroll
  for (int u = 0; u < 8 / R; ++u)
#pragma unroll
    for (int r = 0; r < R; ++r)
      out[u * R + r] = r ? tw[fct_fft_tw_index<N>(t + u * FctFftShape<N>::T, r, R, Ns)]
                         : FctCplx{1.0f, 0.0f};
}

// Exchange ahead of pass P (P >= 1) through LDS.  With two buffers the
// exchanges alternate between them in the order they happen (E counts the
// exchanges before this one in the line pair; a pair has an even number), so
// a store never targets the buffer the previous exchange is still loading
// from and one barrier per exchange suffices.
template <int N, int P, bool kPingPong, int E>
__device__ __forceinline__ void fft_exchange(FctCplx* v, FctCplx* lds, int t) {
  constexpr int kSlots = N + N / 8;
  FctCplx* buf = kPingPong ? lds + (E & 1) * kSlots : lds;
  if (!kPingPong) __syncthreads();   // earlier loads from the single buffer are done
  fct_fft_pass_store<N, P - 1>(v, buf, t);
  __syncthreads();
  fct_fft_pass_load<N, P>(v, buf, t);
}

// One forward transform of the 8 values per thread in v (natural-order
// inputs t + m T in v[m]; outputs t + m T in v[fct_fft_final_reg<N>(m)]).
// E0: exchanges already made in this line pair (0 for the forward
// transform, kPasses - 1 for the inverse).
template <int N, bool kPingPong, int E0>
__device__ __forceinline__ void fft_forward_regs(FctCplx* v, FctCplx* lds, int t,
                                                 const FctCplx* tw1, const FctCplx* tw2,
                                                 const FctCplx* tw3) {
  fct_fft_pass_compute<N, 0>(v, nullptr);
  fft_exchange<N, 1, kPingPong, E0>(v, lds, t);
  fct_fft_pass_compute<N, 1>(v, tw1);
  fft_exchange<N, 2, kPingPong, E0 + 1>(v, lds, t);
  fct_fft_pass_compute<N, 2>(v, tw2);
  if constexpr (FctFftShape<N>::kPasses == 4) {
    fft_exchange<N, 3, kPingPong, E0 + 2>(v, lds, t);
    fct_fft_pass_compute<N, 3>(v, tw3);
  }
}

template <int N>
__global__ __launch_bounds__(N / 8) __attribute__((amdgpu_waves_per_eu(3, 8)))
void rebin_filter_reg_kernel(RebinFilterArgs a) {
  using S = FctFftShape<N>;
  constexpr int T = S::T;
  constexpr bool kPingPong = N <= 2048;
  constexpr int kSlots = N + N / 8;
  __shared__ FctCplx lds[kPingPong ? 2 * kSlots : kSlots];
  __shared__ float edge[T / 64][4][2];   // lane 0 of each wave: FP16 / BF16 neighbours
  __shared__ uint4 ents[4 * T];          // the thread's packed table entries (no z FFS)
  const int t = threadIdx.x;
  const FctCplx* tw = reinterpret_cast<const FctCplx*>(a.tw);
  // Each thread packs and later reads only its own slots (no barrier).
  if (!a.zshift && !a.lines_in) {
#pragma unroll
    for (int m = 0; m < 4; ++m)
      ents[m * T + t] = pack_entry(a.tab[min(t + m * T, a.ncols - 1)]);
  }

  FctCplx tw1[8], tw2[8], tw3[8];
  load_twiddles<N, 1>(tw1, tw, t);
  load_twiddles<N, 2>(tw2, tw, t);
  if constexpr (S::kPasses == 4) load_twiddles<N, 3>(tw3, tw, t);
  float hn[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) hn[m] = a.Hn[t + m * T];
  float is = 1.0f;   // 1 / storage scale (FP16 / BF16)
  if (a.precision != FCT_PRECISION_FP32 && !a.rebin_only) {
    const float s = storage_scale(a);
    is = 1.0f / s;   // exact: a power of two
    if (blockIdx.x == 0 && t == 0) a.scale[0] = s;
  }

  // Order of the line pairs.  With an even row count a pair is rows
  // (2i, 2i + 1) of one view, and the pairs are taken row pair by row pair,
  // sweeping the views: the blocks running together then read the same two
  // fan rows of neighbouring views, which stay in L2 (a parallel view reads
  // the fan views within about half a fan of its own).  The list is split
  // into eight contiguous ranges, one per XCD (block b runs on XCD b mod 8).
  // Odd row counts keep the line order (pairs straddle views).
  const int npairs = (a.nlines + 1) >> 1;
  const bool by_row = (a.nrows & 1) == 0 && (gridDim.x & 7) == 0;
  const int nv_range = a.nlines / a.nrows, j_first = a.line0 / a.nrows;
  const int xcd = blockIdx.x & 7, nb8 = gridDim.x >> 3;
  const int p_lo = by_row ? (int)(((long long)npairs * xcd) >> 3) : 0;
  const int p_hi = by_row ? (int)(((long long)npairs * (xcd + 1)) >> 3) : npairs;
  const int p_step = by_row ? nb8 : (int)gridDim.x;
  for (int pair = p_lo + (by_row ? (int)(blockIdx.x >> 3) : (int)blockIdx.x); pair < p_hi;
       pair += p_step) {
    int l0;
    if (by_row) {
      const int rp = pair / nv_range;
      l0 = (j_first + (pair - rp * nv_range)) * a.nrows + 2 * rp;
    } else {
      l0 = a.line0 + 2 * pair;
    }
    const int l1 = l0 + 1;
    const bool has1 = l1 < a.line0 + a.nlines;
    const int j0 = l0 / a.nrows, r0 = l0 - j0 * a.nrows;
    const int j1 = has1 ? l1 / a.nrows : 0, r1 = has1 ? l1 - j1 * a.nrows : 0;
    FctCplx v[8];
    if (a.lines_in) {
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int c = t + m * T;
        float va = 0.0f, vb = 0.0f;
        if (c < a.ncols) {
          va = a.lines_in[(size_t)l0 * a.cpitch + c];
          if (has1) vb = a.lines_in[(size_t)l1 * a.cpitch + c];
        }
        v[m] = FctCplx{va, vb};
        v[m + 4] = FctCplx{0.0f, 0.0f};
      }
    } else {
      // The table entries of the thread's four channels and then all their
      // taps in flight together (no branch between them; every tap offset
      // lies inside the scan for every entry).  Channels past ncols read the
      // last channel's entry; those and the entries without a fan sample
      // give 0 (cuda_kernels.cuh:32-34).
      float va[4], vb[4];
      if (a.zshift)
        gather_pair<true>(a, t, T, j0, r0, j1, r1, has1, va, vb, ents);
      else
        gather_pair<false>(a, t, T, j0, r0, j1, r1, has1, va, vb, ents);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        v[m] = FctCplx{va[m], vb[m]};
        v[m + 4] = FctCplx{0.0f, 0.0f};
      }
    }
    if (a.rebin_only) {   // rebinned rows, FP32, no filter (no LDS exchange: the
                          // exchange parity of the next pair is unchanged)
      float* out = static_cast<float*>(a.out);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int c = t + m * T;
        if (c < a.cpitch) {
          st_nt(out + (size_t)l0 * a.cpitch + c, v[m].x);
          if (has1) st_nt(out + (size_t)l1 * a.cpitch + c, v[m].y);
        }
      }
      continue;
    }
    fft_forward_regs<N, kPingPong, 0>(v, lds, t, tw1, tw2, tw3);
    FctCplx w[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const FctCplx z = v[fct_fft_final_reg<N>(m)];
      w[m] = FctCplx{z.x * hn[m], -z.y * hn[m]};   // conj(H/N * Z)
    }
    fft_forward_regs<N, kPingPong, S::kPasses - 1>(w, lds, t, tw1, tw2, tw3);   // y = conj(result)

    const size_t o0 = (size_t)l0 * a.cpitch, o1 = (size_t)l1 * a.cpitch;
    if (a.precision == FCT_PRECISION_FP32) {
      float* out = static_cast<float*>(a.out);
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int c = t + m * T;
        if (c < a.cpitch) {
          const FctCplx y = w[fct_fft_final_reg<N>(m)];
          const bool in = c < a.ncols;
          st_nt(out + o0 + c, in ? y.x : 0.0f);
          if (has1) st_nt(out + o1 + c, in ? -y.y : 0.0f);
        }
      }
      continue;
    }
    float ya[4], yb[4];   // the pair's scaled values at channel t + m T (0 past ncols)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const FctCplx y = w[fct_fft_final_reg<N>(m)];
      const bool in = t + m * T < a.ncols;
      ya[m] = in ? y.x * is : 0.0f;
      yb[m] = in ? -y.y * is : 0.0f;
    }
    if (has1 && (r0 & 1) == 0 && r1 == r0 + 1) {
      // Rows r0 (even) and r0 + 1 of one view: the 2 x 2 quads of row r0 are
      // complete here (one 8-byte store), row r0 + 1 gives the first half of
      // its quads (the second comes from the next pair, or is 0 on the last
      // row) and the second half of row r0 - 1's.  Channel c + 1 is held by
      // lane t + 1 (same m), or across a wave boundary by lane 0 of the next
      // wave, or for t = T - 1 by thread 0 at m + 1 (0 past N / 2 >= cpitch).
      if ((t & 63) == 0) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          edge[t >> 6][m][0] = ya[m];
          edge[t >> 6][m][1] = yb[m];
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const int c = t + m * T;
        float na = __shfl_down(ya[m], 1, 64), nb = __shfl_down(yb[m], 1, 64);
        if ((t & 63) == 63) {
          const int wn = (t >> 6) + 1;
          if (wn < T / 64) {
            na = edge[wn][m][0];
            nb = edge[wn][m][1];
          } else {
            na = m < 3 ? edge[0][m + 1][0] : 0.0f;
            nb = m < 3 ? edge[0][m + 1][1] : 0.0f;
          }
        }
        if (c >= a.cpitch) continue;
        if (a.precision == FCT_PRECISION_FP16)
          quad_store_pair<__half>(a, o0 + c, r0, ya[m], na, yb[m], nb);
        else
          quad_store_pair<__hip_bfloat16>(a, o0 + c, r0, ya[m], na, yb[m], nb);
      }
      continue;
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int c = t + m * T;
      if (c >= a.cpitch) continue;
      if (a.precision == FCT_PRECISION_FP16)
        quad_store_lines<__half>(a, o0, o1, r0, r1, has1, c, ya[m], yb[m]);
      else
        quad_store_lines<__hip_bfloat16>(a, o0, o1, r0, r1, has1, c, ya[m], yb[m]);
    }
  }
}

template <int N>
hipError_t launch_reg(const RebinFilterArgs& a, hipStream_t s) {
  static int per_cu = -1;   // resident blocks per CU (same for every device here)
  if (per_cu < 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, rebin_filter_reg_kernel<N>, N / 8, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    per_cu = nb;
  }
  const int npairs = (a.nlines + 1) / 2;
  int blocks = std::max(1, std::min(npairs, per_cu * std::max(a.ncu, 1)));
  if (blocks >= 8) blocks &= ~7;   // whole XCD groups (the row-pair-major order)
  hipLaunchKernelGGL(rebin_filter_reg_kernel<N>, dim3(blocks), dim3(N / 8), 0, s, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_scale_reset(float* scale, hipStream_t s) {
  hipLaunchKernelGGL(scale_reset_kernel, dim3(1), dim3(64), 0, s, scale);
  return hipGetLastError();
}

hipError_t launch_scale_set(float* scale, float max_abs_raw, hipStream_t s) {
  hipLaunchKernelGGL(scale_set_kernel, dim3(1), dim3(64), 0, s, scale, max_abs_raw);
  return hipGetLastError();
}

hipError_t launch_absmax(const float* x, int64_t n, float* scale, int ncu, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t want = (n / 4 + 255) / 256;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(want, 8LL * std::max(ncu, 1)));
  if ((reinterpret_cast<uintptr_t>(x) & 15) == 0)
    hipLaunchKernelGGL(absmax_kernel<true>, dim3(grid), dim3(256), 0, s, x, n, scale);
  else
    hipLaunchKernelGGL(absmax_kernel<false>, dim3(grid), dim3(256), 0, s, x, n, scale);
  return hipGetLastError();
}

hipError_t launch_scale_update(float* scale, float hsum, hipStream_t s) {
  hipLaunchKernelGGL(scale_update_kernel, dim3(1), dim3(64), 0, s, scale, hsum);
  return hipGetLastError();
}

hipError_t launch_scale_commit(float* scale, hipStream_t s) {
  hipLaunchKernelGGL(scale_commit_kernel, dim3(1), dim3(64), 0, s, scale);
  return hipGetLastError();
}

hipError_t launch_quad_rescale(void* quads, size_t n_entries, int precision, const float* scale,
                               int ncu, hipStream_t s) {
  if (n_entries == 0) return hipSuccess;
  const size_t want = (n_entries + 255) / 256;
  const int grid = (int)std::max<size_t>(1, std::min<size_t>(want, 8 * (size_t)std::max(ncu, 1)));
  if (precision == FCT_PRECISION_FP16)
    hipLaunchKernelGGL(quad_rescale_kernel<__half>, dim3(grid), dim3(256), 0, s,
                       static_cast<uint2*>(quads), n_entries, scale);
  else if (precision == FCT_PRECISION_BF16)
    hipLaunchKernelGGL(quad_rescale_kernel<__hip_bfloat16>, dim3(grid), dim3(256), 0, s,
                       static_cast<uint2*>(quads), n_entries, scale);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_rebin_filter(const RebinFilterArgs& a, hipStream_t s) {
  if (a.nlines <= 0) return hipSuccess;
  if (a.line0 & 1) return hipErrorInvalidValue;
  if (a.precision != FCT_PRECISION_FP32 && !a.rebin_only && !a.skip_scale) {
    if (!a.scale || a.lines_in) return hipErrorInvalidValue;
    hipError_t e = launch_scale_reset(a.scale, s);
    if (e == hipSuccess) e = launch_absmax(a.raw, a.raw_elems, a.scale, a.ncu, s);
    if (e != hipSuccess) return e;
  }
  switch (a.nfft) {
    case 512: return launch_reg<512>(a, s);
    case 1024: return launch_reg<1024>(a, s);
    case 2048: return launch_reg<2048>(a, s);
    case 4096: return launch_reg<4096>(a, s);
  }
  const int blocks = (a.nlines + 1) / 2;
  const size_t lds = 2 * (size_t)a.nfft * sizeof(FctCplx);
  hipLaunchKernelGGL(rebin_filter_kernel, dim3(blocks), dim3(kRfThreads), lds, s, a);
  return hipGetLastError();
}

}  // namespace fct
