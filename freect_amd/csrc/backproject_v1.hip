// backproject_v1.hip -- instruction-lean helical WFBP backprojection.
//
// Same arithmetic and selection as backproject.hip (cuda_kernels.cuh:198-287
// restated with the shared FP32 code of wfbp_math.h), organised to spend few
// VALU instructions per selected (voxel, view) sample -- the kernel is
// VALU-issue-bound:
//   * per (theta~, view parity) each lane's z-interval of |q| <= 1 is
//     [zt - kb - lt, zt - kb + lt]; a DPP wave reduction of its ends gives,
//     per view, the slices any lane of the wave can select, intersected with
//     the view window's exact slice range (a host table built with the same
//     FP32 test); other slices are skipped wave-uniformly and the exact
//     per-sample selection still runs, so selection is unchanged;
//   * the detector row position is one fma per slice, and when the central
//     row is (nrows-1)/2 the cone test |q| <= 1 is implied exactly by the row
//     test, one unsigned compare;
//   * taps are two 8-byte buffer loads with 32-bit byte offsets (the second
//     row through the scalar offset); lanes that select nothing read whatever
//     the range-checked buffer returns and get weight 0 instead of a branch;
//   * FP16 taps are interpolated across channels with v_dot2_f32_f16;
//   * the per-theta~ mean uses v_rcp.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "wfbp_internal.h"

namespace fct {
namespace {

constexpr int kRsrcWord3 = 0x00020000;   // raw buffer, 32-bit data format (gfx9)

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
}

// Wave-wide min / max (all 64 lanes must be active), result uniform.
__device__ __forceinline__ float wave_min(float v) {
  v = fminf(v, dpp<0xB1, 0xF>(v));    // quad_perm [1,0,3,2]
  v = fminf(v, dpp<0x4E, 0xF>(v));    // quad_perm [2,3,0,1]
  v = fminf(v, dpp<0x141, 0xF>(v));   // row_half_mirror
  v = fminf(v, dpp<0x140, 0xF>(v));   // row_mirror
  v = fminf(v, dpp<0x142, 0xA>(v));   // row_bcast:15
  v = fminf(v, dpp<0x143, 0xC>(v));   // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_max(float v) {
  v = fmaxf(v, dpp<0xB1, 0xF>(v));
  v = fmaxf(v, dpp<0x4E, 0xF>(v));
  v = fmaxf(v, dpp<0x141, 0xF>(v));
  v = fmaxf(v, dpp<0x140, 0xF>(v));
  v = fmaxf(v, dpp<0x142, 0xA>(v));
  v = fmaxf(v, dpp<0x143, 0xC>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b & 7, i = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + i;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Read-only table entry at a wave-uniform address, loaded through the
// constant address space so it becomes a scalar load into SGPRs even where
// the kernel also stores to global memory (the persistent batch loop).
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// Per (lane, view) interpolation state across the two channels pf, pf+1.
template <int PREC>
struct ChanInterp;

template <>
struct ChanInterp<FCT_PRECISION_FP32> {
  float wp;
  uint32_t base;   // byte offset of (view, row 0, pf)
  __device__ void init(uint32_t view_bytes, uint32_t pf, float w) {
    wp = w;
    base = view_bytes + pf * 4u;
  }
  __device__ float row(__amdgpu_buffer_rsrc_t r, uint32_t off, int soff) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, soff, 0);
    const float a = __uint_as_float(v[0]), b = __uint_as_float(v[1]);
    return fmaf(wp, b - a, a);
  }
};

template <>
struct ChanInterp<FCT_PRECISION_FP16> {
  half2_t w2;      // (1 - wp, wp)
  uint32_t base;   // 4-byte aligned byte offset of (view, row 0, pf & ~1)
  uint32_t shift;  // 0 or 2: byte position of channel pf inside the loaded pair
  __device__ void init(uint32_t view_bytes, uint32_t pf, float w) {
    w2 = half2_t{(_Float16)(1.0f - w), (_Float16)w};
    base = view_bytes + (pf & ~1u) * 2u;
    shift = (pf & 1u) * 2u;
  }
  __device__ float row(__amdgpu_buffer_rsrc_t r, uint32_t off, int soff) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, soff, 0);
    const uint32_t pair = __builtin_amdgcn_alignbyte(v[1], v[0], shift);
    const half2_t h = __builtin_bit_cast(half2_t, pair);
    return __builtin_amdgcn_fdot2(h, w2, 0.0f, false);
  }
};

template <>
struct ChanInterp<FCT_PRECISION_BF16> {
  float wp;
  uint32_t base;
  uint32_t shift;
  __device__ void init(uint32_t view_bytes, uint32_t pf, float w) {
    wp = w;
    base = view_bytes + (pf & ~1u) * 2u;
    shift = (pf & 1u) * 2u;
  }
  __device__ float row(__amdgpu_buffer_rsrc_t r, uint32_t off, int soff) const {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, soff, 0);
    const uint32_t pair = __builtin_amdgcn_alignbyte(v[1], v[0], shift);
    const float a = __uint_as_float(pair << 16), b = __uint_as_float(pair & 0xFFFF0000u);
    return fmaf(wp, b - a, a);
  }
};

// One wave reconstructs tile `lin` (slab-major order) of the job described by a.
template <int TZ, int PREC, int MODE, bool CENTERED>
__device__ __forceinline__ void bp_v1_tile(const BpArgs& a, int lin) {
  constexpr uint32_t kElem = PREC == FCT_PRECISION_FP32 ? 4u : 2u;
  const int lane = threadIdx.x;
  const int tiles = a.tiles_x * a.tiles_y;
  const int slab = lin / tiles;
  const int tile = lin - slab * tiles;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const int ix = tx * 8 + (lane & 7), iy = ty * 8 + (lane >> 3);
  const bool inb = ix < a.nx && iy < a.ny;
  const float x = inb ? a.xs[ix] : 0.0f;
  const float y = inb ? a.ys[iy] : 0.0f;
  const FctBpConsts& g = a.g;
  const bool active = inb && fct_in_fov(x, y, g);
  const int s0 = slab * TZ;
  const int nz = min(TZ, a.n_native - s0);
  if (!__any(active)) {
    if (inb)
      for (int i = 0; i < nz; ++i) {
        const size_t o = ((size_t)(s0 + i) * a.ny + iy) * a.nx + ix;
        if (MODE == kBpAccumulate || MODE == kBpMatched) a.native[o] = 0.0f;
        if (MODE == kBpCount) a.counts[o] = 0;
      }
    return;
  }
  const int chunk = s0 / fct_tables::kSliceChunk;
  float zs[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) zs[i] = ld_uniform(a.zs + s0 + (i < nz ? i : nz - 1));
  // Slice spacing inside the slab (uniform, FP32; used only for skipping).
  const float zfirst = zs[0], zlast = zs[nz - 1];
  const float dzs = nz > 1 ? (zlast - zfirst) / (float)(nz - 1) : 1.0f;
  const float inv_dzs = 1.0f / dzs;
  const float* ztc = a.zt + (size_t)chunk * a.nviews;
  const int* wrc = reinterpret_cast<const int*>(a.wrange + (size_t)chunk * a.nviews);
  const int* vrp = reinterpret_cast<const int*>(a.slab_vrange + slab);
  const int2 vr = make_int2(ld_uniform(vrp), ld_uniform(vrp + 1));
  const int H = a.H;
  const uint32_t vstride = (uint32_t)g.nrows * (uint32_t)a.cpitch;
  const uint32_t rowB = (uint32_t)a.cpitch * kElem;   // bytes per filtered row

  // Buffer resources based at the slab's first view (keeps offsets < 4 GiB).
  const char* fb = static_cast<const char*>(a.filt) + (size_t)vr.x * vstride * kElem;
  const uint32_t fbytes =
      (uint32_t)min((size_t)(vr.y - vr.x) * vstride * kElem, (size_t)0xFFFFFFF0u);
  const __amdgpu_buffer_rsrc_t rf =
      __builtin_amdgcn_make_buffer_rsrc((void*)fb, 0, fbytes, kRsrcWord3);
  constexpr bool kScaled = PREC != FCT_PRECISION_FP32;
  const void* sb = kScaled ? (const void*)(a.row_scale + (size_t)vr.x * g.nrows) : (const void*)fb;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)sb, 0, kScaled ? (uint32_t)((vr.y - vr.x) * g.nrows * 4) : 0u, kRsrcWord3);

  float acc[TZ];
  int cnt[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) {
    acc[i] = 0.0f;
    cnt[i] = 0;
  }
  const float BIG = 3.0e38f;

  for (int t = 0; t < H; ++t) {
    const int kmin = vr.x <= t ? 0 : (vr.x - t + H - 1) / H;
    const int kmax = vr.y <= t ? 0 : (vr.y - t + H - 1) / H;
    if (kmin >= kmax) continue;
    const float* csp = reinterpret_cast<const float*>(a.cs + t);
    const float2 cs = make_float2(ld_uniform(csp), ld_uniform(csp + 1));
    // fct_bp_column for (c, s) and (-c, -s): p, sqrt and asin are shared.
    const float p = fmaf(x, cs.y, -(y * cs.x));
    const float sq = sqrtf(fmaf(-p, p, g.R2));
    const float f = fmaf(x, cs.x, y * cs.y);
    const float kb = g.kappa * fct_asinf(p * g.invR);
    float num[TZ], den[TZ];
#pragma unroll
    for (int i = 0; i < TZ; ++i) {
      num[i] = 0.0f;
      den[i] = 0.0f;
    }
    int ilo_t = TZ, ihi_t = -1;

#pragma unroll
    for (int par = 0; par < 2; ++par) {
      int k = kmin + (((kmin & 1) != par) ? 1 : 0);
      if (k >= kmax) continue;
      const float pidx = par ? fmaf(-p, g.inv_ds, g.cc) : fmaf(p, g.inv_ds, g.cc);
      const float lt = (par ? sq + f : sq - f) * g.tanc;
      const float invLh = g.hr / lt;
      const float kbs = par ? -kb : kb;
      const float pfl = floorf(pidx);
      const bool chan_ok = active && pfl >= 0.0f && pfl <= (float)(g.ncols - 2);
      const uint32_t pf = chan_ok ? (uint32_t)(int)pfl : 0u;
      const float wp = pidx - pfl;
      // Matched mode: R / l, the voxel-driven footprint's depth correction
      // (a row spans drow l / Dsd of z at distance l from the source).
      const float fR = MODE == kBpMatched ? g.tanc / (lt * g.invR) : 0.0f;
      // z-interval of |q| <= 1 relative to the view's table z, over the wave.
      const float A = wave_min(active ? -kbs - lt : BIG);
      const float B = wave_max(active ? -kbs + lt : -BIG);
      for (; k < kmax; k += 2) {
        const int v = t + k * H;
        const float ztv = ld_uniform(ztc + v);
        // Exact window slice range (native indices).
        const int2 wr = make_int2(ld_uniform(wrc + 2 * v), ld_uniform(wrc + 2 * v + 1));
        float fa = (ztv + A - zfirst) * inv_dzs, fbb = (ztv + B - zfirst) * inv_dzs;
        if (fa > fbb) {
          const float tmp = fa;
          fa = fbb;
          fbb = tmp;
        }
        // (The matched projector pair uses every view that sees the voxel:
        // no view window.)
        const int ilo = max(max(0, (int)floorf(fa) - 1), MODE == kBpMatched ? 0 : wr.x - s0);
        const int ihi = min(min(nz - 1, (int)ceilf(fbb) + 1),
                            MODE == kBpMatched ? nz - 1 : wr.y - s0);
        if (ilo > ihi) continue;
        ilo_t = min(ilo_t, ilo);
        ihi_t = max(ihi_t, ihi);
        const float row0 = fmaf(kbs - ztv, invLh, g.cr);   // fct_bp_row0
        ChanInterp<PREC> ci;
        ci.init((uint32_t)(v - vr.x) * vstride * kElem, pf, wp);
        const uint32_t sbase = (uint32_t)(v - vr.x) * (uint32_t)g.nrows * 4u;
#pragma unroll
        for (int i = 0; i < TZ; ++i) {
          if (i < ilo || i > ihi) continue;
          const float qi = fmaf(zs[i], invLh, row0);       // fct_bp_row
          const float qfl = floorf(qi);
          const int iq = (int)qfl;
          bool sel;
          if (CENTERED)
            sel = chan_ok && (uint32_t)iq <= (uint32_t)(g.nrows - 2);
          else
            sel = chan_ok && !(fabsf(qi - g.cr) > g.hr) && qfl >= 0.0f &&
                  qfl <= (float)(g.nrows - 2);
          if (MODE == kBpCount) {
            cnt[i] += sel ? 1 : 0;
            continue;
          }
          // 24-bit multiply: exact for selected rows (< 2^24); other lanes
          // read an arbitrary in- or out-of-range offset and get weight 0.
          const uint32_t off = ci.base + __umul24((uint32_t)iq, rowB);
          float lo = ci.row(rf, off, 0);
          float hi = ci.row(rf, off, (int)rowB);
          if (kScaled) {
            const auto sc = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)(sbase + (uint32_t)iq * 4u),
                                                                 0, 0);
            lo *= __uint_as_float(sc[0]);
            hi *= __uint_as_float(sc[1]);
          }
          const float val = fmaf(qi - qfl, hi - lo, lo);
          if (MODE == kBpMatched) {   // plain sum over the views, weight R / l
            num[i] = fmaf(sel ? fR : 0.0f, val, num[i]);
          } else {
            const float W = sel ? fct_w_of_row(qi, g) : 0.0f;
            num[i] = fmaf(W, val, num[i]);
            den[i] += W;
          }
        }
      }
    }
    if (MODE == kBpAccumulate) {
#pragma unroll
      for (int i = 0; i < TZ; ++i)
        if (i >= ilo_t && i <= ihi_t)
          acc[i] = fmaf(num[i], __builtin_amdgcn_rcpf(fmaxf(den[i], 1e-30f)), acc[i]);
    } else if (MODE == kBpMatched) {
#pragma unroll
      for (int i = 0; i < TZ; ++i) acc[i] += num[i];
    }
  }
  if (!inb) return;
  if (MODE == kBpAccumulate || MODE == kBpMatched) {
#pragma unroll
    for (int i = 0; i < TZ; ++i)
      if (i < nz) a.native[((size_t)(s0 + i) * a.ny + iy) * a.nx + ix] = acc[i] * a.scale;
  } else if (MODE == kBpCount) {
    unsigned long long tot = 0;
#pragma unroll
    for (int i = 0; i < TZ; ++i)
      if (i < nz) {
        a.counts[((size_t)(s0 + i) * a.ny + iy) * a.nx + ix] = cnt[i];
        tot += (unsigned long long)cnt[i];
      }
    if (tot) atomicAdd(a.total, tot);
  }
}

template <int TZ, int PREC, int MODE, bool CENTERED>
__global__ __launch_bounds__(64) void bp_v1_kernel(BpArgs a) {
  bp_v1_tile<TZ, PREC, MODE, CENTERED>(a, xcd_block(blockIdx.x, gridDim.x));
}

// Persistent backprojection of many jobs in one launch.  The tiles of all
// jobs form one list (job-major, slab-major inside a job) split into eight
// contiguous ranges, one per XCD (workgroup b runs on XCD b mod 8), so that
// neighbouring tiles -- which read the same projection rows -- share an L2.
// A wave takes the next tile of its XCD's range with one atomic and moves on
// to the other ranges when its own is exhausted; every wave exits once all
// eight counters have passed their range ends.  A tile is computed by the
// same code as in bp_v1_kernel, so each job's volume is bitwise identical to
// its own single-job launch with the same tile_z.
template <int TZ, int PREC, bool CENTERED>
__global__ __launch_bounds__(64) void bp_batch_kernel(BpBatchArgs b,
                                                      const BpArgs* __restrict__ jobs,
                                                      const int32_t* __restrict__ tile_start) {
  const int xcd = blockIdx.x & 7;
  for (int step = 0; step < 8; ++step) {
    const int r = (xcd + step) & 7;
    const int lo = (int)(((long long)b.total_tiles * r) >> 3);
    const int hi = (int)(((long long)b.total_tiles * (r + 1)) >> 3);
    while (true) {
      uint32_t got = 0;
      if (threadIdx.x == 0) got = atomicAdd(&b.queue[r], 1u);
      const int tile = lo + (int)__builtin_amdgcn_readfirstlane(__shfl(got, 0, 64));
      if (tile >= hi) break;
      int j = 0;   // job: last j with tile_start[j] <= tile
      for (int span = 32; span > 0; span >>= 1)
        if (j + span < b.n_jobs && ld_uniform(tile_start + j + span) <= tile) j += span;
      bp_v1_tile<TZ, PREC, kBpAccumulate, CENTERED>(jobs[j], tile - ld_uniform(tile_start + j));
    }
  }
}

template <int TZ, int PREC, bool C>
hipError_t launch3(const BpArgs& a, int mode, hipStream_t s) {
  const int blocks = a.tiles_x * a.tiles_y * a.nslabs;
  if (mode == kBpAccumulate)
    hipLaunchKernelGGL((bp_v1_kernel<TZ, PREC, kBpAccumulate, C>), dim3(blocks), dim3(64), 0, s, a);
  else if (mode == kBpCount)
    hipLaunchKernelGGL((bp_v1_kernel<TZ, PREC, kBpCount, C>), dim3(blocks), dim3(64), 0, s, a);
  else if (mode == kBpMatched && PREC == FCT_PRECISION_FP32)
    hipLaunchKernelGGL((bp_v1_kernel<TZ, PREC, kBpMatched, C>), dim3(blocks), dim3(64), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int TZ, int PREC>
hipError_t launch2(const BpArgs& a, int mode, hipStream_t s) {
  const bool centered = a.g.cr == a.g.hr;
  return centered ? launch3<TZ, PREC, true>(a, mode, s) : launch3<TZ, PREC, false>(a, mode, s);
}

template <int TZ>
hipError_t launch1(const BpArgs& a, int precision, int mode, hipStream_t s) {
  switch (precision) {
    case FCT_PRECISION_FP32: return launch2<TZ, FCT_PRECISION_FP32>(a, mode, s);
    case FCT_PRECISION_FP16: return launch2<TZ, FCT_PRECISION_FP16>(a, mode, s);
    case FCT_PRECISION_BF16: return launch2<TZ, FCT_PRECISION_BF16>(a, mode, s);
  }
  return hipErrorInvalidValue;
}

template <int TZ, int PREC>
hipError_t launch_batch2(const BpBatchArgs& b, bool centered, int grid, hipStream_t s) {
  if (centered)
    hipLaunchKernelGGL((bp_batch_kernel<TZ, PREC, true>), dim3(grid), dim3(64), 0, s, b, b.jobs,
                       b.tile_start);
  else
    hipLaunchKernelGGL((bp_batch_kernel<TZ, PREC, false>), dim3(grid), dim3(64), 0, s, b, b.jobs,
                       b.tile_start);
  return hipGetLastError();
}

template <int TZ>
hipError_t launch_batch1(const BpBatchArgs& b, int precision, bool centered, int grid,
                         hipStream_t s) {
  switch (precision) {
    case FCT_PRECISION_FP32: return launch_batch2<TZ, FCT_PRECISION_FP32>(b, centered, grid, s);
    case FCT_PRECISION_FP16: return launch_batch2<TZ, FCT_PRECISION_FP16>(b, centered, grid, s);
    case FCT_PRECISION_BF16: return launch_batch2<TZ, FCT_PRECISION_BF16>(b, centered, grid, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_backproject_batch(const BpBatchArgs& b, int tile_z, int precision,
                                    bool centered, int grid, hipStream_t s) {
  if (b.n_jobs < 1 || b.n_jobs > 64 || grid < 8 || (grid & 7)) return hipErrorInvalidValue;
  switch (tile_z) {
    case 4: return launch_batch1<4>(b, precision, centered, grid, s);
    case 8: return launch_batch1<8>(b, precision, centered, grid, s);
    case 16: return launch_batch1<16>(b, precision, centered, grid, s);
    case 32: return launch_batch1<32>(b, precision, centered, grid, s);
  }
  return hipErrorInvalidValue;
}

hipError_t backproject_batch_waves_per_cu(int tile_z, int precision, int* waves) {
  // Occupancy of the FP32 / FP16 / BF16 instantiations (the centred and
  // general forms use the same registers to within a few).
  const void* f = nullptr;
#define FCT_BATCH_K(TZ)                                                             \
  case TZ:                                                                          \
    f = precision == FCT_PRECISION_FP32                                             \
            ? (const void*)bp_batch_kernel<TZ, FCT_PRECISION_FP32, false>           \
            : precision == FCT_PRECISION_FP16                                       \
                  ? (const void*)bp_batch_kernel<TZ, FCT_PRECISION_FP16, false>     \
                  : (const void*)bp_batch_kernel<TZ, FCT_PRECISION_BF16, false>;    \
    break;
  switch (tile_z) {
    FCT_BATCH_K(4)
    FCT_BATCH_K(8)
    FCT_BATCH_K(16)
    FCT_BATCH_K(32)
    default: return hipErrorInvalidValue;
  }
#undef FCT_BATCH_K
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(waves, f, 64, 0);
}

hipError_t launch_backproject_v1(const BpArgs& a, int tile_z, int precision, int mode,
                                 hipStream_t s) {
  switch (tile_z) {
    case 4: return launch1<4>(a, precision, mode, s);
    case 8: return launch1<8>(a, precision, mode, s);
    case 16: return launch1<16>(a, precision, mode, s);
    case 32: return launch1<32>(a, precision, mode, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fct
