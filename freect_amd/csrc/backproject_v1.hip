// backproject_v1.hip -- instruction-lean helical WFBP backprojection.
//
// Same arithmetic and selection as backproject.hip (cuda_kernels.cuh:198-287
// restated with the shared FP32 code of wfbp_math.h), reorganised to spend
// fewer VALU instructions per selected (voxel, view) sample:
//   * per (theta~, view parity) each lane's z-interval of |q| <= 1 is
//     [zt - kb - l tan, zt - kb + l tan]; a DPP wave reduction of its ends
//     gives, per view, the slices any lane of the wave can select, and the
//     other slices are skipped wave-uniformly (never changes a selection:
//     the exact per-sample test still runs, with a one-slice margin);
//   * when the central row is (nrows-1)/2, |q| <= 1 is implied exactly by the
//     row test 0 <= floor(qidx) <= nrows-2 (fma is monotone), which becomes
//     one unsigned compare;
//   * taps are read through a buffer resource with 32-bit offsets (adjacent
//     channels in one 8-byte load), out-of-range lanes read the 0 the buffer
//     unit returns, and the lane's weight is zeroed instead of branching;
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

// Two adjacent taps (channel pf and pf+1) of one filtered row, as floats.
template <int PREC>
__device__ __forceinline__ float2 load_pair(__amdgpu_buffer_rsrc_t r, uint32_t el);

template <>
__device__ __forceinline__ float2 load_pair<FCT_PRECISION_FP32>(__amdgpu_buffer_rsrc_t r,
                                                               uint32_t el) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(el * 4u), 0, 0);
  return make_float2(__uint_as_float(v[0]), __uint_as_float(v[1]));
}

template <>
__device__ __forceinline__ float2 load_pair<FCT_PRECISION_FP16>(__amdgpu_buffer_rsrc_t r,
                                                               uint32_t el) {
  const uint32_t b = el * 2u;
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(b & ~3u), 0, 0);
  const uint32_t w = __builtin_amdgcn_alignbyte(v[1], v[0], b & 2u);
  const __half2 h = *reinterpret_cast<const __half2*>(&w);
  return __half22float2(h);
}

template <>
__device__ __forceinline__ float2 load_pair<FCT_PRECISION_BF16>(__amdgpu_buffer_rsrc_t r,
                                                               uint32_t el) {
  const uint32_t b = el * 2u;
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)(b & ~3u), 0, 0);
  const uint32_t w = __builtin_amdgcn_alignbyte(v[1], v[0], b & 2u);
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xFFFF0000u));
}

template <int TZ, int PREC, int MODE, bool CENTERED>
__global__ __launch_bounds__(64) void bp_v1_kernel(BpArgs a) {
  constexpr int kElem = PREC == FCT_PRECISION_FP32 ? 4 : 2;
  const int lane = threadIdx.x;
  const int lin = xcd_block(blockIdx.x, gridDim.x);
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
        if (MODE == kBpAccumulate) a.native[o] = 0.0f;
        if (MODE == kBpCount) a.counts[o] = 0;
      }
    return;
  }
  const int chunk = s0 / fct_tables::kSliceChunk;
  float zs[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) zs[i] = a.zs[s0 + (i < nz ? i : nz - 1)];
  // Slice spacing inside the slab (uniform, FP32; used only for skipping).
  const float zfirst = zs[0], zlast = zs[nz - 1];
  const float dzs = nz > 1 ? (zlast - zfirst) / (float)(nz - 1) : 1.0f;
  const float inv_dzs = 1.0f / dzs;
  const float* ztc = a.zt + (size_t)chunk * a.nviews;
  const int2 vr = a.slab_vrange[slab];
  const int H = a.H;
  const uint32_t vstride = (uint32_t)g.nrows * (uint32_t)a.cpitch;

  // Buffer resources based at the slab's first view (keeps offsets < 4 GiB).
  const char* fb = static_cast<const char*>(a.filt) + (size_t)vr.x * vstride * kElem;
  const uint32_t fbytes = (uint32_t)min((size_t)(vr.y - vr.x) * vstride * kElem, (size_t)0xFFFFFFFFu);
  const __amdgpu_buffer_rsrc_t rf = __builtin_amdgcn_make_buffer_rsrc((void*)fb, 0, fbytes, kRsrcWord3);
  const bool scaled = PREC != FCT_PRECISION_FP32;
  const void* sb = scaled ? (const void*)(a.row_scale + (size_t)vr.x * g.nrows) : (const void*)fb;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)sb, 0, scaled ? (uint32_t)((vr.y - vr.x) * g.nrows * 4) : 0u, kRsrcWord3);

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
    const float2 cs = a.cs[t];
    // fct_bp_column for (c, s) and (-c, -s).
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
      const float Lt = (par ? sq + f : sq - f) * g.tanc;
      const float invL = 1.0f / Lt;
      const float kbs = par ? -kb : kb;
      const float pfl = floorf(pidx);
      const bool chan_ok = active && pfl >= 0.0f && pfl <= (float)(g.ncols - 2);
      const uint32_t pf = chan_ok ? (uint32_t)(int)pfl : 0u;
      const float wp = pidx - pfl;
      // z-interval of |q| <= 1 relative to the view's table z, over the wave.
      const float A = wave_min(active ? -kbs - Lt : BIG);
      const float B = wave_max(active ? -kbs + Lt : -BIG);
      for (; k < kmax; k += 2) {
        const int v = t + k * H;
        const float ztv = ztc[v];
        const float qoff = (kbs - ztv) * invL;
        // Slices any lane can select (window and q), one-slice margin.
        const float za = ztv + fmaxf(A, -g.zrot), zb = ztv + fminf(B, g.zrot);
        float fa = (za - zfirst) * inv_dzs, fbb = (zb - zfirst) * inv_dzs;
        if (fa > fbb) {
          const float tmp = fa;
          fa = fbb;
          fbb = tmp;
        }
        const int ilo = max(0, (int)floorf(fa) - 1);
        const int ihi = min(nz - 1, (int)ceilf(fbb) + 1);
        if (ilo > ihi) continue;
        ilo_t = min(ilo_t, ilo);
        ihi_t = max(ihi_t, ihi);
        const uint32_t vbase = (uint32_t)(v - vr.x) * vstride + pf;
        const uint32_t sbase = (uint32_t)(v - vr.x) * (uint32_t)g.nrows;
#pragma unroll
        for (int i = 0; i < TZ; ++i) {
          if (i < ilo || i > ihi) continue;
          if (!fct_bp_in_window(ztv, zs[i], g)) continue;
          const float q = fct_bp_q(zs[i], invL, qoff);
          const float qi = fct_bp_qidx(q, g);
          const float qfl = floorf(qi);
          const int iq = (int)qfl;
          bool sel;
          if (CENTERED) {
            sel = chan_ok && (uint32_t)iq <= (uint32_t)(g.nrows - 2);
          } else {
            sel = chan_ok && !(q > 1.0f || q < -1.0f) && qfl >= 0.0f &&
                  qfl <= (float)(g.nrows - 2);
          }
          if (MODE == kBpCount) {
            cnt[i] += sel ? 1 : 0;
            continue;
          }
          const uint32_t el = sel ? vbase + (uint32_t)iq * (uint32_t)a.cpitch : 0xFFFFFFF0u / kElem;
          float2 t0 = load_pair<PREC>(rf, el);
          float2 t1 = load_pair<PREC>(rf, el + (uint32_t)a.cpitch);
          float lo = fmaf(wp, t0.y - t0.x, t0.x);
          float hi = fmaf(wp, t1.y - t1.x, t1.x);
          if (PREC != FCT_PRECISION_FP32) {
            const auto sc = __builtin_amdgcn_raw_buffer_load_b64(
                rs, (int)((sel ? sbase + (uint32_t)iq : 0u) * 4u), 0, 0);
            lo *= __uint_as_float(sc[0]);
            hi *= __uint_as_float(sc[1]);
          }
          const float wq = qi - qfl;
          const float val = fmaf(wq, hi - lo, lo);
          const float W = sel ? fct_w_of_q(q) : 0.0f;
          num[i] = fmaf(W, val, num[i]);
          den[i] += W;
        }
      }
    }
    if (MODE == kBpAccumulate) {
#pragma unroll
      for (int i = 0; i < TZ; ++i)
        if (i >= ilo_t && i <= ihi_t)
          acc[i] = fmaf(num[i], __builtin_amdgcn_rcpf(fmaxf(den[i], 1e-30f)), acc[i]);
    }
  }
  if (!inb) return;
  if (MODE == kBpAccumulate) {
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

template <int TZ, int PREC, bool C>
hipError_t launch3(const BpArgs& a, int mode, hipStream_t s) {
  const int blocks = a.tiles_x * a.tiles_y * a.nslabs;
  if (mode == kBpAccumulate)
    hipLaunchKernelGGL((bp_v1_kernel<TZ, PREC, kBpAccumulate, C>), dim3(blocks), dim3(64), 0, s, a);
  else if (mode == kBpCount)
    hipLaunchKernelGGL((bp_v1_kernel<TZ, PREC, kBpCount, C>), dim3(blocks), dim3(64), 0, s, a);
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

}  // namespace

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
