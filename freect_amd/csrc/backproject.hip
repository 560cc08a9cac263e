// backproject.hip -- voxel-driven helical WFBP backprojection.
//
// Replaces backproject_kernel (cuda_kernels.cuh:198-287, launched at
// ReconstructionMachine.cu:345-358).  For every native voxel and every
// theta~ in [0, ppr/2), the views theta~ + k*ppr/2 that select the voxel are
// combined as sum_k W(q_k) P_k / sum_k W(q_k) and the means are summed over
// theta~ with weight 2 pi / ppr.  A theta~ with no selecting view contributes
// nothing (the reference would produce 0/0).
//
// Work decomposition: one 64-lane wave per tile of 8x8 voxel columns x TZ
// slices.  The slices of a tile are accumulated in registers over all views,
// so every voxel is written exactly once.  Per (column, theta~) the geometry
// p, sqrt(R^2 - p^2), asin(p/R) is computed once and serves both the even
// views (angle theta) and the odd views (angle theta + pi); per view only the
// q offset changes; per slice q is one fma.  The view window test depends
// only on (view, slice) and is therefore wave-uniform.
//
// Selection arithmetic is the shared FP32 code of wfbp_math.h, which the CPU
// restatement uses too (bit-exact selection).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "wfbp_internal.h"

namespace fct {
namespace {

template <int PREC>
struct Elem;
template <>
struct Elem<FCT_PRECISION_FP32> {
  using T = float;
  static __device__ __forceinline__ float f(T v) { return v; }
};
template <>
struct Elem<FCT_PRECISION_FP16> {
  using T = __half;
  static __device__ __forceinline__ float f(T v) { return __half2float(v); }
};
template <>
struct Elem<FCT_PRECISION_BF16> {
  using T = __hip_bfloat16;
  static __device__ __forceinline__ float f(T v) { return __bfloat162float(v); }
};

// Logical block id such that each XCD (block b runs on XCD b % 8) receives a
// contiguous range of tiles, which keeps neighbouring tiles -- and the
// projection rows they share -- in one L2.
__device__ __forceinline__ int xcd_linear_block(int b, int nb) {
  const int x = b & 7, i = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + i;
}


template <int TZ, int PREC, int MODE>
__global__ __launch_bounds__(64) void bp_kernel(BpArgs a) {
  using E = Elem<PREC>;
  using T = typename E::T;
  const int lane = threadIdx.x;
  int tile, slab;
  if (MODE == kBpSelect) {
    tile = (a.sel_iy / 8) * a.tiles_x + a.sel_ix / 8;
    slab = a.sel_is / TZ;
  } else {
    const int lin = a.slab0 * a.tiles_x * a.tiles_y + xcd_linear_block(blockIdx.x, gridDim.x);
    const int tiles = a.tiles_x * a.tiles_y;
    slab = lin / tiles;
    tile = lin - slab * tiles;
  }
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
    if (MODE == kBpAccumulate && inb)
      for (int i = 0; i < nz; ++i) a.native[((size_t)(s0 + i) * a.ny + iy) * a.nx + ix] = 0.0f;
    if (MODE == kBpCount && inb)
      for (int i = 0; i < nz; ++i) a.counts[((size_t)(s0 + i) * a.ny + iy) * a.nx + ix] = 0;
    return;
  }
  const int chunk = s0 / fct_tables::kSliceChunk;
  float zs[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) zs[i] = i < nz ? a.zs[s0 + i] : 0.0f;
  const float* ztc = a.zt + (size_t)chunk * a.nviews;
  const int2 vr = a.slab_vrange[slab];
  const int H = a.H;
  const size_t vstride = (size_t)g.nrows * a.cpitch;
  const T* filt = static_cast<const T*>(a.filt);

  float acc[TZ];
  int cnt[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) {
    acc[i] = 0.0f;
    cnt[i] = 0;
  }

  for (int t = 0; t < H; ++t) {
    const int kmin = vr.x <= t ? 0 : (vr.x - t + H - 1) / H;
    const int kmax = vr.y <= t ? 0 : (vr.y - t + H - 1) / H;
    if (kmin >= kmax) continue;
    const float2 cs = a.cs[t];
    // fct_bp_column for (c, s) and (-c, -s), sharing p, sqrt and asin.
    const float p = fmaf(x, cs.y, -(y * cs.x));
    const float sq = sqrtf(fmaf(-p, p, g.R2));
    const float f = fmaf(x, cs.x, y * cs.y);
    const float kb = g.kappa * fct_asinf(p * g.invR);
    const float pidx_e = fmaf(p, g.inv_ds, g.cc);
    const float pidx_o = fmaf(-p, g.inv_ds, g.cc);
    const float invL_e = g.hr / ((sq - f) * g.tanc);   // FctBpColumn::invLh
    const float invL_o = g.hr / ((sq + f) * g.tanc);

    float num[TZ], den[TZ];
#pragma unroll
    for (int i = 0; i < TZ; ++i) {
      num[i] = 0.0f;
      den[i] = 0.0f;
    }
    for (int k = kmin; k < kmax; ++k) {
      const int v = t + k * H;
      const bool odd = k & 1;
      const float pidx = odd ? pidx_o : pidx_e;
      const float invL = odd ? invL_o : invL_e;
      const float kbs = odd ? -kb : kb;
      const float ztv = ztc[v];
      const float row0 = fmaf(kbs - ztv, invL, g.cr);   // fct_bp_row0
      const float pfl = floorf(pidx);
      const bool chan_ok = active && pfl >= 0.0f && pfl <= (float)(g.ncols - 2);
      const int pf = chan_ok ? (int)pfl : 0;
      // (FP16: the channel weight of the lean kernel's FP16 interpolation.)
      const float wp = PREC == FCT_PRECISION_FP16 ? fp16_channel_weight(pidx - pfl) : pidx - pfl;
      const T* vbase = filt + (size_t)v * vstride + pf;   // FP32 rows
      const T* qbase = filt + ((size_t)v * vstride + pf) * 4;   // FP16 / BF16 quads
#pragma unroll
      for (int i = 0; i < TZ; ++i) {
        if (i < nz && fct_bp_in_window(ztv, zs[i], g)) {
          const float qi = fmaf(zs[i], invL, row0);         // fct_bp_row
          const float qfl = floorf(qi);
          const bool sel = chan_ok && !(fabsf(qi - g.cr) > g.hr) && qfl >= 0.0f &&
                           qfl <= (float)(g.nrows - 2);
          if (MODE == kBpSelect) {
            if (s0 + i == a.sel_is && ix == a.sel_ix && iy == a.sel_iy)
              a.sel[v] = sel ? ((int)qfl * 65536 + pf) : -1;
          } else if (MODE == kBpCount) {
            cnt[i] += sel ? 1 : 0;
          } else if (sel) {
            const int qf = (int)qfl;
            const float wq = qi - qfl;
            float y00, y01, y10, y11;
            if (PREC == FCT_PRECISION_FP32) {
              const T* r0 = vbase + (size_t)qf * a.cpitch;
              const T* r1 = r0 + a.cpitch;
              y00 = E::f(r0[0]);
              y01 = E::f(r0[1]);
              y10 = E::f(r1[0]);
              y11 = E::f(r1[1]);
            } else {   // quad entry (qf, pf)
              const T* qd = qbase + (size_t)qf * a.cpitch * 4;
              y00 = E::f(qd[0]);
              y01 = E::f(qd[1]);
              y10 = E::f(qd[2]);
              y11 = E::f(qd[3]);
            }
            const float lo = fmaf(wp, y01 - y00, y00);
            const float hi = fmaf(wp, y11 - y10, y10);
            const float val = fmaf(wq, hi - lo, lo);
            const float W = fct_w_of_row(qi, g);
            num[i] = fmaf(W, val, num[i]);
            den[i] += W;
          }
        }
      }
    }
    if (MODE == kBpAccumulate) {
#pragma unroll
      for (int i = 0; i < TZ; ++i)
        if (den[i] > 0.0f) acc[i] += num[i] / den[i];
    }
  }
  if (!inb) return;
  if (MODE == kBpAccumulate) {
    // FP16 / BF16 projections carry one storage scale per scan.
    const float out_scale = PREC == FCT_PRECISION_FP32 ? a.scale : a.scale * a.filt_scale[0];
#pragma unroll
    for (int i = 0; i < TZ; ++i)
      if (i < nz) a.native[((size_t)(s0 + i) * a.ny + iy) * a.nx + ix] = acc[i] * out_scale;
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

template <int TZ, int PREC>
hipError_t launch_tz_prec(const BpArgs& a, int mode, hipStream_t s) {
  const int blocks =
      mode == kBpSelect ? 1 : a.tiles_x * a.tiles_y * (a.slab_count ? a.slab_count : a.nslabs);
  if (mode == kBpAccumulate)
    hipLaunchKernelGGL((bp_kernel<TZ, PREC, kBpAccumulate>), dim3(blocks), dim3(64), 0, s, a);
  else if (mode == kBpCount)
    hipLaunchKernelGGL((bp_kernel<TZ, PREC, kBpCount>), dim3(blocks), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((bp_kernel<TZ, PREC, kBpSelect>), dim3(blocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <int TZ>
hipError_t launch_tz(const BpArgs& a, int precision, int mode, hipStream_t s) {
  switch (precision) {
    case FCT_PRECISION_FP32: return launch_tz_prec<TZ, FCT_PRECISION_FP32>(a, mode, s);
    case FCT_PRECISION_FP16: return launch_tz_prec<TZ, FCT_PRECISION_FP16>(a, mode, s);
    case FCT_PRECISION_BF16: return launch_tz_prec<TZ, FCT_PRECISION_BF16>(a, mode, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_backproject(const BpArgs& a, int tile_z, int variant, int precision, int mode,
                              hipStream_t s) {
  // Slabs [slab0, slab0 + slab_count) inside the volume (slab_count 0: every slab).
  if (a.slab0 < 0 || a.slab_count < 0 || (a.slab_count == 0 && a.slab0 != 0) ||
      a.slab0 + a.slab_count > a.nslabs)
    return hipErrorInvalidValue;
  if (variant != FCT_BP_DIRECT && mode != kBpSelect)
    return launch_backproject_v1(a, tile_z, precision, mode, s);
  switch (tile_z) {
    case 4: return launch_tz<4>(a, precision, mode, s);
    case 8: return launch_tz<8>(a, precision, mode, s);
    case 16: return launch_tz<16>(a, precision, mode, s);
    case 32: return launch_tz<32>(a, precision, mode, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fct
