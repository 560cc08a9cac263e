// adjoint.hip -- exact transpose of the FP32 WFBP chain (a matched helical
// forward projector): native volume -> filtered parallel rows -> raw fan data.
//
// The forward chain (rebin, filter, W-weighted backprojection,
// cuda_kernels.cuh:11-93, 198-287) is linear in the raw data for a fixed
// geometry: W(q) and its per-theta~ sum depend on geometry only.  Transposes:
//   backprojection^T: per voxel and theta~, the weight sum over the selected
//     views is formed first, then every selected sample scatters
//     scale * xbar * W / sum W times its four bilinear weights (FP32 hardware
//     atomics; summation order, not values, is non-deterministic);
//   filter^T: the filter (symmetric Toeplitz), run by rebin_filter.hip's
//     filter-only mode;
//   rebin^T: every parallel sample scatters its four bilinear weights to the
//     clamped fan samples it read.
// Selection and the FP32 weights are computed by the same code as the forward
// kernels, so the transpose is exact up to FP32 summation order.
#include "wfbp_internal.h"

namespace fct {
namespace {

// MATCHED: the transpose of backproject_v1.hip's matched mode instead --
// every view that sees the voxel, weight scale * R / l, no W normalisation
// and no view window: a voxel-driven helical projector of the native volume
// onto the parallel rows (with the rebinning transpose, onto the fan data).
template <int TZ, bool CENTERED, bool MATCHED>
__global__ __launch_bounds__(64) void bp_adjoint_kernel(BpAdjArgs a) {
  const BpArgs& b = a.bp;
  const int lane = threadIdx.x;
  const int tiles = b.tiles_x * b.tiles_y;
  const int slab = blockIdx.x / tiles;
  const int tile = blockIdx.x - slab * tiles;
  const int tx = tile % b.tiles_x, ty = tile / b.tiles_x;
  const int ix = tx * 8 + (lane & 7), iy = ty * 8 + (lane >> 3);
  const bool inb = ix < b.nx && iy < b.ny;
  const float x = inb ? b.xs[ix] : 0.0f;
  const float y = inb ? b.ys[iy] : 0.0f;
  const FctBpConsts& g = b.g;
  const bool active = inb && fct_in_fov(x, y, g);
  if (!__any(active)) return;
  const int s0 = slab * TZ;
  const int nz = min(TZ, b.n_native - s0);
  const int chunk = s0 / fct_tables::kSliceChunk;
  float zs[TZ], xb[TZ];
#pragma unroll
  for (int i = 0; i < TZ; ++i) {
    zs[i] = i < nz ? b.zs[s0 + i] : 0.0f;
    xb[i] = (i < nz && active) ? a.xbar[((size_t)(s0 + i) * b.ny + iy) * b.nx + ix] * b.scale
                               : 0.0f;
  }
  const float* ztc = b.zt + (size_t)chunk * b.nviews;
  const int2 vr = b.slab_vrange[slab];
  const int H = b.H;
  const size_t vstride = (size_t)g.nrows * b.cpitch;

  for (int t = 0; t < H; ++t) {
    const int kmin = vr.x <= t ? 0 : (vr.x - t + H - 1) / H;
    const int kmax = vr.y <= t ? 0 : (vr.y - t + H - 1) / H;
    if (kmin >= kmax) continue;
    const float2 cs = b.cs[t];
    const float p = fmaf(x, cs.y, -(y * cs.x));
    const float sq = sqrtf(fmaf(-p, p, g.R2));
    const float f = fmaf(x, cs.x, y * cs.y);
    const float kb = g.kappa * fct_asinf(p * g.invR);
    const float pidx_e = fmaf(p, g.inv_ds, g.cc), pidx_o = fmaf(-p, g.inv_ds, g.cc);
    const float invL_e = g.hr / ((sq - f) * g.tanc), invL_o = g.hr / ((sq + f) * g.tanc);
    const float fR_e = g.tanc / ((sq - f) * g.tanc * g.invR);   // R / l (matched mode)
    const float fR_o = g.tanc / ((sq + f) * g.tanc * g.invR);
    float den[TZ];
#pragma unroll
    for (int i = 0; i < TZ; ++i) den[i] = 0.0f;
    // Pass 0: weight sums; pass 1: scatter.
    for (int pass = MATCHED ? 1 : 0; pass < 2; ++pass) {
      float rden[TZ];
      if (pass == 1) {
#pragma unroll
        for (int i = 0; i < TZ; ++i) rden[i] = den[i] > 0.0f ? xb[i] / den[i] : 0.0f;
      }
      for (int k = kmin; k < kmax; ++k) {
        const int v = t + k * H;
        const bool odd = k & 1;
        const float pidx = odd ? pidx_o : pidx_e;
        const float invL = odd ? invL_o : invL_e;
        const float ztv = ztc[v];
        const float row0 = fmaf((odd ? -kb : kb) - ztv, invL, g.cr);   // fct_bp_row0
        const float pfl = floorf(pidx);
        const bool chan_ok = active && pfl >= 0.0f && pfl <= (float)(g.ncols - 2);
        const int pf = chan_ok ? (int)pfl : 0;
        const float wp = pidx - pfl;
        float* vb = a.gbar + (size_t)v * vstride + pf;
#pragma unroll
        for (int i = 0; i < TZ; ++i) {
          if (i >= nz || (!MATCHED && !fct_bp_in_window(ztv, zs[i], g))) continue;
          const float qi = fmaf(zs[i], invL, row0);   // fct_bp_row
          const float qfl = floorf(qi);
          bool sel;
          if (CENTERED)
            sel = chan_ok && (uint32_t)(int)qfl <= (uint32_t)(g.nrows - 2);
          else
            sel = chan_ok && !(fabsf(qi - g.cr) > g.hr) && qfl >= 0.0f &&
                  qfl <= (float)(g.nrows - 2);
          if (!sel) continue;
          if (MATCHED) {
            const float c = xb[i] * (odd ? fR_o : fR_e);
            const float wq = qi - qfl;
            float* r0 = vb + (size_t)(int)qfl * b.cpitch;
            float* r1 = r0 + b.cpitch;
            unsafeAtomicAdd(r0, c * (1.0f - wq) * (1.0f - wp));
            unsafeAtomicAdd(r0 + 1, c * (1.0f - wq) * wp);
            unsafeAtomicAdd(r1, c * wq * (1.0f - wp));
            unsafeAtomicAdd(r1 + 1, c * wq * wp);
            continue;
          }
          const float W = fct_w_of_row(qi, g);
          if (pass == 0) {
            den[i] += W;
            continue;
          }
          const float c = W * rden[i];
          if (c == 0.0f) continue;
          const float wq = qi - qfl;
          float* r0 = vb + (size_t)(int)qfl * b.cpitch;
          float* r1 = r0 + b.cpitch;
          unsafeAtomicAdd(r0, c * (1.0f - wq) * (1.0f - wp));
          unsafeAtomicAdd(r0 + 1, c * (1.0f - wq) * wp);
          unsafeAtomicAdd(r1, c * wq * (1.0f - wp));
          unsafeAtomicAdd(r1 + 1, c * wq * wp);
        }
      }
    }
  }
}

// Transpose of the rebinning: one block per parallel line (j, effective row
// q), threads over its channels; the same clamped (focal state, projection,
// raw row, fan channel) addressing as rebin_filter.hip's gather.
__global__ __launch_bounds__(256) void rebin_adjoint_kernel(RebinAdjArgs a) {
  const int line = blockIdx.x;             // j * nrows + q
  const int j = line / a.nrows, q = line - j * a.nrows;
  const int row = q >> a.zshift;
  const RebinDev* tab = a.tab + (size_t)(q & a.zshift) * a.ncols;
  for (int c = threadIdx.x; c < a.ncols; c += 256) {
    const RebinDev e = tab[c];
    if (!e.state) continue;
    const float y = a.gbar[(size_t)line * a.cpitch + c];
    if (y == 0.0f) continue;
    const int rho = e.state - 1;
    const int va = min(max(j + e.aoff, 0), e.mmax) * a.states + rho;
    const int vb = min(max(j + e.aoff + 1, 0), e.mmax) * a.states + rho;
    float* pa = a.pbar + ((size_t)va * a.raw_nrows + row) * a.raw_ncols;
    float* pb = a.pbar + ((size_t)vb * a.raw_nrows + row) * a.raw_ncols;
    unsafeAtomicAdd(pa + e.b0, (1.0f - e.fb) * (1.0f - e.fa) * y);
    unsafeAtomicAdd(pb + e.b0, (1.0f - e.fb) * e.fa * y);
    unsafeAtomicAdd(pa + e.b1, e.fb * (1.0f - e.fa) * y);
    unsafeAtomicAdd(pb + e.b1, e.fb * e.fa * y);
  }
}

__global__ __launch_bounds__(256) void fill_kernel(float* p, size_t n, float v) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = v;
}

template <int TZ>
hipError_t launch_tz(const BpAdjArgs& a, hipStream_t s, bool matched) {
  const int blocks = a.bp.tiles_x * a.bp.tiles_y * a.bp.nslabs;
  const bool c = a.bp.g.cr == a.bp.g.hr;
  if (matched) {
    if (c)
      hipLaunchKernelGGL((bp_adjoint_kernel<TZ, true, true>), dim3(blocks), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL((bp_adjoint_kernel<TZ, false, true>), dim3(blocks), dim3(64), 0, s, a);
  } else if (c) {
    hipLaunchKernelGGL((bp_adjoint_kernel<TZ, true, false>), dim3(blocks), dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL((bp_adjoint_kernel<TZ, false, false>), dim3(blocks), dim3(64), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_backproject_adjoint(const BpAdjArgs& a, int tile_z, hipStream_t s,
                                      bool matched) {
  switch (tile_z) {
    case 4: return launch_tz<4>(a, s, matched);
    case 8: return launch_tz<8>(a, s, matched);
    case 16: return launch_tz<16>(a, s, matched);
    case 32: return launch_tz<32>(a, s, matched);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rebin_adjoint(const RebinAdjArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(rebin_adjoint_kernel, dim3(a.nviews * a.nrows), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s) {
  const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s, p,
                     n, v);
  return hipGetLastError();
}

}  // namespace fct
