// backproject_v1.hip -- instruction-lean helical WFBP backprojection.
//
// Same arithmetic and selection as backproject.hip (cuda_kernels.cuh:198-287
// restated with the shared FP32 code of wfbp_math.h), organised to spend few
// instructions per selected (voxel, view) sample -- the kernel is bound by
// instruction issue (VALU and vector-memory), not by HBM:
//   * per (theta~, view parity) each lane's z-interval of |q| <= 1 is
//     [zt - kb - lt, zt - kb + lt]; the extremes of its ends over the tile
//     (DPP wave reductions, tabulated once per plan by bp_zbounds_kernel)
//     give, per view, the slices any lane of the wave can select, intersected
//     with the view window's exact slice range (a host table built with the
//     same FP32 test); other slices are skipped wave-uniformly and the exact
//     per-sample selection still runs, so selection is unchanged;
//   * the views of one theta~ and parity move monotonically in z, so the
//     view loop ends at the first view whose slice range lies past the slab;
//   * when a view covers the whole slab (the common case) the slices are
//     processed four at a time without a per-slice branch, the taps of the
//     four in flight together (with FP16 taps also for views covering part of
//     the slab, the slices outside the view's range masked);
//   * the detector row position is one fma per slice, and when the central
//     row is (nrows-1)/2 the cone test |q| <= 1 is implied exactly by the row
//     test, one unsigned compare;
//   * taps: FP32 rows are two 8-byte buffer loads, each at its own vector
//     offset; FP16 / BF16 projections are stored as 2 x 2 quads, one 8-byte
//     load per sample, and scaled once per scan at the end; FP16 quads are
//     interpolated across channels by two v_dot2_f32_f16 with the channel
//     weights packed as FP16 (exact products, FP32 sums); lanes that select
//     nothing read whatever the range-checked buffer returns (filtered
//     values of the slab, or 0 outside it) and get weight 0 instead of a
//     branch;
//   * the per-theta~ mean uses v_rcp.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "wfbp_internal.h"

namespace fct {
namespace {

constexpr int kRsrcWord3 = 0x00020000;   // raw buffer, 32-bit data format (gfx9)

// Lane move; rows outside ROW_MASK are left unspecified (the reductions
// below only read lane 63, whose inputs are always in enabled rows).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, ROW_MASK, 0xF, true);
}

// Order-preserving map of a float's bits to a signed int (negative floats get
// their magnitude bits flipped); the map is its own inverse.  Integer min /
// max fuse with the DPP lane moves, one instruction per step.
__device__ __forceinline__ int fkey(float x) {
  const int b = __float_as_int(x);
  return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float fkey_inv(int k) {
  return __int_as_float(k ^ ((k >> 31) & 0x7FFFFFFF));
}

// Wave-wide min / max (all 64 lanes must be active), result uniform.
__device__ __forceinline__ float wave_min(float x) {
  int v = fkey(x);
  v = min(v, dpp<0xB1, 0xF>(v));    // quad_perm [1,0,3,2]
  v = min(v, dpp<0x4E, 0xF>(v));    // quad_perm [2,3,0,1]
  v = min(v, dpp<0x141, 0xF>(v));   // row_half_mirror
  v = min(v, dpp<0x140, 0xF>(v));   // row_mirror
  v = min(v, dpp<0x142, 0xA>(v));   // row_bcast:15
  v = min(v, dpp<0x143, 0xC>(v));   // row_bcast:31
  return fkey_inv(__builtin_amdgcn_readlane(v, 63));
}
__device__ __forceinline__ float wave_max(float x) {
  int v = fkey(x);
  v = max(v, dpp<0xB1, 0xF>(v));
  v = max(v, dpp<0x4E, 0xF>(v));
  v = max(v, dpp<0x141, 0xF>(v));
  v = max(v, dpp<0x140, 0xF>(v));
  v = max(v, dpp<0x142, 0xA>(v));
  v = max(v, dpp<0x143, 0xC>(v));
  return fkey_inv(__builtin_amdgcn_readlane(v, 63));
}

__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b & 7, i = b >> 3;
  const int q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + i;
}


// Read-only table entry at a wave-uniform address, loaded through the
// constant address space so it becomes a scalar load into SGPRs even where
// the kernel also stores to global memory (the persistent batch loop).
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p) {
  return *(const __attribute__((address_space(4))) T*)p;
}

// Per (lane, view) bilinear interpolation: fetch() issues the loads of rows
// iq and iq + 1 at channels pf, pf + 1 from the byte offset of (view, iq, pf);
// value() interpolates them across the channels and then between the rows
// with the row fraction fq.
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// The filtered projections of a slab as a buffer resource based at the
// slab's first view.
struct BpRsrc {
  __amdgpu_buffer_rsrc_t rows;
};

template <int PREC>
struct ChanInterp;

template <>
struct ChanInterp<FCT_PRECISION_FP32> {   // plain rows: two 8-byte loads
  static constexpr uint32_t kEntry = 4;     // bytes per (row, channel)
  struct Raw {
    u32x2 a, b;
  };
  float wp;
  uint32_t base;   // byte offset of (view, row 0, pf)
  __device__ void init(uint32_t view_bytes, uint32_t pf, float w) {
    wp = w;
    base = view_bytes + pf * kEntry;
  }
  // Row iq + 1 at the vector offset off + rowB: both loads are range-checked
  // (the hardware's check covers the vector offset, not the scalar one), so
  // lanes that select nothing never read past the slab's projections.
  __device__ static Raw fetch(const BpRsrc& r, uint32_t off, uint32_t rowB) {
    return Raw{__builtin_amdgcn_raw_buffer_load_b64(r.rows, (int)off, 0, 0),
               __builtin_amdgcn_raw_buffer_load_b64(r.rows, (int)(off + rowB), 0, 0)};
  }
  __device__ float value(const Raw& t, float fq) const {
    const uint32_t a0 = t.a.x, a1 = t.a.y, b0 = t.b.x, b1 = t.b.y;
    const float fa0 = __uint_as_float(a0), fa1 = __uint_as_float(a1);
    const float fb0 = __uint_as_float(b0), fb1 = __uint_as_float(b1);
    const float r0 = fmaf(wp, fa1 - fa0, fa0), r1 = fmaf(wp, fb1 - fb0, fb0);
    return fmaf(fq, r1 - r0, r0);
  }
};

// a.lo * b.lo + a.hi * b.hi + c: two FP16 products (exact in FP32) and an
// FP32 sum in one dot instruction.  (Through the builtin, not inline asm: a
// dot result read by another VALU instruction needs wait states that only
// the compiler's hazard recognizer inserts.)
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float dot2_f16(uint32_t a, uint32_t b, float c) {
  return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, a), __builtin_bit_cast(half2_t, b), c,
                                false);
}

template <>
struct ChanInterp<FCT_PRECISION_FP16> {   // quad entries: one 8-byte load
  static constexpr uint32_t kEntry = 8;
  typedef u32x2 Raw;
  // Channel weights (1 - wq, wq) as packed FP16 and their negatives, with
  // wq = fp16_channel_weight(wp): both weights are exact in FP16 and sum to 1
  // (a position error below 2^-12 channel).
  uint32_t w, nw;
  uint32_t base;
  __device__ void init(uint32_t view_bytes, uint32_t pf, float wp) {
    const float wq = fp16_channel_weight(wp);
    w = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(1.0f - wq, wq));
    nw = w ^ 0x80008000u;
    base = view_bytes + pf * kEntry;
  }
  __device__ static Raw fetch(const BpRsrc& r, uint32_t off, uint32_t) {
    return __builtin_amdgcn_raw_buffer_load_b64(r.rows, (int)off, 0, 0);
  }
  __device__ float value(const Raw& t, float fq) const {
    // -r0 and r1 - r0 from the two rows' FP16 pairs, then r0 + fq (r1 - r0).
    // (Through named scalars: hipcc (ROCm 7.2) folds __builtin_bit_cast of a
    // vector element such as t[1] to element 0.)
    const uint32_t q0 = t.x, q1 = t.y;
    const float nr0 = dot2_f16(q0, nw, 0.0f);
    const float d = dot2_f16(q1, w, nr0);
    return fmaf(fq, d, -nr0);
  }
  // The same for four samples, as one block of eight v_dot2_f32_f16 in the
  // VOP3P form (the builtin selects the two-address v_dot2c form and copies
  // the accumulator).  Every dot result is read at least three instructions
  // after it is written (the wait a dot result needs before another VALU
  // instruction reads it): the second dot of a sample reads the first four
  // instructions later, and s_nop 2 ends the block.
  __device__ void values4(const Raw* t, const float* fq, float* val) const {
    const uint32_t a0 = t[0].x, a1 = t[1].x, a2 = t[2].x, a3 = t[3].x;
    const uint32_t b0 = t[0].y, b1 = t[1].y, b2 = t[2].y, b3 = t[3].y;
    float n0, n1, n2, n3, d0, d1, d2, d3;
    asm("v_dot2_f32_f16 %0, %8, %16, 0\n\t"
        "v_dot2_f32_f16 %1, %9, %16, 0\n\t"
        "v_dot2_f32_f16 %2, %10, %16, 0\n\t"
        "v_dot2_f32_f16 %3, %11, %16, 0\n\t"
        "v_dot2_f32_f16 %4, %12, %17, %0\n\t"
        "v_dot2_f32_f16 %5, %13, %17, %1\n\t"
        "v_dot2_f32_f16 %6, %14, %17, %2\n\t"
        "v_dot2_f32_f16 %7, %15, %17, %3\n\t"
        "s_nop 2"
        : "=&v"(n0), "=&v"(n1), "=&v"(n2), "=&v"(n3), "=&v"(d0), "=&v"(d1), "=&v"(d2),
          "=&v"(d3)
        : "v"(a0), "v"(a1), "v"(a2), "v"(a3), "v"(b0), "v"(b1), "v"(b2), "v"(b3), "v"(nw),
          "v"(w));
    val[0] = fmaf(fq[0], d0, -n0);
    val[1] = fmaf(fq[1], d1, -n1);
    val[2] = fmaf(fq[2], d2, -n2);
    val[3] = fmaf(fq[3], d3, -n3);
  }
};

template <>
struct ChanInterp<FCT_PRECISION_BF16> {   // quad entries: one 8-byte load
  static constexpr uint32_t kEntry = 8;
  typedef u32x2 Raw;
  float wp;
  uint32_t base;
  __device__ void init(uint32_t view_bytes, uint32_t pf, float w) {
    wp = w;
    base = view_bytes + pf * kEntry;
  }
  __device__ static Raw fetch(const BpRsrc& r, uint32_t off, uint32_t) {
    return __builtin_amdgcn_raw_buffer_load_b64(r.rows, (int)off, 0, 0);
  }
  __device__ float value(const Raw& t, float fq) const {
    const uint32_t q0 = t.x, q1 = t.y;
    const float a0 = __uint_as_float(q0 << 16), a1 = __uint_as_float(q0 & 0xFFFF0000u);
    const float b0 = __uint_as_float(q1 << 16), b1 = __uint_as_float(q1 & 0xFFFF0000u);
    const float r0 = fmaf(wp, a1 - a0, a0), r1 = fmaf(wp, b1 - b0, b0);
    return fmaf(fq, r1 - r0, r0);
  }
};

// Per (voxel column, theta~): p, sqrt(R^2 - p^2), the in-plane offset f and
// the cone-angle z term kb (fct_bp_column for the views of both parities,
// which share p, sqrt and asin).
struct BpColumn {
  float p, sq, f, kb;
};

// sqrt(x) correctly rounded, as sqrtf, for x >= 2^-96 (R^2 - p^2 here is
// of the order of R^2): the hardware square root (within 1 ulp) moved to
// its neighbour below or above when the residual x - s' s of that neighbour
// says the true root lies beyond it.  sqrtf's expansion adds a rescaling of
// tiny inputs and the zero / infinity cases, which cannot occur here.
__device__ __forceinline__ float sqrt_rn_large(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sdn = __int_as_float(__float_as_int(s) - 1);
  const float sup = __int_as_float(__float_as_int(s) + 1);
  float r = fmaf(-sdn, s, x) <= 0.0f ? sdn : s;
  r = fmaf(-sup, s, x) > 0.0f ? sup : r;
  return r;
}

// n / d correctly rounded, as the IEEE division, for operands far from the
// range limits (here n = hr < 2^15 and d = lt, a cone length of millimetres
// times tan(cone / 2)): the hardware reciprocal refined once, the quotient
// corrected twice by its exact residual -- the division's own sequence
// without its operand scaling and special-case fix-up, identities for such
// operands.
__device__ __forceinline__ float div_rn_normal(float n, float d) {
  float r = __builtin_amdgcn_rcpf(d);
  r = fmaf(fmaf(-d, r, 1.0f), r, r);
  float q = n * r;
  q = fmaf(fmaf(-d, q, n), r, q);
  return fmaf(fmaf(-d, q, n), r, q);
}
__device__ __forceinline__ BpColumn bp_column(float x, float y, float2 cs, const FctBpConsts& g) {
  BpColumn c;
  c.p = fmaf(x, cs.y, -(y * cs.x));
  c.sq = sqrt_rn_large(fmaf(-c.p, c.p, g.R2));
  c.f = fmaf(x, cs.x, y * cs.y);
  c.kb = g.kappa * fct_asinf(c.p * g.invR);
  return c;
}

// z-interval of |q| <= 1 relative to a view's table z for the voxels of a
// wave's tile in the FOV: [min(-kbs - lt), max(-kbs + lt)] with kbs = kb,
// lt = (sq - f) tan for parity 0 and kbs = -kb, lt = (sq + f) tan for parity
// 1, returned as (min, max, min, max).  It depends on the tile and theta~
// only, so bp_zbounds_kernel tabulates it once per plan.
__device__ __forceinline__ float4 tile_zbounds(const BpColumn& c, bool active,
                                               const FctBpConsts& g) {
  const float BIG = 3.0e38f;
  const float lt0 = (c.sq - c.f) * g.tanc, lt1 = (c.sq + c.f) * g.tanc;
  float4 r;
  r.x = wave_min(active ? -c.kb - lt0 : BIG);
  r.y = wave_max(active ? -c.kb + lt0 : -BIG);
  r.z = wave_min(active ? c.kb - lt1 : BIG);
  r.w = wave_max(active ? c.kb + lt1 : -BIG);
  return r;
}

// One wave per tile (all 64 lanes active, as the reductions need).
__global__ __launch_bounds__(64) void bp_zbounds_kernel(BpArgs a, float4* out) {
  const int tile = blockIdx.x;
  const int lane = threadIdx.x;
  const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
  const int ix = tx * 8 + (lane & 7), iy = ty * 8 + (lane >> 3);
  const bool inb = ix < a.nx && iy < a.ny;
  const float x = inb ? a.xs[ix] : 0.0f;
  const float y = inb ? a.ys[iy] : 0.0f;
  const bool active = inb && fct_in_fov(x, y, a.g);
  for (int t = 0; t < a.H; ++t) {
    const float4 r = tile_zbounds(bp_column(x, y, a.cs[t], a.g), active, a.g);
    if (lane == 0) out[(size_t)tile * a.H + t] = r;
  }
}

// One (voxel, view, slice) sample in two steps: bp_fetch computes the row
// position and the exact selection and issues the tap loads; bp_add
// interpolates and accumulates into (num, den) or cnt.  Branch-free: lanes
// that select nothing read whatever the range-checked buffer returns at their
// offset and get weight 0.
template <int PREC>
struct BpTap {
  float qi, fq;   // row position and its fraction qi - floor(qi)
  bool sel;
  typename ChanInterp<PREC>::Raw raw;
};

// floor(x) as an int in one instruction (hipcc emits v_floor + v_cvt for
// (int)floorf(x)); same value for every finite x in int range.
__device__ __forceinline__ int floor_int(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

template <int PREC, int MODE, bool CENTERED>
__device__ __forceinline__ BpTap<PREC> bp_fetch(float zsi, float invLh, float row0,
                                                const ChanInterp<PREC>& ci,
                                                const BpRsrc& rf, uint32_t rowB,
                                                bool chan_ok, const FctBpConsts& g) {
  BpTap<PREC> s;
  s.qi = fmaf(zsi, invLh, row0);   // fct_bp_row
  const int iq = floor_int(s.qi);
  // qi - floor(qi), exact for qi >= 0 (the selected rows); v_fract clamps
  // below 1 where that difference would round up, which only happens for
  // tiny negative qi (not selected).
  s.fq = __builtin_amdgcn_fractf(s.qi);
  if (CENTERED)   // |q| <= 1 is implied by the row test when cr = hr
    s.sel = chan_ok && (uint32_t)iq <= (uint32_t)(g.nrows - 2);
  else
    s.sel = chan_ok && !(fabsf(s.qi - g.cr) > g.hr) && iq >= 0 && iq <= g.nrows - 2;
  // 24-bit multiply: exact for selected rows (< 2^24).
  if (MODE != kBpCount)
    s.raw = ChanInterp<PREC>::fetch(rf, ci.base + __umul24((uint32_t)iq, rowB), rowB);
  return s;
}

// Weight of a sample (0 when it is not selected): W(q) for the WFBP sum,
// R / l for the matched projector's plain sum over the views.
//
// FREE (centred rows, cr == hr, WFBP sum): no selection mask.  The ramp
// argument t is clamped to [0, 1] by the fma's clamp modifier; the rows a
// sample may use, qi in [0, nrows - 1), are those with |qi - cr| < hr, and
// every other qi has |qi - cr| >= hr, where t <= 0 exactly (wk is rounded
// up), so W = 0 without a compare and a select.  Lanes whose column misses
// the detector or the FOV enter with row0 = -1e30 and rows per mm 0 (the
// caller's view setup): t = 0 for them too.  Their tap offsets stay inside
// the slab (the 24-bit row product of floor(-1e30) is 0) and the taps they
// read are finite, so num gains 0 * value = 0.
template <int PREC, int MODE, bool FREE = false>
__device__ __forceinline__ float bp_weight(const BpTap<PREC>& s, float fR, const FctBpConsts& g) {
  if (MODE == kBpMatched) return s.sel ? fR : 0.0f;
  if (FREE)
    return fct_w_of_t(
        __builtin_amdgcn_fmed3f(fmaf(-fabsf(s.qi - g.cr), g.wk, FCT_WC), 0.0f, 1.0f));
  const float w = fct_w_of_row(s.qi, g);
  return s.sel ? w : 0.0f;
}

// Accumulation of a sample of weight W (bp_weight) and interpolated value val.
template <int MODE>
__device__ __forceinline__ void bp_acc(float W, float val, float& num, float& den) {
  num = fmaf(W, val, num);
  if (MODE != kBpMatched) den += W;
}

template <int PREC, int MODE, bool FREE = false>
__device__ __forceinline__ void bp_add(const BpTap<PREC>& s, const ChanInterp<PREC>& ci, float fR,
                                       const FctBpConsts& g, float& num, float& den, int& cnt) {
  if (MODE == kBpCount) {
    cnt += s.sel ? 1 : 0;
    return;
  }
  bp_acc<MODE>(bp_weight<PREC, MODE, FREE>(s, fR, g), ci.value(s.raw, s.fq), num, den);
}

// Four samples of one view (the whole-slab path): FP16 through the dot block.
template <int PREC, int MODE, bool FREE = false>
__device__ __forceinline__ void bp_add4(const BpTap<PREC>* s, const ChanInterp<PREC>& ci,
                                        float fR, const FctBpConsts& g, float* num, float* den,
                                        int* cnt) {
  if constexpr (PREC == FCT_PRECISION_FP16 && MODE != kBpCount) {
    // The weights first: their arithmetic overlaps the tap loads.
    typename ChanInterp<PREC>::Raw t[4];
    float fq[4], W[4], val[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      t[i] = s[i].raw;
      fq[i] = s[i].fq;
      W[i] = bp_weight<PREC, MODE, FREE>(s[i], fR, g);
    }
    ci.values4(t, fq, val);
#pragma unroll
    for (int i = 0; i < 4; ++i) bp_acc<MODE>(W[i], val[i], num[i], den[i]);
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      bp_add<PREC, MODE, FREE>(s[i], ci, fR, g, num[i], den[i], cnt[i]);
  }
}

// Order of the tiles of a slab: square blocks of kSuper x kSuper tiles
// (row-major blocks, row-major tiles inside a block; partial blocks at the
// right and bottom edges), so that the tiles an XCD runs at the same time
// cover compact patches of the slice rather than a band of whole tile rows:
// a narrower detector footprint per view in the XCD's L2 (FP16 full
// workload 52.5 -> 48.6 ms against row-major order; blocks of 8 or 16
// tiles alike, 4 and 32 slower).  Only the order changes: every tile is
// computed as before.
constexpr int kSuper = 16;
__device__ __forceinline__ int2 tile_of(int l, int tiles_x, int tiles_y) {
  const int band = kSuper * tiles_x;                 // tiles of a full block row
  const int sy = l / band, r = l - sy * band;
  const int h = min(kSuper, tiles_y - sy * kSuper);  // rows of this block row
  const int sx = r / (kSuper * h), r2 = r - sx * kSuper * h;
  const int w = min(kSuper, tiles_x - sx * kSuper);  // columns of this block
  return make_int2(sx * kSuper + r2 % w, sy * kSuper + r2 / w);
}

// One wave reconstructs tile `lin` (slab-major order) of the job described by a.
template <int TZ, int PREC, int MODE, bool CENTERED>
__device__ __forceinline__ void bp_v1_tile(const BpArgs& a, int lin) {
  constexpr uint32_t kEntry = ChanInterp<PREC>::kEntry;
  const int lane = threadIdx.x;
  const int tiles = a.tiles_x * a.tiles_y;
  const int slab = lin / tiles;
  const int2 txy = tile_of(lin - slab * tiles, a.tiles_x, a.tiles_y);
  const int tx = txy.x, ty = txy.y;
  const int tile = ty * a.tiles_x + tx;
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
  const float inv_dzs = 1.0f / dzs, ninv_dzs = -inv_dzs;
  // Along k (views t + k H of one theta~ and parity) the view z moves
  // monotonically, and with it every bound of a view's slice range [ilo, ihi]
  // below, all in the same direction: up (in slice index) when the table feed
  // and the slice step have the same sign.  A view whose range lies entirely
  // past the slab in that direction is followed only by such views.
  const bool up = (g.kappa > 0.0f) == (dzs > 0.0f);
  // Centred WFBP sums: the whole-slab samples take their weight without a
  // selection mask (bp_weight), the lanes whose column misses the detector
  // parked at row -1e30 with 0 rows per mm (the per-slice masks of the
  // partial-slab paths already exclude those lanes).
  constexpr bool kFree = CENTERED && MODE == kBpAccumulate;
  const bool can_stop = nz > 1;
  const float* ztc = a.zt + (size_t)chunk * a.nviews;
  const int* wrc = reinterpret_cast<const int*>(a.wrange + (size_t)chunk * a.nviews);
  const int* vrp = reinterpret_cast<const int*>(a.slab_vrange + slab);
  const int2 vr = make_int2(ld_uniform(vrp), ld_uniform(vrp + 1));
  const int H = a.H;
  const uint32_t vstride = (uint32_t)g.nrows * (uint32_t)a.cpitch;   // entries per view
  const uint32_t rowB = (uint32_t)a.cpitch * kEntry;                 // bytes per row

  // Buffer resource based at the slab's first view (keeps offsets < 4 GiB;
  // the planner checks that a slab's views fit).
  const char* fb = static_cast<const char*>(a.filt) + (size_t)vr.x * vstride * kEntry;
  const uint32_t fbytes =
      (uint32_t)min((size_t)(vr.y - vr.x) * vstride * kEntry, (size_t)0xFFFFFFF0u);
  BpRsrc rf;
  rf.rows = __builtin_amdgcn_make_buffer_rsrc((void*)fb, 0, fbytes, kRsrcWord3);

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
    const float* csp = reinterpret_cast<const float*>(a.cs + t);
    const float2 cs = make_float2(ld_uniform(csp), ld_uniform(csp + 1));
    const BpColumn col = bp_column(x, y, cs, g);
    const float p = col.p, sq = col.sq, f = col.f, kb = col.kb;
    // The tile's z-interval extremes for this theta~ (tabulated per plan).
    const float* zbp = reinterpret_cast<const float*>(a.zbounds + (size_t)tile * H + t);
    const float4 zb =
        make_float4(ld_uniform(zbp), ld_uniform(zbp + 1), ld_uniform(zbp + 2), ld_uniform(zbp + 3));
    float num[TZ], den[TZ];
#pragma unroll
    for (int i = 0; i < TZ; ++i) {
      // 1e-30 is below half an ulp of any nonzero weight sum (W >= ~1e-13
      // when a sample is selected), so den ends as the exact sum when a
      // view contributes and 1e-30 (num = 0, no change to acc) otherwise.
      // (Through an empty asm: as literals the compiler materialised both
      // start values again at the view loop's entry, 32 instead of 14 moves
      // per theta~.)
      float zero = 0.0f, tiny = 1e-30f;
      asm("" : "+v"(zero));
      asm("" : "+v"(tiny));
      num[i] = zero;
      den[i] = tiny;
    }

#pragma unroll
    for (int par = 0; par < 2; ++par) {
      int k = kmin + (((kmin & 1) != par) ? 1 : 0);
      if (k >= kmax) continue;
      const float pidx = par ? fmaf(-p, g.inv_ds, g.cc) : fmaf(p, g.inv_ds, g.cc);
      const float lt = (par ? sq + f : sq - f) * g.tanc;
      const float kbs = par ? -kb : kb;
      const float pfl = floorf(pidx);
      const bool chan_ok = active && pfl >= 0.0f && pfl <= (float)(g.ncols - 2);
      // Rows per mm of z; 0 for the lanes parked off the detector (kFree).
      const float invLh = kFree && !chan_ok ? 0.0f : div_rn_normal(g.hr, lt);
      const uint32_t pf = chan_ok ? (uint32_t)(int)pfl : 0u;
      const float wp = pidx - pfl;
      // Matched mode: R / l, the voxel-driven footprint's depth correction
      // (a row spans drow l / Dsd of z at distance l from the source).
      const float fR = MODE == kBpMatched ? g.tanc / (lt * g.invR) : 0.0f;
      // z-interval of |q| <= 1 relative to the view's table z, over the wave.
      const float A = par ? zb.z : zb.x;
      const float B = par ? zb.w : zb.y;
      // Slice positions of the interval ends for a view at table z ztv:
      // ztv * inv_dzs + c, the lower end first (A <= B, so the order is
      // that of the slice step's sign).
      const float ca = (A - zfirst) * inv_dzs, cb = (B - zfirst) * inv_dzs;
      const float clo = inv_dzs > 0.0f ? ca : cb, nchi = -(inv_dzs > 0.0f ? cb : ca);
      for (; k < kmax; k += 2) {
        const int v = t + k * H;
        const float ztv = ld_uniform(ztc + v);
        // Exact window slice range (native indices).
        const int2 wr = make_int2(ld_uniform(wrc + 2 * v), ld_uniform(wrc + 2 * v + 1));
        // One slice of margin on each side covers the rounding of these
        // positions; the per-sample test below is exact.
        const int flo = floor_int(fmaf(ztv, inv_dzs, clo));
        const int fhi = floor_int(fmaf(ztv, ninv_dzs, nchi));   // -ceil(upper end)
        // (The matched projector pair uses every view that sees the voxel:
        // no view window.)
        const int ilo = max(max(0, flo - 1), MODE == kBpMatched ? 0 : wr.x - s0);
        const int ihi = min(min(nz - 1, 1 - fhi), MODE == kBpMatched ? nz - 1 : wr.y - s0);
        if (ilo > ihi) {
          if (can_stop && (up ? ilo > nz - 1 : ihi < 0)) break;
          continue;
        }
        const float row0 =                                 // fct_bp_row0
            kFree && !chan_ok ? -1.0e30f : fmaf(kbs - ztv, invLh, g.cr);
        ChanInterp<PREC> ci;
        ci.init((uint32_t)(v - vr.x) * vstride * kEntry, pf, wp);
        if (ilo == 0 && ihi == nz - 1) {
          // Every slice of the slab (the common case): no per-slice branch,
          // and the taps of four slices are in flight before the first is
          // used (four, not TZ: more would cost waves per SIMD).  With
          // nz < TZ the extra slices repeat the last one and are not stored.
          // Centred WFBP sums take the weight without a selection mask
          // (bp_weight; the lanes off the detector are parked at row -1e30).
#pragma unroll
          for (int h = 0; h < TZ; h += 4) {
            BpTap<PREC> tap[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
              tap[i] = bp_fetch<PREC, MODE, CENTERED>(zs[h + i], invLh, row0, ci, rf, rowB,
                                                      chan_ok, g);
            // FP16: the four loads are issued before the weight arithmetic
            // that waits for them (the scheduler would sink them otherwise).
            if (PREC == FCT_PRECISION_FP16) __builtin_amdgcn_sched_barrier(0);
            bp_add4<PREC, MODE, kFree>(tap, ci, fR, g, num + h, den + h, cnt + h);
          }
        } else if constexpr (PREC == FCT_PRECISION_FP16) {
          // Part of the slab, FP16: groups of four slices that meet
          // [ilo, ihi] through the dot block, the slices outside the range
          // masked like unselected samples (cheaper than per-slice samples
          // with the two-address dot form; for FP32 / BF16 the masked slices
          // cost more than they save).
#pragma unroll
          for (int h = 0; h < TZ; h += 4) {
            if (ilo > h + 3 || ihi < h) continue;
            BpTap<PREC> tap[4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
              tap[i] = bp_fetch<PREC, MODE, CENTERED>(zs[h + i], invLh, row0, ci, rf, rowB,
                                                      chan_ok && h + i >= ilo && h + i <= ihi, g);
            __builtin_amdgcn_sched_barrier(0);
            bp_add4<PREC, MODE>(tap, ci, fR, g, num + h, den + h, cnt + h);
          }
        } else {
#pragma unroll
          for (int i = 0; i < TZ; ++i)
            if (i >= ilo && i <= ihi) {
              const BpTap<PREC> tp =
                  bp_fetch<PREC, MODE, CENTERED>(zs[i], invLh, row0, ci, rf, rowB, chan_ok, g);
              bp_add<PREC, MODE>(tp, ci, fR, g, num[i], den[i], cnt[i]);
            }
        }
      }
    }
    if (MODE == kBpAccumulate) {
#pragma unroll
      for (int i = 0; i < TZ; ++i)
        acc[i] = fmaf(num[i], __builtin_amdgcn_rcpf(den[i]), acc[i]);
    } else if (MODE == kBpMatched) {
#pragma unroll
      for (int i = 0; i < TZ; ++i) acc[i] += num[i];
    }
  }
  if (!inb) return;
  if (MODE == kBpAccumulate || MODE == kBpMatched) {
    // FP16 / BF16 projections carry one storage scale per scan.
    const float out_scale =
        PREC == FCT_PRECISION_FP32 ? a.scale : a.scale * ld_uniform(a.filt_scale);
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

// Occupancy floor for the 16-bit kernels at TZ <= 8: at least seven waves
// per SIMD (the few spills this costs sit in the per-theta~ setup, outside
// the sample loop).  At TZ = 16 five (96 VGPRs; the allocator's own choice
// is 97, one over); a floor of six spills inside the view loop.
template <int TZ, int PREC>
constexpr int bp_min_waves() {
  return PREC == FCT_PRECISION_FP32 ? 1 : TZ <= 8 ? 7 : TZ == 16 ? 5 : 1;
}

template <int TZ, int PREC, int MODE, bool CENTERED>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(TZ >= 8 ? bp_min_waves<TZ, PREC>() : 1, 8)))
void bp_v1_kernel(BpArgs a) {
  bp_v1_tile<TZ, PREC, MODE, CENTERED>(
      a, a.slab0 * a.tiles_x * a.tiles_y + xcd_block(blockIdx.x, gridDim.x));
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
// (The occupancy floors of bp_v1_kernel, also at TZ = 4.)
template <int TZ, int PREC, bool CENTERED>
__global__ __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(bp_min_waves<TZ, PREC>(), 8)))
void bp_batch_kernel(BpBatchArgs b,
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
  const int blocks = a.tiles_x * a.tiles_y * (a.slab_count ? a.slab_count : a.nslabs);
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

hipError_t launch_bp_zbounds(const BpArgs& a, float4* zbounds, hipStream_t s) {
  const int tiles = a.tiles_x * a.tiles_y;
  if (tiles <= 0 || a.H <= 0) return hipSuccess;
  hipLaunchKernelGGL(bp_zbounds_kernel, dim3(tiles), dim3(64), 0, s, a, zbounds);
  return hipGetLastError();
}

hipError_t launch_backproject_v1(const BpArgs& a, int tile_z, int precision, int mode,
                                 hipStream_t s) {
  // Slabs [slab0, slab0 + slab_count) inside the volume (slab_count 0: every slab).
  if (a.slab0 < 0 || a.slab_count < 0 || (a.slab_count == 0 && a.slab0 != 0) ||
      a.slab0 + a.slab_count > a.nslabs)
    return hipErrorInvalidValue;
  switch (tile_z) {
    case 4: return launch1<4>(a, precision, mode, s);
    case 8: return launch1<8>(a, precision, mode, s);
    case 16: return launch1<16>(a, precision, mode, s);
    case 32: return launch1<32>(a, precision, mode, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace fct
