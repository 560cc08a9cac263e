// wfbp_math.h -- FP32 index-selection arithmetic shared by the HIP kernels and
// the FP64 CPU restatement, plus the FP64 -> FP32 table recipes they both read.
//
// Every quantity that decides WHICH sample is used (view window, fan/parallel
// channel, detector row, in-FOV test) is computed here, in FP32, with one fixed
// operation order, explicit fmaf, and correctly rounded sqrt and division.
// The device code is compiled with -ffp-contract=off and the host code with
// g++ -ffp-contract=off, so both produce the same bits and therefore the same
// floor() results.  Weights and accumulation are free to use other precisions.
//
// Geometry convention (one convention, no channel/row flips, no offset kluge):
//   fan view v      : tube angle  a_v = angle0 + v*dtheta, dtheta = 2*pi/ppr
//                     source z    z_v = z0 + v*dz,          dz = feed/ppr (signed)
//   fan channel c   : fan angle  beta = (c - cc)*dgamma,  dgamma = dch/Dsd
//   detector row r  : height at the detector  w = (r - cr)*drow
//   ray (v,r,c)     : S = (R cos a, R sin a, z_v); in-plane direction
//                     -(cos(a+beta), sin(a+beta)); it reaches the detector after
//                     in-plane distance Dsd, at height z_v + w.
//   parallel view j : theta_j = angle0 + j*dtheta (same count as fan views)
//   parallel chan c : t_c = (c - cc)*ds, ds = R*dgamma; a point (x,y) projects to
//                     p = x sin(theta) - y cos(theta).
//   The reference (cuda_kernels.cuh:235-251, ReconstructionMachine.cu:150-162)
//   reaches the same formulas through host channel/row flips plus a +1.75
//   channel offset; here the convention is stated once and used everywhere.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define FCT_HD __host__ __device__ __forceinline__
#else
#define FCT_HD static inline
#endif

#ifdef __cplusplus
extern "C++" {
#endif

// ---------------------------------------------------------------------------
// asin for |x| <= 1 in FP32, bit-identical on host and device.
// Method: Maclaurin series of arcsin (Abramowitz & Stegun 4.4.40) for
// |x| <= 0.5, evaluated by Horner with fmaf; for |x| > 0.5 the half-angle
// identity asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) (A&S 4.4.31 form) maps the
// argument back to |u| <= 0.5.  The function is exactly odd.
// Series coefficients c_n = (2n)! / (4^n (n!)^2 (2n+1)), n = 1..13.
// Truncation error at |x| = 0.5 is below 3e-11 relative.
// ---------------------------------------------------------------------------
FCT_HD float fct_asin_core(float x) {  // |x| <= 0.5
  const float z = x * x;
  float p = 0.0057400377f;              // n=13
  p = fmaf(p, z, 0.0064472103f);        // n=12
  p = fmaf(p, z, 0.0073125259f);        // n=11
  p = fmaf(p, z, 0.0083903358f);        // n=10
  p = fmaf(p, z, 0.0097616095f);        // n=9
  p = fmaf(p, z, 0.011551801f);         // n=8
  p = fmaf(p, z, 0.013964844f);         // n=7
  p = fmaf(p, z, 0.017352764f);         // n=6
  p = fmaf(p, z, 0.022372159f);         // n=5
  p = fmaf(p, z, 0.030381944f);         // n=4
  p = fmaf(p, z, 0.044642857f);         // n=3
  p = fmaf(p, z, 0.075000000f);         // n=2
  p = fmaf(p, z, 0.16666667f);          // n=1
  return fmaf(x * z, p, x);
}

FCT_HD float fct_asinf(float x) {
  const float ax = fabsf(x);
  float r;
  if (ax <= 0.5f) {
    r = fct_asin_core(ax);
  } else {
    const float u = sqrtf((1.0f - ax) * 0.5f);
    r = fmaf(-2.0f, fct_asin_core(u), 1.5707964f);
  }
  return x < 0.0f ? -r : r;
}

// ---------------------------------------------------------------------------
// Backprojection geometry constants (FP32), derived once in FP64 on the host.
// Mirrors GPUPrecompute (recon_structs.h:52-83) with the signed table feed.
// ---------------------------------------------------------------------------
struct FctBpConsts {
  float R2;            // R^2                              (recon_structs.h:75)
  float invR;          // 1/R                              (recon_structs.h:78)
  float kappa;         // feed/(2 pi), signed              (recon_structs.h:76)
  float tanc;          // tan(theta_cone/2)                (recon_structs.h:79)
  float inv_ds;        // Dsd/(R*dch) = 1/ds               (recon_structs.h:80)
  float cc;            // central channel
  float hr;            // (nrows-1)/2: q -> row scale      (cuda_kernels.cuh:248)
  float cr;            // central row
  float zrot;          // |feed|: view window half width   (cuda_kernels.cuh:232)
  float half_fov_sq;   // (acq_fov/2)^2                    (recon_structs.h:72)
  float wk;            // 1 / (hr (1 - Q)): weight ramp slope per row
  int32_t ncols;
  int32_t nrows;
};

// In-FOV test of a voxel column (cuda_kernels.cuh:220-221).
FCT_HD bool fct_in_fov(float x, float y, const FctBpConsts& g) {
  return !(fmaf(x, x, y * y) > g.half_fov_sq);
}

// Per (voxel column, parallel view) quantities.  (c,s) = (cos, sin) of the
// parallel view angle; the complementary view (angle + pi) is (-c,-s).
// The cone coordinate q = (z - zt + kb) / (l tan(cone/2)) of
// cuda_kernels.cuh:239 enters only through the detector row position
// qi = q * hr + cr (cuda_kernels.cuh:248 with an explicit central row), which
// is computed directly: qi = z * invLh + (kb - zt) * invLh + cr.
struct FctBpColumn {
  float pidx;   // parallel channel position               (cuda_kernels.cuh:246)
  float lt;     // l * tan(cone/2): |q| <= 1 <=> |z - zt + kb| <= lt
  float invLh;  // hr / (l * tan(cone/2)): rows per mm of z
  float kb;     // kappa * asin(p/R): source-z shift of this ray
};

// cuda_kernels.cuh:235-239, 246: p, l and the channel position.
FCT_HD FctBpColumn fct_bp_column(float x, float y, float c, float s,
                                 const FctBpConsts& g) {
  FctBpColumn o;
  const float p = fmaf(x, s, -(y * c));
  const float sq = sqrtf(fmaf(-p, p, g.R2));
  const float l = sq - fmaf(x, c, y * s);
  o.pidx = fmaf(p, g.inv_ds, g.cc);
  o.lt = l * g.tanc;
  o.invLh = g.hr / o.lt;
  o.kb = g.kappa * fct_asinf(p * g.invR);
  return o;
}

// Per view: row position at z = 0 (slice z are offsets from a chunk
// reference, like zt).
FCT_HD float fct_bp_row0(const FctBpColumn& col, float zt, const FctBpConsts& g) {
  return fmaf(col.kb - zt, col.invLh, g.cr);
}

// Per (column, view, slice): detector row position of the voxel.
FCT_HD float fct_bp_row(float zs, const FctBpColumn& col, float row0) {
  return fmaf(zs, col.invLh, row0);
}

// Per (view, slice): view window (cuda_kernels.cuh:232-233).
FCT_HD bool fct_bp_in_window(float zt, float zs, const FctBpConsts& g) {
  return !(fabsf(zt - zs) > g.zrot);
}

// Selection of the 2x2 bilinear footprint: |q| <= 1 (cuda_kernels.cuh:243),
// i.e. |qi - cr| <= hr, and the whole footprint on the detector.  When
// cr == hr the row test alone implies the cone test (exactly, in FP32).
FCT_HD bool fct_bp_select(float qi, float pidx, const FctBpConsts& g, int* qf, int* pf) {
  if (fabsf(qi - g.cr) > g.hr) return false;
  const float fq = floorf(qi);
  const float fp = floorf(pidx);
  if (!(fq >= 0.0f) || !(fp >= 0.0f)) return false;
  if (fq > (float)(g.nrows - 2) || fp > (float)(g.ncols - 2)) return false;
  *qf = (int)fq;
  *pf = (int)fp;
  return true;
}

// Stierstorfer helical weight, Q = 0.6 (cuda_kernels.cuh:79-93).
#define FCT_WQ 0.6f

// W(q) = 1 for |q| < Q, cos^2(pi/2 (|q|-Q)/(1-Q)) for Q <= |q| < 1, written
// as sin^2(pi/2 t) with t = min((1-|q|)/(1-Q), 1) for selected samples
// (|q| <= 1).  sin(pi/2 t) = t * S(t^2) with S a cubic fitted for minimax
// relative error of sin(pi/2 t) / t on [0, 1] under the constraint S(1) = 1
// (Lawson-weighted least squares; the constant term then moved by one ulp so
// that the FP32 evaluation gives S(1) = 1): FP32 relative error of W below
// 2.5e-6 on [0, 1], including near W -> 0 where an absolute-accuracy cosine
// would not be, and W = 1 exactly in the flat part (t = 1).  W only weights
// the views of a theta~ against each other; it is not part of index
// selection.
FCT_HD float fct_w_of_t(float t) {
  const float u = t * t;
  float s = -0.0043512345f;
  s = fmaf(s, u, 0.07947272f);
  s = fmaf(s, u, -0.6459161f);
  s = fmaf(s, u, 1.5707947f);
  return u * (s * s);
}

// 1 / (1 - Q) as the FP32 expression 1 / (1 - 0.6f) = 2.5000002: t >= 1, so
// W = 1 exactly, for every FP32 |q| <= 0.6f.
#define FCT_WC (1.0f / (1.0f - FCT_WQ))

FCT_HD float fct_w_of_q(float q) {
  return fct_w_of_t(fminf(fmaf(-fabsf(q), FCT_WC, FCT_WC), 1.0f));
}

// The same weight from the row position: |q| = |qi - cr| / hr, so
// t = min(1/(1-Q) - |qi - cr| * wk, 1) with wk = 1 / (hr (1-Q)), rounded up
// so that hr * wk >= FCT_WC: t <= 0 wherever |qi - cr| >= hr.
FCT_HD float fct_w_of_row(float qi, const FctBpConsts& g) {
  return fct_w_of_t(fminf(fmaf(-fabsf(qi - g.cr), g.wk, FCT_WC), 1.0f));
}

#ifdef __cplusplus
}
#endif

// ---------------------------------------------------------------------------
// Host-only FP64 -> FP32 table recipes (used by the planner and the oracle).
// ---------------------------------------------------------------------------
#ifdef __cplusplus
#include <cmath>
namespace fct_tables {

// Parallel view angle of index t (only t < ppr/2 is tabulated; the view
// t + k*ppr/2 uses (-1)^k (cos, sin)).
inline void view_cos_sin(double angle0, double dtheta, int t, float* c,
                         float* s) {
  const double th = angle0 + dtheta * (double)t;
  *c = (float)std::cos(th);
  *s = (float)std::sin(th);
}

// Source z of parallel view v relative to a chunk reference, rounded once.
inline float view_z_rel(double z0, double dz, int v, double zref) {
  return (float)((z0 + dz * (double)v) - zref);
}

// Voxel centre (cuda_kernels.cuh:215-216): (i - (n-1)/2) * fov/n + origin.
inline float voxel_coord(int i, int n, double fov, double origin) {
  return (float)(((double)i - 0.5 * ((double)n - 1.0)) * (fov / (double)n) +
                 origin);
}

// Fan -> parallel rebin table for parallel channel c (cuda_kernels.cuh:17-36).
// Exact: the view position j - beta/dtheta is split into an integer offset and
// a fraction that do not depend on j.
struct RebinEntry {
  int32_t valid;   // fan channel position inside [0, ncols)
  int32_t b0, b1;  // fan channels (b1 clamped like the texture)
  int32_t aoff;    // fan view offset: views j+aoff, j+aoff+1
  double fb;       // channel fraction
  double fa;       // view fraction
};

inline RebinEntry rebin_entry(int c, double cc, double dgamma, double dtheta,
                              int ncols) {
  RebinEntry e{};
  const double sb = ((double)c - cc) * dgamma;
  if (!(std::fabs(sb) <= 1.0)) { e.valid = 0; return e; }
  const double beta = std::asin(sb);
  const double bidx = beta / dgamma + cc;
  if (!(bidx >= 0.0) || !(bidx < (double)ncols)) { e.valid = 0; return e; }
  e.valid = 1;
  const double bf = std::floor(bidx);
  e.b0 = (int32_t)bf;
  e.b1 = e.b0 + 1 < ncols ? e.b0 + 1 : ncols - 1;
  e.fb = bidx - bf;
  const double nd = -beta / dtheta;
  const double af = std::floor(nd);
  e.aoff = (int32_t)af;
  e.fa = nd - af;
  return e;
}

// FP32 backprojection constants from FP64 geometry (recon_structs.h:69-83,
// ReconstructionMachine.cu:62-69).  tan(theta_cone/2) is taken in closed form:
// theta_cone = 2 atan(hr * csw / R) gives tan(theta_cone/2) = hr * csw / R.
inline FctBpConsts bp_consts(int ncols, int nrows, double dch, double drow,
                             double dsd, double r, double cc, double cr,
                             double feed) {
  FctBpConsts g{};
  const double pi = 3.14159265358979323846;
  const double csw = r * drow / dsd;
  const double hr = 0.5 * ((double)nrows - 1.0);
  const double ds = dch * r / dsd;
  const double acq = 2.0 * r * std::sin(0.5 * ((double)ncols - 1.0) * dch / dsd);
  g.R2 = (float)(r * r);
  g.invR = (float)(1.0 / r);
  g.kappa = (float)(feed / (2.0 * pi));
  g.tanc = (float)(hr * csw / r);
  g.inv_ds = (float)(1.0 / ds);
  g.cc = (float)cc;
  g.hr = (float)hr;
  g.cr = (float)cr;
  g.zrot = (float)std::fabs(feed);
  g.half_fov_sq = (float)(0.25 * acq * acq);
  // Rounded up to the smallest FP32 wk with hr * wk >= FCT_WC (the product
  // is exact in FP64): the weight ramp then reaches t <= 0 exactly at
  // |qi - cr| = hr.
  g.wk = (float)(1.0 / (hr * (1.0 - 0.6)));
  while ((double)g.hr * (double)g.wk < (double)FCT_WC) g.wk = std::nextafter(g.wk, HUGE_VALF);
  g.ncols = ncols;
  g.nrows = nrows;
  return g;
}

// Slice grid count: the values start + i*delta (i >= 0) strictly before end,
// computed from FP64 start and step (ReconstructionMachine.cu:110-132 counts
// the same set by FP32 accumulation).  start == end gives one slice.
inline int slice_count(double start, double end, double step) {
  if (start == end) return 1;
  if (!(step > 0.0)) return 0;
  const double x = std::fabs(end - start) / step;
  const double n = std::ceil(x - 1e-9 * (x > 1.0 ? x : 1.0));
  return n < 1.0 ? 1 : (int)n;
}

inline double slice_z(double start, double end, double step, int i) {
  const double sgn = end >= start ? 1.0 : -1.0;
  return start + sgn * step * (double)i;
}

// Slices are grouped in chunks of this many; each chunk carries its own FP64
// reference z so that the kernels see small FP32 offsets.
constexpr int kSliceChunk = 64;

}  // namespace fct_tables
#endif
