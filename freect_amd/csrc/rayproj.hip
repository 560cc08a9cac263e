// rayproj.hip -- ray-driven helical projector of the native volume onto the
// raw fan data (Joseph's method) and its exact transpose.  No counterpart in
// the reference: the data-consistency refinement needs a forward model of
// the measured line integrals, and the WFBP chain's own transpose (a filtered
// sinogram) or the voxel-driven matched projector (which leaves channels
// between the footprints of voxels larger than a channel empty) are not one.
//
// A ray is that of the synthetic scan generator (aux_kernels.hip): the
// (deflected) focal spot of raw projection v to its detector cell, in the
// plan's frame (a clockwise scan as its mirror image, like the voxel
// tables).  Joseph's method (P. M. Joseph, "An improved algorithm for
// reprojecting rays through pixel images", IEEE TMI 1 (1982) 192): along the
// in-plane axis the ray is most aligned with, at every voxel plane of that
// axis, the volume is interpolated bilinearly in the other in-plane axis and
// z (zero outside the volume) and the samples are summed times the path
// length per plane, |step / u_axis|.  The transpose scatters each ray's
// value with the same weights (FP32 atomics).
#include "wfbp_internal.h"

namespace fct {
namespace {

struct Ray {
  float sx, sy, sz, ux, uy, uz;
};

// FP64 ray of (view, row, channel), then FP32 for the march (positions of a
// few hundred mm: 3e-5 mm resolution).
__device__ Ray make_ray(const RayProjArgs& a, int v, int row, int c) {
  const double ang = a.angle0 + a.dtheta * (double)v;
  const double ca = cos(ang), sa = sin(ang);
  double sx = a.r * ca, sy = a.r * sa, sz = a.z0 + a.dz * (double)v;
  const double w = ((double)row - a.cr) * a.drow;
  const double beta = ((double)c - a.cc) * a.dgamma;
  double dx = -a.dsd * cos(ang + beta), dy = -a.dsd * sin(ang + beta), dz = w;
  if (a.states > 1) {
    const double* o = a.ffs_off[v % a.states];
    const double ox = o[0] * ca - o[1] * sa, oy = o[0] * sa + o[1] * ca;
    sx += ox;
    sy += oy;
    sz += o[2];
    dx -= ox;
    dy -= oy;
    dz -= o[2];
  }
  const double len = sqrt(dx * dx + dy * dy + dz * dz);
  return Ray{(float)sx, (float)sy, (float)sz, (float)(dx / len), (float)(dy / len),
             (float)(dz / len)};
}

// Index range [lo, hi] of the planes whose sample f = A + i B lies in
// (-1, n) (the bilinear footprint touches the volume).
__device__ __forceinline__ void plane_range(float A, float B, int n, int* lo, int* hi) {
  if (B == 0.0f) {
    if (!(A > -1.0f && A < (float)n)) *lo = 1, *hi = 0;
    return;
  }
  float t0 = (-1.0f - A) / B, t1 = ((float)n - A) / B;
  if (t0 > t1) {
    const float t = t0;
    t0 = t1;
    t1 = t;
  }
  *lo = max(*lo, (int)floorf(t0));
  *hi = min(*hi, (int)ceilf(t1));
}

// Walk of a ray: calls f(i_primary, j, k, weight) for every volume voxel a
// plane's bilinear sample touches (the weight without the path length).
template <class F>
__device__ __forceinline__ float walk(const RayProjArgs& a, const Ray& r, F&& f) {
  const bool xmajor = fabsf(r.ux) >= fabsf(r.uy);
  // Primary axis p (planes), secondary s (in-plane), and z.
  const float p0 = xmajor ? a.gx0 : a.gy0, dp = xmajor ? a.gdx : a.gdy;
  const float q0 = xmajor ? a.gy0 : a.gx0, dq = xmajor ? a.gdy : a.gdx;
  const float sp = xmajor ? r.sx : r.sy, sq = xmajor ? r.sy : r.sx;
  const float up = xmajor ? r.ux : r.uy, uq = xmajor ? r.uy : r.ux;
  const int np = xmajor ? a.nx : a.ny, nq = xmajor ? a.ny : a.nx;
  // t at plane i: (p0 + i dp - sp) / up; sample positions linear in i.
  const float t0 = (p0 - sp) / up, dt = dp / up;
  const float Aq = (sq + t0 * uq - q0) / dq, Bq = dt * uq / dq;
  const float Az = (r.sz + t0 * r.uz - a.gz0) / a.gdz, Bz = dt * r.uz / a.gdz;
  int lo = 0, hi = np - 1;
  plane_range(Aq, Bq, nq, &lo, &hi);
  plane_range(Az, Bz, a.nz, &lo, &hi);
  for (int i = lo; i <= hi; ++i) {
    const float fq = fmaf((float)i, Bq, Aq), fz = fmaf((float)i, Bz, Az);
    const float jq = floorf(fq), kz = floorf(fz);
    const float wq = fq - jq, wz = fz - kz;
    const int j = (int)jq, k = (int)kz;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk)
#pragma unroll
      for (int dj = 0; dj < 2; ++dj) {
        const int jj = j + dj, kk = k + dk;
        if (jj < 0 || jj >= nq || kk < 0 || kk >= a.nz) continue;
        const float wgt = (dj ? wq : 1.0f - wq) * (dk ? wz : 1.0f - wz);
        const int ix = xmajor ? i : jj, iy = xmajor ? jj : i;
        f(((size_t)kk * a.ny + iy) * a.nx + ix, wgt);
      }
  }
  return fabsf(dt);   // path length per plane (u is a unit vector)
}

__global__ __launch_bounds__(256) void ray_project_kernel(RayProjArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int row = blockIdx.y, v = a.v0 + blockIdx.z;
  if (c >= a.ncols) return;
  const Ray r = make_ray(a, v, row, c);
  float acc = 0.0f;
  const float len = walk(a, r, [&](size_t o, float w) { acc = fmaf(w, a.vol_in[o], acc); });
  a.raw_out[((size_t)(v - a.v0) * a.nrows + row) * a.ncols + c] = acc * len;
}

__global__ __launch_bounds__(256) void ray_backproject_kernel(RayProjArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int row = blockIdx.y, v = a.v0 + blockIdx.z;
  if (c >= a.ncols) return;
  const float val = a.raw_in[((size_t)(v - a.v0) * a.nrows + row) * a.ncols + c];
  if (val == 0.0f) return;
  const Ray r = make_ray(a, v, row, c);
  // The path length is known before the walk: |dp / u_primary|.
  const bool xmajor = fabsf(r.ux) >= fabsf(r.uy);
  const float len = fabsf((xmajor ? a.gdx : a.gdy) / (xmajor ? r.ux : r.uy));
  const float s = val * len;
  walk(a, r, [&](size_t o, float w) { atomicAdd(a.vol_out + o, w * s); });
}

}  // namespace

hipError_t launch_ray_project(const RayProjArgs& a, bool transpose, hipStream_t s) {
  if (a.v1 <= a.v0) return hipSuccess;
  dim3 grid((a.ncols + 255) / 256, a.nrows, a.v1 - a.v0);
  if (transpose)
    hipLaunchKernelGGL(ray_backproject_kernel, grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(ray_project_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace fct
