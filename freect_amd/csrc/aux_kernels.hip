// aux_kernels.hip -- slice thickening and the synthetic scan generator.
#include "wfbp_internal.h"

namespace fct {
namespace {

// thicken_slices (cuda_kernels.cuh:289-320): out[s] = sum_i w_i native[i] /
// sum_i w_i with triangular w_i.  The host supplies, per requested slice, the
// native slices with nonzero weight and the weights already divided by their
// sum (FP64 -> FP32).
__global__ __launch_bounds__(256) void thicken_kernel(ThickenArgs a) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int s = a.s0 + blockIdx.y;
  if (p >= a.plane) return;
  float acc = 0.0f;
  for (int k = a.off[s]; k < a.off[s + 1]; ++k)
    acc = fmaf(a.w[k], a.native[(size_t)a.idx[k] * a.plane + p], acc);
  // Hounsfield units on request (the reference stores mm^-1 only,
  // ReconstructionMachine.cu:436-455).
  if (a.hu) acc = fmaf(acc, a.hu_scale, -1000.0f);
  a.out[(size_t)s * a.plane + p] = acc;
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014) as a counter-based hash.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct RayRng {
  unsigned long long key, ctr;
  __device__ double uniform() {   // (0, 1)
    const unsigned long long r = mix64(key ^ mix64(ctr++));
    return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
};

// Poisson sample.  lambda < 10: multiplication (Knuth) method.  Otherwise
// PTRS, the transformed rejection method with squeeze of W. Hormann, "The
// transformed rejection method for generating Poisson random variables",
// Insurance: Mathematics and Economics 12 (1993).
__device__ double poisson(double lam, RayRng& rng) {
  if (lam <= 0.0) return 0.0;
  if (lam < 10.0) {
    const double L = exp(-lam);
    double k = 0.0, prod = rng.uniform();
    while (prod > L && k < 1000.0) {
      prod *= rng.uniform();
      k += 1.0;
    }
    return k;
  }
  const double slam = sqrt(lam), loglam = log(lam);
  const double b = 0.931 + 2.53 * slam;
  const double aa = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (int it = 0; it < 1000; ++it) {
    const double U = rng.uniform() - 0.5;
    const double V = rng.uniform();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * aa / us + b) * U + lam + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(aa / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
      return k;
  }
  return floor(lam + 0.5);
}

// One ray per thread: source S = (R cos a, R sin a, z_v), fan angle
// beta = (c - cc) dgamma, detector cell at in-plane distance Dsd and height
// (r - cr) drow.  A clockwise scan traces the counter-clockwise ray at
// angle -a_v and mirrors it (y -> -y).  Analytic chords of an infinite water cylinder along z and of
// rotated ellipsoids, in FP64.
__global__ __launch_bounds__(256) void phantom_kernel(PhantomArgs a) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int row = blockIdx.y;
  const int v = a.v0 + blockIdx.z;
  if (c >= a.ncols) return;
  const double ang = a.angle0 + a.dtheta * (double)v;
  double sx = a.r * cos(ang), sy = a.r * sin(ang);
  double sz = a.z0 + a.dz * (double)v;
  const double w = ((double)row - a.cr) * a.drow;
  const double beta = ((double)c - a.cc) * a.dgamma;
  const double dx = -a.dsd * cos(ang + beta);
  const double dy = -a.dsd * sin(ang + beta);
  double ux, uy, uz;
  if (a.states == 1) {
    const double len = sqrt(a.dsd * a.dsd + w * w);
    ux = dx / len;
    uy = dy / len;
    uz = w / len;
  } else {
    // Flying focal spot: the focal spot of state v mod S moves by
    // (dr, da, dz) along (e_w, e_u, e_z); the detector cell stays where the
    // undeflected source puts it (generate_system_matrix_ffs.cpp:126-149).
    const double* o = a.ffs_off[v % a.states];
    const double ca = cos(ang), sa = sin(ang);
    const double ox = o[0] * ca - o[1] * sa, oy = o[0] * sa + o[1] * ca;
    sx += ox;
    sy += oy;
    sz += o[2];
    const double ex = dx - ox, ey = dy - oy, ez = w - o[2];
    const double len = sqrt(ex * ex + ey * ey + ez * ez);
    ux = ex / len;
    uy = ey / len;
    uz = ez / len;
  }
  if (a.mirror) {   // clockwise: the mirror image of the ray traced above
    sy = -sy;
    uy = -uy;
  }
  double acc = 0.0;
  {
    const double A = ux * ux + uy * uy;
    const double B = 2.0 * (sx * ux + sy * uy);
    const double C = sx * sx + sy * sy - a.cyl_r * a.cyl_r;
    const double disc = B * B - 4.0 * A * C;
    if (A > 0.0 && disc > 0.0) acc += a.cyl_mu * sqrt(disc) / A;
  }
  for (int i = 0; i < a.n_ell; ++i) {
    const fct_ellipsoid& E = a.ell[i];
    const double px = sx - E.c[0], py = sy - E.c[1], pz = sz - E.c[2];
    const double lx = (E.rot[0] * px + E.rot[3] * py + E.rot[6] * pz) / E.a[0];
    const double ly = (E.rot[1] * px + E.rot[4] * py + E.rot[7] * pz) / E.a[1];
    const double lz = (E.rot[2] * px + E.rot[5] * py + E.rot[8] * pz) / E.a[2];
    const double mx = (E.rot[0] * ux + E.rot[3] * uy + E.rot[6] * uz) / E.a[0];
    const double my = (E.rot[1] * ux + E.rot[4] * uy + E.rot[7] * uz) / E.a[1];
    const double mz = (E.rot[2] * ux + E.rot[5] * uy + E.rot[8] * uz) / E.a[2];
    const double A = mx * mx + my * my + mz * mz;
    const double B = 2.0 * (lx * mx + ly * my + lz * mz);
    const double C = lx * lx + ly * ly + lz * lz - 1.0;
    const double disc = B * B - 4.0 * A * C;
    if (disc > 0.0) acc += E.dmu * sqrt(disc) / A;
  }
  if (a.i0 > 0.0) {
    const unsigned long long ray = ((unsigned long long)v * a.nrows + row) * a.ncols + c;
    RayRng rng{mix64(a.seed), ray << 20};
    const double n = poisson(a.i0 * exp(-acc), rng);
    acc = log(a.i0 / fmax(n, 0.5));
  }
  a.out[((size_t)blockIdx.z * a.nrows + row) * a.ncols + c] = (float)acc;
}

}  // namespace

hipError_t launch_thicken(const ThickenArgs& a, hipStream_t s) {
  if (a.n_req <= 0) return hipSuccess;
  dim3 grid((a.plane + 255) / 256, a.n_req);
  hipLaunchKernelGGL(thicken_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_phantom(const PhantomArgs& a, hipStream_t s) {
  dim3 grid((a.ncols + 255) / 256, a.nrows, a.v1 - a.v0);
  hipLaunchKernelGGL(phantom_kernel, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace fct
