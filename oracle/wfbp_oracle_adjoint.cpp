// wfbp_oracle_adjoint.cpp -- TEST INFRASTRUCTURE ONLY: FP64 transpose of the
// restated WFBP chain (rebin -> filter -> W-weighted backprojection), the
// checker for the GPU adjoint.  The chain is linear in the raw data for a
// fixed geometry (the weights W and their per-theta~ sums depend on geometry
// only), so its transpose is:
//   backprojection^T: every selected (voxel, view) sample scatters
//                     scale * W / sum W * bilinear weights * xbar;
//   filter^T        : the filter itself (h[n] = h[-n], so the Toeplitz matrix
//                     h[i-j] is symmetric);
//   rebin^T         : every parallel sample scatters its four bilinear weights
//                     to the (clamped) fan samples it read.
// Selection uses the shared FP32 code (freect_amd/csrc/wfbp_math.h), weights
// are FP64, exactly as in the forward restatement (wfbp_oracle.cpp).
#include <algorithm>
#include <cmath>
#include <vector>

#include "../freect_amd/csrc/ffs_geometry.h"
#include "../freect_amd/csrc/wfbp_math.h"
#include "wfbp_oracle.h"

namespace {
constexpr double kPi = 3.14159265358979323846;

double w_of_q(double q) {
  const double Q = 0.6;
  q = std::fabs(q);
  if (q < Q) return 1.0;
  if (q < 1.0) {
    const double c = std::cos(0.5 * kPi * (q - Q) / (1.0 - Q));
    return c * c;
  }
  return 0.0;
}
}  // namespace

extern "C" {

// xbar: native slices [s0, s1) ([s-s0][y][x]); gbar: parallel views [v0, v1)
// of [view][row][ch], accumulated into (caller zeroes it).
int or_backproject_adjoint(const or_scan* s, const or_recon* r, const double* xbar, int v0,
                           int v1, int s0, int s1, double* gbar) {
  or_derived d;
  or_geometry(s, r, &d);
  const int nn = d.n_slices_native;
  if (s0 < 0 || s1 > nn || s0 >= s1) return -1;
  if (s->views_per_rotation % 2) return -2;
  if (s->rotation < 0) return -3;   // counter-clockwise only (pyoracle mirrors)
  const int ncols = s->n_channels, nrows = s->n_rows, nviews = s->n_views;
  const int H = s->views_per_rotation / 2, ns = s1 - s0, nx = r->nx, ny = r->ny;
  const double dtheta = 2.0 * kPi / s->views_per_rotation, dz = s->feed / s->views_per_rotation;
  const double kappa = s->feed / (2.0 * kPi), hr = 0.5 * (nrows - 1.0);
  const double tanc = hr * d.csw / s->r;
  const FctBpConsts g = fct_tables::bp_consts(ncols, nrows, s->channel_spacing, s->row_spacing,
                                              s->dsd, s->r, s->central_channel, s->central_row,
                                              s->feed);
  const int chunk = fct_tables::kSliceChunk;
  std::vector<double> zn(nn);
  for (int i = 0; i < nn; ++i) zn[i] = fct_tables::slice_z(r->start_pos, r->end_pos, d.csw, i);
  std::vector<float> xs(nx), ys(ny), c32(H), s32(H), zs32(ns);
  std::vector<double> x64(nx), y64(ny), c64(H), s64(H);
  for (int i = 0; i < nx; ++i) {
    xs[i] = fct_tables::voxel_coord(i, nx, r->recon_fov, r->x_origin);
    x64[i] = ((double)i - 0.5 * ((double)nx - 1.0)) * (r->recon_fov / nx) + r->x_origin;
  }
  for (int i = 0; i < ny; ++i) {
    ys[i] = fct_tables::voxel_coord(i, ny, r->recon_fov, r->y_origin);
    y64[i] = ((double)i - 0.5 * ((double)ny - 1.0)) * (r->recon_fov / ny) + r->y_origin;
  }
  for (int t = 0; t < H; ++t) {
    fct_tables::view_cos_sin(s->angle0, dtheta, t, &c32[t], &s32[t]);
    c64[t] = std::cos(s->angle0 + dtheta * t);
    s64[t] = std::sin(s->angle0 + dtheta * t);
  }
  for (int i = 0; i < ns; ++i) zs32[i] = (float)(zn[s0 + i] - zn[((s0 + i) / chunk) * chunk]);
  const double scale = 2.0 * kPi / s->views_per_rotation;
  const size_t vstride = (size_t)nrows * ncols;

#pragma omp parallel for schedule(dynamic, 4)
  for (int col = 0; col < nx * ny; ++col) {
    const int ix = col % nx, iy = col / nx;
    if (!fct_in_fov(xs[ix], ys[iy], g)) continue;
    struct Sample {
      int v, i, qf, pf;
      double wq, wp, W;
    };
    std::vector<Sample> smp;
    std::vector<double> den(ns);
    for (int t = 0; t < H; ++t) {
      smp.clear();
      std::fill(den.begin(), den.end(), 0.0);
      for (int k = 0;; ++k) {
        const int v = t + k * H;
        if (v >= nviews) break;
        if (v < v0 || v >= v1) continue;
        const float sg = (k & 1) ? -1.0f : 1.0f;
        const FctBpColumn cl = fct_bp_column(xs[ix], ys[iy], sg * c32[t], sg * s32[t], g);
        const double sgn = (k & 1) ? -1.0 : 1.0;
        const double cc = sgn * c64[t], ss = sgn * s64[t];
        const double p = x64[ix] * ss - y64[iy] * cc;
        const double l = std::sqrt(s->r * s->r - p * p) - (x64[ix] * cc + y64[iy] * ss);
        const double zsrc = s->z0 + dz * v - kappa * std::asin(p / s->r);
        const double pidx64 = p / d.ds + s->central_channel;
        for (int i = 0; i < ns; ++i) {
          const double zref = zn[((s0 + i) / chunk) * chunk];
          const float zt = fct_tables::view_z_rel(s->z0, dz, v, zref);
          if (!fct_bp_in_window(zt, zs32[i], g)) continue;
          const float qi32 = fct_bp_row(zs32[i], cl, fct_bp_row0(cl, zt, g));
          int qf, pf;
          if (!fct_bp_select(qi32, cl.pidx, g, &qf, &pf)) continue;
          const double q64 = (zn[s0 + i] - zsrc) / (l * tanc);
          const double W = w_of_q(q64);
          smp.push_back({v, i, qf, pf, q64 * hr + s->central_row - qf, pidx64 - pf, W});
          den[i] += W;
        }
      }
      for (const Sample& e : smp) {
        if (!(den[e.i] > 0.0)) continue;
        const double c = scale * xbar[(size_t)e.i * nx * ny + col] * e.W / den[e.i];
        double* g0 = gbar + (size_t)(e.v - v0) * vstride + (size_t)e.qf * ncols + e.pf;
        double* g1 = g0 + ncols;
        const double a00 = c * (1 - e.wq) * (1 - e.wp), a01 = c * (1 - e.wq) * e.wp;
        const double a10 = c * e.wq * (1 - e.wp), a11 = c * e.wq * e.wp;
#pragma omp atomic
        g0[0] += a00;
#pragma omp atomic
        g0[1] += a01;
#pragma omp atomic
        g1[0] += a10;
#pragma omp atomic
        g1[1] += a11;
      }
    }
  }
  return 0;
}

// Transpose of or_rebin for parallel views [v0, v1): pbar is the whole raw
// array [view][row][ch] and is accumulated into (caller zeroes it).
void or_rebin_adjoint(const or_scan* s, const double* gbar, double* pbar, int v0, int v1) {
  // Transpose of or_rebin: parallel rows [j - v0][q][k] of the interleaved
  // detector scatter into raw fan data [view][row][ch].  Raw rows are split
  // over threads; both effective rows 2r, 2r+1 of raw row r stay in one.
  const fct_ffs::Geometry F = fct_ffs::make_geometry(
      s->ffs_mode, s->n_channels, s->n_rows, s->n_views, s->views_per_rotation,
      s->channel_spacing, s->row_spacing, s->central_channel, s->central_row, s->r, s->dsd,
      s->anode_angle);
  const int zpar = fct_ffs::has_z(F.mode) ? 2 : 1;
  std::vector<fct_ffs::Entry> tab((size_t)zpar * F.e_ncols);
  for (int zb = 0; zb < zpar; ++zb)
    for (int k = 0; k < F.e_ncols; ++k) tab[(size_t)zb * F.e_ncols + k] = fct_ffs::rebin_entry(F, k, zb);
  const size_t rvs = (size_t)F.nrows * F.ncols;
  const size_t evs = (size_t)F.e_nrows * F.e_ncols;
#pragma omp parallel for schedule(static)
  for (int row = 0; row < F.nrows; ++row) {
    for (int zb = 0; zb < zpar; ++zb) {
      const int q = zpar == 2 ? 2 * row + zb : row;
      for (int j = v0; j < v1; ++j) {
        const double* gl = gbar + (size_t)(j - v0) * evs + (size_t)q * F.e_ncols;
        for (int k = 0; k < F.e_ncols; ++k) {
          const fct_ffs::Entry& e = tab[(size_t)zb * F.e_ncols + k];
          if (!e.valid) continue;
          const int va = std::min(std::max(j + e.aoff, 0), e.mmax) * F.S + e.rho;
          const int vb = std::min(std::max(j + e.aoff + 1, 0), e.mmax) * F.S + e.rho;
          double* pa = pbar + (size_t)va * rvs + (size_t)row * F.ncols;
          double* pb = pbar + (size_t)vb * rvs + (size_t)row * F.ncols;
          const double y = gl[k];
          pa[e.b0] += (1 - e.fb) * (1 - e.fa) * y;
          pb[e.b0] += (1 - e.fb) * e.fa * y;
          pa[e.b1] += e.fb * (1 - e.fa) * y;
          pb[e.b1] += e.fb * e.fa * y;
        }
      }
    }
  }
}

// Whole transpose for native slices [s0, s1): xbar -> pbar (all raw views).
int or_adjoint_native(const or_scan* raw_scan, const or_recon* r, const double* xbar, int s0,
                      int s1, double* pbar) {
  or_scan es;
  or_effective_scan(raw_scan, &es);
  const or_scan* s = &es;
  int v0, v1;
  or_view_range(s, r, s0, s1, &v0, &v1);
  or_derived d;
  or_geometry(s, r, &d);
  if (v1 <= v0) {   // no view reaches these slices: the transpose is zero
    std::fill(pbar, pbar + (size_t)raw_scan->n_views * raw_scan->n_rows * raw_scan->n_channels,
              0.0);
    return 0;
  }
  const size_t vs = (size_t)s->n_rows * s->n_channels;
  std::vector<double> gbar((size_t)(v1 - v0) * vs, 0.0);
  int st = or_backproject_adjoint(s, r, xbar, v0, v1, s0, s1, gbar.data());
  if (st) return st;
  std::vector<double> h(d.nfft);
  or_filter_kernel(d.nfft, d.ds, r->filter_kind, r->filter_c, r->filter_a, h.data());
  or_filter_rows(gbar.data(), (int64_t)(v1 - v0) * s->n_rows, s->n_channels, h.data(), d.nfft, 1);
  std::fill(pbar, pbar + (size_t)raw_scan->n_views * raw_scan->n_rows * raw_scan->n_channels, 0.0);
  or_rebin_adjoint(raw_scan, gbar.data(), pbar, v0, v1);
  return 0;
}

}  // extern "C"
