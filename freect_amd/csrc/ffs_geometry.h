// ffs_geometry.h -- flying focal spot (FFS): focal-spot deflections, the
// interleaved ("effective") detector and the per-state rebinning table.
//
// Raw FFS data are projections i = 0..nviews-1 at tube angle
// angle0 + i * 2 pi / ppr whose focal spot cycles through S states
// (rho = i mod S; S = 2 for phi or z FFS, 4 for both).  State rho moves the
// focal spot by dr_rho along e_w (radial), da_rho along e_u (anti-radial) and
// dz_rho along z, with the amplitudes and the order of
// generate_system_matrix_ffs.cpp:66-67 and 242-290; the detector stays fixed
// relative to the undeflected source (:136-149).  One unit is used for the
// anode angle throughout (radians), where that file mixes degrees (:66) and
// radians (:267-282).
//
// Interleave: every state is rebinned on its own into parallel geometry, and
// the states are merged on a finer grid:
//   * phi FFS: parallel channel spacing ds/2; effective channel k is served by
//     the state deflected by +da for even k and by -da for odd k (the +da rays
//     lie a quarter channel below the undeflected ones at the isocentre, the
//     -da rays a quarter above), central channel 2 cc + 1/2.  Each state is
//     rebinned exactly at the effective channel's p (its own fan-channel
//     position from the deflected-source ray geometry), so the in-plane
//     interleave has no approximation beyond the bilinear interpolation;
//   * z FFS: row spacing drow/2; effective row 2 r + b takes raw row r of the
//     state deflected by -dz (b = 0) or +dz (b = 1), central row 2 cr + 1/2.
//     The backprojection then treats the effective rows as one detector seen
//     from the undeflected source: exact at the isocentre (the deflection
//     shifts the rays there by a quarter row) and approximate elsewhere, the
//     usual z-FFS interleave.
// Parallel view j (angle angle0 + j S 2pi/ppr, table z z0 + j S feed/ppr)
// collects, for every state, raw projections rho + S m around m = j.
#pragma once

#include <cmath>
#include <cstdint>

#include "wfbp_math.h"

namespace fct_ffs {

enum Mode { kNone = 0, kPhi = 1, kZ = 2, kPhiZ = 3 };

inline int n_states(int mode) { return mode == kPhiZ ? 4 : mode == kNone ? 1 : 2; }
inline bool has_phi(int mode) { return (mode & 1) != 0; }
inline bool has_z(int mode) { return (mode & 2) != 0; }

struct Deflection {
  double dr, da, dz;   // along e_w, e_u, e_z (mm)
};

// Deflection amplitudes (generate_system_matrix_ffs.cpp:66-67):
// dr = Dsd csw / (4 (Dsd - R) tan(anode)), da = Dsd R sin(dgamma) / (4 (Dsd - R)).
inline void amplitudes(double dsd, double r, double csw, double dgamma, double anode,
                       double* dr, double* da) {
  *dr = dsd * csw / (4.0 * (dsd - r) * std::tan(anode));
  *da = dsd * r * std::sin(dgamma) / (4.0 * (dsd - r));
}

// Offset of state rho (generate_ffs_offset, generate_system_matrix_ffs.cpp:250-283);
// the z component of a radial move dr is dr tan(anode).
inline Deflection state_offset(int mode, int rho, double dr, double da, double anode) {
  const double dz = dr * std::tan(anode);
  switch (mode) {
    case kPhi:   // :253-259
      return rho == 0 ? Deflection{0.0, da, 0.0} : Deflection{0.0, -da, 0.0};
    case kZ:     // :264-269
      return rho == 0 ? Deflection{-dr, 0.0, -dz} : Deflection{dr, 0.0, dz};
    case kPhiZ: {   // :273-282
      const double sr = rho < 2 ? -1.0 : 1.0, sa = (rho & 1) ? 1.0 : -1.0;
      return Deflection{sr * dr, sa * da, sr * dz};
    }
    default:
      return Deflection{0.0, 0.0, 0.0};
  }
}

// State serving effective channel k (phi bit) and effective-row parity zb.
inline int state_of(int mode, int k, int zb) {
  const int plus_da = (k & 1) == 0;   // +da for even k
  switch (mode) {
    case kPhi: return plus_da ? 0 : 1;            // phi order: rho 0 is +da
    case kZ: return zb;                           // z order: rho 0 is -dz
    case kPhiZ: return (plus_da ? 1 : 0) + 2 * zb;   // both: rho & 1 is +da, rho >> 1 is +dz
    default: return 0;
  }
}

// In-plane parallel coordinates of the ray from the deflected focal spot
// (R + dr) e_w + da e_u to the detector point at fan angle beta (fixed
// detector, R e_w - Dsd (cos beta e_w + sin beta e_u)): its angle relative to
// the view, beta' = atan2(da + Dsd sin beta, dr + Dsd cos beta), and its
// signed distance from the isocentre, p = (R + dr) sin beta' - da cos beta'.
// With no deflection beta' = beta and p = R sin beta.
inline void ray_parallel(double beta, double r, double dsd, double dr, double da, double* bp,
                         double* p) {
  *bp = std::atan2(da + dsd * std::sin(beta), dr + dsd * std::cos(beta));
  *p = (r + dr) * std::sin(*bp) - da * std::cos(*bp);
}

// Fan angle beta whose deflected ray has parallel distance p (p increases
// with beta on the detector): bisection to FP64 resolution.
inline bool solve_beta(double p, double r, double dsd, double dr, double da, double* beta) {
  if (!(std::fabs(p) < r + std::fabs(dr))) return false;
  double lo = -1.5, hi = 1.5, bp, v;
  ray_parallel(lo, r, dsd, dr, da, &bp, &v);
  if (v > p) return false;
  ray_parallel(hi, r, dsd, dr, da, &bp, &v);
  if (v < p) return false;
  for (int it = 0; it < 80 && hi - lo > 0.0; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    ray_parallel(mid, r, dsd, dr, da, &bp, &v);
    (v < p ? lo : hi) = mid;
  }
  *beta = 0.5 * (lo + hi);
  return true;
}

// Raw and effective (interleaved) geometry.
struct Geometry {
  int mode;
  int S;                         // focal states
  // raw
  int ncols, nrows, nviews, ppr;
  double dch, drow, cc, cr;      // mm at the detector, fractional indices
  double r, dsd, anode;
  double dr, da;                 // deflection amplitudes
  // effective
  int e_ncols, e_nrows, e_nviews, e_ppr;
  double e_dch, e_drow, e_cc, e_cr;
};

inline Geometry make_geometry(int mode, int ncols, int nrows, int nviews, int ppr, double dch,
                              double drow, double cc, double cr, double r, double dsd,
                              double anode) {
  Geometry g{};
  g.mode = mode;
  g.S = n_states(mode);
  g.ncols = ncols;
  g.nrows = nrows;
  g.nviews = nviews;
  g.ppr = ppr;
  g.dch = dch;
  g.drow = drow;
  g.cc = cc;
  g.cr = cr;
  g.r = r;
  g.dsd = dsd;
  g.anode = anode;
  if (mode != kNone) amplitudes(dsd, r, r * drow / dsd, dch / dsd, anode, &g.dr, &g.da);
  const bool phi = has_phi(mode), z = has_z(mode);
  g.e_ncols = phi ? 2 * ncols : ncols;
  g.e_dch = phi ? 0.5 * dch : dch;
  g.e_cc = phi ? 2.0 * cc + 0.5 : cc;
  g.e_nrows = z ? 2 * nrows : nrows;
  g.e_drow = z ? 0.5 * drow : drow;
  g.e_cr = z ? 2.0 * cr + 0.5 : cr;
  g.e_ppr = ppr / g.S;
  g.e_nviews = nviews / g.S;
  return g;
}

// Rebinning entry of effective channel k for effective-row parity zb: state
// rho, fan channels b0/b1 and fraction fb on the raw detector, and the
// j-independent split of the state's projection position m = j + nd into
// aoff + fa.  mmax is the last projection index m of state rho (clamped
// like the texture at the ends of the scan).  Mode none is exactly
// fct_tables::rebin_entry.
struct Entry {
  int32_t valid, rho, b0, b1, aoff, mmax;
  double fb, fa;
};

inline Entry rebin_entry(const Geometry& g, int k, int zb) {
  Entry e{};
  const double dtheta = 2.0 * 3.14159265358979323846 / (double)g.ppr;
  e.rho = state_of(g.mode, k, zb);
  e.mmax = (g.nviews - e.rho + g.S - 1) / g.S - 1;
  if (g.mode == kNone) {
    const fct_tables::RebinEntry b =
        fct_tables::rebin_entry(k, g.cc, g.dch / g.dsd, dtheta, g.ncols);
    e.valid = b.valid;
    e.b0 = b.b0;
    e.b1 = b.b1;
    e.aoff = b.aoff;
    e.fb = b.fb;
    e.fa = b.fa;
    return e;
  }
  const Deflection d = state_offset(g.mode, e.rho, g.dr, g.da, g.anode);
  const double dgamma = g.dch / g.dsd;
  const double p = ((double)k - g.e_cc) * (g.e_dch * g.r / g.dsd);
  double beta = 0.0, bprime = 0.0;
  if (d.dr == 0.0 && d.da == 0.0) {
    const double sb = p / g.r;
    if (!(std::fabs(sb) <= 1.0)) return e;
    beta = std::asin(sb);
    bprime = beta;
  } else {
    if (!solve_beta(p, g.r, g.dsd, d.dr, d.da, &beta)) return e;
    double pp;
    ray_parallel(beta, g.r, g.dsd, d.dr, d.da, &bprime, &pp);
  }
  const double bidx = beta / dgamma + g.cc;
  if (!(bidx >= 0.0) || !(bidx < (double)g.ncols)) return e;
  if (e.mmax < 0) return e;
  e.valid = 1;
  const double bf = std::floor(bidx);
  e.b0 = (int32_t)bf;
  e.b1 = e.b0 + 1 < g.ncols ? e.b0 + 1 : g.ncols - 1;
  e.fb = bidx - bf;
  const double nd = -(bprime / dtheta + (double)e.rho) / (double)g.S;
  const double af = std::floor(nd);
  e.aoff = (int32_t)af;
  e.fa = nd - af;
  return e;
}

}  // namespace fct_ffs
