// wfbp_oracle.h -- TEST INFRASTRUCTURE ONLY.
//
// FP64 CPU restatement of FreeCT's WFBP path (fct_wfbp), used by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
// The product (freect_amd/) never links, loads or calls anything here.
//
// What follows which reference lines (paths relative to the FreeCT tree):
//   or_geometry        ReconstructionMachine.cu:44-74, recon_structs.h:69-83
//   or_slice_grid      ReconstructionMachine.cu:107-137
//   or_filter_kernel   ReconstructionMachine.cu:478-513, resources/ramp.m:1-15
//   or_rebin           cuda_kernels.cuh:11-39; FFS interleave of
//                      freect_amd/csrc/ffs_geometry.h (deflections and order
//                      of fct_icd generate_system_matrix_ffs.cpp:66-67, 242-290)
//   or_filter_rows     cuda_kernels.cuh:41-77 + ReconstructionMachine.cu:215-290
//                      (FFT product == linear convolution, see below)
//   or_backproject     cuda_kernels.cuh:79-93, 95-196, 198-287
//   or_thicken         cuda_kernels.cuh:289-320
//   or_project_phantom new (synthetic scan generator; analytic line integrals)
//   or_adaptive_filter new (adaptive filtration; the key is parsed at
//                      parse_config.cpp:57 and never applied)
//
// Pins (tests/test_oracle_*.py): the nine reference filter vectors
// (resources/filters/*.txt), Ram-Lak / Shepp-Logan closed forms, closed-form
// geometry, W(q) special values, direct-vs-FFT convolution, rebinning of a
// centred water cylinder against 2*mu*sqrt(a^2-t^2), FBP of a uniform cylinder
// against mu, triangular thickening weights, analytic chord lengths.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n_channels, n_rows, n_views, views_per_rotation;
  double channel_spacing;   // mm at the detector (arc length)
  double row_spacing;       // mm at the detector
  double dsd, r;            // source-detector, source-isocentre distance (mm)
  double central_channel, central_row;
  double angle0;            // tube angle of view 0 (rad)
  double z0;                // source z of view 0 (mm)
  double feed;              // table feed per rotation (mm, signed)
  int32_t ffs_mode;         // 0 none, 1 phi, 2 z, 3 phi+z (focal state = view mod S)
  int32_t rotation;         // +1 (or 0) counter-clockwise, -1 clockwise
  double anode_angle;       // rad
} or_scan;

typedef struct {
  double start_pos, end_pos, recon_fov, slice_thickness, slice_pitch;
  int32_t nx, ny;
  double x_origin, y_origin;
  int32_t filter_kind;      // 0: generalised-Hamming ramp (c, a); 1: Shepp-Logan
  double filter_c, filter_a;
} or_recon;

typedef struct {
  int32_t nfft;
  double csw, zrot, theta_cone, acq_fov, ds;
  double half_acq_fov_sq, z_rot_over_2pi, tan_half_cone, pixel_scale;
  double two_pi_over_ppr;
  int32_t views_per_half_turn, n_half_turns;
  int32_t n_slices_native, n_slices_requested;
} or_derived;

typedef struct {
  double c[3];       // centre (mm)
  double a[3];       // semi-axes (mm)
  double rot[9];     // row-major R; body axis i is column i of R
  double dmu;        // additive attenuation (1/mm)
} or_ellipsoid;

// A clockwise scan (rotation -1) is the mirror image y -> -y of the
// counter-clockwise scan with tube angles -a_v: or_project_phantom traces its
// rays directly; the reconstruction functions take counter-clockwise scans
// (pyoracle mirrors a clockwise one: angle0 -> -angle0, y_origin -> -y_origin,
// rows of the result reversed).
// The interleaved detector an FFS scan is reconstructed on (ffs_mode 0);
// the scan itself without FFS.  Every function below except or_rebin,
// or_rebin_adjoint, or_project_phantom, or_reconstruct_native and
// or_adjoint_native takes the interleaved (parallel-domain) geometry.
void or_effective_scan(const or_scan* s, or_scan* out);
// Shared FP32 helical weight fct_w_of_q and the FP64 definition at q[0..n).
void or_weights(const float* q, int n, float* w32, double* w64);
// The shared FP32 weight from the detector row position qi[0..n) for a
// detector of nrows centred rows (bp_consts): w_sel = fct_w_of_row where the
// selection rule takes the row pair (floor(qi) in [0, nrows - 2]), else 0;
// w_clamp = the unmasked form with t clamped to [0, 1] (the backprojection's
// whole-slab path for centred rows).
void or_row_weights(int nrows, const float* qi, int n, float* w_sel, float* w_clamp);
int or_geometry(const or_scan* s, const or_recon* r, or_derived* d);
int or_slice_grid(double start, double end, double step, double* z, int cap);
void or_filter_kernel(int n, double ds, int kind, double c, double a, double* h);
void or_rebin(const or_scan* s, const float* raw, double* out, int v0, int v1);
// Ray-driven projector (new; the refinement's line-integral model): Joseph's
// method along every raw ray of the physical scan (FFS and clockwise rays
// traced directly) through the native volume [slice][y][x], zero outside;
// raw [view][row][ch] overwritten.  or_ray_backproject: its exact transpose
// (vol overwritten).
int or_ray_project(const or_scan* s, const or_recon* r, const double* vol, double* raw);
int or_ray_backproject(const or_scan* s, const or_recon* r, const double* raw, double* vol);
// Adaptive filtration of the raw scan with strength s (new; the FreeCT key
// adaptive_filtration_s is parsed but unused): per view, samples above
// (1 - s) vmax are blended with their 3 x 3 (row, channel) mean.
void or_adaptive_filter(const or_scan* s, double strength, const float* raw, float* out);
// The rebinning table (fct_ffs::rebin_entry, as the plan uploads it):
// ient[e][valid, rho, b0, b1, aoff, mmax], fent[e][fb, fa]; returns the entry
// count (call with NULL buffers to size them).
int or_rebin_entries(const or_scan* s, int32_t* ient, double* fent);
void or_filter_rows(double* rows, int64_t nlines, int ncols, const double* h,
                    int nfft, int use_fft);
// filtered: parallel views [v0, v1) of [view][row][ch] FP64.  Computes native
// slices [s0, s1) into vol (layout [s - s0][y][x]); counts selected samples.
int or_backproject(const or_scan* s, const or_recon* r, const double* filtered,
                   int v0, int v1, int s0, int s1, double* vol,
                   int64_t* n_selected);
// FP32-accumulation instantiation of the same loop (CPU baseline only).
int or_backproject_f32(const or_scan* s, const or_recon* r,
                       const float* filtered, int v0, int v1, int s0, int s1,
                       float* vol, int64_t* n_selected);
int or_thicken(const or_scan* s, const or_recon* r, const double* native,
               double* out);
// Selected-sample count per native voxel of slices [s0, s1) (layout like vol).
int or_bp_counts(const or_scan* s, const or_recon* r, int s0, int s1,
                 int32_t* counts, int64_t* n_selected);
// Coverage per native voxel of slices [s0, s1): min over theta~ of the FP64
// weight sum sum_k W(q_k) of the selected views (0 where some theta~ is empty).
int or_bp_coverage(const or_scan* s, const or_recon* r, int s0, int s1,
                   double* minden);
// One (voxel column, slice) of the selection rule, for bit-exact comparison:
// writes, for every parallel view, qf*65536+pf or -1 when unused.
int or_bp_selection(const or_scan* s, const or_recon* r, int ix, int iy,
                    int is, int32_t* sel);
void or_project_phantom(const or_scan* s, double cyl_radius, double cyl_mu,
                        const or_ellipsoid* e, int ne, float* out, int v0,
                        int v1);
// Whole chain for native slices [s0, s1): raw fan data -> native volume.
int or_reconstruct_native(const or_scan* s, const or_recon* r, const float* raw,
                          int s0, int s1, double* vol, int64_t* n_selected);
// View range [v0, v1) that native slices [s0, s1) can select.
void or_view_range(const or_scan* s, const or_recon* r, int s0, int s1,
                   int* v0, int* v1);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif
// Transpose of the chain (wfbp_oracle_adjoint.cpp).  gbar / pbar are
// accumulated into.
int or_backproject_adjoint(const or_scan* s, const or_recon* r, const double* xbar, int v0,
                           int v1, int s0, int s1, double* gbar);
void or_rebin_adjoint(const or_scan* s, const double* gbar, double* pbar, int v0, int v1);
int or_adjoint_native(const or_scan* s, const or_recon* r, const double* xbar, int s0, int s1,
                      double* pbar);
#ifdef __cplusplus
}
#endif
