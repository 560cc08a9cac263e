// wfbp_internal.h -- plan layout and kernel launcher interface (not exported).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fct_wfbp_gpu.h"
#include "ffs_geometry.h"
#include "wfbp_math.h"

namespace fct {

// Device-side rebinning entry per parallel channel and effective-row parity
// (FP32 weights from the FP64 recipes of wfbp_math.h / ffs_geometry.h).
// state: 0 = no fan sample (output 0), else 1 + the focal state rho; the
// state's projection m = j + aoff (+1) is clamped to [0, mmax] and read at
// raw view m * states + rho.
struct RebinDev {
  int32_t b0, b1, aoff, state;
  float fb, fa;
  int32_t mmax;   // >= 0 (0 for entries without a fan sample)
  uint32_t boff;  // byte offset of (focal state, raw row 0, fan channel b0) in a view group
};

// Filtered projections in HBM.
//  * FP32: plain rows, [view][row][cpitch] floats.
//  * FP16 / BF16: bilinear "quad" entries, [view][row][cpitch][4] 16-bit
//    values, entry (r, c) = (y[r][c], y[r][c+1], y[r+1][c], y[r+1][c+1]) / s,
//    so that one 8-byte load fetches the 2 x 2 footprint of a sample.  Slots
//    beyond the detector (c + 1 = cpitch, r + 1 = nrows) hold 0.  s is one
//    power of two per scan, chosen from a bound on |y| so that |y| / s < 2^15
//    (no overflow); it is kept in scale[0] of the scan's scale buffer.
// [0] s, [1] max |raw| bits (rebin+filter scratch), [2] rescale factor of the
// streaming reconstruction, [3] the s its backprojections read.
constexpr int kScaleElems = 4;

// Channel weight of the FP16 tap interpolation: the fraction wp rounded to a
// multiple of 2^-11, so that it and 1 - it are exact FP16 values summing to 1
// (the backprojection multiplies the FP16 taps by them in v_dot2_f32_f16).
__device__ __forceinline__ float fp16_channel_weight(float wp) {
  return __builtin_rintf(wp * 2048.0f) * (1.0f / 2048.0f);
}

// Everything the rebin + filter kernel reads.
struct RebinFilterArgs {
  const float* raw;          // [view][row][ncols]
  void* out;                 // filtered projections (layout above)
  float* scale;              // [kScaleElems] (FP16 / BF16) or null
  float hsum;                // sum |h_k| of the filter taps (bound on |y| / max|x|)
  int64_t raw_elems;         // raw floats (the abs-max pass)
  const RebinDev* tab;       // [ncols]
  const float* Hn;           // [nfft] real filter spectrum / nfft
  const float2* tw;          // [nfft] exp(-2 pi i k / nfft)
  int32_t nviews, nrows, ncols, cpitch, nfft, log2n;
  int32_t precision;
  int32_t nlines;            // nviews * nrows
  const float* lines_in;     // filter-only mode: [line][cpitch] FP32 input (no rebinning)
  int32_t ncu;               // compute units of the device (persistent grid size)
  int32_t raw_nrows, raw_ncols;   // raw detector (differs from nrows / ncols with FFS)
  int32_t states;            // focal states S (1 without FFS)
  int32_t zshift;            // 1 with z FFS: effective row q = 2 r + parity
  int32_t rebin_only;        // 1: write the rebinned rows (FP32) without filtering
  // Lines [line0, line0 + nlines) of the scan (line0 even, so the line pairs
  // are those of a whole-scan launch); the streaming reconstruction filters
  // views as they arrive.
  int32_t line0;
  int32_t skip_scale;        // 1: the FP16 / BF16 bound is already in scale[1]
  // The gather's raw window (rebinning): bytes per view group (S focal states
  // x raw rows x raw channels), and the smallest aoff and mmax of the table.
  uint32_t group_bytes;
  int32_t aoff_min, mmax_min;
};

// Everything the backprojection kernel reads.
struct BpArgs {
  const void* filt;          // filtered projections (layout above)
  const float* filt_scale;   // scale buffer of FP16 / BF16 projections
  float* native;             // [slice][y][x]
  int32_t* counts;           // count mode: [slice][y][x]
  unsigned long long* total; // count mode
  int32_t* sel;              // selection mode: [nviews]
  const float* xs;           // [nx]
  const float* ys;           // [ny]
  const float2* cs;          // [H] (cos, sin)
  const float* zs;           // [n_native] slice z relative to its chunk reference
  const float* zt;           // [nchunks][nviews] view z relative to the chunk reference
  const int2* slab_vrange;   // [nslabs] views [x, y) a slab can select
  const int2* wrange;        // [nchunks][nviews] native slices [x, y] inside the view window
  // [tiles_y * tiles_x][H] per 8 x 8 voxel tile and theta~: the extremes over
  // the tile's voxels in the FOV of the z-offsets from a view's table z at
  // which |q| = 1, (min, max) for the view parity 0 then 1
  // (launch_bp_zbounds).
  const float4* zbounds;
  FctBpConsts g;
  float scale;               // 2 pi / ppr
  int32_t nx, ny, n_native, nviews, H, cpitch;
  int32_t tiles_x, tiles_y, nslabs;
  int32_t sel_ix, sel_iy, sel_is;
  // Slabs [slab0, slab0 + slab_count) of a launch (slab_count 0: all); the
  // streaming reconstruction backprojects groups of slabs as their views
  // arrive.  Unused by the persistent batch kernel.
  int32_t slab0, slab_count;
};

// Persistent multi-job backprojection (backproject_v1.hip).
struct BpBatchArgs {
  const BpArgs* jobs;        // [n_jobs] device copy of each job's arguments
  const int32_t* tile_start; // [n_jobs + 1] first global tile of each job
  uint32_t* queue;           // [8] per-range tile counters, zero at launch
  int32_t n_jobs;            // <= 64
  int32_t total_tiles;
};

struct ThickenArgs {
  const float* native;
  float* out;
  const int32_t* off;        // [n_req + 1] into idx/w
  const int32_t* idx;        // native slice indices
  const float* w;            // normalised weights
  int32_t plane;             // nx*ny
  int32_t n_req;             // slices [s0, s0 + n_req) of the output
  int32_t s0;
  int32_t hu;                // 1: store 1000 (mu - mu_water) / mu_water = mu hu_scale - 1000
  float hu_scale;            // 1000 / mu_water
};

struct PhantomArgs {
  float* out;
  const fct_ellipsoid* ell;  // device copy
  int32_t n_ell;
  int32_t ncols, nrows, v0, v1;
  double r, dsd, dgamma, drow, cc, cr, angle0, dtheta, z0, dz;
  double cyl_r, cyl_mu, i0;
  unsigned long long seed;
  int32_t states;            // focal states (1: no FFS)
  int32_t mirror;            // 1: clockwise scan, rays mirrored y -> -y
  double ffs_off[4][3];      // per state (dr, da, dz) along (e_w, e_u, e_z)
};

// Ray-driven projector of the native volume onto the raw fan data
// (rayproj.hip), in the plan's frame (a clockwise scan mirrored).
struct RayProjArgs {
  const float* vol_in;       // project: native [slice][y][x]
  float* vol_out;            // transpose: native, accumulated into
  const float* raw_in;       // transpose: [view - v0][row][ch]
  float* raw_out;            // project: [view - v0][row][ch]
  int32_t ncols, nrows, v0, v1, nx, ny, nz, states;
  double r, dsd, dgamma, drow, cc, cr, angle0, dtheta, z0, dz;
  double ffs_off[4][3];      // per focal state (dr, da, dz) along (e_w, e_u, e_z)
  float gx0, gdx, gy0, gdy, gz0, gdz;   // voxel-centre grids (FP32 of FP64)
};

struct BpAdjArgs {
  BpArgs bp;                 // geometry tables (filt/native unused)
  const float* xbar;         // [slice][y][x] native-volume cotangent
  float* gbar;               // [view][row][cpitch] FP32, accumulated into
};

struct RebinAdjArgs {
  const float* gbar;         // [view][row][cpitch] (parallel, effective detector)
  float* pbar;               // [raw view][raw row][raw ch], accumulated into
  const RebinDev* tab;
  int32_t nviews, nrows, ncols, cpitch;   // parallel views, effective rows / channels
  int32_t raw_nrows, raw_ncols, states, zshift;
};

hipError_t launch_rebin_filter(const RebinFilterArgs& a, hipStream_t s);
// Ray-driven projection of vol_in onto raw_out (views [v0, v1)), or with
// transpose its exact transpose from raw_in into vol_out (accumulated).
hipError_t launch_ray_project(const RayProjArgs& a, bool transpose, hipStream_t s);
// Adaptive filtration of raw views [v0, v1): per view the largest finite
// |p| (vmax), then every sample blended with its 3 x 3 (row, channel) mean by
// lambda = clamp((p - (1 - s) vmax) / (s vmax), 0, 1).  out != in.
hipError_t launch_adaptive_filter(const float* in, float* out, float* vmax, int v0, int v1,
                                  int nrows, int ncols, float s, hipStream_t st);
// FP16 / BF16 storage bound in scale[1]: cleared, raised to max|x| over n
// floats (accumulating over calls), or set from a caller's max|raw|.
hipError_t launch_scale_reset(float* scale, hipStream_t s);
hipError_t launch_absmax(const float* x, int64_t n, float* scale, int ncu, hipStream_t s);
hipError_t launch_scale_set(float* scale, float max_abs_raw, hipStream_t s);
// Streaming FP16 / BF16: scale[0] from the bound so far, scale[2] the factor
// for the quads already stored; then those n_entries quads rescaled by it.
hipError_t launch_scale_update(float* scale, float hsum, hipStream_t s);
hipError_t launch_scale_commit(float* scale, hipStream_t s);
hipError_t launch_quad_rescale(void* quads, size_t n_entries, int precision, const float* scale,
                               int ncu, hipStream_t s);
hipError_t launch_backproject_adjoint(const BpAdjArgs& a, int tile_z, hipStream_t s,
                                      bool matched = false);
hipError_t launch_rebin_adjoint(const RebinAdjArgs& a, hipStream_t s);
hipError_t launch_fill(float* p, size_t n, float v, hipStream_t s);
hipError_t launch_backproject(const BpArgs& a, int tile_z, int variant, int precision, int mode,
                              hipStream_t s);
// Fills a.zbounds (a BpArgs whose zbounds member is the output) for the
// plan's tiles and theta~ (backproject_v1.hip).
hipError_t launch_bp_zbounds(const BpArgs& a, float4* zbounds, hipStream_t s);
hipError_t launch_backproject_v1(const BpArgs& a, int tile_z, int precision, int mode,
                                 hipStream_t s);
hipError_t launch_backproject_batch(const BpBatchArgs& b, int tile_z, int precision,
                                    bool centered, int grid, hipStream_t s);
hipError_t backproject_batch_waves_per_cu(int tile_z, int precision, int* waves);
hipError_t launch_thicken(const ThickenArgs& a, hipStream_t s);
hipError_t launch_phantom(const PhantomArgs& a, hipStream_t s);

// kBpMatched: the plain R / l-weighted sum over every view that sees the
// voxel (no view window, no W normalisation) -- the transpose of the matched
// projector of adjoint.hip's matched mode.
enum BpMode { kBpAccumulate = 0, kBpCount = 1, kBpSelect = 2, kBpMatched = 3 };

}  // namespace fct

struct fct_wfbp_gpu_plan;

namespace fct {
// Plan helpers shared by capi.hip and batch.hip.
BpArgs plan_bp_args(const fct_wfbp_gpu_plan* p);   // tables only (no buffers)
int plan_set_tile_z(fct_wfbp_gpu_plan* p, int tz);  // rebuilds the slab view ranges
// Rebin + filter of the lines some slab reads (the whole-chain entry points).
int chain_rebin_filter(fct_wfbp_gpu_plan* p, const float* raw, void* filtered, float* scale);
}  // namespace fct

struct fct_wfbp_gpu_ctx {
  int device = 0;
  std::string last_error;
  // The streaming chain's copy stream and two backprojection streams,
  // shared by the context's plans (created on first use): with a few
  // hardware queues per process (GPU_MAX_HW_QUEUES), streams of every plan
  // would share queues with the plans' own streams, and a copy queued
  // behind a kernel waits for it.
  hipStream_t copy_stream = nullptr;
  hipStream_t bp_stream[2] = {nullptr, nullptr};
};

struct fct_wfbp_gpu_plan {
  fct_wfbp_gpu_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  fct_scan_geometry scan{};         // interleaved (effective) geometry, ffs_mode none
  fct_scan_geometry raw_scan{};     // as given (raw detector, FFS mode)
  fct_ffs::Geometry ffs{};
  fct_recon_config recon{};
  fct_exec_options opts{};

  int ncols = 0, nrows = 0, nviews = 0, ppr = 0, H = 0, nfft = 0, log2n = 0, cpitch = 0;
  int n_native = 0, n_req = 0, nchunks = 0, tile_z = 16, variant = 0, nslabs = 0;
  int ncu = 0;                      // compute units of the plan's device
  double dgamma = 0, ds = 0, dtheta = 0, dz = 0, csw = 0, acq = 0;
  FctBpConsts g{};
  float scale = 0;
  int out_hu = 0;                   // thickened volume in HU (fct_wfbp_gpu_set_output_units)
  double mu_water = 0;
  int af_on = 0;                    // adaptive filtration of the raw scan before rebinning
  float af_s = 0;                   // its strength s in (0, 1]

  // Device tables owned by the plan.
  float* d_xs = nullptr;
  float* d_ys = nullptr;
  float2* d_cs = nullptr;
  float* d_zs = nullptr;
  float* d_zt = nullptr;
  int2* d_slab_vrange = nullptr;
  int2* d_wrange = nullptr;
  float4* d_zbounds = nullptr;
  fct::RebinDev* d_rebin = nullptr;
  float* d_Hn = nullptr;
  float2* d_tw = nullptr;
  int32_t* d_thk_off = nullptr;
  int32_t* d_thk_idx = nullptr;
  float* d_thk_w = nullptr;

  // Intermediate buffers for fct_wfbp_gpu_reconstruct (allocated on first use).
  void* d_filtered = nullptr;
  float* d_scale = nullptr;
  float* d_native = nullptr;
  float* d_raw_stage = nullptr;
  float* d_vol_stage = nullptr;
  float* d_gbar = nullptr;          // adjoint scratch, [view][row][cpitch] FP32
  int32_t* d_thkT_off = nullptr;    // transposed thickening table
  int32_t* d_thkT_idx = nullptr;
  float* d_thkT_w = nullptr;
  float* d_af_raw = nullptr;        // adaptively filtered raw scan (allocated on first use)
  float* d_af_vmax = nullptr;       // per raw view max |p|

  // Streaming reconstruction (fct_wfbp_gpu_reconstruct_stream and the host
  // path of fct_wfbp_gpu_reconstruct), created on first use.
  hipEvent_t ev_compute = nullptr;
  float* h_stage[2] = {nullptr, nullptr};   // pinned staging for a view reader
  int stage_views = 0;
  hipEvent_t ev_bp[2] = {nullptr, nullptr};   // after the latest piece on each stream
  hipEvent_t ev_rebin = nullptr;         // after the latest rebinning launch
  hipEvent_t ev_ratio = nullptr;         // after the storage-scale ratio's copy to h_ratio
  float* h_ratio = nullptr;              // pinned: scale[2] of the latest scale update
  std::vector<int2> h_slab_vrange;      // host copy of d_slab_vrange
  std::vector<fct::RebinDev> h_rebin;   // host copy of d_rebin
  std::vector<int32_t> h_thk_need;      // per requested slice: 1 + its last native slice (0: none)
  std::vector<hipEvent_t> ev_chunk;     // per chunk: its copy has landed
  int rebin_reach = 0;                  // parallel view j reads raw projections < j S + reach
  int rebin_aoff_min = 0, rebin_mmax_min = 0;   // over the rebinning table (the gather window)

  int64_t raw_elems() const {
    return (int64_t)raw_scan.n_views * raw_scan.n_rows * raw_scan.n_channels;
  }
  size_t filtered_entry_bytes() const {   // FP32 value, or 4 x 16-bit quad
    return opts.precision == FCT_PRECISION_FP32 ? 4 : 8;
  }
  size_t filtered_bytes() const {
    return (size_t)nviews * nrows * cpitch * filtered_entry_bytes();
  }
  size_t scale_elems() const {
    return opts.precision == FCT_PRECISION_FP32 ? 0 : (size_t)fct::kScaleElems;
  }
  double hsum = 0;                  // sum |h_k| of the filter taps
};
