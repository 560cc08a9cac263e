/* fct_wfbp_gpu.h -- C ABI of the MI355X-native WFBP engine.
 *
 * Drop-in boundary for FreeCT's fct_wfbp reconstruction path.  The reference
 * has no FFI; its boundary is the recon YAML (parse_config.cpp:31-62), the
 * fct::RawDataSet getters (RawDataSet.h:91-139), the raw float32 output volume
 * (ReconstructionMachine.cu:436-455) and the fct::ReconstructionMachine
 * methods (ReconstructionMachine.h:16-21).  Every entry point here takes plain
 * pointers and sizes, returns an fct_status, never exits the process and never
 * throws.  Device buffers are hipMalloc'd by the caller unless stated.
 *
 * Threading: a plan is bound to one device and one stream and is not
 * re-entrant; distinct plans may be driven from distinct host threads.
 */
#ifndef FCT_WFBP_GPU_H
#define FCT_WFBP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCT_API __attribute__((visibility("default")))

typedef enum {
  FCT_OK = 0,
  FCT_ERR_INVALID_ARGUMENT = 1,
  FCT_ERR_UNSUPPORTED = 2,
  FCT_ERR_OUT_OF_MEMORY = 3,
  FCT_ERR_DEVICE = 4,
  FCT_ERR_LIMIT = 5
} fct_status;

/* Precision of the filtered projections handed from rebin_filter to
 * backproject.  Accumulation is FP32 in every mode. */
typedef enum {
  FCT_PRECISION_FP32 = 0,
  FCT_PRECISION_FP16 = 1,   /* FP16 quads x one power-of-two FP32 scale per scan */
  FCT_PRECISION_BF16 = 2    /* BF16 quads x one power-of-two FP32 scale per scan */
} fct_precision;

/* Scan description: what RawDataSet.h:91-121 exposes, with the per-view tube
 * angle and table position as the linear law angle0 + rotation*v*2pi/ppr and
 * z0 + v*feed/ppr (the reference's rebinning already assumes uniform views,
 * cuda_kernels.cuh:15-19; its backprojection reads the per-view tables,
 * ReconstructionMachine.cu:96-102, which the C++ host checks against this law
 * before it builds a plan). */
typedef struct {
  int32_t n_channels;          /* getDetectorChannels */
  int32_t n_rows;              /* getDetectorRows */
  int32_t n_views;             /* getTotalNumProjections */
  int32_t views_per_rotation;  /* getProjectionsPerRotation (must be even) */
  double channel_spacing;      /* getDetectorTransverseSpacing, mm at detector */
  double row_spacing;          /* getDetectorAxialSpacing, mm at detector */
  double dsd;                  /* getDistSourceToDetector */
  double r;                    /* getDistSourceToIsocenter */
  double central_channel;      /* getDetectorCentralChannel */
  double central_row;          /* getDetectorCentralRow */
  double angle0;               /* getTubeAngle(0) */
  double z0;                   /* getTablePosition(0) */
  double feed;                 /* getTablePosition(ppr) - getTablePosition(0) */
  /* Flying focal spot (getFlyingFocalSpotMode, RawDataSet.h:102): the focal
   * state of projection v is v mod S (S = 2, or 4 for phi+z) with the
   * deflections and order of fct_icd's generate_system_matrix_ffs.cpp:66-67,
   * 242-290.  With FFS the plan works on the interleaved detector
   * (fct_plan_info.n_filtered_*): phi doubles the channels, z the rows, and
   * there are n_views / S parallel views. */
  int32_t ffs_mode;            /* fct_ffs_mode */
  /* Direction of the tube's rotation: +1 (or 0) counter-clockwise, the tube
   * angle growing with the view; -1 clockwise.  The fan angle of a channel,
   * (c - central_channel) * channel_spacing / dsd, and the phi-FFS
   * deflection are measured in the direction of rotation, so a clockwise
   * scan is the mirror image (y -> -y) of a counter-clockwise one. */
  int32_t rotation;
  double anode_angle;          /* rad; z-FFS deflection dr tan(anode) */
} fct_scan_geometry;

typedef enum {
  FCT_FFS_NONE = 0,
  FCT_FFS_PHI = 1,             /* in-plane (anti-radial) focal spot wobble */
  FCT_FFS_Z = 2,               /* z (radial + axial) wobble */
  FCT_FFS_PHI_Z = 3            /* both: a 4-state cycle */
} fct_ffs_mode;

/* Recon configuration: the keys of parse_config.cpp:39-57 plus filter
 * selection.  recon_kernel, tube_angle_offset and adaptive_filtration_s are
 * carried but unused, as in the reference (parse_config.cpp:53,56,57). */
typedef struct {
  double start_pos, end_pos;   /* mm, same frame as the table positions */
  double recon_fov;            /* mm */
  double slice_thickness;      /* mm */
  double slice_pitch;          /* mm; 0 -> slice_thickness */
  int32_t nx, ny;
  double x_origin, y_origin;   /* mm */
  int32_t recon_kernel;
  int32_t filter_kind;         /* 0: generalised-Hamming ramp (c, a); 1: Shepp-Logan */
  double filter_c, filter_a;   /* ramp cutoff and Hamming a (ramp.m:7-8) */
  double tube_angle_offset;
  double adaptive_filtration_s;
} fct_recon_config;

/* Backprojection kernel variants (identical selection and arithmetic). */
typedef enum {
  FCT_BP_AUTO = 0,             /* tuned default */
  FCT_BP_DIRECT = 1,           /* one wave per 8x8xTZ tile, taps read per sample */
  FCT_BP_LEAN = 2              /* + wave-uniform slice skipping, buffer loads */
} fct_bp_variant;

typedef struct {
  int32_t precision;           /* fct_precision */
  int32_t tile_z;              /* slices per backprojection tile: 0 = tuned default, or 8/16/32 */
  int32_t bp_variant;          /* fct_bp_variant */
  /* 1: adaptive filtration of the raw scan with strength
   * recon.adaptive_filtration_s in (0, 1] before the rebinning (the
   * reference parses the key and never applies it, parse_config.cpp:57,
   * README.md:14, so a reference recon file alone does not turn it on):
   * per view, with vmax the largest |p| of the view, every sample p is
   * blended with the mean of its 3 x 3 (row, channel) neighbourhood by
   * lambda = clamp((p - (1 - s) vmax) / (s vmax), 0, 1), smoothing only the
   * most attenuated (noisiest) rays after the idea of Kachelriess, Watzke and
   * Kalender, Med. Phys. 28 (2001) 475.  The plan then holds a raw-sized
   * buffer (not allowed by fct_wfbp_gpu_reconstruct_noalloc). */
  int32_t adaptive_filtration;
} fct_exec_options;

/* Units of the thickened volume (fct_wfbp_gpu_set_output_units). */
typedef enum {
  FCT_UNITS_MU = 0,            /* linear attenuation, 1/mm (the reference's output) */
  FCT_UNITS_HU = 1             /* Hounsfield: 1000 (mu - mu_water) / mu_water */
} fct_units;

typedef struct {
  int32_t n_slices_native;     /* slices at the collimated width (ReconstructionMachine.cu:110-137) */
  int32_t n_slices_requested;  /* slices at slice_pitch */
  int32_t nfft;                /* 2^ceil(log2(2*ncols)) (ReconstructionMachine.cu:54) */
  int32_t channel_pitch;       /* padded channels per filtered row */
  int64_t raw_elems;           /* n_views*n_rows*n_channels floats, [view][row][ch] */
  int64_t filtered_bytes;      /* bytes of the filtered buffer: FP32 [view][row][channel_pitch];
                                  FP16/BF16 [view][row][channel_pitch][4] bilinear quads */
  int64_t scale_elems;         /* floats of the FP16/BF16 scale buffer (0 for FP32) */
  int64_t native_elems;        /* n_slices_native*ny*nx floats, [slice][y][x] */
  int64_t volume_elems;        /* n_slices_requested*ny*nx floats, [slice][y][x] */
  double collimated_slice_width;
  double acquisition_fov;
  double z_rot;
  int32_t tile_z;
  int32_t bp_variant;
  /* Filtered (parallel, interleaved) dimensions: equal to n_views, n_rows,
   * n_channels without FFS. */
  int32_t n_parallel_views;
  int32_t n_filtered_rows;
  int32_t n_filtered_channels;
  int32_t ffs_states;          /* focal states S (1 without FFS) */
  /* Raw views [first, end) some native slice reads: the views the streaming
   * chain uploads (and requests from a view reader). */
  int32_t read_views_first;
  int32_t read_views_end;
} fct_plan_info;

typedef struct {
  double c[3];                 /* centre, mm */
  double a[3];                 /* semi-axes, mm */
  double rot[9];               /* row-major rotation; body axis i = column i */
  double dmu;                  /* additive attenuation, 1/mm */
} fct_ellipsoid;

typedef struct fct_wfbp_gpu_ctx fct_wfbp_gpu_ctx;
typedef struct fct_wfbp_gpu_plan fct_wfbp_gpu_plan;

/* Device selection and validation; replaces validate_selected_device
 * (util.cu:6-35) and main.cpp's -d option (main.cpp:43-44).  The context
 * owns the streaming chain's copy and backprojection streams, shared by its
 * plans (streaming calls of several plans of one context queue on them). */
FCT_API int fct_wfbp_gpu_create(int32_t device, fct_wfbp_gpu_ctx** ctx);
FCT_API int fct_wfbp_gpu_destroy(fct_wfbp_gpu_ctx* ctx);
/* Last error message of the context (the reference prints and exits,
 * gpu_error_check.h:38-58). */
FCT_API const char* fct_wfbp_gpu_last_error(const fct_wfbp_gpu_ctx* ctx);

/* Setup: ConfigureCTGeometry + AllocateKeyArrays + generate_filter
 * (ReconstructionMachine.cu:44-139, 478-544, recon_structs.h:69-83).
 * FP64 view tables and slice grids, FP32 per-chunk offsets, filter spectrum,
 * rebinning tables.  stream is the hipStream_t the plan's work is queued on;
 * NULL is the device's null stream.  FCT_ERR_LIMIT for scans outside the
 * rebinning's addressing: a raw view group (focal states x rows x channels
 * floats) of 16 MiB or more, a fan spanning 4 GiB of raw views, fan view
 * offsets beyond -2048..2047 or 2^20 or more projections per focal state. */
FCT_API int fct_wfbp_gpu_plan_create(fct_wfbp_gpu_ctx* ctx, const fct_scan_geometry* scan,
                                     const fct_recon_config* recon,
                                     const fct_exec_options* opts, void* stream,
                                     fct_wfbp_gpu_plan** plan);
FCT_API int fct_wfbp_gpu_plan_destroy(fct_wfbp_gpu_plan* plan);
FCT_API int fct_wfbp_gpu_plan_info(const fct_wfbp_gpu_plan* plan, fct_plan_info* info);
/* Moves the plan to another stream after its queued work on the current one
 * (plan set-up included) has completed. */
FCT_API int fct_wfbp_gpu_set_stream(fct_wfbp_gpu_plan* plan, void* stream);

/* Rebin + ramp filter (ReconstructionMachine.cu:141-306, cuda_kernels.cuh:11-77).
 * raw: device [view][row][ch] float32 fan data.  filtered: device buffer of
 * filtered_bytes.  FP32 stores the filtered rows y[view][row][ch]; FP16/BF16
 * store, per (view, row r, channel c), the quad (y[r][c], y[r][c+1],
 * y[r+1][c], y[r+1][c+1]) / s (0 beyond the detector) with one power-of-two s
 * per scan that keeps |y| / s < 2^15, and write s to scale[0].  scale: device
 * buffer of scale_elems floats (NULL for FP32).  Asynchronous on the plan's
 * stream. */
FCT_API int fct_wfbp_gpu_rebin_filter(fct_wfbp_gpu_plan* plan, const float* raw,
                                      void* filtered, float* scale);

/* Helical-weighted backprojection into native slices
 * (ReconstructionMachine.cu:308-385, cuda_kernels.cuh:79-93, 198-287). */
FCT_API int fct_wfbp_gpu_backproject(fct_wfbp_gpu_plan* plan, const void* filtered,
                                     const float* scale, float* native);

/* Triangular slice thickening (ReconstructionMachine.cu:387-434,
 * cuda_kernels.cuh:289-320). */
FCT_API int fct_wfbp_gpu_thicken(fct_wfbp_gpu_plan* plan, const float* native, float* volume);

/* Units of the volume the thickening (and every whole-chain entry point)
 * writes: FCT_UNITS_MU (default, as the reference: no HU conversion
 * anywhere, ReconstructionMachine.cu:436-455) or FCT_UNITS_HU with
 * mu_water in 1/mm, fused into the thickening kernel. */
FCT_API int fct_wfbp_gpu_set_output_units(fct_wfbp_gpu_plan* plan, int32_t units, double mu_water);

/* Whole chain (RunReconstruction, ReconstructionMachine.cu:76-82).  raw and
 * volume are device pointers unless host_buffers != 0, in which case they are
 * host pointers, the scan is streamed to the device in chunks of views as in
 * fct_wfbp_gpu_reconstruct_stream (copies on the plan's copy stream,
 * overlapped with rebin+filter and backprojection of what has arrived; a
 * pinned raw array, e.g. from hipHostMalloc, makes the copies asynchronous
 * DMA) and the call returns after the volume is copied back.  The plan owns
 * the intermediate buffers. */
FCT_API int fct_wfbp_gpu_reconstruct(fct_wfbp_gpu_plan* plan, const float* raw, float* volume,
                                     int32_t host_buffers);

/* Whole chain on device buffers (raw [view][row][ch], volume
 * [slice][y][x]) that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as fct_wfbp_gpu_reconstruct, with the filtered projections, scale
 * and native volume in a fixed 256 MiB device array that each GPU holds for
 * the life of the process.  FCT_ERR_LIMIT when the plan's filtered + native
 * buffers exceed it (e.g. 888 x 64 x 9840 views); calls that overlap on one
 * GPU race (they share the array). */
FCT_API int fct_wfbp_gpu_reconstruct_noalloc(fct_wfbp_gpu_plan* plan, const float* raw,
                                             float* volume);

/* Source of raw views for the streaming chain: writes views [v0, v0 + n) of
 * the scan, [view][row][ch] float32, to dst (pinned host memory owned by the
 * plan) and returns 0, or nonzero to abort.  Called in view order on the
 * thread that called fct_wfbp_gpu_reconstruct_stream, while the previous
 * chunk is being copied and processed, for the views the volume's slices
 * read only (the leading and trailing views of a scan longer than the
 * volume needs are not requested).  It takes the place of
 * RawDataSet::copyProjection (RawDataSet.h:136-139) as AllocateKeyArrays
 * uses it (ReconstructionMachine.cu:88-93). */
typedef int (*fct_view_reader)(void* user, int32_t v0, int32_t n, float* dst);

/* Streaming whole chain: RunReconstruction with the reference's host copy of
 * every frame, its single-threaded flip and its synchronous per-row copies
 * (ReconstructionMachine.cu:88-93, 150-162, 246) replaced by chunked,
 * double-buffered pinned staging and asynchronous H2D on a copy stream.  On
 * the plan's stream the parallel views whose raw views have all arrived are
 * rebinned and filtered, and groups of slabs whose views are all filtered are
 * backprojected, while later chunks are read and copied.  views_per_chunk: 0
 * for 256.  max_abs_raw: for FP16/BF16 plans a bound >= max|raw| fixes the
 * storage scale up front (an underestimate overflows the 16-bit values); 0
 * uses the running abs-max of the views so far and rescales the stored
 * projections by a power of two when it grows.  The volume (same values as
 * fct_wfbp_gpu_reconstruct) is copied to volume_host [slice][y][x] before the
 * call returns. */
FCT_API int fct_wfbp_gpu_reconstruct_stream(fct_wfbp_gpu_plan* plan, fct_view_reader read,
                                            void* user, int32_t views_per_chunk,
                                            float max_abs_raw, float* volume_host);

/* Adjoint (no reference counterpart): the exact transpose of the FP32 chain
 * rebin -> filter -> backproject, i.e. a matched helical forward projector.
 * Buffers are device pointers in the forward layouts; outputs are
 * overwritten.  Stage transposes:
 *   backproject_adjoint: native [slice][y][x] -> FP32 filtered [view][row][channel_pitch]
 *   filter:              FP32 filtered rows -> filtered rows (self-adjoint; in place allowed)
 *   rebin_adjoint:       FP32 filtered rows -> raw [view][row][ch]
 *   thicken_adjoint:     volume [slice][y][x] -> native [slice][y][x]
 * adjoint runs the first three with plan-owned scratch. */
FCT_API int fct_wfbp_gpu_backproject_adjoint(fct_wfbp_gpu_plan* plan, const float* native_bar,
                                             float* filtered_bar);
FCT_API int fct_wfbp_gpu_filter(fct_wfbp_gpu_plan* plan, const float* lines_in, float* lines_out);
/* Rebinning alone (rebin_kernel, cuda_kernels.cuh:11-39, without the filter):
 * raw fan data -> FP32 parallel rows [view][row][channel_pitch].  With
 * fct_wfbp_gpu_backproject on an FP32 plan it forms the unfiltered chain
 * B R whose transpose R^T B^T (fct_wfbp_gpu_backproject_adjoint then
 * fct_wfbp_gpu_rebin_adjoint) is the matched helical projector. */
FCT_API int fct_wfbp_gpu_rebin(fct_wfbp_gpu_plan* plan, const float* raw, float* lines_out);
/* Matched helical projector pair (no reference counterpart): the
 * voxel-driven projection of a native volume onto FP32 parallel rows, every
 * view that sees a voxel contributing its bilinear footprint with weight
 * (2 pi / ppr) R / l (l the source-voxel distance in the plane, correcting
 * the footprint's depth), and its exact transpose, the plain R / l-weighted
 * sum over the same samples (the W-weighted backprojection without the
 * weights' normalisation or the view window).  With fct_wfbp_gpu_rebin_adjoint
 * / fct_wfbp_gpu_rebin this projects onto the fan data; up to one scale it
 * approximates line integrals.  lines is zero-filled first. */
FCT_API int fct_wfbp_gpu_matched_project(fct_wfbp_gpu_plan* plan, const float* native,
                                         float* lines);
FCT_API int fct_wfbp_gpu_matched_backproject(fct_wfbp_gpu_plan* plan, const float* lines,
                                             float* native);
FCT_API int fct_wfbp_gpu_rebin_adjoint(fct_wfbp_gpu_plan* plan, const float* filtered_bar,
                                       float* raw_bar);
/* Ray-driven projector (no reference counterpart; the forward model of the
 * data-consistency refinement): line integrals of the native volume
 * [slice][y][x] along every raw ray (the focal spot of projection v, with
 * its FFS deflection, to its detector cell) by Joseph's method -- at each
 * voxel plane of the in-plane axis the ray is most aligned with, the volume
 * interpolated bilinearly in the other in-plane axis and z, zero outside --
 * into raw [view][row][ch] (overwritten); and its exact transpose from raw
 * rays into the native volume (overwritten). */
FCT_API int fct_wfbp_gpu_ray_project(fct_wfbp_gpu_plan* plan, const float* native, float* raw);
FCT_API int fct_wfbp_gpu_ray_backproject(fct_wfbp_gpu_plan* plan, const float* raw,
                                         float* native);
FCT_API int fct_wfbp_gpu_thicken_adjoint(fct_wfbp_gpu_plan* plan, const float* volume_bar,
                                         float* native_bar);
FCT_API int fct_wfbp_gpu_adjoint(fct_wfbp_gpu_plan* plan, const float* native_bar, float* raw_bar);

/* Exact count of selected (voxel, view) samples, the throughput unit.
 * per_voxel (optional, device, native_elems int32) receives the count of
 * every native voxel; *total the sum. Synchronous. */
FCT_API int fct_wfbp_gpu_count_voxel_views(fct_wfbp_gpu_plan* plan, int32_t* per_voxel,
                                           int64_t* total);

/* Selection of one native voxel for every parallel view: qf*65536+pf or -1
 * (device buffer of n_views int32).  Synchronous.  Test hook for bit-exact
 * index selection. */
FCT_API int fct_wfbp_gpu_voxel_selection(fct_wfbp_gpu_plan* plan, int32_t ix, int32_t iy,
                                         int32_t islice, int32_t* sel);

FCT_API int fct_wfbp_gpu_synchronize(fct_wfbp_gpu_plan* plan);

/* Seeded synthetic scan (no reference counterpart): analytic FP64 line
 * integrals of a water cylinder plus ellipsoids through the fan geometry, and,
 * when i0 > 0, Poisson noise N ~ P(i0 exp(-p)) with p = ln(i0 / max(N, 1/2)).
 * ellipsoids is a HOST array.  raw: device [view][row][ch] float32.  Views
 * [v0, v1) are written at raw[(v - v0)]. */
FCT_API int fct_wfbp_gpu_phantom_project(fct_wfbp_gpu_ctx* ctx, const fct_scan_geometry* scan,
                                         double cyl_radius, double cyl_mu,
                                         const fct_ellipsoid* ellipsoids, int32_t n_ellipsoids,
                                         double i0, uint64_t seed, int32_t v0, int32_t v1,
                                         float* raw, void* stream);

/* ---- Batched reconstruction of uneven jobs --------------------------------
 * No counterpart in the reference: main.cpp:59-66 runs one
 * ReconstructionMachine (ReconstructionMachine.h:16-21) per process, one
 * volume at a time.  A batch owns one plan per job (same device, same
 * precision and tile_z, any geometry / matrix / slice range / filter) plus
 * every intermediate buffer.  One run is: rebin_filter of every job, ONE
 * persistent backprojection launch over the tiles of all jobs (work queue per
 * XCD; a batch of one job launches the single-plan grid kernel instead),
 * thicken of every job -- recorded once into a hipGraph and replayed.
 * Each job's volume is bitwise identical to fct_wfbp_gpu_reconstruct on a
 * plan with the same options. */
typedef struct fct_wfbp_gpu_batch fct_wfbp_gpu_batch;

typedef struct {
  int32_t n_jobs;
  int32_t tile_z;          /* current slices per wave tile (set by autotune) */
  int32_t total_tiles;     /* backprojection tiles over all jobs */
  int32_t grid;            /* persistent workgroups of the backprojection */
  int64_t device_bytes;    /* filtered + scale + native buffers held by the batch */
} fct_batch_info;

typedef enum {
  FCT_BATCH_RUN_GRAPH = 0,   /* replay the recorded hipGraph (record on first use) */
  FCT_BATCH_RUN_EAGER = 1    /* launch the same kernels directly */
} fct_batch_run_mode;

/* opts: precision and tile_z for every job (bp_variant is ignored: the batch
 * uses the lean kernel).  stream: NULL creates a stream owned by the batch
 * (hipGraph capture needs one).  At most 64 jobs. */
FCT_API int fct_wfbp_gpu_batch_create(fct_wfbp_gpu_ctx* ctx, int32_t n_jobs,
                                      const fct_scan_geometry* scans,
                                      const fct_recon_config* recons,
                                      const fct_exec_options* opts, void* stream,
                                      fct_wfbp_gpu_batch** batch);
FCT_API int fct_wfbp_gpu_batch_destroy(fct_wfbp_gpu_batch* batch);
/* Bind job's device input raw [view][row][ch] and output volume
 * [slice_req][y][x].  Rebinding invalidates the recorded graph. */
FCT_API int fct_wfbp_gpu_batch_submit(fct_wfbp_gpu_batch* batch, int32_t job, const float* raw,
                                      float* volume);
/* Enqueue one run on the batch stream (asynchronous). */
FCT_API int fct_wfbp_gpu_batch_run(fct_wfbp_gpu_batch* batch, int32_t mode);
FCT_API int fct_wfbp_gpu_batch_synchronize(fct_wfbp_gpu_batch* batch);
/* Enqueue only the backprojection of every job (eager), reading
 * the filtered projections the last run left in the batch's buffers and
 * rewriting the native volumes.  For timing the kernel on the batch stream. */
FCT_API int fct_wfbp_gpu_batch_backproject(fct_wfbp_gpu_batch* batch);
/* Autotuner: runs rebin_filter of every job once, then times the
 * backprojection `reps` times for each candidate tile_z (HIP events on the
 * batch stream), keeps the fastest and returns each candidate's mean ms in
 * ms_out (host, n entries, may be NULL).  Synchronous; needs every job bound. */
FCT_API int fct_wfbp_gpu_batch_autotune(fct_wfbp_gpu_batch* batch, const int32_t* tile_z,
                                        int32_t n, int32_t reps, float* ms_out);
FCT_API int fct_wfbp_gpu_batch_info(const fct_wfbp_gpu_batch* batch, fct_batch_info* info);
/* The plan of one job (owned by the batch), for plan_info / counting. */
FCT_API int fct_wfbp_gpu_batch_job_plan(fct_wfbp_gpu_batch* batch, int32_t job,
                                        fct_wfbp_gpu_plan** plan);

#ifdef __cplusplus
}
#endif
#endif /* FCT_WFBP_GPU_H */
