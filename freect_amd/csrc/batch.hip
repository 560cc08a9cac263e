// batch.hip -- many uneven reconstructions per run: one plan per job, one
// persistent backprojection launch over the tiles of every job, the whole run
// recorded once into a hipGraph, and an autotuner for the tile depth.
//
// No counterpart in the reference (main.cpp:59-66 reconstructs one volume per
// process).  Per job the stages are the single-plan ones
// (fct_wfbp_gpu_rebin_filter / _backproject / _thicken); only the
// backprojection is merged, because it is the stage whose per-job launches
// leave the GPU idle in their tails (a job's last wave of tiles) and whose
// cost differs most between jobs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <new>
#include <vector>

#include "wfbp_internal.h"

struct fct_wfbp_gpu_batch {
  fct_wfbp_gpu_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  fct_exec_options opts{};
  std::vector<fct_wfbp_gpu_plan*> plans;
  std::vector<const float*> raw;     // bound inputs
  std::vector<float*> volume;        // bound outputs
  fct::BpArgs* d_jobs = nullptr;     // [n_jobs]
  int32_t* d_tile_start = nullptr;   // [n_jobs + 1]
  uint32_t* d_queue = nullptr;       // [8]
  int total_tiles = 0;
  int grid = 0;
  int64_t device_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void set_err(fct_wfbp_gpu_ctx* ctx, const char* msg, hipError_t e = hipSuccess) {
  if (!ctx) return;
  char buf[512];
  if (e != hipSuccess)
    snprintf(buf, sizeof buf, "%s: %s", msg, hipGetErrorString(e));
  else
    snprintf(buf, sizeof buf, "%s", msg);
  ctx->last_error = buf;
}

int hip_status(fct_wfbp_gpu_ctx* ctx, const char* what, hipError_t e) {
  if (e == hipSuccess) return FCT_OK;
  set_err(ctx, what, e);
  return e == hipErrorOutOfMemory ? FCT_ERR_OUT_OF_MEMORY : FCT_ERR_DEVICE;
}

#define BATCH_TRY(b, expr)                                       \
  do {                                                           \
    const int st_ = hip_status((b)->ctx, #expr, (expr));         \
    if (st_ != FCT_OK) return st_;                               \
  } while (0)

void drop_graph(fct_wfbp_gpu_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Job records and the global tile list for the plans' current tile_z.
int upload_jobs(fct_wfbp_gpu_batch* b) {
  const int n = (int)b->plans.size();
  std::vector<fct::BpArgs> jobs(n);
  std::vector<int32_t> start(n + 1, 0);
  long long total = 0;
  for (int j = 0; j < n; ++j) {
    fct_wfbp_gpu_plan* p = b->plans[j];
    jobs[j] = fct::plan_bp_args(p);
    jobs[j].filt = p->d_filtered;
    jobs[j].filt_scale = p->d_scale;
    jobs[j].native = p->d_native;
    start[j] = (int32_t)total;
    total += (long long)jobs[j].tiles_x * jobs[j].tiles_y * jobs[j].nslabs;
  }
  if (total >= (1LL << 30)) return set_err(b->ctx, "too many tiles in one batch"), FCT_ERR_LIMIT;
  start[n] = (int32_t)total;
  b->total_tiles = (int)total;
  BATCH_TRY(b, hipMemcpyAsync(b->d_jobs, jobs.data(), n * sizeof(fct::BpArgs),
                              hipMemcpyHostToDevice, b->stream));
  BATCH_TRY(b, hipMemcpyAsync(b->d_tile_start, start.data(), (n + 1) * sizeof(int32_t),
                              hipMemcpyHostToDevice, b->stream));
  BATCH_TRY(b, hipStreamSynchronize(b->stream));
  int per_cu = 0;
  BATCH_TRY(b, fct::backproject_batch_waves_per_cu(b->plans[0]->tile_z, b->opts.precision,
                                                    &per_cu));
  const int ncu = std::max(b->plans[0]->ncu, 1);
  b->grid = std::max(8, (std::max(per_cu, 1) * ncu) / 8 * 8);
  return FCT_OK;
}

bool all_centered(const fct_wfbp_gpu_batch* b) {
  for (const fct_wfbp_gpu_plan* p : b->plans)
    if (p->g.cr != p->g.hr) return false;
  return true;
}

// The persistent backprojection of every job (queue counters cleared by a
// kernel, so the sequence can be recorded into a graph).  A batch of one job
// has no tiles of other jobs to balance: it launches the single-plan grid
// kernel, which runs the same tile code (the volume is bitwise the same) and
// has no queue to walk.
int enqueue_backproject(fct_wfbp_gpu_batch* b) {
  if (b->plans.size() == 1) {
    fct_wfbp_gpu_plan* p = b->plans[0];
    fct::BpArgs a = fct::plan_bp_args(p);
    a.filt = p->d_filtered;
    a.filt_scale = p->d_scale;
    a.native = p->d_native;
    BATCH_TRY(b, fct::launch_backproject_v1(a, p->tile_z, b->opts.precision, fct::kBpAccumulate,
                                            b->stream));
    return FCT_OK;
  }
  BATCH_TRY(b, fct::launch_fill(reinterpret_cast<float*>(b->d_queue), 8, 0.0f, b->stream));
  fct::BpBatchArgs a{};
  a.jobs = b->d_jobs;
  a.tile_start = b->d_tile_start;
  a.queue = b->d_queue;
  a.n_jobs = (int32_t)b->plans.size();
  a.total_tiles = b->total_tiles;
  BATCH_TRY(b, fct::launch_backproject_batch(a, b->plans[0]->tile_z, b->opts.precision,
                                             all_centered(b), b->grid, b->stream));
  return FCT_OK;
}

int enqueue_rebin_filter(fct_wfbp_gpu_batch* b) {
  for (size_t j = 0; j < b->plans.size(); ++j) {
    fct_wfbp_gpu_plan* p = b->plans[j];
    const int st = fct::chain_rebin_filter(p, b->raw[j], p->d_filtered, p->d_scale);
    if (st != FCT_OK) return st;
  }
  return FCT_OK;
}

int enqueue_run(fct_wfbp_gpu_batch* b) {
  int st = enqueue_rebin_filter(b);
  if (st != FCT_OK) return st;
  st = enqueue_backproject(b);
  if (st != FCT_OK) return st;
  for (size_t j = 0; j < b->plans.size(); ++j) {
    fct_wfbp_gpu_plan* p = b->plans[j];
    st = fct_wfbp_gpu_thicken(p, p->d_native, b->volume[j]);
    if (st != FCT_OK) return st;
  }
  return FCT_OK;
}

bool all_bound(const fct_wfbp_gpu_batch* b) {
  for (size_t j = 0; j < b->plans.size(); ++j)
    if (!b->raw[j] || !b->volume[j]) return false;
  return true;
}

void destroy(fct_wfbp_gpu_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  for (fct_wfbp_gpu_plan* p : b->plans)
    if (p) fct_wfbp_gpu_plan_destroy(p);
  if (b->d_jobs) (void)hipFree(b->d_jobs);
  if (b->d_tile_start) (void)hipFree(b->d_tile_start);
  if (b->d_queue) (void)hipFree(b->d_queue);
  if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

}  // namespace

extern "C" {

FCT_API int fct_wfbp_gpu_batch_create(fct_wfbp_gpu_ctx* ctx, int32_t n_jobs,
                                      const fct_scan_geometry* scans,
                                      const fct_recon_config* recons,
                                      const fct_exec_options* opts, void* stream,
                                      fct_wfbp_gpu_batch** out) {
  if (!ctx || !scans || !recons || !out || n_jobs < 1) return FCT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  if (n_jobs > 64) return set_err(ctx, "at most 64 jobs per batch"), FCT_ERR_LIMIT;
  const hipError_t e0 = hipSetDevice(ctx->device);
  if (e0 != hipSuccess) return hip_status(ctx, "hipSetDevice", e0);
  auto* b = new (std::nothrow) fct_wfbp_gpu_batch();
  if (!b) return FCT_ERR_OUT_OF_MEMORY;
  b->ctx = ctx;
  b->device = ctx->device;
  if (opts) b->opts = *opts;
  b->opts.bp_variant = FCT_BP_LEAN;
  if (b->opts.tile_z == 0) b->opts.tile_z = 8;
  auto fail = [&](int st) {
    destroy(b);
    return st;
  };
  if (stream) {
    b->stream = static_cast<hipStream_t>(stream);
  } else {
    const hipError_t e = hipStreamCreate(&b->stream);
    if (e != hipSuccess) return fail(hip_status(ctx, "hipStreamCreate", e));
    b->own_stream = true;
  }
  b->plans.assign(n_jobs, nullptr);
  b->raw.assign(n_jobs, nullptr);
  b->volume.assign(n_jobs, nullptr);
  for (int j = 0; j < n_jobs; ++j) {
    int st = fct_wfbp_gpu_plan_create(ctx, &scans[j], &recons[j], &b->opts, b->stream,
                                      &b->plans[j]);
    if (st != FCT_OK) return fail(st);
    fct_wfbp_gpu_plan* p = b->plans[j];
    const size_t nat = (size_t)p->n_native * p->recon.ny * p->recon.nx * sizeof(float);
    hipError_t e = hipMalloc(&p->d_filtered, p->filtered_bytes());
    if (e == hipSuccess && p->scale_elems())
      e = hipMalloc(reinterpret_cast<void**>(&p->d_scale), p->scale_elems() * 4);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->d_native), nat);
    if (e != hipSuccess) return fail(hip_status(ctx, "batch buffers", e));
    b->device_bytes += (int64_t)(p->filtered_bytes() + p->scale_elems() * 4 + nat);
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b->d_jobs), n_jobs * sizeof(fct::BpArgs));
  if (e == hipSuccess)
    e = hipMalloc(reinterpret_cast<void**>(&b->d_tile_start), (n_jobs + 1) * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&b->d_queue), 8 * sizeof(uint32_t));
  if (e != hipSuccess) return fail(hip_status(ctx, "batch tables", e));
  const int st = upload_jobs(b);
  if (st != FCT_OK) return fail(st);
  *out = b;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_destroy(fct_wfbp_gpu_batch* b) {
  destroy(b);
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_submit(fct_wfbp_gpu_batch* b, int32_t job, const float* raw,
                                      float* volume) {
  if (!b || job < 0 || job >= (int32_t)b->plans.size() || !raw || !volume)
    return FCT_ERR_INVALID_ARGUMENT;
  if (b->raw[job] != raw || b->volume[job] != volume) drop_graph(b);
  b->raw[job] = raw;
  b->volume[job] = volume;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_run(fct_wfbp_gpu_batch* b, int32_t mode) {
  if (!b) return FCT_ERR_INVALID_ARGUMENT;
  if (!all_bound(b)) return set_err(b->ctx, "every job needs raw and volume bound"), FCT_ERR_INVALID_ARGUMENT;
  BATCH_TRY(b, hipSetDevice(b->device));
  if (mode == FCT_BATCH_RUN_EAGER) return enqueue_run(b);
  if (mode != FCT_BATCH_RUN_GRAPH) return FCT_ERR_INVALID_ARGUMENT;
  if (!b->exec) {
    BATCH_TRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int st = enqueue_run(b);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b->stream, &g);
    if (st != FCT_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    BATCH_TRY(b, ee);
    b->graph = g;
    const hipError_t ei = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      drop_graph(b);
      return hip_status(b->ctx, "hipGraphInstantiate", ei);
    }
  }
  BATCH_TRY(b, hipGraphLaunch(b->exec, b->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_backproject(fct_wfbp_gpu_batch* b) {
  if (!b) return FCT_ERR_INVALID_ARGUMENT;
  BATCH_TRY(b, hipSetDevice(b->device));
  return enqueue_backproject(b);
}

FCT_API int fct_wfbp_gpu_batch_synchronize(fct_wfbp_gpu_batch* b) {
  if (!b) return FCT_ERR_INVALID_ARGUMENT;
  BATCH_TRY(b, hipSetDevice(b->device));
  BATCH_TRY(b, hipStreamSynchronize(b->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_autotune(fct_wfbp_gpu_batch* b, const int32_t* tile_z, int32_t n,
                                        int32_t reps, float* ms_out) {
  if (!b || !tile_z || n < 1 || reps < 1) return FCT_ERR_INVALID_ARGUMENT;
  if (!all_bound(b)) return set_err(b->ctx, "every job needs raw and volume bound"), FCT_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n; ++i)
    if (tile_z[i] != 4 && tile_z[i] != 8 && tile_z[i] != 16 && tile_z[i] != 32)
      return set_err(b->ctx, "tile_z candidates must be 4, 8, 16 or 32"), FCT_ERR_INVALID_ARGUMENT;
  BATCH_TRY(b, hipSetDevice(b->device));
  int st = enqueue_rebin_filter(b);
  if (st != FCT_OK) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  BATCH_TRY(b, hipEventCreate(&e0));
  BATCH_TRY(b, hipEventCreate(&e1));
  float best_ms = 0.0f;
  int best = -1;
  for (int i = 0; i < n && st == FCT_OK; ++i) {
    for (fct_wfbp_gpu_plan* p : b->plans)
      if ((st = fct::plan_set_tile_z(p, tile_z[i])) != FCT_OK) break;
    if (st == FCT_OK) st = upload_jobs(b);
    if (st == FCT_OK) st = enqueue_backproject(b);   // warm-up
    if (st != FCT_OK) break;
    (void)hipEventRecord(e0, b->stream);
    for (int r = 0; r < reps && st == FCT_OK; ++r) st = enqueue_backproject(b);
    (void)hipEventRecord(e1, b->stream);
    if (st != FCT_OK) break;
    float ms = 0.0f;
    st = hip_status(b->ctx, "autotune timing", hipEventSynchronize(e1));
    if (st == FCT_OK) st = hip_status(b->ctx, "autotune timing", hipEventElapsedTime(&ms, e0, e1));
    ms /= (float)reps;
    if (ms_out) ms_out[i] = ms;
    if (best < 0 || ms < best_ms) {
      best_ms = ms;
      best = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  drop_graph(b);
  if (st != FCT_OK) return st;
  for (fct_wfbp_gpu_plan* p : b->plans)
    if ((st = fct::plan_set_tile_z(p, tile_z[best])) != FCT_OK) return st;
  b->opts.tile_z = tile_z[best];
  for (fct_wfbp_gpu_plan* p : b->plans) p->opts.tile_z = tile_z[best];
  return upload_jobs(b);
}

FCT_API int fct_wfbp_gpu_batch_info(const fct_wfbp_gpu_batch* b, fct_batch_info* info) {
  if (!b || !info) return FCT_ERR_INVALID_ARGUMENT;
  info->n_jobs = (int32_t)b->plans.size();
  info->tile_z = b->plans[0]->tile_z;
  info->total_tiles = b->total_tiles;
  info->grid = b->grid;
  info->device_bytes = b->device_bytes;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_batch_job_plan(fct_wfbp_gpu_batch* b, int32_t job,
                                        fct_wfbp_gpu_plan** plan) {
  if (!b || !plan || job < 0 || job >= (int32_t)b->plans.size()) return FCT_ERR_INVALID_ARGUMENT;
  *plan = b->plans[job];
  return FCT_OK;
}

}  // extern "C"
