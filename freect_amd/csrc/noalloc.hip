// noalloc.hip -- the whole chain without any device allocation, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid it).
//
// fct_wfbp_gpu_reconstruct keeps its filtered projections, FP16 / BF16
// scale and native volume in plan-owned buffers that it allocates on first
// use.  fct_wfbp_gpu_reconstruct_noalloc runs the same stage kernels with
// those three buffers carved out of one file-scope device array, `scratch`,
// at 256-byte aligned offsets.  Every GPU the library runs on holds its own
// copy of the array (kScratchBytes) until the process ends; its address on
// a device is looked up once with hipGetSymbolAddress while that device is
// current and kept per device.  Calls on one GPU share that array and
// overlapping calls on one GPU race; nothing in it has to start at zero (the
// rebinning writes every filtered entry, the backprojection every native
// voxel, and the abs-max slot of the scale is cleared by a kernel).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "wfbp_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)256 << 20;   // 256 MiB per GPU
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current): the
// host-side name is only a stand-in for the symbol, never a device address.
hipError_t scratch_on(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

int fail(fct_wfbp_gpu_plan* p, int code, const char* fmt, size_t a, size_t b) {
  char buf[256];
  std::snprintf(buf, sizeof buf, fmt, a, b);
  p->ctx->last_error = buf;
  return code;
}

}  // namespace

extern "C" {

FCT_API int fct_wfbp_gpu_reconstruct_noalloc(fct_wfbp_gpu_plan* p, const float* raw,
                                             float* volume) {
  if (!p || !raw || !volume) return FCT_ERR_INVALID_ARGUMENT;
  if (p->af_on) {
    p->ctx->last_error =
        "reconstruct_noalloc: adaptive filtration needs a raw-sized buffer (use "
        "fct_wfbp_gpu_reconstruct)";
    return FCT_ERR_UNSUPPORTED;
  }
  const size_t fbytes = align256(p->filtered_bytes());
  const size_t sbytes = align256(p->scale_elems() * sizeof(float));
  const size_t nbytes = align256((size_t)p->n_native * p->recon.ny * p->recon.nx * sizeof(float));
  if (fbytes + sbytes + nbytes > kScratchBytes)
    return fail(p, FCT_ERR_LIMIT,
                "reconstruct_noalloc needs %zu bytes of scratch; the allocation-free path holds "
                "%zu per GPU (use fct_wfbp_gpu_reconstruct)",
                fbytes + sbytes + nbytes, kScratchBytes);
  hipError_t e = hipSetDevice(p->device);
  void* base = nullptr;
  if (e == hipSuccess) e = scratch_on(p->device, &base);
  if (e != hipSuccess)
    return fail(p, FCT_ERR_DEVICE, "reconstruct_noalloc: scratch lookup failed (%zu, %zu)",
                (size_t)e, (size_t)p->device);
  char* b = static_cast<char*>(base);
  void* filtered = b;
  float* scale = sbytes ? reinterpret_cast<float*>(b + fbytes) : nullptr;
  float* native = reinterpret_cast<float*>(b + fbytes + sbytes);
  int st = fct::chain_rebin_filter(p, raw, filtered, scale);
  if (st) return st;
  st = fct_wfbp_gpu_backproject(p, filtered, scale, native);
  if (st) return st;
  return fct_wfbp_gpu_thicken(p, native, volume);
}

}  // extern "C"
