// capi.hip -- host side of the C ABI: device context, plan setup (geometry,
// FP64 tables, filter spectrum) and the stage entry points.
//
// Setup follows ConfigureCTGeometry / AllocateKeyArrays / generate_filter
// (ReconstructionMachine.cu:44-139, 478-544) and GPUPrecompute
// (recon_structs.h:69-83), with three changes of form: view angles, table
// positions and slice positions are computed in FP64 and handed to the
// kernels as FP32 values relative to a per-chunk FP64 reference; the filter
// spectrum is computed in FP64 on the host; every failure is a status code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "wfbp_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

void set_err(fct_wfbp_gpu_ctx* ctx, const char* fmt, ...) {
  if (!ctx) return;
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  ctx->last_error = buf;
}

#define HIP_TRY(ctx, expr)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      set_err(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));                \
      return e_ == hipErrorOutOfMemory ? FCT_ERR_OUT_OF_MEMORY : FCT_ERR_DEVICE;  \
    }                                                                             \
  } while (0)

template <typename T>
hipError_t upload(T** dst, const std::vector<T>& src, hipStream_t s) {
  if (src.empty()) {
    *dst = nullptr;
    return hipSuccess;
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), src.size() * sizeof(T));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

void fft64(std::vector<std::complex<double>>& a) {
  const int n = (int)a.size();
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int i = 0; i < n; i += len)
      for (int k = 0; k < len / 2; ++k) {
        const double ang = -2.0 * kPi * (double)k / (double)len;
        const std::complex<double> w(std::cos(ang), std::sin(ang));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// Spatial filter taps h[i + n/2], i in [-n/2, n/2).  kind 0: the generalised
// Hamming ramp of generate_filter (ReconstructionMachine.cu:493-506, ramp.m);
// kind 1: the Shepp-Logan kernel 2 / (pi^2 ds (1 - 4 i^2)).
std::vector<double> filter_taps(int n, double ds, int kind, double c, double a) {
  auto r = [](double t) {
    if (t == 0.0) return 0.5;
    return std::sin(t) / t + (std::cos(t) - 1.0) / (t * t);
  };
  std::vector<double> h(n);
  for (int i = -n / 2; i < n / 2; ++i) {
    double v;
    if (kind == 1) {
      v = 2.0 / (kPi * kPi * ds * (1.0 - 4.0 * (double)i * (double)i));
    } else {
      v = (c * c / (2.0 * ds)) * (a * r(c * kPi * i) + ((1.0 - a) / 2.0) * r(kPi * c * i + kPi) +
                                   ((1.0 - a) / 2.0) * r(kPi * c * i - kPi));
    }
    h[i + n / 2] = v;
  }
  return h;
}

int validate(fct_wfbp_gpu_ctx* ctx, const fct_scan_geometry* s, const fct_recon_config* r,
             const fct_exec_options* o) {
  if (!s || !r) return set_err(ctx, "null geometry or config"), FCT_ERR_INVALID_ARGUMENT;
  if (s->n_channels < 2 || s->n_rows < 2 || s->n_views < 1)
    return set_err(ctx, "detector needs >= 2 channels, >= 2 rows and >= 1 view"),
           FCT_ERR_INVALID_ARGUMENT;
  if (s->n_channels > 2048)
    return set_err(ctx, "n_channels %d exceeds 2048 (filter length limit)", s->n_channels),
           FCT_ERR_LIMIT;
  if (s->views_per_rotation < 2 || (s->views_per_rotation & 1))
    return set_err(ctx, "views_per_rotation must be even and >= 2 (got %d)",
                   s->views_per_rotation),
           FCT_ERR_UNSUPPORTED;
  if (!(s->dsd > s->r) || !(s->r > 0) || !(s->channel_spacing > 0) || !(s->row_spacing > 0))
    return set_err(ctx, "invalid distances or spacings"), FCT_ERR_INVALID_ARGUMENT;
  if (!(std::fabs(s->feed) > 0))
    return set_err(ctx, "table feed must be nonzero (helical scans only)"), FCT_ERR_UNSUPPORTED;
  if (s->ffs_mode < FCT_FFS_NONE || s->ffs_mode > FCT_FFS_PHI_Z)
    return set_err(ctx, "unknown ffs_mode %d", s->ffs_mode), FCT_ERR_INVALID_ARGUMENT;
  if (s->rotation < -1 || s->rotation > 1)
    return set_err(ctx, "rotation must be +1 (or 0) or -1 (got %d)", s->rotation),
           FCT_ERR_INVALID_ARGUMENT;
  if (s->ffs_mode != FCT_FFS_NONE) {
    const int S = fct_ffs::n_states(s->ffs_mode);
    if (s->views_per_rotation % (2 * S))
      return set_err(ctx, "with FFS, views_per_rotation must be a multiple of %d", 2 * S),
             FCT_ERR_UNSUPPORTED;
    if (s->n_views < S) return set_err(ctx, "fewer views than focal states"), FCT_ERR_INVALID_ARGUMENT;
    if (!(s->anode_angle > 0.0 && s->anode_angle < 1.5))
      return set_err(ctx, "anode_angle must be in (0, 1.5) rad"), FCT_ERR_INVALID_ARGUMENT;
    if (fct_ffs::has_phi(s->ffs_mode) && 2 * s->n_channels > 2048)
      return set_err(ctx, "interleaved channels %d exceed 2048", 2 * s->n_channels), FCT_ERR_LIMIT;
  }
  if (r->nx < 1 || r->ny < 1 || r->nx > 8192 || r->ny > 8192)
    return set_err(ctx, "nx, ny must be in [1, 8192]"), FCT_ERR_INVALID_ARGUMENT;
  if (!(r->recon_fov > 0) || !(r->slice_thickness > 0))
    return set_err(ctx, "recon_fov and slice_thickness must be > 0"), FCT_ERR_INVALID_ARGUMENT;
  if (r->filter_kind < 0 || r->filter_kind > 1)
    return set_err(ctx, "unknown filter_kind %d", r->filter_kind), FCT_ERR_INVALID_ARGUMENT;
  if (o && o->adaptive_filtration &&
      !(r->adaptive_filtration_s > 0.0 && r->adaptive_filtration_s <= 1.0))
    return set_err(ctx, "adaptive filtration needs adaptive_filtration_s in (0, 1] (got %g)",
                   r->adaptive_filtration_s),
           FCT_ERR_INVALID_ARGUMENT;
  if (o) {
    if (o->precision < 0 || o->precision > 2)
      return set_err(ctx, "unknown precision %d", o->precision), FCT_ERR_INVALID_ARGUMENT;
    if (o->tile_z != 0 && o->tile_z != 4 && o->tile_z != 8 && o->tile_z != 16 && o->tile_z != 32)
      return set_err(ctx, "tile_z must be 0, 4, 8, 16 or 32"), FCT_ERR_INVALID_ARGUMENT;
    if (o->bp_variant < FCT_BP_AUTO || o->bp_variant > FCT_BP_LEAN)
      return set_err(ctx, "unknown bp_variant %d", o->bp_variant), FCT_ERR_INVALID_ARGUMENT;
  }
  return FCT_OK;
}

void free_plan_buffers(fct_wfbp_gpu_plan* p) {
  void* bufs[] = {p->d_xs,       p->d_ys,        p->d_cs,        p->d_zs,       p->d_zt,
                  p->d_slab_vrange, p->d_rebin,  p->d_Hn,        p->d_tw,       p->d_thk_off,
                  p->d_thk_idx,  p->d_thk_w,     p->d_filtered,  p->d_scale, p->d_native,
                  p->d_raw_stage, p->d_vol_stage, p->d_gbar, p->d_thkT_off, p->d_thkT_idx,
                  p->d_thkT_w, p->d_wrange, p->d_zbounds, p->d_af_raw, p->d_af_vmax};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (int b = 0; b < 2; ++b)
    if (p->h_stage[b]) (void)hipHostFree(p->h_stage[b]);
  if (p->ev_compute) (void)hipEventDestroy(p->ev_compute);
  for (int b = 0; b < 2; ++b)
    if (p->ev_bp[b]) (void)hipEventDestroy(p->ev_bp[b]);
  for (hipEvent_t e : p->ev_chunk) (void)hipEventDestroy(e);
  if (p->ev_rebin) (void)hipEventDestroy(p->ev_rebin);
  if (p->ev_ratio) (void)hipEventDestroy(p->ev_ratio);
  if (p->h_ratio) (void)hipHostFree(p->h_ratio);
}

// Views a slab of tz slices can select: table z within one feed of the
// slab's slices (a superset of the view window), with a margin of views.
// False if one slab's views exceed the 4 GiB a buffer resource addresses.
bool slab_view_ranges(const fct_wfbp_gpu_plan* p, const std::vector<double>& zn, int tz,
                      std::vector<int2>* out) {
  const fct_scan_geometry& s = p->scan;
  const int nslabs = (p->n_native + tz - 1) / tz;
  std::vector<int2>& vr = *out;
  vr.assign(nslabs, make_int2(0, 0));
  for (int sl = 0; sl < nslabs; ++sl) {
    const int a0 = sl * tz, a1 = std::min(p->n_native, a0 + tz) - 1;
    const double zlo = std::min(zn[a0], zn[a1]) - std::fabs(s.feed);
    const double zhi = std::max(zn[a0], zn[a1]) + std::fabs(s.feed);
    double va = (zlo - s.z0) / p->dz, vb = (zhi - s.z0) / p->dz;
    if (va > vb) std::swap(va, vb);
    const double lo = std::floor(va) - 2, hi = std::ceil(vb) + 3;
    vr[sl].x = (int)std::max(0.0, std::min(lo, (double)p->nviews));
    vr[sl].y = (int)std::max(0.0, std::min(hi, (double)p->nviews));
    if ((size_t)(vr[sl].y - vr[sl].x) * p->nrows * p->cpitch * p->filtered_entry_bytes() >=
        ((size_t)1 << 32))
      return false;
  }
  return true;
}

std::vector<double> native_slice_z(const fct_wfbp_gpu_plan* p) {
  std::vector<double> zn(p->n_native);
  for (int i = 0; i < p->n_native; ++i)
    zn[i] = fct_tables::slice_z(p->recon.start_pos, p->recon.end_pos, p->csw, i);
  return zn;
}

}  // namespace

namespace fct {

int plan_set_tile_z(fct_wfbp_gpu_plan* p, int tz) {
  if (tz != 4 && tz != 8 && tz != 16 && tz != 32) return FCT_ERR_INVALID_ARGUMENT;
  if (tz == p->tile_z) return FCT_OK;
  std::vector<int2> vr;
  if (!slab_view_ranges(p, native_slice_z(p), tz, &vr))
    return set_err(p->ctx, "a slab's view range exceeds 4 GiB of filtered rows"), FCT_ERR_LIMIT;
  int2* d = nullptr;
  HIP_TRY(p->ctx, upload(&d, vr, p->stream));
  if (p->d_slab_vrange) (void)hipFree(p->d_slab_vrange);
  p->d_slab_vrange = d;
  p->h_slab_vrange = vr;
  p->tile_z = tz;
  p->nslabs = (int)vr.size();
  return FCT_OK;
}

BpArgs plan_bp_args(const fct_wfbp_gpu_plan* p) {
  BpArgs a{};
  a.xs = p->d_xs;
  a.ys = p->d_ys;
  a.cs = p->d_cs;
  a.zs = p->d_zs;
  a.zt = p->d_zt;
  a.slab_vrange = p->d_slab_vrange;
  a.wrange = p->d_wrange;
  a.zbounds = p->d_zbounds;
  a.g = p->g;
  a.scale = p->scale;
  a.nx = p->recon.nx;
  a.ny = p->recon.ny;
  a.n_native = p->n_native;
  a.nviews = p->nviews;
  a.H = p->H;
  a.cpitch = p->cpitch;
  a.tiles_x = (p->recon.nx + 7) / 8;
  a.tiles_y = (p->recon.ny + 7) / 8;
  a.nslabs = p->nslabs;
  return a;
}

}  // namespace fct

namespace {

fct::BpArgs bp_args(const fct_wfbp_gpu_plan* p) { return fct::plan_bp_args(p); }

// The ray-driven projector's geometry: the raw scan's rays and the native
// voxel grid in the plan's frame (a clockwise scan as its mirror image: tube
// angles -a_v counter-clockwise, y -> -y, like the voxel tables).
fct::RayProjArgs ray_args(const fct_wfbp_gpu_plan* p) {
  const fct_scan_geometry& s = p->raw_scan;
  const fct_recon_config& r = p->recon;
  fct::RayProjArgs a{};
  a.ncols = s.n_channels;
  a.nrows = s.n_rows;
  a.v0 = 0;
  a.v1 = s.n_views;
  a.nx = r.nx;
  a.ny = r.ny;
  a.nz = p->n_native;
  a.r = s.r;
  a.dsd = s.dsd;
  a.dgamma = s.channel_spacing / s.dsd;
  a.drow = s.row_spacing;
  a.cc = s.central_channel;
  a.cr = s.central_row;
  const bool cw = s.rotation < 0;
  a.angle0 = cw ? -s.angle0 : s.angle0;
  a.dtheta = 2.0 * kPi / (double)s.views_per_rotation;
  a.z0 = s.z0;
  a.dz = s.feed / (double)s.views_per_rotation;
  a.states = p->ffs.S;
  for (int rho = 0; rho < p->ffs.S; ++rho) {
    const fct_ffs::Deflection d = fct_ffs::state_offset(p->ffs.mode, rho, p->ffs.dr, p->ffs.da,
                                                        p->ffs.anode);
    a.ffs_off[rho][0] = d.dr;
    a.ffs_off[rho][1] = d.da;
    a.ffs_off[rho][2] = d.dz;
  }
  const double dx = r.recon_fov / r.nx, dy = r.recon_fov / r.ny;
  const double x0 = -0.5 * (r.nx - 1) * dx + r.x_origin, y0 = -0.5 * (r.ny - 1) * dy + r.y_origin;
  a.gx0 = (float)x0;
  a.gdx = (float)dx;
  a.gy0 = (float)(cw ? -y0 : y0);
  a.gdy = (float)(cw ? -dy : dy);
  const double z0 = fct_tables::slice_z(r.start_pos, r.end_pos, p->csw, 0);
  const double z1 = fct_tables::slice_z(r.start_pos, r.end_pos, p->csw, 1);
  a.gz0 = (float)z0;
  a.gdz = (float)(z1 - z0);
  return a;
}

// Thickening of requested slices [s0, s1) from the native slices.
fct::ThickenArgs thicken_args(const fct_wfbp_gpu_plan* p, const float* native, float* volume,
                              int s0, int s1) {
  fct::ThickenArgs a{};
  a.native = native;
  a.out = volume;
  a.off = p->d_thk_off;
  a.idx = p->d_thk_idx;
  a.w = p->d_thk_w;
  a.plane = p->recon.nx * p->recon.ny;
  a.s0 = s0;
  a.n_req = s1 - s0;
  a.hu = p->out_hu;
  a.hu_scale = p->out_hu ? (float)(1000.0 / p->mu_water) : 1.0f;
  return a;
}

// The rebinning gather's raw window (RebinFilterArgs).
void set_gather_window(const fct_wfbp_gpu_plan* p, fct::RebinFilterArgs* a) {
  a->group_bytes = (uint32_t)((int64_t)p->ffs.S * p->raw_scan.n_rows * p->raw_scan.n_channels *
                              (int64_t)sizeof(float));
  a->aoff_min = p->rebin_aoff_min;
  a->mmax_min = p->rebin_mmax_min;
}

// Rebin + filter of every line of the scan (line0 / nlines narrow it).
fct::RebinFilterArgs rf_args(const fct_wfbp_gpu_plan* p, const float* raw, void* out,
                             float* scale) {
  fct::RebinFilterArgs a{};
  a.raw = raw;
  a.out = out;
  a.scale = scale;
  a.hsum = (float)p->hsum;
  a.raw_elems = p->raw_elems();
  a.tab = p->d_rebin;
  a.Hn = p->d_Hn;
  a.tw = p->d_tw;
  a.nviews = p->nviews;
  a.nrows = p->nrows;
  a.ncols = p->ncols;
  a.cpitch = p->cpitch;
  a.nfft = p->nfft;
  a.log2n = p->log2n;
  a.precision = p->opts.precision;
  a.nlines = p->nviews * p->nrows;
  a.ncu = p->ncu;
  a.raw_nrows = p->raw_scan.n_rows;
  a.raw_ncols = p->raw_scan.n_channels;
  a.states = p->ffs.S;
  a.zshift = fct_ffs::has_z(p->ffs.mode) ? 1 : 0;
  set_gather_window(p, &a);
  return a;
}

// ---- Streaming whole chain ------------------------------------------------
// The reference copies every frame into one host array
// (ReconstructionMachine.cu:88-93), flips it on one thread (:150-162) and
// copies each row sheet synchronously before rebinning it (:246).  Here the
// views arrive in chunks: each chunk is copied to the device on the plan's
// copy stream (from the caller's array, or from pinned staging a view
// reader fills while the previous chunk is in flight), and on the plan's
// stream the parallel views whose raw views have all arrived are rebinned
// and filtered, and groups of slabs whose views are all filtered are
// backprojected, while later chunks are still being copied.  Each line pair
// and each tile is computed by the same code as in the one-shot path.
// Only the views some slab reads travel: the parallel views of the union of
// the slabs' view ranges and the raw views their rebinning reads (a scan
// usually extends past the volume by more than the cone reaches).  Slab
// groups are backprojected on the context's two backprojection streams in
// turn, each waiting for the rebinning of its last line, so that they
// overlap each other's last waves, the rebinning of later chunks and the
// copies.
// FP16 / BF16 storage needs a bound on max|raw| before the first line is
// stored: a caller's bound, or the running abs-max of the views so far.  In
// the latter case each chunk's scale update is read back (a 4-byte copy the
// host waits for while the next chunk is in flight) and only when the
// storage scale grows are the quads already stored rescaled, after the
// backprojections in flight (which read them) have finished.
struct ViewSource {
  const float* host;       // caller's scan, [view][row][ch]
  fct_view_reader read;    // or a reader filling pinned staging
  void* user;
};

constexpr int kStreamViews = 256;       // default views per chunk
constexpr int kStreamMinSlabs = 4;      // smallest backprojection group

// Parallel views [par.x, par.y) some slab reads, and raw views [raw.x, raw.y)
// the rebinning of those reads (rebin_sample: projections j + aoff and
// j + aoff + 1 of the entry's focal state, clamped to [0, mmax]).
void needed_views(const fct_wfbp_gpu_plan* p, int2* par, int2* raw) {
  int lo = p->nviews, hi = 0;
  for (const int2 v : p->h_slab_vrange)
    if (v.x < v.y) {
      lo = std::min(lo, v.x);
      hi = std::max(hi, v.y);
    }
  if (lo >= hi) {
    *par = make_int2(0, 0);
    *raw = make_int2(0, 0);
    return;
  }
  if (p->nrows & 1) lo &= ~1;   // line pairs start at an even line
  const int S = p->ffs.S;
  int rlo = p->raw_scan.n_views, rhi = 0;
  for (const fct::RebinDev& e : p->h_rebin) {
    if (e.state == 0) continue;
    const int rho = e.state - 1;
    const int ma = std::max(std::min(lo + e.aoff, e.mmax), 0);
    const int mb = std::max(std::min(hi - 1 + e.aoff + 1, e.mmax), 0);
    rlo = std::min(rlo, ma * S + rho);
    rhi = std::max(rhi, mb * S + rho + 1);
  }
  if (rlo >= rhi) rlo = rhi = 0;
  *par = make_int2(lo, hi);
  *raw = make_int2(rlo, std::min(rhi, p->raw_scan.n_views));
}

}  // namespace

namespace fct {

// Rebin + filter of the whole chain: the lines some slab reads only, the
// FP16 / BF16 bound from the raw views they read (the stage call
// fct_wfbp_gpu_rebin_filter filters every line of the scan).
int chain_rebin_filter(fct_wfbp_gpu_plan* p, const float* raw, void* filtered, float* scale) {
  int2 par, rawr;
  needed_views(p, &par, &rawr);
  const size_t per_view = (size_t)p->raw_scan.n_rows * p->raw_scan.n_channels;
  if (p->af_on) {   // adaptive filtration of the views read, into the plan's buffer
    HIP_TRY(p->ctx, fct::launch_adaptive_filter(raw, p->d_af_raw, p->d_af_vmax, rawr.x, rawr.y,
                                                p->raw_scan.n_rows, p->raw_scan.n_channels,
                                                p->af_s, p->stream));
    raw = p->d_af_raw;
  }
  fct::RebinFilterArgs a = rf_args(p, raw, filtered, scale);
  if (p->opts.precision != FCT_PRECISION_FP32) {
    HIP_TRY(p->ctx, fct::launch_scale_reset(scale, p->stream));
    HIP_TRY(p->ctx, fct::launch_absmax(raw + (size_t)rawr.x * per_view,
                                       (int64_t)(rawr.y - rawr.x) * (int64_t)per_view, scale,
                                       p->ncu, p->stream));
    // scale[0] also when no line is left to filter.
    HIP_TRY(p->ctx, fct::launch_scale_update(scale, (float)p->hsum, p->stream));
    a.skip_scale = 1;
  }
  a.line0 = par.x * p->nrows;
  a.nlines = (par.y - par.x) * p->nrows;
  HIP_TRY(p->ctx, fct::launch_rebin_filter(a, p->stream));
  return FCT_OK;
}

}  // namespace fct

namespace {

// Page-locked host memory (hipHostMalloc, or registered): copies from and to
// it are asynchronous DMA; from pageable memory they block the calling thread.
bool is_pinned(const void* ptr) {
  hipPointerAttribute_t at{};
  const hipError_t e = hipPointerGetAttributes(&at, ptr);
  (void)hipGetLastError();   // clear the error some runtimes report for pageable memory
  return e == hipSuccess && at.type == hipMemoryTypeHost;
}

int reconstruct_streaming(fct_wfbp_gpu_plan* p, const ViewSource& src, int vpc,
                          float max_abs_raw, float* volume_host) {
  fct_wfbp_gpu_ctx* ctx = p->ctx;
  if (vpc <= 0) vpc = kStreamViews;
  const int nv = p->raw_scan.n_views;
  vpc = std::min(vpc, nv);
  const size_t per_view = (size_t)p->raw_scan.n_rows * p->raw_scan.n_channels;
  const size_t plane = (size_t)p->recon.ny * p->recon.nx;
  const size_t nat_bytes = (size_t)p->n_native * plane * sizeof(float);
  const size_t vol_bytes = (size_t)p->n_req * plane * sizeof(float);
  int2 par, rawr;
  needed_views(p, &par, &rawr);
  const int nchunks = rawr.y > rawr.x ? (rawr.y - rawr.x + vpc - 1) / vpc : 0;
  if (!p->d_filtered) HIP_TRY(ctx, hipMalloc(&p->d_filtered, p->filtered_bytes()));
  if (p->scale_elems() && !p->d_scale)
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&p->d_scale), p->scale_elems() * sizeof(float)));
  if (!p->d_native) HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&p->d_native), nat_bytes));
  if (!p->d_raw_stage)
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&p->d_raw_stage),
                           (size_t)p->raw_elems() * sizeof(float)));
  if (!p->d_vol_stage) HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&p->d_vol_stage), vol_bytes));
  if (!ctx->copy_stream)
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b) {
    if (!ctx->bp_stream[b])
      HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->bp_stream[b], hipStreamNonBlocking));
    if (!p->ev_bp[b]) HIP_TRY(ctx, hipEventCreateWithFlags(&p->ev_bp[b], hipEventDisableTiming));
  }
  hipStream_t copy_stream = ctx->copy_stream;
  while ((int)p->ev_chunk.size() < nchunks) {
    hipEvent_t e = nullptr;
    HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
    p->ev_chunk.push_back(e);
  }
  if (!p->ev_compute) HIP_TRY(ctx, hipEventCreateWithFlags(&p->ev_compute, hipEventDisableTiming));
  if (!p->ev_rebin) HIP_TRY(ctx, hipEventCreateWithFlags(&p->ev_rebin, hipEventDisableTiming));
  if (!p->ev_ratio) HIP_TRY(ctx, hipEventCreateWithFlags(&p->ev_ratio, hipEventDisableTiming));
  if (!p->h_ratio)
    HIP_TRY(ctx, hipHostMalloc(reinterpret_cast<void**>(&p->h_ratio), sizeof(float),
                               hipHostMallocDefault));
  if (src.read && p->stage_views != vpc) {
    for (int b = 0; b < 2; ++b) {
      if (p->h_stage[b]) HIP_TRY(ctx, hipHostFree(p->h_stage[b]));
      p->h_stage[b] = nullptr;
    }
    p->stage_views = 0;
    for (int b = 0; b < 2; ++b)
      HIP_TRY(ctx, hipHostMalloc(reinterpret_cast<void**>(&p->h_stage[b]),
                                 (size_t)vpc * per_view * sizeof(float), hipHostMallocDefault));
    p->stage_views = vpc;
  }
  // The copies overwrite the raw buffer, and the backprojections the native
  // slices, only after the plan's earlier work (which may still read them)
  // has finished.
  HIP_TRY(ctx, hipEventRecord(p->ev_compute, p->stream));
  HIP_TRY(ctx, hipStreamWaitEvent(copy_stream, p->ev_compute, 0));
  for (int b = 0; b < 2; ++b) HIP_TRY(ctx, hipStreamWaitEvent(ctx->bp_stream[b], p->ev_compute, 0));

  const int prec = p->opts.precision;
  const bool q16 = prec != FCT_PRECISION_FP32;
  const bool grow = q16 && !(max_abs_raw > 0.0f);
  if (q16) {
    HIP_TRY(ctx, fct::launch_fill(p->d_scale, fct::kScaleElems, 0.0f, p->stream));
    if (!grow) {
      HIP_TRY(ctx, fct::launch_scale_set(p->d_scale, max_abs_raw, p->stream));
      HIP_TRY(ctx, fct::launch_scale_update(p->d_scale, (float)p->hsum, p->stream));
      HIP_TRY(ctx, fct::launch_scale_commit(p->d_scale, p->stream));
    }
  }
  // With adaptive filtration the rebinning reads the filtered copy of each chunk.
  const float* rebin_src = p->af_on ? p->d_af_raw : p->d_raw_stage;
  fct::RebinFilterArgs rf = rf_args(p, rebin_src, p->d_filtered, p->d_scale);
  rf.skip_scale = 1;
  fct::BpArgs bp = bp_args(p);
  bp.filt = p->d_filtered;
  // scale[3]: scale[0] may change (growth) while a group is in flight.
  bp.filt_scale = q16 ? p->d_scale + 3 : nullptr;
  bp.native = p->d_native;
  const int S = p->ffs.S, nrows = p->nrows;
  const size_t entries_per_view = (size_t)nrows * p->cpitch;
  int j_done = par.x, s_done = 0, t_done = 0, groups = 0;
  bool bp_busy[2] = {false, false};
  // With a pinned source every upload is queued up front, and then volume
  // ranges go to a pinned destination on the copy stream as they are
  // thickened; otherwise the whole volume is copied at the end (a
  // pageable destination synchronously).
  const bool out_pinned = is_pinned(volume_host);
  const bool src_pinned = !src.read && is_pinned(src.host);
  bool out_early = out_pinned && src_pinned;

  auto rebin_upto = [&](int j1) -> hipError_t {
    if (j1 <= j_done) return hipSuccess;
    rf.line0 = j_done * nrows;
    rf.nlines = (j1 - j_done) * nrows;
    hipError_t e = fct::launch_rebin_filter(rf, p->stream);
    if (e == hipSuccess) e = hipEventRecord(p->ev_rebin, p->stream);
    j_done = j1;
    return e;
  };
  // Requested slices [t_done, t1) whose native slices are all backprojected
  // are thickened on stream st (and, with out_early, copied to the host
  // after it on the copy stream, whose uploads are all queued by then).
  auto thicken_upto = [&](hipStream_t st, int t1) -> hipError_t {
    if (t1 <= t_done) return hipSuccess;
    hipError_t e = fct::launch_thicken(thicken_args(p, p->d_native, p->d_vol_stage, t_done, t1), st);
    if (out_early && e == hipSuccess)
      e = hipMemcpyAsync(volume_host + (size_t)t_done * plane, p->d_vol_stage + (size_t)t_done * plane,
                         (size_t)(t1 - t_done) * plane * sizeof(float), hipMemcpyDeviceToHost, st);
    t_done = t1;
    return e;
  };
  auto bp_ready = [&](bool final) -> hipError_t {
    int s1 = s_done;
    if (final) {
      s1 = p->nslabs;
    } else {
      while (s1 < p->nslabs) {
        const int2 v = p->h_slab_vrange[s1];
        if (v.y > j_done && v.x < v.y) break;
        ++s1;
      }
      if (s1 - s_done < kStreamMinSlabs) return hipSuccess;
    }
    if (s1 <= s_done) return hipSuccess;   // every slab already launched
    // One group of every ready slab per step, on the two backprojection
    // streams in turn (a group's last waves overlap the next one's first;
    // splitting the last group so that part of the volume is copied out
    // earlier made its pieces share the CUs and each other's L2: slower).
    const int b = groups++ & 1;
    hipStream_t bs = ctx->bp_stream[b];
    hipError_t e = hipStreamWaitEvent(bs, p->ev_rebin, 0);
    bp.slab0 = s_done;
    bp.slab_count = s1 - s_done;
    s_done = s1;
    if (e == hipSuccess)
      e = fct::launch_backproject(bp, p->tile_z, p->variant, prec, fct::kBpAccumulate, bs);
    if (e == hipSuccess) e = hipEventRecord(p->ev_bp[b], bs);
    bp_busy[b] = true;
    if (e != hipSuccess || !out_early) return e;
    // The slices this group completes, after every group so far.
    int t1 = t_done;
    const int native_done = std::min(p->n_native, s_done * p->tile_z);
    if (s_done == p->nslabs)
      t1 = p->n_req;
    else
      while (t1 < p->n_req && p->h_thk_need[t1] <= native_done) ++t1;
    if (t1 > t_done)
      for (int c = 0; c < 2 && e == hipSuccess; ++c)
        if (bp_busy[c]) e = hipStreamWaitEvent(copy_stream, p->ev_bp[c], 0);
    if (e == hipSuccess) e = thicken_upto(copy_stream, t1);
    return e;
  };
  // The plan's stream waits for the backprojections issued so far.
  auto join_bp = [&]() -> hipError_t {
    for (int b = 0; b < 2; ++b)
      if (bp_busy[b]) {
        const hipError_t e = hipStreamWaitEvent(p->stream, p->ev_bp[b], 0);
        if (e != hipSuccess) return e;
      }
    return hipSuccess;
  };
  auto drain = [&]() {
    (void)hipStreamSynchronize(copy_stream);
    for (int b = 0; b < 2; ++b) (void)hipStreamSynchronize(ctx->bp_stream[b]);
    (void)hipStreamSynchronize(p->stream);
  };
  auto fail = [&](hipError_t e) {
    drain();
    set_err(ctx, "streaming reconstruction: %s", hipGetErrorString(e));
    return FCT_ERR_DEVICE;
  };

  // Chunk k covers raw views [rawr.x + k vpc, ...) up to rawr.y.  A caller's
  // pinned array is copied all at once (the copies queue on the copy
  // stream); a view reader fills the two staging buffers in turn, and a
  // pageable array (whose copies block) is copied, one chunk ahead.
  auto chunk_v0 = [&](int k) { return rawr.x + k * vpc; };
  auto chunk_n = [&](int k) { return std::min(vpc, rawr.y - chunk_v0(k)); };
  int issued = 0;
  auto issue = [&](int k) -> int {
    const int v0 = chunk_v0(k), n = chunk_n(k), b = k & 1;
    float* dst = p->d_raw_stage + (size_t)v0 * per_view;
    const size_t bytes = (size_t)n * per_view * sizeof(float);
    hipError_t e = hipSuccess;
    if (src.read) {
      if (k >= 2) e = hipEventSynchronize(p->ev_chunk[k - 2]);   // staging b free again
      if (e == hipSuccess && src.read(src.user, v0, n, p->h_stage[b]) != 0) {
        drain();
        set_err(ctx, "the view reader failed for views [%d, %d)", v0, v0 + n);
        return FCT_ERR_INVALID_ARGUMENT;
      }
      if (e == hipSuccess)
        e = hipMemcpyAsync(dst, p->h_stage[b], bytes, hipMemcpyHostToDevice, copy_stream);
    } else {
      e = hipMemcpyAsync(dst, src.host + (size_t)v0 * per_view, bytes, hipMemcpyHostToDevice,
                         copy_stream);
    }
    if (e == hipSuccess) e = hipEventRecord(p->ev_chunk[k], copy_stream);
    if (e != hipSuccess) return fail(e);
    issued = k + 1;
    return FCT_OK;
  };

  // Each step takes every chunk that has landed (at least the next one):
  // one abs-max / scale update / rebinning launch for all of them, then the
  // slab groups they make ready.  The scale read-back can wait behind a
  // running group (whose waves hold the CUs); the chunks that land
  // meanwhile join the step.
  bool committed = false;
  for (int k = 0; k < nchunks;) {
    const int ahead = src_pinned ? nchunks : std::min(nchunks, k + 2);
    while (issued < ahead) {
      const int st = issue(issued);
      if (st) return st;
    }
    hipError_t e = hipEventSynchronize(p->ev_chunk[k]);
    int k1 = k, arrived = chunk_v0(k);
    for (;;) {
      const int ka = k1;
      k1 = ka + 1;
      while (e == hipSuccess && k1 < issued && hipEventQuery(p->ev_chunk[k1]) == hipSuccess) ++k1;
      if (e == hipSuccess) e = hipStreamWaitEvent(p->stream, p->ev_chunk[k1 - 1], 0);
      const int va = arrived;
      arrived = chunk_v0(k1 - 1) + chunk_n(k1 - 1);
      if (e == hipSuccess && p->af_on)
        e = fct::launch_adaptive_filter(p->d_raw_stage, p->d_af_raw, p->d_af_vmax, va, arrived,
                                        p->raw_scan.n_rows, p->raw_scan.n_channels, p->af_s,
                                        p->stream);
      if (e != hipSuccess || !grow) break;
      e = fct::launch_absmax(rebin_src + (size_t)va * per_view,
                             (int64_t)(arrived - va) * (int64_t)per_view, p->d_scale, p->ncu,
                             p->stream);
      if (e == hipSuccess) e = fct::launch_scale_update(p->d_scale, (float)p->hsum, p->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(p->h_ratio, p->d_scale + 2, sizeof(float), hipMemcpyDeviceToHost,
                           p->stream);
      if (e == hipSuccess) e = hipEventRecord(p->ev_ratio, p->stream);
      if (e == hipSuccess) e = hipEventSynchronize(p->ev_ratio);
      const bool grew = *p->h_ratio != 1.0f;
      if (e == hipSuccess && grew) {
        // Rescale the stored lines and publish the new scale once no
        // backprojection reads either.
        e = join_bp();
        if (e == hipSuccess && j_done > par.x)
          e = fct::launch_quad_rescale(
              static_cast<char*>(p->d_filtered) +
                  (size_t)par.x * entries_per_view * p->filtered_entry_bytes(),
              (size_t)(j_done - par.x) * entries_per_view, prec, p->d_scale, p->ncu, p->stream);
      }
      if (e == hipSuccess && (grew || !committed)) {
        e = fct::launch_scale_commit(p->d_scale, p->stream);
        committed = true;
      }
      if (e != hipSuccess || k1 >= issued || hipEventQuery(p->ev_chunk[k1]) != hipSuccess) break;
    }
    // Parallel views j with j S + reach <= arrived have every raw view they read.
    int j1 = par.y;
    if (arrived < rawr.y) {
      const int d = arrived - p->rebin_reach;
      j1 = d < 0 ? 0 : std::min(par.y, d / S + 1);
      if (nrows & 1) j1 &= ~1;   // line pairs must not straddle a chunk end
    }
    if (e == hipSuccess) e = rebin_upto(j1);
    if (e == hipSuccess) e = bp_ready(false);
    if (e != hipSuccess) return fail(e);
    k = k1;
  }
  hipError_t e = rebin_upto(par.y);
  if (e == hipSuccess) e = bp_ready(true);
  if (e == hipSuccess) e = join_bp();
  if (e == hipSuccess && t_done < p->n_req) {
    // Without out_early (or slices no piece completed): thickened here, and
    // the whole volume copied below.
    out_early = false;
    e = thicken_upto(p->stream, p->n_req);
  }
  if (e == hipSuccess && !out_early && out_pinned)
    e = hipMemcpyAsync(volume_host, p->d_vol_stage, vol_bytes, hipMemcpyDeviceToHost, p->stream);
  if (e != hipSuccess) return fail(e);
  HIP_TRY(ctx, hipStreamSynchronize(copy_stream));
  for (int b = 0; b < 2; ++b) HIP_TRY(ctx, hipStreamSynchronize(ctx->bp_stream[b]));
  HIP_TRY(ctx, hipStreamSynchronize(p->stream));
  if (!out_pinned)
    HIP_TRY(ctx, hipMemcpy(volume_host, p->d_vol_stage, vol_bytes, hipMemcpyDeviceToHost));
  return FCT_OK;
}

}  // namespace

extern "C" {

FCT_API int fct_wfbp_gpu_create(int32_t device, fct_wfbp_gpu_ctx** out) {
  if (!out) return FCT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto* ctx = new (std::nothrow) fct_wfbp_gpu_ctx();
  if (!ctx) return FCT_ERR_OUT_OF_MEMORY;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    delete ctx;
    return FCT_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    delete ctx;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    delete ctx;
    return FCT_ERR_DEVICE;
  }
  ctx->device = device;
  *out = ctx;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_destroy(fct_wfbp_gpu_ctx* ctx) {
  if (ctx && (ctx->copy_stream || ctx->bp_stream[0] || ctx->bp_stream[1])) {
    (void)hipSetDevice(ctx->device);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    for (int b = 0; b < 2; ++b)
      if (ctx->bp_stream[b]) (void)hipStreamDestroy(ctx->bp_stream[b]);
  }
  delete ctx;
  return FCT_OK;
}

FCT_API const char* fct_wfbp_gpu_last_error(const fct_wfbp_gpu_ctx* ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

FCT_API int fct_wfbp_gpu_plan_create(fct_wfbp_gpu_ctx* ctx, const fct_scan_geometry* scan,
                                     const fct_recon_config* recon, const fct_exec_options* opts,
                                     void* stream, fct_wfbp_gpu_plan** out) {
  if (!ctx || !out) return FCT_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int st = validate(ctx, scan, recon, opts);
  if (st != FCT_OK) return st;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  auto* p = new (std::nothrow) fct_wfbp_gpu_plan();
  if (!p) return FCT_ERR_OUT_OF_MEMORY;
  p->ctx = ctx;
  p->device = ctx->device;
  p->raw_scan = *scan;
  p->ffs = fct_ffs::make_geometry(scan->ffs_mode, scan->n_channels, scan->n_rows, scan->n_views,
                                  scan->views_per_rotation, scan->channel_spacing,
                                  scan->row_spacing, scan->central_channel, scan->central_row,
                                  scan->r, scan->dsd, scan->anode_angle);
  // Everything downstream of the rebinning sees the interleaved detector.
  p->scan = *scan;
  p->scan.n_channels = p->ffs.e_ncols;
  p->scan.n_rows = p->ffs.e_nrows;
  p->scan.n_views = p->ffs.e_nviews;
  p->scan.views_per_rotation = p->ffs.e_ppr;
  p->scan.channel_spacing = p->ffs.e_dch;
  p->scan.row_spacing = p->ffs.e_drow;
  p->scan.central_channel = p->ffs.e_cc;
  p->scan.central_row = p->ffs.e_cr;
  p->scan.ffs_mode = FCT_FFS_NONE;
  p->recon = *recon;
  if (opts) p->opts = *opts;
  p->af_on = opts && opts->adaptive_filtration ? 1 : 0;
  p->af_s = (float)recon->adaptive_filtration_s;
  p->stream = static_cast<hipStream_t>(stream);   // NULL: the null stream
  if (hipDeviceGetAttribute(&p->ncu, hipDeviceAttributeMultiprocessorCount, p->device) !=
      hipSuccess)
    p->ncu = 1;
  auto fail = [&](int code) {
    free_plan_buffers(p);
    delete p;
    return code;
  };
#define PLAN_TRY(expr)                                                             \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_err(ctx, "%s failed: %s", #expr, hipGetErrorString(e_));                 \
      return fail(e_ == hipErrorOutOfMemory ? FCT_ERR_OUT_OF_MEMORY : FCT_ERR_DEVICE); \
    }                                                                              \
  } while (0)

  const fct_scan_geometry& s = p->scan;
  const fct_recon_config& r = *recon;
  p->ncols = s.n_channels;
  p->nrows = s.n_rows;
  p->nviews = s.n_views;
  p->ppr = s.views_per_rotation;
  p->H = p->ppr / 2;
  int n = 1, lg = 0;
  while (n < 2 * p->ncols) {
    n <<= 1;
    ++lg;
  }
  p->nfft = n;                                    // ReconstructionMachine.cu:54
  p->log2n = lg;
  p->cpitch = (p->ncols + 63) / 64 * 64;
  p->dgamma = s.channel_spacing / s.dsd;
  p->ds = s.channel_spacing * s.r / s.dsd;        // ReconstructionMachine.cu:491
  p->dtheta = 2.0 * kPi / (double)p->ppr;
  p->dz = s.feed / (double)p->ppr;
  p->csw = s.r * s.row_spacing / s.dsd;           // ReconstructionMachine.cu:62
  p->acq = 2.0 * s.r * std::sin(0.5 * ((double)p->ncols - 1.0) * p->dgamma);
  p->g = fct_tables::bp_consts(p->ncols, p->nrows, s.channel_spacing, s.row_spacing, s.dsd, s.r,
                               s.central_channel, s.central_row, s.feed);
  p->scale = (float)(2.0 * kPi / (double)p->ppr);
  p->tile_z = (opts && opts->tile_z) ? opts->tile_z : 8;
  p->variant = (opts && opts->bp_variant) ? opts->bp_variant : FCT_BP_LEAN;
  if ((size_t)p->nviews * p->nrows * p->cpitch >= ((size_t)1 << 31))
    return fail((set_err(ctx, "filtered projections exceed 2^31 elements"), FCT_ERR_LIMIT));
  if (p->nfft > 4096) return fail((set_err(ctx, "nfft > 4096 unsupported"), FCT_ERR_LIMIT));

  // Slice grids (ReconstructionMachine.cu:107-137) in FP64.
  const double pitch = r.slice_pitch > 0 ? r.slice_pitch : r.slice_thickness;
  p->n_native = fct_tables::slice_count(r.start_pos, r.end_pos, p->csw);
  p->n_req = fct_tables::slice_count(r.start_pos, r.end_pos, pitch);
  if (p->n_native < 1 || p->n_req < 1) return fail((set_err(ctx, "empty slice grid"), FCT_ERR_INVALID_ARGUMENT));
  std::vector<double> zn(p->n_native), zr(p->n_req);
  for (int i = 0; i < p->n_native; ++i) zn[i] = fct_tables::slice_z(r.start_pos, r.end_pos, p->csw, i);
  for (int i = 0; i < p->n_req; ++i) zr[i] = fct_tables::slice_z(r.start_pos, r.end_pos, pitch, i);
  const int chunk = fct_tables::kSliceChunk;
  p->nchunks = (p->n_native + chunk - 1) / chunk;
  p->nslabs = (p->n_native + p->tile_z - 1) / p->tile_z;

  // A clockwise scan is the mirror image y -> -y of the counter-clockwise
  // scan with tube angles -a_v: the kernels reconstruct that one at the
  // mirrored voxel centres (-y), which is the clockwise volume in place.
  const bool cw = s.rotation < 0;
  std::vector<float> xs(r.nx), ys(r.ny), zs(p->n_native), zt((size_t)p->nchunks * p->nviews);
  for (int i = 0; i < r.nx; ++i) xs[i] = fct_tables::voxel_coord(i, r.nx, r.recon_fov, r.x_origin);
  for (int i = 0; i < r.ny; ++i) {
    ys[i] = fct_tables::voxel_coord(i, r.ny, r.recon_fov, r.y_origin);
    if (cw) ys[i] = -ys[i];
  }
  for (int i = 0; i < p->n_native; ++i) zs[i] = (float)(zn[i] - zn[(i / chunk) * chunk]);
  for (int c = 0; c < p->nchunks; ++c)
    for (int v = 0; v < p->nviews; ++v)
      zt[(size_t)c * p->nviews + v] = fct_tables::view_z_rel(s.z0, p->dz, v, zn[c * chunk]);
  std::vector<float2> cs(p->H);
  const double angle0 = cw ? -s.angle0 : s.angle0;
  for (int t = 0; t < p->H; ++t) fct_tables::view_cos_sin(angle0, p->dtheta, t, &cs[t].x, &cs[t].y);

  std::vector<int2> vr;
  if (!slab_view_ranges(p, zn, p->tile_z, &vr))
    return fail((set_err(ctx, "a slab's view range exceeds 4 GiB of filtered rows"),
                 FCT_ERR_LIMIT));

  // Exact view-window slice ranges per (chunk, view): the slices s of the
  // chunk with fct_bp_in_window(zt, zs) true form one interval (the test is
  // monotone in zs); the same FP32 values and test as the kernels' per-sample
  // form, evaluated once here.  A window that holds no slice of the chunk is
  // stored as the empty interval just before (a0, a0 - 1) or just after
  // (a1, a1 - 1) the chunk, on the side where it lies, so that both ends stay
  // monotone in the view index (the lean kernel ends its view loops on that).
  const double slice_dir = p->recon.end_pos >= p->recon.start_pos ? 1.0 : -1.0;
  std::vector<int2> wrange((size_t)p->nchunks * p->nviews);
  for (int c = 0; c < p->nchunks; ++c) {
    const int a0 = c * chunk, a1 = std::min(p->n_native, a0 + chunk);
    for (int v = 0; v < p->nviews; ++v) {
      int lo = p->n_native, hi = -1;
      const float ztv = zt[(size_t)c * p->nviews + v];
      for (int sidx = a0; sidx < a1; ++sidx)
        if (fct_bp_in_window(ztv, zs[sidx], p->g)) {
          lo = std::min(lo, sidx);
          hi = std::max(hi, sidx);
        }
      if (hi < 0) {
        const bool before = ((double)ztv - (double)zs[a0]) * slice_dir < 0.0;
        lo = before ? a0 : a1;
        hi = lo - 1;
      }
      wrange[(size_t)c * p->nviews + v] = make_int2(lo, hi);
    }
  }

  // Rebinning table (cuda_kernels.cuh:17-36), one per effective-row parity
  // with z FFS; without FFS exactly fct_tables::rebin_entry.
  const int zpar = fct_ffs::has_z(p->ffs.mode) ? 2 : 1;
  std::vector<fct::RebinDev> tab((size_t)zpar * p->ncols);
  const int64_t raw_plane = (int64_t)p->raw_scan.n_rows * p->raw_scan.n_channels;
  int aoff_max = 0, mmax_max = 0;
  p->rebin_aoff_min = 0;
  p->rebin_mmax_min = INT32_MAX;
  for (int zb = 0; zb < zpar; ++zb)
    for (int c = 0; c < p->ncols; ++c) {
      const fct_ffs::Entry e = fct_ffs::rebin_entry(p->ffs, c, zb);
      // Entries without a fan sample read (state 0, row, channel 0) of a
      // clamped view and are discarded: their mmax is taken >= 0.
      const int mmax = std::max(e.mmax, 0);
      const uint32_t boff =
          e.valid ? (uint32_t)(((int64_t)e.rho * raw_plane + e.b0) * (int64_t)sizeof(float)) : 0u;
      tab[(size_t)zb * p->ncols + c] =
          fct::RebinDev{e.b0, e.b1, e.aoff, e.valid ? e.rho + 1 : 0, (float)e.fb, (float)e.fa,
                        mmax, boff};
      p->rebin_aoff_min = std::min(p->rebin_aoff_min, e.aoff);
      aoff_max = std::max(aoff_max, e.aoff);
      p->rebin_mmax_min = std::min(p->rebin_mmax_min, mmax);
      mmax_max = std::max(mmax_max, mmax);
      // Parallel view j reads raw projections up to (j + aoff + 1) S + rho.
      if (e.valid)
        p->rebin_reach = std::max(p->rebin_reach, (e.aoff + 1) * p->ffs.S + e.rho + 1);
    }
  // The gather addresses a line pair's taps by 32-bit offsets from the first
  // view group they can reach (rebin_filter.hip): the groups of the table's
  // aoff span (+ 4: the pair's second view, the interpolation's next view,
  // the clamp at the last view and the group itself) must fit 4 GiB, and a
  // view group 16 MiB (24-bit products).
  {
    const int64_t group_bytes = (int64_t)p->ffs.S * raw_plane * (int64_t)sizeof(float);
    if (group_bytes >= (int64_t)1 << 24 ||
        (int64_t)(aoff_max - p->rebin_aoff_min + 4) * group_bytes >= (int64_t)1 << 32)
      return fail((set_err(ctx, "a raw view group (%lld bytes) exceeds the rebinning's "
                                "addressing (16 MiB, 4 GiB per fan)",
                           (long long)group_bytes),
                   FCT_ERR_LIMIT));
    // The packed 16-byte entries of the gather (aoff in 12 bits, mmax in 20).
    if (p->rebin_aoff_min < -2048 || aoff_max > 2047 || mmax_max >= (1 << 20))
      return fail((set_err(ctx, "the rebinning table exceeds its packed form (fan view "
                                "offsets %d..%d, last projection %d; limits -2048..2047, "
                                "2^20 - 1)",
                           p->rebin_aoff_min, aoff_max, mmax_max),
                   FCT_ERR_LIMIT));
  }

  // Filter spectrum (generate_filter, ReconstructionMachine.cu:478-544):
  // taps -> fftshift -> DFT in FP64; the spectrum of the symmetric kernel is
  // real.  1/nfft (cuda_kernels.cuh:56-57) is folded in.
  const std::vector<double> h =
      filter_taps(p->nfft, p->ds, r.filter_kind, r.filter_c, r.filter_a);
  p->hsum = 0.0;   // bound |y| <= max|x| * sum|h| of the FP16 / BF16 storage scale
  for (double v : h) p->hsum += std::fabs(v);
  std::vector<std::complex<double>> Hc(p->nfft);
  for (int k = 0; k < p->nfft; ++k) {
    const int lag = k < p->nfft / 2 ? k : k - p->nfft;
    Hc[k] = h[lag + p->nfft / 2];
  }
  fft64(Hc);
  std::vector<float> Hn(p->nfft);
  for (int k = 0; k < p->nfft; ++k) Hn[k] = (float)(Hc[k].real() / (double)p->nfft);
  std::vector<float2> tw(p->nfft);
  for (int m = 0; m < p->nfft; ++m) {
    const double ang = -2.0 * kPi * (double)m / (double)p->nfft;
    tw[m] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }

  // Thickening table (cuda_kernels.cuh:298-316): weights normalised in FP64.
  std::vector<int32_t> toff(p->n_req + 1, 0), tidx;
  std::vector<float> tw8;
  for (int so = 0; so < p->n_req; ++so) {
    double wsum = 0.0;
    std::vector<std::pair<int, double>> ws;
    for (int i = 0; i < p->n_native; ++i) {
      const double w = std::max(0.0, 1.0 - std::fabs(zn[i] - zr[so]) / r.slice_thickness);
      if (w != 0.0) {
        ws.emplace_back(i, w);
        wsum += w;
      }
    }
    for (auto& pr : ws) {
      tidx.push_back(pr.first);
      tw8.push_back((float)(pr.second / wsum));
    }
    toff[so + 1] = (int32_t)tidx.size();
  }
  p->h_thk_need.assign(p->n_req, 0);
  for (int so = 0; so < p->n_req; ++so)
    for (int k = toff[so]; k < toff[so + 1]; ++k)
      p->h_thk_need[so] = std::max(p->h_thk_need[so], tidx[k] + 1);
  // Transposed table: per native slice, the requested slices and weights.
  std::vector<int32_t> tToff(p->n_native + 1, 0), tTidx;
  std::vector<float> tTw;
  {
    std::vector<std::vector<std::pair<int, float>>> per(p->n_native);
    for (int so = 0; so < p->n_req; ++so)
      for (int k = toff[so]; k < toff[so + 1]; ++k) per[tidx[k]].emplace_back(so, tw8[k]);
    for (int i = 0; i < p->n_native; ++i) {
      for (auto& pr : per[i]) {
        tTidx.push_back(pr.first);
        tTw.push_back(pr.second);
      }
      tToff[i + 1] = (int32_t)tTidx.size();
    }
  }
  if (tidx.empty()) {
    tidx.push_back(0);
    tw8.push_back(0.0f);
  }
  if (tTidx.empty()) {
    tTidx.push_back(0);
    tTw.push_back(0.0f);
  }

  PLAN_TRY(upload(&p->d_xs, xs, p->stream));
  PLAN_TRY(upload(&p->d_ys, ys, p->stream));
  PLAN_TRY(upload(&p->d_cs, cs, p->stream));
  PLAN_TRY(upload(&p->d_zs, zs, p->stream));
  PLAN_TRY(upload(&p->d_zt, zt, p->stream));
  PLAN_TRY(upload(&p->d_slab_vrange, vr, p->stream));
  p->h_slab_vrange = vr;
  if (p->af_on) {   // at set-up, so that no stage call (or captured graph) allocates
    PLAN_TRY(hipMalloc(reinterpret_cast<void**>(&p->d_af_raw),
                       (size_t)p->raw_elems() * sizeof(float)));
    PLAN_TRY(hipMalloc(reinterpret_cast<void**>(&p->d_af_vmax),
                       (size_t)p->raw_scan.n_views * sizeof(float)));
  }
  PLAN_TRY(upload(&p->d_wrange, wrange, p->stream));
  PLAN_TRY(upload(&p->d_rebin, tab, p->stream));
  p->h_rebin = tab;
  PLAN_TRY(upload(&p->d_Hn, Hn, p->stream));
  PLAN_TRY(upload(&p->d_tw, tw, p->stream));
  PLAN_TRY(upload(&p->d_thk_off, toff, p->stream));
  PLAN_TRY(upload(&p->d_thk_idx, tidx, p->stream));
  PLAN_TRY(upload(&p->d_thk_w, tw8, p->stream));
  PLAN_TRY(upload(&p->d_thkT_off, tToff, p->stream));
  PLAN_TRY(upload(&p->d_thkT_idx, tTidx, p->stream));
  PLAN_TRY(upload(&p->d_thkT_w, tTw, p->stream));
  {
    // Per-(tile, theta~) z-interval extremes of the backprojection (from the
    // tables above, on the plan's stream).
    const fct::BpArgs a = fct::plan_bp_args(p);
    const size_t n = (size_t)a.tiles_x * a.tiles_y * a.H;
    PLAN_TRY(hipMalloc(reinterpret_cast<void**>(&p->d_zbounds), n * sizeof(float4)));
    PLAN_TRY(fct::launch_bp_zbounds(a, p->d_zbounds, p->stream));
  }
#undef PLAN_TRY
  *out = p;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_plan_destroy(fct_wfbp_gpu_plan* p) {
  if (!p) return FCT_OK;
  (void)hipSetDevice(p->device);
  (void)hipStreamSynchronize(p->stream);
  free_plan_buffers(p);
  delete p;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_plan_info(const fct_wfbp_gpu_plan* p, fct_plan_info* info) {
  if (!p || !info) return FCT_ERR_INVALID_ARGUMENT;
  std::memset(info, 0, sizeof(*info));
  info->n_slices_native = p->n_native;
  info->n_slices_requested = p->n_req;
  info->nfft = p->nfft;
  info->channel_pitch = p->cpitch;
  info->raw_elems = p->raw_elems();
  info->filtered_bytes = (int64_t)p->filtered_bytes();
  info->scale_elems = (int64_t)p->scale_elems();
  info->native_elems = (int64_t)p->n_native * p->recon.ny * p->recon.nx;
  info->volume_elems = (int64_t)p->n_req * p->recon.ny * p->recon.nx;
  info->collimated_slice_width = p->csw;
  info->acquisition_fov = p->acq;
  info->z_rot = std::fabs(p->scan.feed);
  info->tile_z = p->tile_z;
  info->bp_variant = p->variant;
  info->n_parallel_views = p->nviews;
  info->n_filtered_rows = p->nrows;
  info->n_filtered_channels = p->ncols;
  info->ffs_states = p->ffs.S;
  int2 par, raw;
  needed_views(p, &par, &raw);
  info->read_views_first = raw.x;
  info->read_views_end = raw.y;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_set_stream(fct_wfbp_gpu_plan* p, void* stream) {
  if (!p) return FCT_ERR_INVALID_ARGUMENT;
  // Work already queued on the old stream (the set-up tables of plan
  // creation, earlier stages) completes before the plan moves.
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));
  p->stream = static_cast<hipStream_t>(stream);
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_rebin_filter(fct_wfbp_gpu_plan* p, const float* raw, void* filtered,
                                      float* scale) {
  if (!p || !raw || !filtered) return FCT_ERR_INVALID_ARGUMENT;
  if (p->opts.precision != FCT_PRECISION_FP32 && !scale)
    return set_err(p->ctx, "the scale buffer is required for FP16/BF16"),
           FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  if (p->af_on) {   // adaptive filtration of the whole scan into the plan's buffer
    HIP_TRY(p->ctx, fct::launch_adaptive_filter(raw, p->d_af_raw, p->d_af_vmax, 0,
                                                p->raw_scan.n_views, p->raw_scan.n_rows,
                                                p->raw_scan.n_channels, p->af_s, p->stream));
    raw = p->d_af_raw;
  }
  HIP_TRY(p->ctx, fct::launch_rebin_filter(rf_args(p, raw, filtered, scale), p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_backproject(fct_wfbp_gpu_plan* p, const void* filtered,
                                     const float* scale, float* native) {
  if (!p || !filtered || !native) return FCT_ERR_INVALID_ARGUMENT;
  if (p->opts.precision != FCT_PRECISION_FP32 && !scale)
    return set_err(p->ctx, "the scale buffer is required for FP16/BF16"),
           FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::BpArgs a = bp_args(p);
  a.filt = filtered;
  a.filt_scale = scale;
  a.native = native;
  HIP_TRY(p->ctx, fct::launch_backproject(a, p->tile_z, p->variant, p->opts.precision,
                                          fct::kBpAccumulate, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_thicken(fct_wfbp_gpu_plan* p, const float* native, float* volume) {
  if (!p || !native || !volume) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, fct::launch_thicken(thicken_args(p, native, volume, 0, p->n_req), p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_ray_project(fct_wfbp_gpu_plan* p, const float* native, float* raw) {
  if (!p || !native || !raw) return FCT_ERR_INVALID_ARGUMENT;
  if (p->n_native < 2)
    return set_err(p->ctx, "the ray projector needs at least two native slices"),
           FCT_ERR_UNSUPPORTED;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::RayProjArgs a = ray_args(p);
  a.vol_in = native;
  a.raw_out = raw;
  HIP_TRY(p->ctx, fct::launch_ray_project(a, false, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_ray_backproject(fct_wfbp_gpu_plan* p, const float* raw, float* native) {
  if (!p || !native || !raw) return FCT_ERR_INVALID_ARGUMENT;
  if (p->n_native < 2)
    return set_err(p->ctx, "the ray projector needs at least two native slices"),
           FCT_ERR_UNSUPPORTED;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::RayProjArgs a = ray_args(p);
  a.raw_in = raw;
  a.vol_out = native;
  HIP_TRY(p->ctx, fct::launch_fill(native, (size_t)p->n_native * p->recon.ny * p->recon.nx, 0.0f,
                                   p->stream));
  HIP_TRY(p->ctx, fct::launch_ray_project(a, true, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_set_output_units(fct_wfbp_gpu_plan* p, int32_t units, double mu_water) {
  if (!p) return FCT_ERR_INVALID_ARGUMENT;
  if (units == FCT_UNITS_MU) {
    p->out_hu = 0;
    return FCT_OK;
  }
  if (units != FCT_UNITS_HU) return set_err(p->ctx, "unknown units %d", units), FCT_ERR_INVALID_ARGUMENT;
  if (!(mu_water > 0.0) || !(mu_water < 1e3))
    return set_err(p->ctx, "mu_water must be > 0 (1/mm)"), FCT_ERR_INVALID_ARGUMENT;
  p->out_hu = 1;
  p->mu_water = mu_water;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_rebin(fct_wfbp_gpu_plan* p, const float* raw, float* lines_out) {
  if (!p || !raw || !lines_out) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::RebinFilterArgs a{};
  a.raw = raw;
  a.out = lines_out;
  a.tab = p->d_rebin;
  a.Hn = p->d_Hn;
  a.tw = p->d_tw;
  a.nviews = p->nviews;
  a.nrows = p->nrows;
  a.ncols = p->ncols;
  a.cpitch = p->cpitch;
  a.nfft = p->nfft;
  a.log2n = p->log2n;
  a.precision = FCT_PRECISION_FP32;
  a.nlines = p->nviews * p->nrows;
  a.ncu = p->ncu;
  a.raw_nrows = p->raw_scan.n_rows;
  a.raw_ncols = p->raw_scan.n_channels;
  a.states = p->ffs.S;
  a.zshift = fct_ffs::has_z(p->ffs.mode) ? 1 : 0;
  set_gather_window(p, &a);
  a.rebin_only = 1;
  HIP_TRY(p->ctx, fct::launch_rebin_filter(a, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_filter(fct_wfbp_gpu_plan* p, const float* lines_in, float* lines_out) {
  if (!p || !lines_in || !lines_out) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::RebinFilterArgs a{};
  a.lines_in = lines_in;
  a.out = lines_out;
  a.Hn = p->d_Hn;
  a.tw = p->d_tw;
  a.tab = p->d_rebin;
  a.nviews = p->nviews;
  a.nrows = p->nrows;
  a.ncols = p->ncols;
  a.cpitch = p->cpitch;
  a.nfft = p->nfft;
  a.log2n = p->log2n;
  a.precision = FCT_PRECISION_FP32;
  a.nlines = p->nviews * p->nrows;
  a.ncu = p->ncu;
  a.raw_nrows = p->raw_scan.n_rows;
  a.raw_ncols = p->raw_scan.n_channels;
  a.states = p->ffs.S;
  a.zshift = fct_ffs::has_z(p->ffs.mode) ? 1 : 0;
  HIP_TRY(p->ctx, fct::launch_rebin_filter(a, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_backproject_adjoint(fct_wfbp_gpu_plan* p, const float* native_bar,
                                             float* filtered_bar) {
  if (!p || !native_bar || !filtered_bar) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, fct::launch_fill(filtered_bar, (size_t)p->nviews * p->nrows * p->cpitch, 0.0f,
                                   p->stream));
  fct::BpAdjArgs a{};
  a.bp = bp_args(p);
  a.xbar = native_bar;
  a.gbar = filtered_bar;
  HIP_TRY(p->ctx, fct::launch_backproject_adjoint(a, p->tile_z, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_matched_project(fct_wfbp_gpu_plan* p, const float* native,
                                         float* lines) {
  if (!p || !native || !lines) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, fct::launch_fill(lines, (size_t)p->nviews * p->nrows * p->cpitch, 0.0f,
                                   p->stream));
  fct::BpAdjArgs a{};
  a.bp = bp_args(p);
  a.xbar = native;
  a.gbar = lines;
  HIP_TRY(p->ctx, fct::launch_backproject_adjoint(a, p->tile_z, p->stream, true));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_matched_backproject(fct_wfbp_gpu_plan* p, const float* lines,
                                             float* native) {
  if (!p || !lines || !native) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::BpArgs a = bp_args(p);
  a.filt = lines;
  a.native = native;
  HIP_TRY(p->ctx, fct::launch_backproject(a, p->tile_z, FCT_BP_LEAN, FCT_PRECISION_FP32,
                                          fct::kBpMatched, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_rebin_adjoint(fct_wfbp_gpu_plan* p, const float* filtered_bar,
                                       float* raw_bar) {
  if (!p || !filtered_bar || !raw_bar) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, fct::launch_fill(raw_bar, (size_t)p->raw_elems(), 0.0f, p->stream));
  fct::RebinAdjArgs a{};
  a.gbar = filtered_bar;
  a.pbar = raw_bar;
  a.tab = p->d_rebin;
  a.nviews = p->nviews;
  a.nrows = p->nrows;
  a.ncols = p->ncols;
  a.cpitch = p->cpitch;
  a.raw_nrows = p->raw_scan.n_rows;
  a.raw_ncols = p->raw_scan.n_channels;
  a.states = p->ffs.S;
  a.zshift = fct_ffs::has_z(p->ffs.mode) ? 1 : 0;
  HIP_TRY(p->ctx, fct::launch_rebin_adjoint(a, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_thicken_adjoint(fct_wfbp_gpu_plan* p, const float* volume_bar,
                                         float* native_bar) {
  if (!p || !volume_bar || !native_bar) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  fct::ThickenArgs a{};
  a.native = volume_bar;
  a.out = native_bar;
  a.off = p->d_thkT_off;
  a.idx = p->d_thkT_idx;
  a.w = p->d_thkT_w;
  a.plane = p->recon.nx * p->recon.ny;
  a.n_req = p->n_native;
  HIP_TRY(p->ctx, fct::launch_thicken(a, p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_adjoint(fct_wfbp_gpu_plan* p, const float* native_bar, float* raw_bar) {
  if (!p || !native_bar || !raw_bar) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  if (!p->d_gbar)
    HIP_TRY(p->ctx, hipMalloc(reinterpret_cast<void**>(&p->d_gbar),
                              (size_t)p->nviews * p->nrows * p->cpitch * sizeof(float)));
  int st = fct_wfbp_gpu_backproject_adjoint(p, native_bar, p->d_gbar);
  if (!st) st = fct_wfbp_gpu_filter(p, p->d_gbar, p->d_gbar);
  if (!st) st = fct_wfbp_gpu_rebin_adjoint(p, p->d_gbar, raw_bar);
  return st;
}

FCT_API int fct_wfbp_gpu_reconstruct(fct_wfbp_gpu_plan* p, const float* raw, float* volume,
                                     int32_t host_buffers) {
  if (!p || !raw || !volume) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  if (host_buffers) return reconstruct_streaming(p, ViewSource{raw, nullptr, nullptr}, 0, 0.0f,
                                                 volume);
  const size_t nat_bytes = (size_t)p->n_native * p->recon.ny * p->recon.nx * sizeof(float);
  if (!p->d_filtered) HIP_TRY(p->ctx, hipMalloc(&p->d_filtered, p->filtered_bytes()));
  if (p->scale_elems() && !p->d_scale)
    HIP_TRY(p->ctx, hipMalloc(reinterpret_cast<void**>(&p->d_scale),
                              p->scale_elems() * sizeof(float)));
  if (!p->d_native) HIP_TRY(p->ctx, hipMalloc(reinterpret_cast<void**>(&p->d_native), nat_bytes));
  int st = fct::chain_rebin_filter(p, raw, p->d_filtered, p->d_scale);
  if (st) return st;
  st = fct_wfbp_gpu_backproject(p, p->d_filtered, p->d_scale, p->d_native);
  if (st) return st;
  return fct_wfbp_gpu_thicken(p, p->d_native, volume);
}

FCT_API int fct_wfbp_gpu_reconstruct_stream(fct_wfbp_gpu_plan* p, fct_view_reader read,
                                            void* user, int32_t views_per_chunk,
                                            float max_abs_raw, float* volume_host) {
  if (!p || !read || !volume_host) return FCT_ERR_INVALID_ARGUMENT;
  if (views_per_chunk < 0)
    return set_err(p->ctx, "views_per_chunk must be >= 0"), FCT_ERR_INVALID_ARGUMENT;
  if (!(max_abs_raw >= 0.0f) || !(max_abs_raw <= 3.0e38f))
    return set_err(p->ctx, "max_abs_raw must be finite and >= 0"), FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  return reconstruct_streaming(p, ViewSource{nullptr, read, user}, views_per_chunk, max_abs_raw,
                               volume_host);
}

FCT_API int fct_wfbp_gpu_count_voxel_views(fct_wfbp_gpu_plan* p, int32_t* per_voxel,
                                           int64_t* total) {
  if (!p || !total) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  const size_t nat = (size_t)p->n_native * p->recon.ny * p->recon.nx;
  int32_t* counts = per_voxel;
  unsigned long long* d_tot = nullptr;
  bool own_counts = false;
  if (!counts) {
    HIP_TRY(p->ctx, hipMalloc(reinterpret_cast<void**>(&counts), nat * sizeof(int32_t)));
    own_counts = true;
  }
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_tot), sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemsetAsync(d_tot, 0, sizeof(unsigned long long), p->stream);
  fct::BpArgs a = bp_args(p);
  a.counts = counts;
  a.total = d_tot;
  if (e == hipSuccess)
    e = fct::launch_backproject(a, p->tile_z, p->variant, FCT_PRECISION_FP32, fct::kBpCount, p->stream);
  unsigned long long tot = 0;
  if (e == hipSuccess)
    e = hipMemcpyAsync(&tot, d_tot, sizeof tot, hipMemcpyDeviceToHost, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  if (d_tot) (void)hipFree(d_tot);
  if (own_counts) (void)hipFree(counts);
  if (e != hipSuccess) {
    set_err(p->ctx, "count_voxel_views: %s", hipGetErrorString(e));
    return FCT_ERR_DEVICE;
  }
  *total = (int64_t)tot;
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_voxel_selection(fct_wfbp_gpu_plan* p, int32_t ix, int32_t iy,
                                         int32_t islice, int32_t* sel) {
  if (!p || !sel) return FCT_ERR_INVALID_ARGUMENT;
  if (ix < 0 || iy < 0 || islice < 0 || ix >= p->recon.nx || iy >= p->recon.ny ||
      islice >= p->n_native)
    return set_err(p->ctx, "voxel outside the native grid"), FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, hipMemsetAsync(sel, 0xFF, (size_t)p->nviews * sizeof(int32_t), p->stream));
  fct::BpArgs a = bp_args(p);
  a.sel = sel;
  a.sel_ix = ix;
  a.sel_iy = iy;
  a.sel_is = islice;
  HIP_TRY(p->ctx, fct::launch_backproject(a, p->tile_z, p->variant, FCT_PRECISION_FP32,
                                          fct::kBpSelect, p->stream));
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_synchronize(fct_wfbp_gpu_plan* p) {
  if (!p) return FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(p->ctx, hipSetDevice(p->device));
  HIP_TRY(p->ctx, hipStreamSynchronize(p->stream));
  return FCT_OK;
}

FCT_API int fct_wfbp_gpu_phantom_project(fct_wfbp_gpu_ctx* ctx, const fct_scan_geometry* scan,
                                         double cyl_radius, double cyl_mu,
                                         const fct_ellipsoid* ellipsoids, int32_t n_ellipsoids,
                                         double i0, uint64_t seed, int32_t v0, int32_t v1,
                                         float* raw, void* stream) {
  if (!ctx || !scan || !raw) return FCT_ERR_INVALID_ARGUMENT;
  if (n_ellipsoids < 0 || (n_ellipsoids > 0 && !ellipsoids) || n_ellipsoids > 4096)
    return set_err(ctx, "bad ellipsoid list"), FCT_ERR_INVALID_ARGUMENT;
  if (v0 < 0 || v1 > scan->n_views || v0 >= v1 || v1 - v0 > 65535)
    return set_err(ctx, "bad view range [%d, %d)", v0, v1), FCT_ERR_INVALID_ARGUMENT;
  if (scan->views_per_rotation < 1 || scan->n_channels < 1 || scan->n_rows < 1 ||
      scan->n_rows > 65535)
    return set_err(ctx, "bad scan geometry"), FCT_ERR_INVALID_ARGUMENT;
  if (scan->ffs_mode < FCT_FFS_NONE || scan->ffs_mode > FCT_FFS_PHI_Z ||
      (scan->ffs_mode != FCT_FFS_NONE && !(scan->anode_angle > 0.0 && scan->anode_angle < 1.5)))
    return set_err(ctx, "bad ffs_mode / anode_angle"), FCT_ERR_INVALID_ARGUMENT;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t s = static_cast<hipStream_t>(stream);
  fct_ellipsoid* d_ell = nullptr;
  if (n_ellipsoids > 0) {
    HIP_TRY(ctx, hipMalloc(reinterpret_cast<void**>(&d_ell), n_ellipsoids * sizeof(fct_ellipsoid)));
    hipError_t e = hipMemcpyAsync(d_ell, ellipsoids, n_ellipsoids * sizeof(fct_ellipsoid),
                                  hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      (void)hipFree(d_ell);
      set_err(ctx, "ellipsoid upload: %s", hipGetErrorString(e));
      return FCT_ERR_DEVICE;
    }
  }
  fct::PhantomArgs a{};
  a.out = raw;
  a.ell = d_ell;
  a.n_ell = n_ellipsoids;
  a.ncols = scan->n_channels;
  a.nrows = scan->n_rows;
  a.v0 = v0;
  a.v1 = v1;
  a.r = scan->r;
  a.dsd = scan->dsd;
  a.dgamma = scan->channel_spacing / scan->dsd;
  a.drow = scan->row_spacing;
  a.cc = scan->central_channel;
  a.cr = scan->central_row;
  // Clockwise: the rays of the counter-clockwise scan at angles -a_v,
  // mirrored y -> -y by the kernel.
  a.mirror = scan->rotation < 0 ? 1 : 0;
  a.angle0 = a.mirror ? -scan->angle0 : scan->angle0;
  a.dtheta = 2.0 * kPi / (double)scan->views_per_rotation;
  a.z0 = scan->z0;
  a.dz = scan->feed / (double)scan->views_per_rotation;
  a.cyl_r = cyl_radius;
  a.cyl_mu = cyl_mu;
  a.i0 = i0;
  a.seed = seed;
  const fct_ffs::Geometry fg = fct_ffs::make_geometry(
      scan->ffs_mode, scan->n_channels, scan->n_rows, scan->n_views, scan->views_per_rotation,
      scan->channel_spacing, scan->row_spacing, scan->central_channel, scan->central_row,
      scan->r, scan->dsd, scan->anode_angle);
  a.states = fg.S;
  for (int rho = 0; rho < fg.S; ++rho) {
    const fct_ffs::Deflection d = fct_ffs::state_offset(fg.mode, rho, fg.dr, fg.da, fg.anode);
    a.ffs_off[rho][0] = d.dr;
    a.ffs_off[rho][1] = d.da;
    a.ffs_off[rho][2] = d.dz;
  }
  hipError_t e = fct::launch_phantom(a, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (d_ell) (void)hipFree(d_ell);
  if (e != hipSuccess) {
    set_err(ctx, "phantom_project: %s", hipGetErrorString(e));
    return FCT_ERR_DEVICE;
  }
  return FCT_OK;
}

}  // extern "C"
