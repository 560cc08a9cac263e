// fct_wfbp command-line driver (main.cpp:20-66 of FreeCT's fct_wfbp):
//   fct_wfbp [-v] [-t] [-d <device>] [-o <file>] [--precision fp32|fp16|bf16] recon.yaml
// plus the synthetic-scan writer that stands in for fct_read_util:
//   fct_wfbp --write-synthetic <small|full|adjoint> <dir> [--seed n]
//   fct_wfbp --print-phantom <small|full|adjoint> [--seed n]
//   fct_wfbp --print-config recon.yaml      (parsed recon keys, JSON)
//   fct_wfbp --inspect <freect or DICOM-CT-PD dir>  (reader + derived scan law, JSON)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "fct_host.h"

namespace {

void usage() {
  std::cout << "Usage: fct_wfbp [options] reconstruction_configuration.yaml\n"
               "           -v: verbose\n"
               "     -d <int>: GPU device to utilize\n"
               "    -o <path>: output file (default reconstruction.dat)\n"
               "  --precision fp32|fp16|bf16: filtered-projection storage\n"
               "       fct_wfbp --write-synthetic <small|full|adjoint> <dir> [--seed n] [-d n]\n"
               "       fct_wfbp --print-phantom <small|full|adjoint> [--seed n]\n";
}

int fail(const std::string& what, const std::string& msg) {
  std::cout << "ERROR: " << what << (msg.empty() ? "" : ": " + msg) << std::endl;
  return 1;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    usage();
    return 1;
  }
  bool verbose = false;
  int device = 0;
  std::string out = "reconstruction.dat", precision, synth, synth_dir, phantom, print_cfg, inspect;
  uint64_t seed = 0;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* what) -> std::string {
      if (i + 1 >= argc) {
        std::cout << "ERROR: " << what << " needs a value" << std::endl;
        std::exit(1);
      }
      return argv[++i];
    };
    if (a == "-v") verbose = true;
    else if (a == "-t") continue;
    else if (a == "-d") device = std::stoi(next("-d"));
    else if (a == "-o") out = next("-o");
    else if (a == "--precision") precision = next("--precision");
    else if (a == "--seed") seed = std::stoull(next("--seed"));
    else if (a == "--write-synthetic") {
      synth = next("--write-synthetic");
      synth_dir = next("--write-synthetic");
    } else if (a == "--print-phantom") phantom = next("--print-phantom");
    else if (a == "--print-config") print_cfg = next("--print-config");
    else if (a == "--inspect") inspect = next("--inspect");
    else if (!a.empty() && a[0] == '-') {
      std::cout << "ERROR: Unrecognized argument \"" << a << "\"" << std::endl;
      usage();
      return 1;
    } else pos.push_back(a);
  }

  if (!phantom.empty()) {
    fct_scan_geometry g;
    fct_recon_config rc;
    double i0;
    std::vector<fct_ellipsoid> ell;
    uint64_t used;
    if (!fct::fct_synthetic_workload(phantom, seed, &g, &rc, &i0, &ell, &used))
      return fail("unknown workload", phantom);
    std::printf("{\"seed\": %llu, \"ellipsoids\": [", (unsigned long long)used);
    for (size_t k = 0; k < ell.size(); ++k) {
      const fct_ellipsoid& e = ell[k];
      std::printf("%s{\"c\": [%.17g, %.17g, %.17g], \"a\": [%.17g, %.17g, %.17g], \"rot\": [",
                  k ? ", " : "", e.c[0], e.c[1], e.c[2], e.a[0], e.a[1], e.a[2]);
      for (int j = 0; j < 9; ++j) std::printf("%s%.17g", j ? ", " : "", e.rot[j]);
      std::printf("], \"dmu\": %.17g}", e.dmu);
    }
    std::printf("]}\n");
    return 0;
  }

  if (!print_cfg.empty()) {
    ReconConfig rp;
    std::string err;
    if (parse_config(print_cfg, rp, verbose, &err)) return fail("parse_config", err);
    std::printf("{\"raw_data_dir\": \"%s\", \"output_dir\": \"%s\", \"output_file\": \"%s\", "
                "\"start_pos\": %.9g, \"end_pos\": %.9g, \"recon_fov\": %.9g, "
                "\"slice_thickness\": %.9g, \"slice_pitch\": %.9g, \"nx\": %zu, \"ny\": %zu, "
                "\"recon_kernel\": %d, \"x_origin\": %.9g, \"y_origin\": %.9g, "
                "\"tube_angle_offset\": %.9g, \"adaptive_filtration_s\": %.9g, "
                "\"filter\": \"%s\", \"precision\": \"%s\", \"anode_angle\": %.9g, "
                "\"adaptive_filtration\": %d, \"output_units\": \"%s\", \"mu_water\": %.9g}\n",
                rp.raw_data_dir.c_str(), rp.output_dir.c_str(), rp.output_file.c_str(),
                rp.start_pos, rp.end_pos, rp.recon_fov, rp.slice_thickness, rp.slice_pitch, rp.nx,
                rp.ny, rp.recon_kernel, rp.x_origin, rp.y_origin, rp.tube_angle_offset,
                rp.adaptive_filtration_s, rp.filter.c_str(), rp.precision.c_str(), rp.anode_angle,
                rp.adaptive_filtration, rp.output_units.c_str(), rp.mu_water);
    return 0;
  }

  if (!inspect.empty()) {
    auto ds = fct::open_data_set(inspect, device);
    ds->setPath(inspect);
    if (ds->initialize() || ds->readAll()) return fail("FreeCTDataSet", ds->m_last_error);
    double sum = 0.0;
    std::vector<float> buf((size_t)ds->getDetectorRows() * ds->getDetectorChannels());
    for (int v = 0; v < ds->getTotalNumProjections(); ++v) {
      ds->copyProjection(v, buf.data());
      for (float x : buf) sum += x;
    }
    fct::ReconstructionMachine rm;
    rm.config().filter = "ramp";
    rm.setDataSet(ds);
    if (rm.ConfigureCTGeometry()) return fail("ConfigureCTGeometry", rm.lastError());
    const fct_scan_geometry& g = rm.geometry();
    std::printf("{\"n_views\": %d, \"n_rows\": %d, \"n_channels\": %d, \"ppr\": %d, "
                "\"angle0\": %.17g, \"z0\": %.17g, \"feed\": %.17g, \"dsd\": %.17g, \"r\": %.17g, "
                "\"channel_spacing\": %.17g, \"row_spacing\": %.17g, \"central_channel\": %.17g, "
                "\"central_row\": %.17g, \"sum\": %.17g, \"angle_last\": %.9g, \"z_last\": %.9g, "
                "\"rotation\": %d, \"ffs_mode\": %d, \"anode_angle\": %.17g, "
                "\"manufacturer\": \"%s\", \"detector_shape\": \"%s\"}\n",
                g.n_views, g.n_rows, g.n_channels, g.views_per_rotation, g.angle0, g.z0, g.feed,
                g.dsd, g.r, g.channel_spacing, g.row_spacing, g.central_channel, g.central_row,
                sum, ds->getTubeAngle(g.n_views - 1), ds->getTablePosition(g.n_views - 1),
                g.rotation, g.ffs_mode, g.anode_angle, ds->getManufacturer().c_str(),
                ds->getDetectorShape().c_str());
    return 0;
  }

  if (!synth.empty()) {
    fct::SyntheticDataSet ds(device);
    ds.setPath("synthetic:" + synth + (seed ? ":seed=" + std::to_string(seed) : ""));
    if (ds.initialize() || ds.readAll()) return fail("synthetic scan", ds.m_last_error);
    if (ds.writeAll(synth_dir)) return fail("writing", synth_dir);
    // Recon YAML for the written scan (writeReconFile keys + this engine's).
    const fct_recon_config& r = ds.suggestedRecon();
    FILE* f = std::fopen((synth_dir + "/recon.yaml").c_str(), "w");
    if (!f) return fail("writing", synth_dir + "/recon.yaml");
    std::fprintf(f,
                 "raw_data_dir: %s\noutput_dir: %s\nstart_pos: %.9g\nend_pos: %.9g\n"
                 "recon_fov: %.9g\nslice_thickness: %.9g\nslice_pitch: %.9g\nnx: %d\nny: %d\n"
                 "recon_kernel: 3\nx_origin: 0\ny_origin: 0\ntube_angle_offset: 0\n"
                 "adaptive_filtration_s: 0.2\nfilter: %s\n",
                 synth_dir.c_str(), synth_dir.c_str(), r.start_pos, r.end_pos, r.recon_fov,
                 r.slice_thickness, r.slice_pitch, r.nx, r.ny,
                 r.filter_kind == 1 ? "shepp-logan" : "ramp");
    std::fclose(f);
    if (verbose) ds.printMetadata();
    return 0;
  }

  if (pos.size() != 1) {
    usage();
    return 1;
  }
  fct::ReconstructionMachine rm;
  rm.setVerbose(verbose);
  rm.setDevice(device);
  if (rm.LoadReconConfiguration(pos[0])) return fail("LoadReconConfiguration", rm.lastError());
  if (!precision.empty()) rm.config().precision = precision;
  if (rm.LoadRawData()) return fail("LoadRawData", rm.lastError());
  if (rm.ConfigureCTGeometry()) return fail("ConfigureCTGeometry", rm.lastError());
  if (verbose) rm.PrintCTGeometry();
  if (rm.RunReconstruction()) return fail("RunReconstruction", rm.lastError());
  if (rm.SaveReconstruction(out)) return fail("SaveReconstruction", rm.lastError());
  return 0;
}
