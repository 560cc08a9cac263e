// fct_host.cpp -- host mirror of FreeCT's fct_wfbp path above the C ABI.
#include "fct_host.h"

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <sys/stat.h>

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kMuWater = 0.0193;

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

struct FilterSpec {
  int kind;
  double c, a;
};

bool filter_spec(const std::string& name, FilterSpec* f) {
  // ReconstructionMachine.cu:493-506 / ramp.m parameter sets; sharp, medium and
  // smooth are those of FreeCT's resources/filters vectors.
  if (name == "ramp" || name == "sharp") *f = {0, 1.0, 1.0};
  else if (name == "medium") *f = {0, 0.65, 0.5};
  else if (name == "smooth") *f = {0, 0.4, 0.5};
  else if (name == "hann") *f = {0, 1.0, 0.5};
  else if (name == "hamming") *f = {0, 1.0, 0.54};
  else if (name == "shepp-logan") *f = {1, 1.0, 1.0};
  else return false;
  return true;
}

int precision_code(const std::string& p) {
  if (p == "fp32") return FCT_PRECISION_FP32;
  if (p == "fp16") return FCT_PRECISION_FP16;
  if (p == "bf16") return FCT_PRECISION_BF16;
  return -1;
}

// SplitMix64 (Steele, Lea, Flood 2014).
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform(double lo = 0.0, double hi = 1.0) {
    return lo + (hi - lo) * ((double)(next() >> 11) * (1.0 / 9007199254740992.0));
  }
};

void matmul3(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

}  // namespace

std::map<std::string, std::string> fct_read_flat_yaml(const std::string& path, std::string* err) {
  std::map<std::string, std::string> m;
  std::ifstream in(path);
  if (!in) {
    if (err) *err = "cannot open " + path;
    return m;
  }
  std::string line;
  while (std::getline(in, line)) {
    bool inq = false;
    char q = 0;
    for (size_t i = 0; i < line.size(); ++i) {
      const char ch = line[i];
      if ((ch == '"' || ch == '\'') && (!inq || ch == q)) {
        inq = !inq;
        q = ch;
      } else if (ch == '#' && !inq) {
        line = line.substr(0, i);
        break;
      }
    }
    const size_t c = line.find(':');
    if (c == std::string::npos) continue;
    std::string k = trim(line.substr(0, c)), v = trim(line.substr(c + 1));
    if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front())
      v = v.substr(1, v.size() - 2);
    if (!k.empty()) m[k] = v;
  }
  return m;
}

int parse_config(const std::string& config_file, ReconConfig& rp, bool verbose, std::string* err) {
  std::string e;
  auto m = fct_read_flat_yaml(config_file, &e);
  if (!e.empty()) {
    if (err) *err = e;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  if (verbose) std::cout << "Config File: " << config_file << "\n" << std::endl;
  auto report = [&](const std::string& key, bool found, const std::string& val) {
    if (!verbose) return;
    std::cout << key << ": " << std::string(key.size() < 34 ? 34 - key.size() : 1, ' ')
              << (found ? "FOUND (" + val + ")" : std::string("NOT FOUND")) << std::endl;
  };
  auto get = [&](const char* key, auto& field) -> bool {
    auto it = m.find(key);
    if (it == m.end()) {
      report(key, false, "");
      return true;
    }
    std::istringstream is(it->second);
    using T = std::decay_t<decltype(field)>;
    T v{};
    if constexpr (std::is_same_v<T, std::string>) {
      v = it->second;
    } else {
      double d;
      if (!(is >> d)) {
        if (err) *err = std::string("bad value for ") + key + ": " + it->second;
        return false;
      }
      v = (T)d;
    }
    field = v;
    report(key, true, it->second);
    return true;
  };
  bool ok = get("raw_data_dir", rp.raw_data_dir) && get("output_dir", rp.output_dir) &&
            get("start_pos", rp.start_pos) && get("end_pos", rp.end_pos) &&
            get("recon_fov", rp.recon_fov) && get("slice_thickness", rp.slice_thickness) &&
            get("slice_pitch", rp.slice_pitch) && get("nx", rp.nx) && get("ny", rp.ny) &&
            get("recon_kernel", rp.recon_kernel) && get("x_origin", rp.x_origin) &&
            get("y_origin", rp.y_origin) && get("tube_angle_offset", rp.tube_angle_offset) &&
            get("adaptive_filtration_s", rp.adaptive_filtration_s) && get("filter", rp.filter) &&
            get("precision", rp.precision) && get("anode_angle", rp.anode_angle) &&
            get("adaptive_filtration", rp.adaptive_filtration) &&
            get("output_units", rp.output_units) && get("mu_water", rp.mu_water);
  if (!ok) return FCT_ERR_INVALID_ARGUMENT;
  rp.output_file = config_file + "_recon.dat";        // parse_config.cpp:43
  if (rp.slice_pitch == 0) rp.slice_pitch = rp.slice_thickness;   // parse_config.cpp:59-60
  return FCT_OK;
}

namespace fct {

// ---------------------------------------------------------------------------
// RawDataSet
// ---------------------------------------------------------------------------
bool RawDataSet::projectionIsAvailable(int i) const {
  return i >= 0 && i < (int)m_total_num_projections && i < (int)m_is_loaded.size() &&
         m_is_loaded[i];
}

bool RawDataSet::headerIsAvailable(int i) const {
  return projectionIsAvailable(i) ||
         (i >= 0 && i < (int)m_total_num_projections && i < (int)m_has_header.size() &&
          m_has_header[i]);
}

float RawDataSet::getTablePosition(int i) const {
  return headerIsAvailable(i) ? m_data[i]->m_dfc_axial_position : NAN;
}

float RawDataSet::getTubeAngle(int i) const {
  return headerIsAvailable(i) ? m_data[i]->m_dfc_angular_position : NAN;
}

void RawDataSet::copyProjection(int i, float* buffer) const {
  std::memcpy(buffer, m_data[i]->m_projection.data(),
              (size_t)m_detector_rows * m_detector_channels * sizeof(float));
}

int RawDataSet::readViews(int v0, int n, float* dst) {
  const size_t per = (size_t)m_detector_rows * m_detector_channels;
  for (int v = v0; v < v0 + n; ++v) {
    if (!projectionIsAvailable(v)) {
      const int st = readProjection(v);
      if (st) return st;
    }
    copyProjection(v, dst + per * (size_t)(v - v0));
  }
  return FCT_OK;
}

void RawDataSet::printMetadata() const {
  std::cout << "manufacturer: " << m_manufacturer << "\n"
            << "detector_rows: " << m_detector_rows << "\n"
            << "detector_channels: " << m_detector_channels << "\n"
            << "detector_transverse_spacing: " << m_detector_transverse_spacing << "\n"
            << "detector_axial_spacing: " << m_detector_axial_spacing << "\n"
            << "detector_shape: " << m_detector_shape << "\n"
            << "distance_source_to_detector: " << m_dist_source_to_detector << "\n"
            << "distance_source_to_isocenter: " << m_dist_source_to_isocenter << "\n"
            << "detector_central_row: " << m_detector_central_row << "\n"
            << "detector_central_channel: " << m_detector_central_channel << "\n"
            << "projection_geometry: " << m_projection_geometry << "\n"
            << "scan_type: " << m_scan_type << "\n"
            << "projections_per_rotation: " << m_projections_per_rotation << "\n"
            << "flying_focal_spot_mode: " << m_flying_focal_spot_mode << "\n"
            << "total_num_projections: " << m_total_num_projections << std::endl;
}

// Starter recon YAML with the keys parse_config reads (RawDataSet.cpp:47-83).
int RawDataSet::writeReconFile(const std::string& filepath, bool single_slice_mode) const {
  if (m_data.empty() || !m_data.front() || !m_data.back()) return FCT_ERR_INVALID_ARGUMENT;
  std::ofstream o(filepath);
  if (!o) return FCT_ERR_INVALID_ARGUMENT;
  const float z0 = m_data.front()->getDFCAxialPosition(), z1 = m_data.back()->getDFCAxialPosition();
  o << std::setprecision(9);
  o << "raw_data_dir: " << m_path << "\n";
  o << "output_dir: ./\n";
  o << "start_pos: " << (single_slice_mode ? 0.5f * (z0 + z1) : z0) << "\n";
  o << "end_pos: " << (single_slice_mode ? 0.5f * (z0 + z1) : z1) << "\n";
  o << "recon_fov: 500\nslice_thickness: 1\nnx: 512\nny: 512\nrecon_kernel: 3\n";
  o << "x_origin: 0\ny_origin: 0\ntube_angle_offset: 0\nadaptive_filtration_s: 0.2\n";
  return o ? FCT_OK : FCT_ERR_INVALID_ARGUMENT;
}

// meta.yaml + projections.dat (RawDataSet.cpp:17-45, 85-107).
int RawDataSet::writeAll(const std::string& dirpath) const {
  std::ofstream meta(dirpath + "/meta.yaml");
  if (!meta) return FCT_ERR_INVALID_ARGUMENT;
  meta << std::setprecision(9);
  meta << "# FREECT READER OUTPUT FOR: " << m_path << "\n\n# DETECTOR\n"
       << "manufacturer: " << m_manufacturer << "\n"
       << "detector_rows: " << m_detector_rows << "\n"
       << "detector_channels: " << m_detector_channels << "\n"
       << "detector_transverse_spacing: " << m_detector_transverse_spacing << "\n"
       << "detector_axial_spacing: " << m_detector_axial_spacing << "\n"
       << "detector_shape: " << m_detector_shape << "\n\n# SCANNER GEOMETRY\n"
       << "distance_source_to_detector: " << m_dist_source_to_detector << "\n"
       << "distance_source_to_isocenter: " << m_dist_source_to_isocenter << "\n"
       << "detector_central_row: " << m_detector_central_row << "\n"
       << "detector_central_channel: " << m_detector_central_channel << "\n"
       << "projection_geometry: " << m_projection_geometry << "\n\n# SCAN\n"
       << "scan_type: " << m_scan_type << "\n"
       << "projections_per_rotation: " << m_projections_per_rotation << "\n"
       << "flying_focal_spot_mode: " << m_flying_focal_spot_mode << "\n"
       << "total_num_projections: " << m_total_num_projections << "\n";
  std::ofstream out(dirpath + "/projections.dat", std::ios::binary | std::ios::trunc);
  if (!out) return FCT_ERR_INVALID_ARGUMENT;
  for (const auto& f : m_data) {
    const float hdr[3] = {f->getDFCAngularPosition() + f->getFFSAngularShift(),
                          f->getDFCAxialPosition() + f->getFFSAxialShift(),
                          f->getDFCRadialPosition() + f->getFFSRadialShift()};
    out.write(reinterpret_cast<const char*>(hdr), sizeof hdr);
    out.write(reinterpret_cast<const char*>(f->m_projection.data()),
              f->m_projection.size() * sizeof(float));
  }
  return out ? FCT_OK : FCT_ERR_INVALID_ARGUMENT;
}

// ---------------------------------------------------------------------------
// FreeCTDataSet
// ---------------------------------------------------------------------------
int FreeCTDataSet::initialize() { return readMetadata(); }

int FreeCTDataSet::readMetadata() {
  std::string e;
  auto m = fct_read_flat_yaml(m_path + "/meta.yaml", &e);
  if (!e.empty()) {
    m_last_error = e;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  const char* keys[] = {"manufacturer", "detector_rows", "detector_channels",
                        "detector_transverse_spacing", "detector_axial_spacing",
                        "detector_shape", "distance_source_to_detector",
                        "distance_source_to_isocenter", "detector_central_row",
                        "detector_central_channel", "projection_geometry", "scan_type",
                        "projections_per_rotation", "flying_focal_spot_mode",
                        "total_num_projections"};
  for (const char* k : keys)
    if (!m.count(k)) {
      m_last_error = std::string("meta.yaml: missing key ") + k;
      return FCT_ERR_INVALID_ARGUMENT;
    }
  m_manufacturer = m["manufacturer"];
  m_detector_rows = (uint16_t)std::stoul(m["detector_rows"]);
  m_detector_channels = (uint16_t)std::stoul(m["detector_channels"]);
  m_detector_transverse_spacing = std::stof(m["detector_transverse_spacing"]);
  m_detector_axial_spacing = std::stof(m["detector_axial_spacing"]);
  m_detector_shape = m["detector_shape"];
  m_dist_source_to_detector = std::stof(m["distance_source_to_detector"]);
  m_dist_source_to_isocenter = std::stof(m["distance_source_to_isocenter"]);
  m_detector_central_row = std::stof(m["detector_central_row"]);
  m_detector_central_channel = std::stof(m["detector_central_channel"]);
  m_projection_geometry = m["projection_geometry"];
  m_scan_type = m["scan_type"];
  m_projections_per_rotation = (uint16_t)std::stoul(m["projections_per_rotation"]);
  m_flying_focal_spot_mode = m["flying_focal_spot_mode"];
  m_total_num_projections = (uint16_t)std::stoul(m["total_num_projections"]);
  m_data.clear();
  m_data.resize(m_total_num_projections);
  m_is_loaded.assign(m_total_num_projections, false);
  m_has_header.assign(m_total_num_projections, false);
  m_in.reset();
  m_next_frame = -1;
  return FCT_OK;
}

int FreeCTDataSet::readHeaders() {
  std::ifstream in(m_path + "/projections.dat", std::ios::binary);
  if (!in) {
    m_last_error = "cannot open " + m_path + "/projections.dat";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  const size_t frame = 3 * sizeof(float) + (size_t)m_detector_rows * m_detector_channels * sizeof(float);
  in.seekg(0, std::ios::end);
  if ((size_t)in.tellg() < frame * m_total_num_projections) {
    m_last_error = "projections.dat is shorter than total_num_projections frames";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  for (int i = 0; i < (int)m_total_num_projections; ++i) {
    float hdr[3];
    in.seekg((std::streamoff)(frame * (size_t)i));
    in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    if (!in) {
      m_last_error = "cannot read the header of frame " + std::to_string(i);
      return FCT_ERR_INVALID_ARGUMENT;
    }
    if (!m_data[i]) m_data[i] = std::make_unique<RawDataFrame>();
    RawDataFrame& f = *m_data[i];
    f.m_detector_rows = m_detector_rows;
    f.m_detector_channels = m_detector_channels;
    f.m_dfc_angular_position = hdr[0];
    f.m_dfc_axial_position = hdr[1];
    f.m_dfc_radial_distance = hdr[2];
    m_has_header[i] = true;
  }
  return FCT_OK;
}

std::ifstream* FreeCTDataSet::stream() {
  if (!m_in) {
    m_in = std::make_unique<std::ifstream>(m_path + "/projections.dat", std::ios::binary);
    m_next_frame = -1;
  }
  return *m_in ? m_in.get() : nullptr;
}

int FreeCTDataSet::readViews(int v0, int n, float* dst) {
  if (v0 < 0 || n < 0 || v0 + n > (int)m_total_num_projections) return FCT_ERR_INVALID_ARGUMENT;
  const size_t per = (size_t)m_detector_rows * m_detector_channels;
  const size_t frame = 3 * sizeof(float) + per * sizeof(float);
  for (int v = v0; v < v0 + n; ++v) {
    float* out = dst + per * (size_t)(v - v0);
    if (projectionIsAvailable(v)) {
      copyProjection(v, out);
      continue;
    }
    std::ifstream* in = stream();
    if (!in) {
      m_last_error = "cannot open " + m_path + "/projections.dat";
      return FCT_ERR_INVALID_ARGUMENT;
    }
    if (m_next_frame != v) in->seekg((std::streamoff)(frame * (size_t)v));
    float hdr[3];
    in->read(reinterpret_cast<char*>(hdr), sizeof hdr);
    in->read(reinterpret_cast<char*>(out), per * sizeof(float));
    if (!*in) {
      m_last_error = "projections.dat is shorter than total_num_projections frames";
      m_in.reset();
      return FCT_ERR_INVALID_ARGUMENT;
    }
    m_next_frame = v + 1;
  }
  return FCT_OK;
}

int FreeCTDataSet::readProjection(int i) {
  if (i < 0 || i >= (int)m_total_num_projections) return FCT_ERR_INVALID_ARGUMENT;
  std::ifstream in(m_path + "/projections.dat", std::ios::binary);
  if (!in) {
    m_last_error = "cannot open " + m_path + "/projections.dat";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  const size_t n = (size_t)m_detector_rows * m_detector_channels;
  const size_t frame = 3 * sizeof(float) + n * sizeof(float);
  in.seekg((std::streamoff)(frame * (size_t)i));
  auto f = std::make_unique<RawDataFrame>();
  float hdr[3];
  f->m_projection.resize(n);
  in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
  in.read(reinterpret_cast<char*>(f->m_projection.data()), n * sizeof(float));
  if (!in) {
    m_last_error = "projections.dat is shorter than total_num_projections frames";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  f->m_detector_rows = m_detector_rows;
  f->m_detector_channels = m_detector_channels;
  f->m_dfc_angular_position = hdr[0];
  f->m_dfc_axial_position = hdr[1];
  f->m_dfc_radial_distance = hdr[2];
  m_data[i] = std::move(f);
  m_is_loaded[i] = true;
  return FCT_OK;
}

int FreeCTDataSet::readAll() {
  std::ifstream in(m_path + "/projections.dat", std::ios::binary);
  if (!in) {
    m_last_error = "cannot open " + m_path + "/projections.dat";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  const size_t n = (size_t)m_detector_rows * m_detector_channels;
  for (int i = 0; i < (int)m_total_num_projections; ++i) {
    auto f = std::make_unique<RawDataFrame>();
    float hdr[3];
    f->m_projection.resize(n);
    in.read(reinterpret_cast<char*>(hdr), sizeof hdr);
    in.read(reinterpret_cast<char*>(f->m_projection.data()), n * sizeof(float));
    if (!in) {
      m_last_error = "projections.dat is shorter than total_num_projections frames";
      return FCT_ERR_INVALID_ARGUMENT;
    }
    f->m_detector_rows = m_detector_rows;
    f->m_detector_channels = m_detector_channels;
    f->m_dfc_angular_position = hdr[0];
    f->m_dfc_axial_position = hdr[1];
    f->m_dfc_radial_distance = hdr[2];
    m_data[i] = std::move(f);
    m_is_loaded[i] = true;
  }
  return FCT_OK;
}

// ---------------------------------------------------------------------------
// Synthetic scans
// ---------------------------------------------------------------------------
void fct_make_phantom(uint64_t seed, double z_lo, double z_hi, int n,
                      std::vector<fct_ellipsoid>* out) {
  SplitMix64 rng{seed};
  out->clear();
  for (int i = 0; i < n; ++i) {
    fct_ellipsoid e{};
    const double rad = 150.0 * std::sqrt(rng.uniform());
    const double ang = rng.uniform(0.0, 2.0 * kPi);
    e.c[0] = rad * std::cos(ang);
    e.c[1] = rad * std::sin(ang);
    e.c[2] = rng.uniform(z_lo, z_hi);
    for (int k = 0; k < 3; ++k) e.a[k] = rng.uniform(5.0, 60.0);
    const double phi = rng.uniform(0, 2 * kPi), th = rng.uniform(0, kPi), psi = rng.uniform(0, 2 * kPi);
    const double Rz1[9] = {std::cos(phi), -std::sin(phi), 0, std::sin(phi), std::cos(phi), 0, 0, 0, 1};
    const double Ry[9] = {std::cos(th), 0, std::sin(th), 0, 1, 0, -std::sin(th), 0, std::cos(th)};
    const double Rz2[9] = {std::cos(psi), -std::sin(psi), 0, std::sin(psi), std::cos(psi), 0, 0, 0, 1};
    double T[9];
    matmul3(Rz1, Ry, T);
    matmul3(T, Rz2, e.rot);
    const double hu = rng.uniform(-900.0, 1500.0);
    e.dmu = kMuWater * hu / 1000.0;
    out->push_back(e);
  }
}

bool fct_synthetic_workload(const std::string& name, uint64_t seed, fct_scan_geometry* g,
                            fct_recon_config* rc, double* i0, std::vector<fct_ellipsoid>* ell,
                            uint64_t* seed_used) {
  double R = 541.0, Dsd = 949.0, drow_iso = 0.625, cc_off = 0.25;
  int nch, nrows, ppr, ffs = 0;
  double dgamma, pitch, rotations, zlo, zhi, fov, thick;
  int nx;
  std::string filt;
  uint64_t s;
  if (name == "ffs") {   // scans.ffs(): Siemens-like, phi + z FFS (definitionas.scanner:5-14)
    R = 595.0; Dsd = 1085.6; drow_iso = 0.6; cc_off = 366.25 - 367.5;
    nch = 736; nrows = 32; ppr = 2304; dgamma = 0.0011844; ffs = FCT_FFS_PHI_Z;
    pitch = 0.8; rotations = 8; zlo = -100; zhi = 100; fov = 500; thick = 1.0; nx = 512;
    filt = "sharp"; s = 1002; *i0 = 2.5e4;
  } else if (name == "small") {
    nch = 736; nrows = 16; ppr = 1152; dgamma = 2.0 * std::asin(250.0 / R) / 735.0;
    pitch = 1.0; rotations = 2; zlo = -20; zhi = 20; fov = 250; thick = 1.25; nx = 256;
    filt = "shepp-logan"; s = 1000; *i0 = 0.0;
  } else if (name == "full") {
    nch = 888; nrows = 64; ppr = 984; dgamma = 0.0010831;
    pitch = 0.984; rotations = 10; zlo = -100; zhi = 100; fov = 500; thick = 0.625; nx = 512;
    filt = "ramp"; s = 1001; *i0 = 5e4;
  } else if (name == "adjoint") {   // scans.adjoint(): the small detector at 984 views/rot
    nch = 736; nrows = 16; ppr = 984; dgamma = 2.0 * std::asin(250.0 / R) / 735.0;
    pitch = 1.0; rotations = 3; zlo = -20; zhi = 20; fov = 500; thick = 0.625; nx = 512;
    filt = "ramp"; s = 1004; *i0 = 0.0;
  } else {
    return false;
  }
  if (seed) s = seed;
  *seed_used = s;
  std::memset(g, 0, sizeof *g);
  g->n_channels = nch;
  g->n_rows = nrows;
  g->views_per_rotation = ppr;
  g->n_views = (int)std::lround(rotations * ppr);
  g->channel_spacing = dgamma * Dsd;
  g->row_spacing = drow_iso * Dsd / R;
  g->dsd = Dsd;
  g->r = R;
  g->central_channel = 0.5 * (nch - 1) + cc_off;
  g->central_row = 0.5 * (nrows - 1);
  g->ffs_mode = ffs;
  g->anode_angle = ffs ? 0.1222 : 0.0;
  const double csw = R * g->row_spacing / Dsd;
  g->feed = pitch * nrows * csw;
  g->z0 = -0.5 * (g->n_views - 1) * g->feed / ppr;
  g->angle0 = 0.0;
  std::memset(rc, 0, sizeof *rc);
  rc->start_pos = zlo;
  rc->end_pos = zhi;
  rc->recon_fov = fov;
  rc->slice_thickness = thick;
  rc->slice_pitch = thick;
  rc->nx = rc->ny = nx;
  FilterSpec f{};
  filter_spec(filt, &f);
  rc->filter_kind = f.kind;
  rc->filter_c = f.c;
  rc->filter_a = f.a;
  fct_make_phantom(s, zlo, zhi, 20, ell);
  return true;
}

int SyntheticDataSet::initialize() {
  // m_path = "synthetic:<workload>[:seed=<n>]"
  std::string rest = m_path.substr(std::string("synthetic:").size());
  uint64_t seed = 0;
  const size_t c = rest.find(':');
  if (c != std::string::npos) {
    const std::string opt = rest.substr(c + 1);
    rest = rest.substr(0, c);
    if (opt.rfind("seed=", 0) == 0) seed = std::stoull(opt.substr(5));
  }
  m_workload = rest;
  if (!fct_synthetic_workload(m_workload, seed, &m_geom, &m_recon, &m_i0, &m_ellipsoids, &m_seed)) {
    m_last_error = "unknown synthetic workload '" + m_workload + "' (small, full, adjoint, ffs)";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  m_manufacturer = "SYNTHETIC";
  // The DICOM-CT-PD names the reader maps back (DicomDataSet.cpp:193, 285).
  m_flying_focal_spot_mode = m_geom.ffs_mode == FCT_FFS_PHI_Z ? "FFSXYZ"
                             : m_geom.ffs_mode == FCT_FFS_PHI ? "FFSXY"
                             : m_geom.ffs_mode == FCT_FFS_Z   ? "FFSZ"
                                                              : "FFSNONE";
  m_detector_rows = (uint16_t)m_geom.n_rows;
  m_detector_channels = (uint16_t)m_geom.n_channels;
  m_detector_transverse_spacing = (float)m_geom.channel_spacing;
  m_detector_axial_spacing = (float)m_geom.row_spacing;
  m_dist_source_to_detector = (float)m_geom.dsd;
  m_dist_source_to_isocenter = (float)m_geom.r;
  m_detector_central_row = (float)m_geom.central_row;
  m_detector_central_channel = (float)m_geom.central_channel;
  m_projections_per_rotation = (uint16_t)m_geom.views_per_rotation;
  m_total_num_projections = (uint16_t)m_geom.n_views;
  m_data.clear();
  m_data.resize(m_total_num_projections);
  m_is_loaded.assign(m_total_num_projections, false);
  return FCT_OK;
}

int SyntheticDataSet::readProjection(int i) {
  if (i < 0 || i >= (int)m_total_num_projections) return FCT_ERR_INVALID_ARGUMENT;
  return readAll();
}

int SyntheticDataSet::readAll() {
  fct_wfbp_gpu_ctx* ctx = nullptr;
  int st = fct_wfbp_gpu_create(m_device, &ctx);
  if (st) {
    m_last_error = "no GPU for the synthetic scan generator";
    return st;
  }
  const size_t per_view = (size_t)m_geom.n_rows * m_geom.n_channels;
  const int chunk = 1024;
  float* d = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&d), per_view * chunk * sizeof(float)) != hipSuccess) {
    fct_wfbp_gpu_destroy(ctx);
    m_last_error = "hipMalloc failed";
    return FCT_ERR_OUT_OF_MEMORY;
  }
  std::vector<float> host(per_view * chunk);
  const double dth = 2.0 * kPi / m_geom.views_per_rotation, dz = m_geom.feed / m_geom.views_per_rotation;
  for (int v0 = 0; v0 < m_geom.n_views && st == FCT_OK; v0 += chunk) {
    const int v1 = std::min(m_geom.n_views, v0 + chunk);
    st = fct_wfbp_gpu_phantom_project(ctx, &m_geom, 150.0, kMuWater, m_ellipsoids.data(),
                                      (int)m_ellipsoids.size(), m_i0, m_seed, v0, v1, d, nullptr);
    if (st) {
      m_last_error = fct_wfbp_gpu_last_error(ctx);
      break;
    }
    if (hipMemcpy(host.data(), d, per_view * (v1 - v0) * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      st = FCT_ERR_DEVICE;
      break;
    }
    for (int v = v0; v < v1; ++v) {
      auto f = std::make_unique<RawDataFrame>();
      f->m_projection.assign(host.begin() + (size_t)(v - v0) * per_view,
                             host.begin() + (size_t)(v - v0 + 1) * per_view);
      f->m_detector_rows = m_detector_rows;
      f->m_detector_channels = m_detector_channels;
      f->m_dfc_angular_position = (float)(m_geom.angle0 + dth * v);
      f->m_dfc_axial_position = (float)(m_geom.z0 + dz * v);
      f->m_dfc_radial_distance = (float)m_geom.r;
      m_data[v] = std::move(f);
      m_is_loaded[v] = true;
    }
  }
  (void)hipFree(d);
  fct_wfbp_gpu_destroy(ctx);
  return st;
}

// ---------------------------------------------------------------------------
// ReconstructionMachine
// ---------------------------------------------------------------------------
ReconstructionMachine::~ReconstructionMachine() = default;

int ReconstructionMachine::LoadReconConfiguration(const std::string& filepath) {
  if (m_verbose) std::cout << "ReconstructionMachine: Loading reconstruction configuration..." << std::endl;
  if (!file_exists(filepath)) {   // ReconstructionMachine.cu:28-31
    m_err = "Recon configuration filepath does not exist! (" + filepath + ")";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  m_config_path = filepath;
  return parse_config(filepath, m_rp, m_verbose, &m_err);
}

// raw_data_dir: "synthetic:<workload>", a FreeCT two-file directory
// (meta.yaml + projections.dat), or a directory of DICOM-CT-PD files.
std::shared_ptr<RawDataSet> open_data_set(const std::string& dir, int device) {
  if (dir.rfind("synthetic:", 0) == 0) return std::make_shared<SyntheticDataSet>(device);
  if (file_exists(dir + "/meta.yaml")) return std::make_shared<FreeCTDataSet>();
  return std::make_shared<DicomDataSet>();
}

int ReconstructionMachine::LoadRawData() {
  if (m_verbose) std::cout << "ReconstructionMachine: Loading raw data..." << std::endl;
  m_org_data_set = open_data_set(m_rp.raw_data_dir, m_device);
  m_org_data_set->setPath(m_rp.raw_data_dir);
  int st = m_org_data_set->initialize();
  if (!st) {
    if (auto* f = dynamic_cast<FreeCTDataSet*>(m_org_data_set.get()))
      st = f->readHeaders();   // projections stream from disk during the upload
    else
      st = m_org_data_set->readAll();
  }
  if (st) m_err = m_org_data_set->m_last_error;
  return st;
}

// The reference reads every view's tube angle and table position into
// tables (ReconstructionMachine.cu:96-102) that its backprojection indexes
// per view (cuda_kernels.cuh:230,235); its rebinning assumes uniform views
// (cuda_kernels.cuh:15-19).  The plan takes the law
// a_v = angle0 + rotation*v*2pi/ppr, z_v = z0 + v*feed/ppr.  This derives it
// from the frame headers and checks it against EVERY view: the direction and
// step from views 0 and S (S focal states; with a flying focal spot the
// headers carry each state's deflection, RawDataSet.cpp:99-101, a constant
// offset per state), the feed from the last view in view 0's state, then
// per focal state the mean residual, and every view's residual from its
// state's mean must stay within 2 % of a view step in angle and 2 % of the
// collimated slice width in z (plus two float32 ulps of the stored values).
// angle0 and z0 are the first view's values plus the mean of the states'
// offsets (the undeflected source).  A scan off that law (jittered or
// non-uniform angles or table positions, a step that is not +-2pi/ppr) is
// rejected with FCT_ERR_UNSUPPORTED rather than reconstructed wrongly.
int ReconstructionMachine::ViewLaw(const RawDataSet& ds, fct_scan_geometry* g, std::string* err) {
  constexpr double kTwoPi = 2.0 * 3.14159265358979323846;
  const int n = g->n_views, ppr = g->views_per_rotation;
  if (n < 1 || ppr < 1) {
    *err = "scan has no views or no projections per rotation";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  for (int v = 0; v < n; ++v)
    if (!ds.headerIsAvailable(v)) {
      *err = "projection " + std::to_string(v) + " is not loaded";
      return FCT_ERR_INVALID_ARGUMENT;
    }
  const int S = g->ffs_mode == FCT_FFS_PHI_Z ? 4 : g->ffs_mode ? 2 : 1;
  const double dth = kTwoPi / ppr;
  const double a0 = ds.getTubeAngle(0), z0 = ds.getTablePosition(0);
  g->rotation = 1;
  g->feed = 0.0;
  const int ref = n > S ? S : (n > 1 ? 1 : 0);
  if (ref > 0) {
    const double step = std::remainder((double)ds.getTubeAngle(ref) - a0, kTwoPi) / ref;
    if (std::fabs(step - dth) <= 0.25 * dth) {
      g->rotation = 1;
    } else if (std::fabs(step + dth) <= 0.25 * dth) {
      g->rotation = -1;
    } else {
      char b[160];
      std::snprintf(b, sizeof b, "tube angle step %.6g rad between views 0 and %d is not "
                    "+-2pi/%d", step, ref, ppr);
      *err = b;
      return FCT_ERR_UNSUPPORTED;
    }
    const int last = (n - 1) / S * S;   // last view in view 0's focal state
    if (last > 0) g->feed = ((double)ds.getTablePosition(last) - z0) * ppr / last;
  }
  // Residuals from the law through view 0, their mean per focal state, and
  // every view's deviation from its state's mean.
  std::vector<double> ra(n), rz(n), ma(S, 0.0), mz(S, 0.0);
  std::vector<int> cnt(S, 0);
  for (int v = 0; v < n; ++v) {
    ra[v] = std::remainder((double)ds.getTubeAngle(v) - (a0 + g->rotation * v * dth), kTwoPi);
    rz[v] = (double)ds.getTablePosition(v) - (z0 + v * g->feed / ppr);
    ma[v % S] += ra[v];
    mz[v % S] += rz[v];
    ++cnt[v % S];
  }
  double oa = 0.0, oz = 0.0;
  int ns = 0;
  for (int k = 0; k < S; ++k)
    if (cnt[k]) {
      ma[k] /= cnt[k];
      mz[k] /= cnt[k];
      oa += ma[k];
      oz += mz[k];
      ++ns;
    }
  g->angle0 = a0 + oa / ns;
  g->z0 = z0 + oz / ns;
  const double csw = g->r * g->row_spacing / g->dsd;
  auto ulp2 = [](float x) { return 2.0 * (double)std::nextafter(std::fabs(x), INFINITY) -
                                   2.0 * (double)std::fabs(x); };
  for (int v = 0; v < n; ++v) {
    const float av = ds.getTubeAngle(v), zv = ds.getTablePosition(v);
    const double da = ra[v] - ma[v % S], dz = rz[v] - mz[v % S];
    const double ta = 0.02 * dth + ulp2(av) + ulp2((float)a0);
    const double tz = 0.02 * csw + ulp2(zv) + ulp2((float)z0);
    if (!(std::fabs(da) <= ta) || !(std::fabs(dz) <= tz)) {
      char b[256];
      std::snprintf(b, sizeof b, "view %d is off the uniform helical law: angle off by %.3g rad "
                    "(tolerance %.3g), table position off by %.3g mm (tolerance %.3g)",
                    v, da, ta, dz, tz);
      *err = b;
      return FCT_ERR_UNSUPPORTED;
    }
  }
  if (n > 1 && !(std::fabs(g->feed) > 0)) {
    *err = "table position does not move (axial scans are not supported)";
    return FCT_ERR_UNSUPPORTED;
  }
  return FCT_OK;
}

int ReconstructionMachine::ConfigureCTGeometry() {
  if (!m_org_data_set) {
    m_err = "no raw data loaded";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  RawDataSet& ds = *m_org_data_set;
  fct_scan_geometry& g = m_geom;
  if (!ds.exactGeometry(&g)) {
    // ReconstructionMachine.cu:48-64: geometry from the getters; the view law
    // from the first view and one rotation later (signed feed).
    std::memset(&g, 0, sizeof g);
    g.n_channels = ds.getDetectorChannels();
    g.n_rows = ds.getDetectorRows();
    g.n_views = ds.getTotalNumProjections();
    g.views_per_rotation = ds.getProjectionsPerRotation();
    g.channel_spacing = ds.getDetectorTransverseSpacing();
    g.row_spacing = ds.getDetectorAxialSpacing();
    g.dsd = ds.getDistSourceToDetector();
    g.r = ds.getDistSourceToIsocenter();
    g.central_channel = ds.getDetectorCentralChannel();
    g.central_row = ds.getDetectorCentralRow();
    // FlyingFocalSpotMode (DicomDataSet.cpp:193, 285): the DICOM-CT-PD values
    // FFSNONE / FFSXY / FFSZ / FFSXYZ.  The anode angle is not part of
    // RawDataSet: the recon key anode_angle (rad), else 7 degrees, the
    // Siemens value.
    const std::string fm = ds.getFlyingFocalSpotMode();
    g.ffs_mode = fm == "FFSXYZ" ? FCT_FFS_PHI_Z
                 : fm == "FFSXY" ? FCT_FFS_PHI
                 : fm == "FFSZ"  ? FCT_FFS_Z
                                 : FCT_FFS_NONE;
    g.anode_angle = !g.ffs_mode ? 0.0
                    : m_rp.anode_angle > 0 ? (double)m_rp.anode_angle
                                           : 7.0 * 3.14159265358979323846 / 180.0;
    const int st = ViewLaw(ds, &g, &m_err);
    if (st) return st;
  }
  FilterSpec f{};
  if (!filter_spec(m_rp.filter, &f)) {
    m_err = "unknown filter '" + m_rp.filter + "'";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  std::memset(&m_rc, 0, sizeof m_rc);
  m_rc.start_pos = m_rp.start_pos;
  m_rc.end_pos = m_rp.end_pos;
  m_rc.recon_fov = m_rp.recon_fov;
  m_rc.slice_thickness = m_rp.slice_thickness;
  m_rc.slice_pitch = m_rp.slice_pitch;
  m_rc.nx = (int32_t)m_rp.nx;
  m_rc.ny = (int32_t)m_rp.ny;
  m_rc.x_origin = m_rp.x_origin;
  m_rc.y_origin = m_rp.y_origin;
  m_rc.recon_kernel = m_rp.recon_kernel;
  m_rc.filter_kind = f.kind;
  m_rc.filter_c = f.c;
  m_rc.filter_a = f.a;
  m_rc.tube_angle_offset = m_rp.tube_angle_offset;
  m_rc.adaptive_filtration_s = m_rp.adaptive_filtration_s;
  return FCT_OK;
}

// ReconstructionMachine.cu:457-474.
void ReconstructionMachine::PrintCTGeometry() const {
  const fct_scan_geometry& g = m_geom;
  const int nfft = [&] { int n = 1; while (n < 2 * g.n_channels) n <<= 1; return n; }();
  const double csw = g.r * g.row_spacing / g.dsd;
  const double cone = 2.0 * std::atan(0.5 * (g.n_rows - 1) * csw / g.r);
  const double acq = 2.0 * g.r * std::sin(0.5 * (g.n_channels - 1) * g.channel_spacing / g.dsd);
  std::cout << "CT Geometry and Scan derived parameters: \n"
            << "===========================================\n"
            << "Num projections per turn:           " << g.views_per_rotation << "\n"
            << "Num detector channels:              " << g.n_channels << "\n"
            << "Num detector channels padded:       " << 2 * g.n_channels << "\n"
            << "Num detector channels padded fft:   " << nfft << "\n"
            << "Num detector rows:                  " << g.n_rows << "\n"
            << "Num detector rows padded:           " << 2 * g.n_rows << "\n"
            << "Radius src->isocenter (mm):         " << g.r << "\n"
            << "Radius src->detector (mm):          " << g.dsd << "\n"
            << "Table feed per rotation (mm):       " << g.feed << "\n"
            << "Theta cone (rad):                   " << cone << "\n"
            << "Central channel:                    " << g.central_channel << "\n"
            << "Central row:                        " << g.central_row << "\n"
            << "Acquisition FOV (mm):               " << acq << "\n"
            << "Collimated slicewidth at isocenter: " << csw << std::endl;
}

int ReconstructionMachine::RunReconstruction() {
  if (!m_org_data_set) {
    m_err = "no raw data loaded";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  const int prec = precision_code(m_rp.precision);
  if (prec < 0) {
    m_err = "unknown precision '" + m_rp.precision + "'";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  if (m_rp.output_units != "mu" && m_rp.output_units != "hu") {
    m_err = "unknown output_units '" + m_rp.output_units + "' (mu or hu)";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  // AllocateKeyArrays copies every frame into one [view][row][ch] host array
  // (ReconstructionMachine.cu:84-105); here the data set's streaming
  // iterator fills the plan's pinned staging chunk by chunk while earlier
  // chunks are copied to and processed on the GPU.
  fct_wfbp_gpu_ctx* ctx = nullptr;
  int st = fct_wfbp_gpu_create(m_device, &ctx);
  if (st) {
    m_err = "cannot open GPU device " + std::to_string(m_device);
    return st;
  }
  fct_exec_options opts{};
  opts.precision = prec;
  opts.adaptive_filtration = m_rp.adaptive_filtration ? 1 : 0;
  fct_wfbp_gpu_plan* plan = nullptr;
  st = fct_wfbp_gpu_plan_create(ctx, &m_geom, &m_rc, &opts, nullptr, &plan);
  if (!st && m_rp.output_units == "hu")
    st = fct_wfbp_gpu_set_output_units(plan, FCT_UNITS_HU, m_rp.mu_water);
  if (!st) st = fct_wfbp_gpu_plan_info(plan, &m_info);
  if (!st) {
    m_reconstructed_data.assign((size_t)m_info.volume_elems, 0.0f);
    RawDataSet* ds = m_org_data_set.get();
    std::string reader_err;
    struct Reader {
      RawDataSet* ds;
      std::string* err;
      static int call(void* u, int32_t v0, int32_t n, float* dst) {
        auto* r = static_cast<Reader*>(u);
        const int st = r->ds->readViews(v0, n, dst);
        if (st) *r->err = r->ds->m_last_error;
        return st;
      }
    } reader{ds, &reader_err};
    const auto t0 = std::chrono::steady_clock::now();
    st = fct_wfbp_gpu_reconstruct_stream(plan, &Reader::call, &reader, 0, 0.0f,
                                         m_reconstructed_data.data());
    m_seconds_gpu = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (m_verbose && !st)
      std::cout << "Reconstruction (incl. host<->device copies): " << m_seconds_gpu << " s"
                << std::endl;
    if (st && !reader_err.empty()) {
      m_err = reader_err;
      fct_wfbp_gpu_plan_destroy(plan);
      fct_wfbp_gpu_destroy(ctx);
      m_org_data_set.reset();
      return st;
    }
  }
  m_org_data_set.reset();   // ReconstructionMachine.cu:105
  if (st) m_err = fct_wfbp_gpu_last_error(ctx);
  fct_wfbp_gpu_plan_destroy(plan);
  fct_wfbp_gpu_destroy(ctx);
  return st;
}

// ReconstructionMachine.cu:436-455; an empty path means output_dir/output_file
// (the reference's empty-path branch writes to an empty path).  The volume
// file itself is the reference's: headerless little-endian float32
// [slice][y][x].  Beside it goes a detached NRRD header (<file>.nhdr; NRRD
// format version 4, readable by ITK / 3D Slicer / teem) that names the file,
// its sizes, voxel spacings and the centre of the first voxel, and carries
// the reconstruction's own keys (units, slice thickness, recon config).
int ReconstructionMachine::SaveReconstruction(const std::string& filepath) {
  std::string path = filepath;
  if (path.empty()) path = (m_rp.output_dir.empty() ? "." : m_rp.output_dir) + "/" + m_rp.output_file;
  if (m_verbose) std::cout << "Saving reconstructed data to " << path << std::endl;
  std::ofstream o(path, std::ios::binary);
  if (!o) {
    m_err = "cannot write " + path;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  o.write(reinterpret_cast<const char*>(m_reconstructed_data.data()),
          m_reconstructed_data.size() * sizeof(float));
  if (!o) {
    m_err = "cannot write " + path;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  o.close();
  return WriteSidecarHeader(path);
}

int ReconstructionMachine::WriteSidecarHeader(const std::string& volume_path) const {
  const std::string hdr = volume_path + ".nhdr";
  std::ofstream h(hdr);
  if (!h) return FCT_ERR_INVALID_ARGUMENT;
  const double dx = m_rc.recon_fov / m_rc.nx, dy = m_rc.recon_fov / m_rc.ny;
  const double pitch = m_rc.slice_pitch > 0 ? m_rc.slice_pitch : m_rc.slice_thickness;
  const double dz = m_rc.end_pos >= m_rc.start_pos ? pitch : -pitch;   // slice step (signed)
  const double x0 = -0.5 * (m_rc.nx - 1) * dx + m_rc.x_origin;
  const double y0 = -0.5 * (m_rc.ny - 1) * dy + m_rc.y_origin;
  std::string base = volume_path;
  const size_t slash = base.find_last_of('/');
  if (slash != std::string::npos) base = base.substr(slash + 1);
  const bool hu = m_rp.output_units == "hu";
  h << std::setprecision(12);
  h << "NRRD0004\n"
    << "# FreeCT WFBP reconstruction: headerless float32 [slice][y][x], x fastest\n"
    << "type: float\n"
    << "dimension: 3\n"
    << "sizes: " << m_rc.nx << " " << m_rc.ny << " " << m_info.n_slices_requested << "\n"
    << "spacings: " << dx << " " << dy << " " << dz << "\n"
    << "axis mins: " << x0 << " " << y0 << " " << m_rc.start_pos << "\n"
    << "units: \"mm\" \"mm\" \"mm\"\n"
    << "kinds: domain domain domain\n"
    << "endian: little\n"
    << "encoding: raw\n"
    << "data file: " << base << "\n"
    << "freect_units:=" << (hu ? "HU" : "1/mm") << "\n";
  if (hu) h << "freect_mu_water:=" << m_rp.mu_water << "\n";
  h << "freect_slice_thickness:=" << m_rc.slice_thickness << "\n"
    << "freect_start_pos:=" << m_rc.start_pos << "\n"
    << "freect_end_pos:=" << m_rc.end_pos << "\n"
    << "freect_recon_fov:=" << m_rc.recon_fov << "\n"
    << "freect_filter:=" << m_rp.filter << "\n"
    << "freect_precision:=" << m_rp.precision << "\n"
    << "freect_adaptive_filtration_s:="
    << (m_rp.adaptive_filtration ? m_rp.adaptive_filtration_s : 0.0f) << "\n"
    << "freect_native_slices:=" << m_info.n_slices_native << "\n";
  return h ? FCT_OK : FCT_ERR_INVALID_ARGUMENT;
}

}  // namespace fct
