// fct_host.h -- C++ host side above the C ABI, mirroring FreeCT's fct_wfbp
// interfaces with the same names and meanings:
//   ReconConfig / parse_config          recon_structs.h:33-50, parse_config.cpp:31-62
//   fct::RawDataFrame / fct::RawDataSet RawDataSet.h:32-198, RawDataSet.cpp:17-107
//   fct::FreeCTDataSet                  FreeCTDataSet.h/.cpp (meta.yaml + projections.dat)
//   fct::ReconstructionMachine          ReconstructionMachine.h:8-50
// Errors are reported as fct_status codes with a message (the reference calls
// exit(); only the command-line driver turns an error into an exit code).
#pragma once

#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/fct_wfbp_gpu.h"

// Recon parameters read from the recon YAML (recon_structs.h:33-50).
struct ReconConfig {
  std::string raw_data_dir;
  std::string output_dir;
  std::string output_file;
  float start_pos = 0;
  float end_pos = 0;
  float recon_fov = 0;
  float slice_thickness = 0;
  float slice_pitch = 0;
  size_t nx = 0;
  size_t ny = 0;
  int recon_kernel = 0;
  float x_origin = 0;
  float y_origin = 0;
  float tube_angle_offset = 0;
  float adaptive_filtration_s = 0;
  // Keys of this engine (absent keys keep these defaults).
  std::string filter = "ramp";     // ramp | sharp | medium | smooth | hann | hamming | shepp-logan
  std::string precision = "fp32";  // fp32 | fp16 | bf16
  float anode_angle = 0;           // rad, z-FFS deflection; 0 -> 7 degrees
  int adaptive_filtration = 0;     // 1: apply adaptive_filtration_s (the reference ignores it)
  std::string output_units = "mu";  // mu (1/mm, the reference) | hu
  double mu_water = 0.0193;        // 1/mm, for output_units hu
};

// Flat "key: value" YAML maps (the only YAML the path reads or writes).
std::map<std::string, std::string> fct_read_flat_yaml(const std::string& path, std::string* err);

// Reads the recon keys of parse_config.cpp:39-57, printing FOUND / NOT FOUND
// per key like the reference; a missing key keeps its default (0).
// slice_pitch == 0 -> slice_thickness (parse_config.cpp:59-60).
int parse_config(const std::string& config_file, ReconConfig& rp, bool verbose, std::string* err);

namespace fct {

class RawDataFrame {
 public:
  int getRows() const { return m_detector_rows; }
  int getCols() const { return m_detector_channels; }
  float getDFCAngularPosition() const { return m_dfc_angular_position; }
  float getDFCAxialPosition() const { return m_dfc_axial_position; }
  float getDFCRadialPosition() const { return m_dfc_radial_distance; }
  float getFFSAngularShift() const { return m_ffs_angular_shift; }
  float getFFSAxialShift() const { return m_ffs_axial_shift; }
  float getFFSRadialShift() const { return m_ffs_radial_shift; }

  std::vector<float> m_projection;   // [row][channel], row-major
  uint16_t m_detector_channels = 0;
  uint16_t m_detector_rows = 0;
  float m_dfc_angular_position = 0;
  float m_dfc_axial_position = 0;
  float m_dfc_radial_distance = 0;
  float m_ffs_angular_shift = 0;
  float m_ffs_axial_shift = 0;
  float m_ffs_radial_shift = 0;
};

class RawDataSet {
 public:
  virtual ~RawDataSet() = default;
  void setPath(const std::string& path) { m_path = path; }

  virtual int initialize() = 0;
  virtual int readMetadata() = 0;
  virtual int readProjection(int projection_idx) = 0;
  virtual int readAll() = 0;

  int writeReconFile(const std::string& filepath, bool single_slice_mode = false) const;
  int writeAll(const std::string& dirpath) const;
  void printMetadata() const;

  std::string getManufacturer() const { return m_manufacturer; }
  uint16_t getDetectorRows() const { return m_detector_rows; }
  uint16_t getDetectorChannels() const { return m_detector_channels; }
  float getDetectorTransverseSpacing() const { return m_detector_transverse_spacing; }
  float getDetectorAxialSpacing() const { return m_detector_axial_spacing; }
  std::string getDetectorShape() const { return m_detector_shape; }
  float getDistSourceToDetector() const { return m_dist_source_to_detector; }
  float getDistSourceToIsocenter() const { return m_dist_source_to_isocenter; }
  float getDetectorCentralRow() const { return m_detector_central_row; }
  float getDetectorCentralChannel() const { return m_detector_central_channel; }
  std::string getScanType() const { return m_scan_type; }
  std::string getProjectionGeometry() const { return m_projection_geometry; }
  std::string getFlyingFocalSpotMode() const { return m_flying_focal_spot_mode; }
  uint16_t getProjectionsPerRotation() const { return m_projections_per_rotation; }
  uint16_t getTotalNumProjections() const { return m_total_num_projections; }
  // NaN when the projection (or, for a header-only load, its frame header)
  // is not available (RawDataSet.h:109-121).
  float getTablePosition(int projection_idx) const;
  float getTubeAngle(int projection_idx) const;
  bool projectionIsAvailable(int projection_idx) const;
  bool headerIsAvailable(int projection_idx) const;
  // rows x channels floats, [row][channel] (RawDataSet.h:136-139).
  void copyProjection(int projection_idx, float* buffer) const;
  // Streaming iterator: views [v0, v0 + n) as [view][row][channel] float32
  // into dst, called in view order by the chunked GPU upload in place of
  // copying every frame into one whole-scan array.  The default copies loaded
  // frames and reads missing ones with readProjection.
  virtual int readViews(int v0, int n, float* dst);

  // Exact scan law when the data set knows it (synthetic scans); otherwise
  // ReconstructionMachine derives it from the per-view angles and positions.
  virtual bool exactGeometry(fct_scan_geometry*) const { return false; }

 protected:
  std::string m_path;
  std::vector<std::unique_ptr<RawDataFrame>> m_data;
  std::vector<bool> m_is_loaded;
  std::vector<bool> m_has_header;   // frame header read (angles, positions)
  std::string m_manufacturer = "SYNTHETIC";
  uint16_t m_detector_rows = 0;
  uint16_t m_detector_channels = 0;
  float m_detector_transverse_spacing = 0;
  float m_detector_axial_spacing = 0;
  std::string m_detector_shape = "CYLINDRICAL";
  float m_dist_source_to_detector = 0;
  float m_dist_source_to_isocenter = 0;
  float m_detector_central_row = 0;
  float m_detector_central_channel = 0;
  std::string m_scan_type = "HELICAL";
  std::string m_projection_geometry = "FANBEAM";
  std::string m_flying_focal_spot_mode = "FFSNONE";
  uint16_t m_projections_per_rotation = 0;
  uint16_t m_total_num_projections = 0;

 public:
  std::string m_last_error;
};

// FreeCT's two-file format as written by RawDataSet::writeAll
// (RawDataSet.cpp:85-107, fct_read_util/README.md:27-48): meta.yaml with the
// keys of RawDataSet.h:155-172, and projections.dat holding, per frame, three
// float32 (angle + ffs shift, axial + ffs shift, radial + ffs shift) followed
// by rows x channels float32.
class FreeCTDataSet : public RawDataSet {
 public:
  int initialize() override;
  int readMetadata() override;
  int readProjection(int projection_idx) override;
  int readAll() override;
  // Only the frame headers (tube angle, table position, radius), so that
  // the projections stream from disk through readViews during the upload.
  int readHeaders();
  int readViews(int v0, int n, float* dst) override;

 private:
  std::ifstream* stream();
  std::unique_ptr<std::ifstream> m_in;
  int64_t m_next_frame = -1;   // frame the open stream is positioned at
};

// DICOM-CT-PD projection files (one projection per file, sorted by name),
// read by a self-written DICOM parser (dicom.cpp) in place of DCMTK:
// DicomDataSet.cpp:56-330 and its private-tag map (:178-211).
class DicomDataSet : public RawDataSet {
 public:
  int initialize() override;   // file list, then the scan metadata of the first file
  int readMetadata() override;
  int readProjection(int projection_idx) override;
  int readAll() override;      // every file, on the host's threads

 private:
  std::vector<std::string> m_file_list;
};

// Seeded synthetic helical scan (water cylinder + ellipsoids, analytic line
// integrals on the GPU, optional Poisson noise).  raw_data_dir
// "synthetic:<workload>[:seed=<n>]" with workload small | full | adjoint.
class SyntheticDataSet : public RawDataSet {
 public:
  explicit SyntheticDataSet(int device = 0) : m_device(device) {}
  int initialize() override;
  int readMetadata() override { return FCT_OK; }
  int readProjection(int projection_idx) override;
  int readAll() override;
  bool exactGeometry(fct_scan_geometry* g) const override {
    *g = m_geom;
    return true;
  }
  const fct_recon_config& suggestedRecon() const { return m_recon; }
  const std::string& workload() const { return m_workload; }

 private:
  int m_device;
  std::string m_workload;
  uint64_t m_seed = 0;
  double m_i0 = 0;
  fct_scan_geometry m_geom{};
  fct_recon_config m_recon{};
  std::vector<fct_ellipsoid> m_ellipsoids;
};

// Synthetic workload presets and phantom parameters (same recipe as
// freect_amd/scans.py).
bool fct_synthetic_workload(const std::string& name, uint64_t seed, fct_scan_geometry* g,
                            fct_recon_config* rc, double* i0, std::vector<fct_ellipsoid>* ell,
                            uint64_t* seed_used);
void fct_make_phantom(uint64_t seed, double z_lo, double z_hi, int n,
                      std::vector<fct_ellipsoid>* out);

// The data set a raw_data_dir names: "synthetic:<workload>", a FreeCT
// two-file directory (meta.yaml present) or a DICOM-CT-PD directory.
std::shared_ptr<RawDataSet> open_data_set(const std::string& raw_data_dir, int device = 0);

class ReconstructionMachine {
 public:
  ReconstructionMachine() = default;
  ~ReconstructionMachine();

  int LoadReconConfiguration(const std::string& filepath);
  // FreeCT-format data: frame headers only; the projections are read from
  // disk chunk by chunk while the reconstruction streams them to the GPU.
  int LoadRawData();
  int ConfigureCTGeometry();
  int RunReconstruction();
  void PrintCTGeometry() const;
  // The uniform helical law of a scan's per-view angles and table positions,
  // checked against every view (FCT_ERR_UNSUPPORTED when a view is off it).
  static int ViewLaw(const RawDataSet& ds, fct_scan_geometry* g, std::string* err);
  int SaveReconstruction(const std::string& filepath = "");
  // <volume>.nhdr: detached NRRD header of the saved volume.
  int WriteSidecarHeader(const std::string& volume_path) const;

  void setDevice(int device) { m_device = device; }
  void setVerbose(bool v) { m_verbose = v; }
  const std::string& lastError() const { return m_err; }
  const std::vector<float>& volume() const { return m_reconstructed_data; }
  const fct_plan_info& planInfo() const { return m_info; }
  ReconConfig& config() { return m_rp; }
  const fct_scan_geometry& geometry() const { return m_geom; }
  void setDataSet(std::shared_ptr<RawDataSet> ds) { m_org_data_set = std::move(ds); }

 private:
  ReconConfig m_rp;
  fct_scan_geometry m_geom{};
  fct_recon_config m_rc{};
  fct_plan_info m_info{};
  std::shared_ptr<RawDataSet> m_org_data_set;
  std::vector<float> m_reconstructed_data;
  std::string m_config_path;
  std::string m_err;
  int m_device = 0;
  bool m_verbose = false;
  double m_seconds_gpu = 0;
};

}  // namespace fct
