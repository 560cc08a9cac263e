// dicom.cpp -- DICOM-CT-PD projection reader (fct::DicomDataSet) without
// DCMTK.  FreeCT reads the TCIA "LDCT and Projection Data" files with DCMTK
// (DicomDataSet.cpp:56-148 per frame, :232-302 for the scan metadata, the
// private-tag dictionary at :178-211); DCMTK is absent here, so this file
// holds a self-written parser of what those files use: DICOM Part 10 files
// (128-byte preamble, "DICM", the explicit little-endian group 0002), a data
// set in explicit or implicit VR little endian, nested sequences of defined
// or undefined length (skipped), and the top-level elements the reader
// needs.  Big-endian and encapsulated (compressed) transfer syntaxes are
// rejected with a message.
//
// Per file (one projection): tube angle, table position and radial distance
// of the detector focal centre (7031,1001-1003), the focal spot's deflection
// (7033,100B-100D), detector rows and channels (7029,1010-1011), the pixel
// data (7FE0,0010, uint16) and its rescale (0028,1052 intercept, 0028,1053
// slope: value = slope * pixel + intercept in float, as the reference's
// Eigen expression), stored rows = detector rows or transposed
// (rows = detector channels), DicomDataSet.cpp:135-148.  Scan metadata comes
// from the first file in name order, DicomDataSet.cpp:232-302: the central
// element (7031,1033) is (channel, row) as the reference reads it.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "fct_host.h"

namespace fct {
namespace {

constexpr uint32_t tag(uint16_t g, uint16_t e) { return ((uint32_t)g << 16) | e; }

// VRs of the implicit-VR elements the reader decodes (the standard tags it
// reads and DicomDataSet.cpp:178-211's private dictionary); others are kept
// as raw bytes.
const char* implicit_vr(uint32_t t) {
  switch (t) {
    case tag(0x0008, 0x0070): return "LO";
    case tag(0x0028, 0x0010): case tag(0x0028, 0x0011): case tag(0x0028, 0x0100):
    case tag(0x0028, 0x0103): return "US";
    case tag(0x0028, 0x1052): case tag(0x0028, 0x1053): return "DS";
    case tag(0x7FE0, 0x0010): return "OW";
    case tag(0x7029, 0x1010): case tag(0x7029, 0x1011): case tag(0x7033, 0x1013):
      return "US";
    case tag(0x7029, 0x1002): case tag(0x7029, 0x1006): case tag(0x7031, 0x1001):
    case tag(0x7031, 0x1002): case tag(0x7031, 0x1003): case tag(0x7031, 0x1031):
    case tag(0x7031, 0x1033): case tag(0x7033, 0x100B): case tag(0x7033, 0x100C):
    case tag(0x7033, 0x100D): return "FL";
    case tag(0x7029, 0x100B): case tag(0x7033, 0x100E): case tag(0x7037, 0x1009):
    case tag(0x7037, 0x100A): return "CS";
  }
  return "UN";
}

bool long_form(const char* vr) {   // explicit VRs with 2 reserved bytes and a 4-byte length
  static const char* k[] = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "SV", "UC", "UN", "UR",
                            "UT", "UV"};
  for (const char* v : k)
    if (vr[0] == v[0] && vr[1] == v[1]) return true;
  return false;
}

struct Element {
  char vr[3];
  size_t off, len;
};

class DicomFile {
 public:
  int load(const std::string& path, std::string* err) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(err, "cannot open " + path);
    in.seekg(0, std::ios::end);
    const std::streamoff n = in.tellg();
    in.seekg(0);
    m_buf.resize((size_t)std::max<std::streamoff>(n, 0));
    if (n > 0) in.read(reinterpret_cast<char*>(m_buf.data()), n);
    if (!in) return fail(err, "cannot read " + path);
    m_path = path;
    size_t pos = 0;
    bool explicit_vr = false;
    if (m_buf.size() >= 132 && std::memcmp(&m_buf[128], "DICM", 4) == 0) {
      pos = 132;
      // File meta information: always explicit VR little endian.
      std::string ts;
      while (pos + 8 <= m_buf.size() && u16(pos) == 0x0002) {
        Element e;
        uint32_t t;
        if (!read_element(&pos, true, &t, &e, err)) return FCT_ERR_INVALID_ARGUMENT;
        if (t == tag(0x0002, 0x0010)) ts = str(e);
      }
      if (ts == "1.2.840.10008.1.2.1" || ts.empty()) {
        explicit_vr = true;
      } else if (ts == "1.2.840.10008.1.2") {
        explicit_vr = false;
      } else {
        return fail(err, m_path + ": transfer syntax " + ts +
                             " is not supported (explicit or implicit VR little endian only)");
      }
    }
    while (pos + 8 <= m_buf.size()) {
      Element e;
      uint32_t t;
      if (!read_element(&pos, explicit_vr, &t, &e, err)) return FCT_ERR_INVALID_ARGUMENT;
      if ((t >> 16) != 0xFFFE) m_el[t] = e;
    }
    return FCT_OK;
  }

  bool has(uint32_t t) const { return m_el.count(t) != 0; }
  bool f32(uint32_t t, float* v, size_t i = 0) const {
    auto it = m_el.find(t);
    if (it == m_el.end() || it->second.len < 4 * (i + 1)) return false;
    std::memcpy(v, &m_buf[it->second.off + 4 * i], 4);
    return true;
  }
  bool u16v(uint32_t t, uint16_t* v) const {
    auto it = m_el.find(t);
    if (it == m_el.end() || it->second.len < 2) return false;
    *v = u16(it->second.off);
    return true;
  }
  bool ds(uint32_t t, double* v) const {
    auto it = m_el.find(t);
    if (it == m_el.end()) return false;
    const std::string s = str(it->second);
    char* end = nullptr;
    *v = std::strtod(s.c_str(), &end);
    return end != s.c_str();
  }
  bool text(uint32_t t, std::string* v) const {
    auto it = m_el.find(t);
    if (it == m_el.end()) return false;
    *v = str(it->second);
    return true;
  }
  const uint8_t* bytes(uint32_t t, size_t* len) const {
    auto it = m_el.find(t);
    if (it == m_el.end()) return nullptr;
    *len = it->second.len;
    return &m_buf[it->second.off];
  }
  const std::string& path() const { return m_path; }

 private:
  static int fail(std::string* err, const std::string& m) {
    if (err) *err = m;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  uint16_t u16(size_t p) const { return (uint16_t)(m_buf[p] | (m_buf[p + 1] << 8)); }
  uint32_t u32(size_t p) const {
    return (uint32_t)m_buf[p] | ((uint32_t)m_buf[p + 1] << 8) | ((uint32_t)m_buf[p + 2] << 16) |
           ((uint32_t)m_buf[p + 3] << 24);
  }
  std::string str(const Element& e) const {
    std::string s(reinterpret_cast<const char*>(&m_buf[e.off]), e.len);
    while (!s.empty() && (s.back() == ' ' || s.back() == '\0')) s.pop_back();
    size_t a = 0;
    while (a < s.size() && s[a] == ' ') ++a;
    return s.substr(a);
  }

  // One element at *pos (tag, VR, length, value); undefined-length values
  // (sequences) are skipped to their delimiter.
  bool read_element(size_t* pos, bool explicit_vr, uint32_t* t, Element* e, std::string* err) {
    size_t p = *pos;
    if (p + 8 > m_buf.size()) return fail(err, m_path + ": truncated element"), false;
    *t = tag(u16(p), u16(p + 2));
    p += 4;
    uint32_t len;
    if ((*t >> 16) == 0xFFFE) {   // item / delimiters carry no VR
      std::strcpy(e->vr, "NA");
      len = u32(p);
      p += 4;
    } else if (explicit_vr) {
      e->vr[0] = (char)m_buf[p];
      e->vr[1] = (char)m_buf[p + 1];
      e->vr[2] = 0;
      p += 2;
      if (long_form(e->vr)) {
        if (p + 6 > m_buf.size()) return fail(err, m_path + ": truncated element"), false;
        len = u32(p + 2);
        p += 6;
      } else {
        len = u16(p);
        p += 2;
      }
    } else {
      std::strcpy(e->vr, implicit_vr(*t));
      len = u32(p);
      p += 4;
    }
    if (len == 0xFFFFFFFFu) {
      if (*t == tag(0x7FE0, 0x0010))
        return fail(err, m_path + ": encapsulated (compressed) pixel data is not supported"), false;
      e->off = p;
      if (!skip_undefined(&p, explicit_vr, err)) return false;
      e->len = 0;
      *pos = p;
      return true;
    }
    if ((*t >> 16) == 0xFFFE && *t != tag(0xFFFE, 0xE000)) len = 0;   // delimiters
    if (p + len > m_buf.size()) return fail(err, m_path + ": element runs past the file end"), false;
    e->off = p;
    e->len = len;
    *pos = p + len;
    return true;
  }

  // Skips a value of undefined length: elements (and items, themselves of
  // defined or undefined length) up to the matching sequence or item
  // delimiter.
  bool skip_undefined(size_t* pos, bool explicit_vr, std::string* err, int depth = 0) {
    if (depth > 32) return fail(err, m_path + ": sequences nested too deeply"), false;
    while (*pos + 8 <= m_buf.size()) {
      const uint32_t t = tag(u16(*pos), u16(*pos + 2));
      if (t == tag(0xFFFE, 0xE0DD) || t == tag(0xFFFE, 0xE00D)) {
        *pos += 8;
        return true;
      }
      if (t == tag(0xFFFE, 0xE000)) {   // item
        const uint32_t len = u32(*pos + 4);
        *pos += 8;
        if (len == 0xFFFFFFFFu) {
          if (!skip_undefined(pos, explicit_vr, err, depth + 1)) return false;
        } else {
          *pos += len;
        }
        continue;
      }
      Element e;
      uint32_t tt;
      if (!read_element(pos, explicit_vr, &tt, &e, err)) return false;
    }
    return fail(err, m_path + ": unterminated sequence"), false;
  }

  std::string m_path;
  std::vector<uint8_t> m_buf;
  std::map<uint32_t, Element> m_el;
};

// One projection of a DICOM-CT-PD file into a frame (DicomDataSet.cpp:56-148).
int read_frame(const std::string& path, RawDataFrame* f, std::string* err) {
  DicomFile d;
  int st = d.load(path, err);
  if (st) return st;
  auto need = [&](bool ok, const char* what) {
    if (!ok && err) *err = path + ": missing or malformed " + what;
    return ok;
  };
  uint16_t rows = 0, cols = 0, drows = 0, dch = 0, bits = 16, repr = 0;
  double slope = 0, icept = 0;
  if (!need(d.f32(tag(0x7031, 0x1001), &f->m_dfc_angular_position),
            "DetectorFocalCenterAngularPosition (7031,1001)") ||
      !need(d.f32(tag(0x7031, 0x1002), &f->m_dfc_axial_position),
            "DetectorFocalCenterAxialPosition (7031,1002)") ||
      !need(d.f32(tag(0x7031, 0x1003), &f->m_dfc_radial_distance),
            "DetectorFocalCenterRadialDistance (7031,1003)") ||
      !need(d.u16v(tag(0x7029, 0x1010), &drows), "NumberofDetectorRows (7029,1010)") ||
      !need(d.u16v(tag(0x7029, 0x1011), &dch), "NumberofDetectorColumns (7029,1011)") ||
      !need(d.ds(tag(0x0028, 0x1052), &icept), "RescaleIntercept (0028,1052)") ||
      !need(d.ds(tag(0x0028, 0x1053), &slope), "RescaleSlope (0028,1053)") ||
      !need(d.u16v(tag(0x0028, 0x0010), &rows), "Rows (0028,0010)") ||
      !need(d.u16v(tag(0x0028, 0x0011), &cols), "Columns (0028,0011)") ||
      !need(d.f32(tag(0x7033, 0x100B), &f->m_ffs_angular_shift),
            "SourceAngularPositionShift (7033,100B)") ||
      !need(d.f32(tag(0x7033, 0x100C), &f->m_ffs_axial_shift),
            "SourceAxialPositionShift (7033,100C)") ||
      !need(d.f32(tag(0x7033, 0x100D), &f->m_ffs_radial_shift),
            "SourceRadialDistanceShift (7033,100D)"))
    return FCT_ERR_INVALID_ARGUMENT;
  d.u16v(tag(0x0028, 0x0100), &bits);
  d.u16v(tag(0x0028, 0x0103), &repr);
  if (bits != 16 || repr != 0) {
    if (err) *err = path + ": pixel data must be 16-bit unsigned";
    return FCT_ERR_UNSUPPORTED;
  }
  size_t len = 0;
  const uint8_t* px = d.bytes(tag(0x7FE0, 0x0010), &len);
  const size_t n = (size_t)drows * dch;
  if (!px || len < 2 * n || (size_t)rows * cols != n) {
    if (err) *err = path + ": pixel data does not hold rows x channels values";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  f->m_detector_rows = drows;
  f->m_detector_channels = dch;
  f->m_projection.resize(n);
  const float fs = (float)slope, fi = (float)icept;
  auto pix = [&](size_t i) { return (float)(uint16_t)(px[2 * i] | (px[2 * i + 1] << 8)); };
  if (rows == drows) {            // [row][channel]
    for (size_t i = 0; i < n; ++i) f->m_projection[i] = fs * pix(i) + fi;
  } else if (rows == dch) {       // stored transposed: image row = channel
    for (int r = 0; r < drows; ++r)
      for (int c = 0; c < dch; ++c)
        f->m_projection[(size_t)r * dch + c] = fs * pix((size_t)c * drows + r) + fi;
  } else {
    if (err) *err = path + ": Rows/Columns match neither the detector rows nor its channels";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  return FCT_OK;
}

}  // namespace

int DicomDataSet::initialize() {
  m_file_list.clear();
  DIR* dir = opendir(m_path.c_str());
  if (!dir) {
    m_last_error = "cannot open DICOM directory " + m_path;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  while (dirent* ent = readdir(dir)) {
    const std::string name = ent->d_name;
    if (name == "." || name == "..") continue;
    const std::string full = m_path + "/" + name;
    struct stat st;
    if (stat(full.c_str(), &st) == 0 && S_ISREG(st.st_mode)) m_file_list.push_back(full);
  }
  closedir(dir);
  std::sort(m_file_list.begin(), m_file_list.end());   // DicomDataSet.cpp:159-161
  if (m_file_list.empty()) {
    m_last_error = "no files in DICOM directory " + m_path;
    return FCT_ERR_INVALID_ARGUMENT;
  }
  if (m_file_list.size() > 65535) {
    m_last_error = "more than 65535 projections";
    return FCT_ERR_LIMIT;
  }
  m_total_num_projections = (uint16_t)m_file_list.size();
  m_data.clear();
  m_data.resize(m_total_num_projections);
  m_is_loaded.assign(m_total_num_projections, false);
  m_has_header.assign(m_total_num_projections, false);
  return readMetadata();
}

int DicomDataSet::readMetadata() {
  DicomFile d;
  int st = d.load(m_file_list[0], &m_last_error);
  if (st) return st;
  std::string s;
  if (d.text(tag(0x0008, 0x0070), &s)) m_manufacturer = s;
  float cen[2] = {0, 0};
  std::string ffs, shape, stype, geom;
  const bool ok =
      d.u16v(tag(0x7029, 0x1010), &m_detector_rows) &&
      d.u16v(tag(0x7029, 0x1011), &m_detector_channels) &&
      d.f32(tag(0x7029, 0x1002), &m_detector_transverse_spacing) &&
      d.f32(tag(0x7029, 0x1006), &m_detector_axial_spacing) &&
      d.text(tag(0x7029, 0x100B), &shape) &&
      d.f32(tag(0x7031, 0x1003), &m_dist_source_to_isocenter) &&
      d.f32(tag(0x7031, 0x1031), &m_dist_source_to_detector) &&
      d.text(tag(0x7033, 0x100E), &ffs) &&
      d.u16v(tag(0x7033, 0x1013), &m_projections_per_rotation) &&
      d.text(tag(0x7037, 0x1009), &stype) && d.text(tag(0x7037, 0x100A), &geom) &&
      d.f32(tag(0x7031, 0x1033), &cen[0], 0) && d.f32(tag(0x7031, 0x1033), &cen[1], 1);
  if (!ok) {
    m_last_error = m_file_list[0] + ": missing DICOM-CT-PD scan metadata (detector, geometry, "
                   "flying focal spot, projections per rotation or central element)";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  m_detector_shape = shape;
  m_flying_focal_spot_mode = ffs;
  m_scan_type = stype;
  m_projection_geometry = geom;
  m_detector_central_channel = cen[0];   // DicomDataSet.cpp:296-297
  m_detector_central_row = cen[1];
  return FCT_OK;
}

int DicomDataSet::readProjection(int i) {
  if (i < 0 || i >= (int)m_total_num_projections) return FCT_ERR_INVALID_ARGUMENT;
  auto f = std::make_unique<RawDataFrame>();
  std::string err;
  const int st = read_frame(m_file_list[i], f.get(), &err);
  if (st) {
    m_last_error = err;
    return st;
  }
  if (f->m_detector_rows != m_detector_rows || f->m_detector_channels != m_detector_channels) {
    m_last_error = m_file_list[i] + ": detector size differs from the first file";
    return FCT_ERR_INVALID_ARGUMENT;
  }
  m_data[i] = std::move(f);
  m_is_loaded[i] = true;
  m_has_header[i] = true;
  return FCT_OK;
}

// All projections on the host's threads (DicomDataSet.cpp:304-330).
int DicomDataSet::readAll() {
  const int n = (int)m_total_num_projections;
  const int nt = std::max(1, std::min<int>(n, (int)std::thread::hardware_concurrency()));
  std::vector<std::string> errs(nt);
  std::atomic<int> next(0);
  std::vector<std::unique_ptr<RawDataFrame>> frames(n);
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t] {
      for (int i = next++; i < n; i = next++) {
        auto f = std::make_unique<RawDataFrame>();
        std::string err;
        if (read_frame(m_file_list[i], f.get(), &err)) {
          if (errs[t].empty()) errs[t] = err;
          continue;
        }
        frames[i] = std::move(f);
      }
    });
  for (auto& th : pool) th.join();
  for (const auto& e : errs)
    if (!e.empty()) {
      m_last_error = e;
      return FCT_ERR_INVALID_ARGUMENT;
    }
  for (int i = 0; i < n; ++i) {
    if (frames[i]->m_detector_rows != m_detector_rows ||
        frames[i]->m_detector_channels != m_detector_channels) {
      m_last_error = m_file_list[i] + ": detector size differs from the first file";
      return FCT_ERR_INVALID_ARGUMENT;
    }
    m_data[i] = std::move(frames[i]);
    m_is_loaded[i] = true;
    m_has_header[i] = true;
  }
  return FCT_OK;
}

}  // namespace fct
