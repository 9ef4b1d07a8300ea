// io.cpp -- on-disk output formats of the reference, host side.
//
//   - VTK ImageData (.vti, ASCII): HydroRun::saveVTK (src/HydroRun.h:502-609)
//   - numpy .npy (1-D double): cnpy::npy_save / create_npy_header
//     (src/cnpy/cnpy.h:129-187,343-376), as used by save_cnpy
//     (src/cnpy/cnpy_io.h:38-62)
//
// Both are written from host memory; the device-side entry points in api.cpp
// download the state first.  HDF5/XDMF output (HydroRun.h:619-845) needs the
// HDF5 library, which this build does not link.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "euler2d_hip.h"
#include "status.hpp"

namespace {

// ostream's default rendering of a real (precision 6, %g)
int put_real(char *buf, size_t n, double x) { return std::snprintf(buf, n, "%g", x); }

class Out {
 public:
  explicit Out(std::FILE *f) : f_(f) { buf_.reserve(1 << 20); }
  void str(const char *s) { buf_.append(s); flush_if(); }
  void str(const std::string &s) { buf_.append(s); flush_if(); }
  void integer(long long v) {
    char b[32];
    std::snprintf(b, sizeof(b), "%lld", v);
    buf_.append(b);
  }
  void real(double x) {
    char b[64];
    put_real(b, sizeof(b), x);
    buf_.append(b);
  }
  void flush_if() {
    if (buf_.size() >= (1u << 20)) flush();
  }
  bool flush() {
    const bool ok = std::fwrite(buf_.data(), 1, buf_.size(), f_) == buf_.size();
    buf_.clear();
    return ok;
  }

 private:
  std::FILE *f_;
  std::string buf_;
};

bool little_endian() {
  const int i = 1;
  return *reinterpret_cast<const char *>(&i) != 0;
}

}  // namespace

extern "C" {

int e2d_write_vti(const e2d_params *p, const double *interior, int use_float32,
                  const char *output_dir, const char *output_prefix, int step) {
  if (!p || !interior || !output_dir || !output_prefix || step < 0)
    return e2d::report(E2D_ERR_ARG, "null argument or negative step");
  // file name: outputDir/outputPrefix_<7-digit step>.vti (HydroRun.h:535-543)
  char num[32];
  std::snprintf(num, sizeof(num), "%07d", step);
  const std::string filename =
      std::string(output_dir) + "/" + output_prefix + "_" + num + ".vti";
  std::FILE *f = std::fopen(filename.c_str(), "wb");
  if (!f) return e2d::report(E2D_ERR_IO, "cannot open " + filename + " for writing");

  static const char *const kVarNames[4] = {"rho", "E", "mx", "my"};  // HydroParams.cpp:17
  Out o(f);
  o.str("<?xml version=\"1.0\"?>\n");
  o.str(little_endian()
            ? "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
            : "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"BigEndian\">\n");
  o.str("  <ImageData WholeExtent=\"0 ");
  o.integer(p->nx);
  o.str(" 0 ");
  o.integer(p->ny);
  o.str(" 0 0\" Origin=\"");
  o.real(p->xmin);
  o.str(" ");
  o.real(p->ymin);
  o.str(" ");
  o.real(0.0);
  o.str("\" Spacing=\"");
  o.real(p->dx);
  o.str(" ");
  o.real(p->dy);
  o.str(" ");
  o.real(0.0);
  o.str("\">\n");
  o.str("  <Piece Extent=\"0 ");
  o.integer(p->nx);
  o.str(" 0 ");
  o.integer(p->ny);
  o.str(" 0 0 \">\n");
  o.str("    <PointData>\n    </PointData>\n    <CellData>\n");
  const size_t plane = (size_t)p->nx * p->ny;
  for (int v = 0; v < 4; v++) {
    o.str("    <DataArray type=\"");
    o.str(use_float32 ? "Float32" : "Float64");
    o.str("\" Name=\"");
    o.str(kVarNames[v]);
    o.str("\" format=\"ascii\" >\n");
    // interior cells, i fastest (HydroRun.h:583-598)
    const double *src = interior + v * plane;
    for (size_t k = 0; k < plane; k++) {
      o.real(use_float32 ? (double)(float)src[k] : src[k]);
      o.str(" ");
    }
    o.str("\n    </DataArray>\n");
  }
  o.str("    </CellData>\n  </Piece>\n  </ImageData>\n</VTKFile>\n");
  const bool ok = o.flush();
  if (std::fclose(f) != 0 || !ok) return e2d::report(E2D_ERR_IO, "write failed: " + filename);
  return E2D_OK;
}

int e2d_write_npy(const char *path, const double *data, int64_t n) {
  if (!path || (!data && n > 0) || n < 0) return e2d::report(E2D_ERR_ARG, "bad argument");
  // header dictionary of create_npy_header<double> for shape (n,), padded
  // with spaces so that 10 + len is a multiple of 16, ending in '\n'
  std::string dict = "{'descr': '";
  dict += little_endian() ? '<' : '>';
  dict += "f8', 'fortran_order': False, 'shape': (";
  dict += std::to_string(n);
  dict += ",), }";
  const size_t pad = 16 - (10 + dict.size()) % 16;
  dict.append(pad, ' ');
  dict.back() = '\n';
  std::vector<unsigned char> head;
  head.push_back(0x93);
  for (const char c : std::string("NUMPY")) head.push_back((unsigned char)c);
  head.push_back(0x01);
  head.push_back(0x00);
  const uint16_t len = (uint16_t)dict.size();
  head.push_back((unsigned char)(len & 0xff));
  head.push_back((unsigned char)(len >> 8));
  head.insert(head.end(), dict.begin(), dict.end());

  std::FILE *f = std::fopen(path, "wb");
  if (!f) return e2d::report(E2D_ERR_IO, std::string("cannot open ") + path + " for writing");
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  if (n > 0) ok = ok && std::fwrite(data, sizeof(double), (size_t)n, f) == (size_t)n;
  if (std::fclose(f) != 0 || !ok) return e2d::report(E2D_ERR_IO, std::string("write failed: ") + path);
  return E2D_OK;
}

}  // extern "C"
