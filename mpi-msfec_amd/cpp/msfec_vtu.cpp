// VTU writer for fine-grid fields of one coarse box (see msfec_vtu.hpp).
#include "msfec_vtu.hpp"

#include <cstdio>
#include <stdexcept>

namespace msfec {

std::string int_to_string(unsigned value, unsigned digits) {
  std::string s = std::to_string(value);
  if (s.size() < digits) s.insert(0, digits - s.size(), '0');
  return s;
}

void write_vtu_box(const std::string& path, int dim, const double* corners, int n_refine,
                   const std::vector<std::string>& names, const std::vector<const double*>& fields) {
  write_vtu_box(path, dim, corners, n_refine, names, fields, {});
}

void write_vtu_box(const std::string& path, int dim, const double* corners, int n_refine,
                   const std::vector<std::string>& names, const std::vector<const double*>& fields,
                   const std::vector<CellField>& cell_fields) {
  if (dim != 2 && dim != 3) throw std::invalid_argument("dim must be 2 or 3");
  if (names.size() != fields.size()) throw std::invalid_argument("one name per field");
  const int m = 1 << n_refine, S = m + 1;
  const long np = (dim == 2) ? (long)S * S : (long)S * S * S;
  const long nc = (dim == 2) ? (long)m * m : (long)m * m * m;
  // corner 0 is the low corner, corner 2^d the high end of direction d
  double lo[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    lo[d] = corners[d];
    h[d] = (corners[(1 << d) * dim + d] - corners[d]) / m;
  }
  std::FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot open " + path + " for writing");
  std::fprintf(f, "<?xml version=\"1.0\"?>\n");
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  std::fprintf(f, "<UnstructuredGrid>\n<Piece NumberOfPoints=\"%ld\" NumberOfCells=\"%ld\">\n", np, nc);
  std::fprintf(f, "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (long p = 0; p < np; ++p) {
    const long i = p % S, j = (p / S) % S, k = (dim == 3) ? p / ((long)S * S) : 0;
    std::fprintf(f, "%.17g %.17g %.17g\n", lo[0] + i * h[0], lo[1] + j * h[1],
                 dim == 3 ? lo[2] + k * h[2] : 0.0);
  }
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n");
  std::fprintf(f, "<DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n");
  for (long c = 0; c < nc; ++c) {
    const long i = c % m, j = (c / m) % m, k = (dim == 3) ? c / ((long)m * m) : 0;
    const long n0 = (k * S + j) * S + i;
    // VTK vertex order: counter-clockwise in the bottom face, then the top face
    const long q[4] = {n0, n0 + 1, n0 + S + 1, n0 + S};
    if (dim == 2) {
      std::fprintf(f, "%ld %ld %ld %ld\n", q[0], q[1], q[2], q[3]);
    } else {
      const long up = (long)S * S;
      std::fprintf(f, "%ld %ld %ld %ld %ld %ld %ld %ld\n", q[0], q[1], q[2], q[3], q[0] + up,
                   q[1] + up, q[2] + up, q[3] + up);
    }
  }
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  const int nv = 1 << dim;
  for (long c = 1; c <= nc; ++c) std::fprintf(f, "%ld\n", c * nv);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (long c = 0; c < nc; ++c) std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);  // VTK_QUAD, VTK_HEXAHEDRON
  std::fprintf(f, "</DataArray>\n</Cells>\n");
  std::fprintf(f, "<PointData%s%s%s>\n", names.empty() ? "" : " Scalars=\"",
               names.empty() ? "" : names[0].c_str(), names.empty() ? "" : "\"");
  for (size_t a = 0; a < names.size(); ++a) {
    std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", names[a].c_str());
    for (long p = 0; p < np; ++p) std::fprintf(f, "%.17g\n", fields[a][p]);
    std::fprintf(f, "</DataArray>\n");
  }
  std::fprintf(f, "</PointData>\n");
  if (!cell_fields.empty()) {
    std::fprintf(f, "<CellData>\n");
    for (const CellField& cf : cell_fields) {
      if (cf.ncomp != 1 && cf.ncomp != 3) {
        std::fclose(f);
        throw std::invalid_argument("cell fields have 1 or 3 components");
      }
      std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"%d\" format=\"ascii\">\n",
                   cf.name.c_str(), cf.ncomp);
      for (long c = 0; c < nc; ++c) {
        for (int k = 0; k < cf.ncomp; ++k) std::fprintf(f, k ? " %.17g" : "%.17g", cf.values[c * cf.ncomp + k]);
        std::fprintf(f, "\n");
      }
      std::fprintf(f, "</DataArray>\n");
    }
    std::fprintf(f, "</CellData>\n");
  }
  std::fprintf(f, "</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) throw std::runtime_error("error writing " + path);
}

}  // namespace msfec
