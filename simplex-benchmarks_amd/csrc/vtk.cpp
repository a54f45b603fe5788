// Legacy VTK reader for tetrahedral meshes (the external-mesh input of
// CG/cells_batched/application.cc:1699-1756, GridIn::read_vtk).
//
// Reads ASCII UNSTRUCTURED_GRID files: POINTS, then CELLS either in the
// classic layout ("CELLS n size" followed by "k i0 .. ik-1" per cell) or in
// the VTK 5 layout ("CELLS n+1 m", then OFFSETS and CONNECTIVITY arrays),
// and CELL_TYPES.  Tetrahedra (type 10) become cells; points, lines and
// triangles (types 1, 3, 5: boundary descriptions) are skipped; any other
// cell type is rejected.  Tetrahedra of negative volume get local vertices 1
// and 2 swapped, so that det J > 0.
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "mesh.h"

namespace smf {

namespace {

struct Tokens {
  std::istream& in;
  std::string tok;
  explicit Tokens(std::istream& s) : in(s) {}
  bool next() { return static_cast<bool>(in >> tok); }
  std::string need(const char* what) {
    if (!next()) throw std::runtime_error(std::string("VTK: unexpected end of file, expected ") + what);
    return tok;
  }
  int64_t integer(const char* what) {
    const std::string t = need(what);
    size_t pos = 0;
    long long v = 0;
    try {
      v = std::stoll(t, &pos);
    } catch (...) {
      pos = 0;
    }
    if (pos != t.size()) throw std::runtime_error(std::string("VTK: expected an integer (") + what + "), got '" + t + "'");
    return v;
  }
  double real(const char* what) {
    const std::string t = need(what);
    size_t pos = 0;
    double v = 0;
    try {
      v = std::stod(t, &pos);
    } catch (...) {
      pos = 0;
    }
    if (pos != t.size()) throw std::runtime_error(std::string("VTK: expected a number (") + what + "), got '" + t + "'");
    return v;
  }
};

std::string upper(std::string s) {
  for (char& c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}

// skips a "NAME type" header of a VTK 5 array (OFFSETS / CONNECTIVITY)
void array_header(Tokens& t, const char* name) {
  if (upper(t.need(name)) != name) throw std::runtime_error(std::string("VTK: expected ") + name + ", got '" + t.tok + "'");
  t.need("array type");
}

}  // namespace

Mesh read_vtk(const std::string& path, int64_t* n_flipped) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("VTK: cannot open '" + path + "'");
  std::string line;
  if (!std::getline(f, line) || line.find("vtk") == std::string::npos)
    throw std::runtime_error("VTK: '" + path + "' has no '# vtk DataFile' header");
  std::getline(f, line);  // title
  Tokens t(f);
  if (upper(t.need("ASCII/BINARY")) != "ASCII")
    throw std::runtime_error("VTK: only ASCII files are read (found " + t.tok + ")");
  if (upper(t.need("DATASET")) != "DATASET" || upper(t.need("dataset type")) != "UNSTRUCTURED_GRID")
    throw std::runtime_error("VTK: dataset must be UNSTRUCTURED_GRID");

  Mesh m;
  std::vector<int64_t> offsets, conn;  // cell c: conn[offsets[c] .. offsets[c+1])
  std::vector<int64_t> types;
  bool have_points = false, have_cells = false;
  while (t.next()) {
    const std::string key = upper(t.tok);
    if (key == "POINTS") {
      const int64_t n = t.integer("number of points");
      t.need("point type");
      if (n <= 0) throw std::runtime_error("VTK: no points");
      m.nv = n;
      m.x.resize(3 * n);
      for (int64_t i = 0; i < 3 * n; ++i) m.x[i] = t.real("point coordinate");
      have_points = true;
    } else if (key == "CELLS") {
      const int64_t a = t.integer("CELLS count"), b = t.integer("CELLS size");
      if (a < 0 || b < 0) throw std::runtime_error("VTK: bad CELLS header");
      // VTK 5 layout when OFFSETS follows, classic layout otherwise
      const std::streampos here = f.tellg();
      std::string peek;
      f >> peek;
      f.seekg(here);
      if (upper(peek) == "OFFSETS") {
        array_header(t, "OFFSETS");
        offsets.resize(a);
        for (int64_t i = 0; i < a; ++i) offsets[i] = t.integer("offset");
        array_header(t, "CONNECTIVITY");
        conn.resize(b);
        for (int64_t i = 0; i < b; ++i) conn[i] = t.integer("connectivity");
        if (a < 1 || offsets[0] != 0 || offsets[a - 1] != b)
          throw std::runtime_error("VTK: OFFSETS do not match CONNECTIVITY");
        // every cell's range [offsets[c], offsets[c+1]) must lie inside CONNECTIVITY
        for (int64_t i = 1; i < a; ++i)
          if (offsets[i] < offsets[i - 1] || offsets[i] > b)
            throw std::runtime_error("VTK: OFFSETS must rise monotonically within [0, " + std::to_string(b) + "]");
      } else {
        if (b < a) throw std::runtime_error("VTK: bad CELLS header");
        offsets.assign(1, 0);
        conn.reserve(b - a);
        for (int64_t c = 0; c < a; ++c) {
          const int64_t k = t.integer("cell vertex count");
          if (k < 1 || k > 64) throw std::runtime_error("VTK: bad cell vertex count");
          for (int64_t i = 0; i < k; ++i) conn.push_back(t.integer("cell vertex"));
          offsets.push_back((int64_t)conn.size());
        }
        if ((int64_t)conn.size() + a != b) throw std::runtime_error("VTK: CELLS size does not match its entries");
      }
      have_cells = true;
    } else if (key == "CELL_TYPES") {
      const int64_t n = t.integer("number of cell types");
      if (n < 0) throw std::runtime_error("VTK: negative CELL_TYPES count");
      types.resize(n);
      for (int64_t i = 0; i < n; ++i) types[i] = t.integer("cell type");
    } else if (key == "FIELD") {
      // FIELD name n, then per array "name components tuples type" and its values
      t.need("field name");
      const int64_t na = t.integer("number of field arrays");
      if (na < 0) throw std::runtime_error("VTK: negative FIELD array count");
      for (int64_t a = 0; a < na; ++a) {
        t.need("field array name");
        const int64_t nc = t.integer("components"), nt = t.integer("tuples");
        if (nc < 0 || nt < 0) throw std::runtime_error("VTK: negative FIELD array size");
        t.need("field array type");
        for (int64_t i = 0; i < nc * nt; ++i) t.need("field value");
      }
    } else if (key == "METADATA") {
      // a metadata block ends at the next empty line
      std::getline(f, line);
      while (std::getline(f, line) && line.find_first_not_of(" \t\r") != std::string::npos) {
      }
    } else if (key == "CELL_DATA" || key == "POINT_DATA") {
      break;  // attribute data is not needed
    } else {
      throw std::runtime_error("VTK: unexpected section '" + t.tok + "'");
    }
  }
  if (!have_points || !have_cells) throw std::runtime_error("VTK: POINTS and CELLS are required");
  const int64_t ncell_in = (int64_t)offsets.size() - 1;
  if ((int64_t)types.size() != ncell_in) throw std::runtime_error("VTK: CELL_TYPES does not match CELLS");
  int64_t flipped = 0;
  for (int64_t c = 0; c < ncell_in; ++c) {
    const int64_t ty = types[c], k = offsets[c + 1] - offsets[c];
    if (ty == 1 || ty == 3 || ty == 5) continue;
    if (ty != 10) throw std::runtime_error("VTK: cell type " + std::to_string(ty) + " (only tetrahedra, type 10)");
    if (k != 4) throw std::runtime_error("VTK: a tetrahedron with " + std::to_string(k) + " vertices");
    int32_t v[4];
    for (int i = 0; i < 4; ++i) {
      const int64_t id = conn[offsets[c] + i];
      if (id < 0 || id >= m.nv) throw std::runtime_error("VTK: cell references a missing point");
      v[i] = (int32_t)id;
    }
    double J[3][3];
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) J[d][e] = m.x[3 * v[e + 1] + d] - m.x[3 * v[0] + d];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (det < 0) {
      std::swap(v[1], v[2]);
      ++flipped;
    }
    m.cells.insert(m.cells.end(), v, v + 4);
  }
  m.nc = (int64_t)m.cells.size() / 4;
  if (m.nv >= (int64_t(1) << 31)) throw std::runtime_error("VTK: more than 2^31 - 1 points");
  if (n_flipped) *n_flipped = flipped;
  return m;
}

}  // namespace smf
