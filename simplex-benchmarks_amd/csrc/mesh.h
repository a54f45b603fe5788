// Tetrahedral meshes: Kuhn (Freudenthal) cube generator with seeded vertex
// perturbation, face/edge connectivity, and the hierarchical cell reordering
// that cuts the mesh into spatially compact patches.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace smf {

struct Mesh {
  int64_t nv = 0, nc = 0;
  std::vector<double> x;         // 3*nv
  std::vector<int32_t> cells;    // 4*nc, local vertex order as given
  // Kuhn lattice information (kuhn_n > 0 only for generated cubes): integer
  // lattice coordinates of each vertex on the unperturbed (kuhn_n)^3 grid.
  int kuhn_n = 0;
  std::vector<int32_t> lattice;  // 3*nv
};

// splitmix64 (Steele, Lea, Flood 2014): the PRNG behind every seeded input.
struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform on [-1, 1)
  double uniform_pm1() { return 2.0 * ((next() >> 11) * 0x1.0p-53) - 1.0; }
};

// Unit cube [0,1]^3 split into n^3 hexes x 6 Kuhn tetrahedra (each hex cut
// along its (0,0,0)-(1,1,1) diagonal).  Vertex (i,j,k) has id
// i + (n+1)(j + (n+1)k); hex (i,j,k) has id i + n(j + n k); its tetrahedron s
// (s = 0..5 over the axis permutations (x,y,z)(x,z,y)(y,x,z)(y,z,x)(z,x,y)(z,y,x))
// is cell 6*hex + s with vertices along the lattice path 000 -> e_a -> e_a+e_b
// -> 111; odd permutations swap local vertices 2 and 3 so that det J > 0.
// Interior vertices are displaced per coordinate by perturb*h*U[-1,1) drawn
// from splitmix64(seed) in vertex-id order (x, y, z per vertex).
Mesh kuhn_cube(int n, double perturb, uint64_t seed);

// Tetrahedral mesh from a legacy ASCII VTK unstructured grid (vtk.cpp);
// n_flipped counts tetrahedra whose local vertices 1 and 2 were swapped to
// make det J > 0.
Mesh read_vtk(const std::string& path, int64_t* n_flipped);

struct FaceInfo {
  // per cell 4 face ids (face f opposite local vertex f)
  std::vector<int64_t> cell_face;              // 4*nc
  // per face: (cell, local face) of both sides; side 1 cell = -1 on the boundary
  std::vector<std::array<int64_t, 2>> face_cell;
  std::vector<std::array<int8_t, 2>> face_lf;
  int64_t n_faces = 0, n_boundary = 0;
};
FaceInfo build_faces(const Mesh& m);

struct EdgeInfo {
  std::vector<int64_t> cell_edge;              // 6*nc (local edge order of fe.h)
  std::vector<std::array<int32_t, 2>> verts;   // global vertex ids, ascending
  int64_t n_edges = 0;
};
EdgeInfo build_edges(const Mesh& m);

// Hierarchical reordering by recursive coordinate bisection of the cell
// centroids: split at the median of the longest bounding-box extent until a
// range holds at most max_patch cells; leaves, visited depth first, are the
// patches.  With reorder == false the generator order is kept and patches are
// consecutive chunks.  perm[new] = old; patch_ptr has n_patches+1 entries.
void reorder_cells(const Mesh& m, bool reorder, int max_patch, std::vector<int64_t>& perm,
                   std::vector<int64_t>& patch_ptr);

// Geometry of an affine tetrahedron: J = [x1-x0, x2-x0, x3-x0] (columns).
struct CellGeom {
  double J[3][3], Jinv[3][3], det;
};
CellGeom cell_geometry(const Mesh& m, int64_t c);

}  // namespace smf
