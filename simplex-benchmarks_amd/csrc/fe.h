// Reference-tetrahedron finite-element data: Lagrange P1..P3 bases in
// closed barycentric form, node layout, quadrature selection, faces.
//
// Conventions (self-defined; the reference delegates them to deal.II's
// FE_SimplexP / QGaussSimplex, CG/cells_batched/application.cc:1609-1614):
//   * unit tetrahedron, vertices v0=(0,0,0) v1=(1,0,0) v2=(0,1,0) v3=(0,0,1);
//     barycentrics l0=1-x-y-z, l1=x, l2=y, l3=z.
//   * node order: 4 vertices, then per edge (0,1)(0,2)(0,3)(1,2)(1,3)(2,3) its
//     p-1 nodes ordered from the lower local vertex to the higher one, then one
//     node per face for p=3 (face f is opposite vertex f).
//   * face f has local vertices kFaceVerts[f] (ascending).
#pragma once
#include <array>
#include <cstdint>
#include <vector>

namespace smf {

constexpr int kEdgeVerts[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr int kFaceVerts[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

inline int n_loc_of(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

struct Quadrature {
  int degree = 0;                  // exactness degree
  std::vector<std::array<double, 3>> pts;  // unit-simplex coordinates (z unused on triangles)
  std::vector<double> w;           // weights (sum = 1/6 tet, 1/2 triangle)
  int size() const { return (int)w.size(); }
};

// Smallest stored symmetric rule exact to `degree` (tet: 1,2,4,5,6; tri: 1,2,4,6).
Quadrature tet_rule(int degree);
Quadrature tri_rule(int degree);

struct LocalNode {
  int kind;     // 0 vertex, 1 edge, 2 face
  int entity;   // local vertex / edge / face number
  int sub;      // index within the entity (edge: 0 = nearer the lower local vertex)
  std::array<int, 4> alpha;  // lattice multi-index, sum = p
};

struct Lagrange {
  int p = 0, n_loc = 0;
  std::vector<LocalNode> nodes;
  explicit Lagrange(int degree);
  // values[n_loc], grads[n_loc*3] (d/dx, d/dy, d/dz on the unit tet) at x.
  void eval(const double x[3], double* values, double* grads) const;
  // Reference coordinate of node i.
  std::array<double, 3> node_point(int i) const;
};

// Per-face affine parametrisation: face f point with triangle coords (s,t)
// is v_a + s (v_b - v_a) + t (v_c - v_a) with (a,b,c) = kFaceVerts[f].
std::array<double, 3> face_to_cell(int f, double s, double t);

}  // namespace smf
