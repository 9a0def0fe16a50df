// Mesh assets for the library's MJCF loader (crates/pippin/src/mesh.rs).
#pragma once
#include <string>
#include <vector>

namespace phh {

// Hulls of tabletop objects: the brute-force construction below is sized for
// small point sets; the device scans every hull vertex for its support.
constexpr int kMaxHullPoints = 64;

struct Hull {
  std::vector<double> vert;               // m x 3, sorted by (x, y, z)
  std::vector<std::vector<int>> faces;    // convex polygons, CCW from outside, smallest index first; sorted
};

// 'v x y z' lines of an OBJ file, each coordinate read as f32 (mesh.rs:42-53),
// widened and scaled per axis.
bool read_obj_vertices(const std::string& text, const double* scale, std::vector<double>* out);
// Triangle corners of a binary or ASCII STL file (f32), scaled per axis.
bool read_stl_vertices(const std::string& bytes, const double* scale, std::vector<double>* out);
// Convex hull of a point set (mesh::convex_hull, mesh.rs:63-69) as a set of
// extreme points and polygonal faces; false (err set) for a degenerate set.
bool convex_hull(const std::vector<double>& pts, Hull* out, std::string* err);
// Volume, centre of mass and inertia about it per unit density
// (mesh::mass_properties, mesh.rs:72-105), over the faces' fan triangles.
void mass_properties(const Hull& h, double* vol, double* com, double* icom9);

}  // namespace phh
