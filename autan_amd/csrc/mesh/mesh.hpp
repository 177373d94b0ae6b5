// Host-side mesh, geometry and gradient-stencil setup for the explicit Euler residual.
//
// This is the set-up half of the path: what the reference's Euler<P> constructor, pack_geometry
// and cell_stencil compute once on the host before any residual is evaluated
// (physics/src/euler.cpp:143-246, codegen/cpp/geometry.cpp:166-307,432-617,
// core/src/geometry.cpp:34-198, core/src/renumber.cpp:109-148). Everything here is FP64 and
// built with -ffp-contract=off so that the values match a restatement compiled the same way.
//
// Periodic meshes: a periodic face is an ordinary interior face whose neighbour sits, as seen
// from the owner, at its own position plus `shift` (3 x n_interior, zero for other faces). Face
// nodes are stored in the owner's frame. The reference reaches the same values through periodic
// ghost images (core/include/autan/core/mesh.hpp, ghost_translation).
#pragma once

#include <array>
#include <initializer_list>
#include <cstdint>
#include <string>
#include <vector>

namespace mx {

using index_t = std::int32_t;
using offset_t = std::int64_t;

/// Boundary condition codes, the values of the reference's BoundaryKind
/// (physics/include/autan/physics/euler.hpp:30-37).
enum BoundaryKind : int { bc_state = 0, bc_farfield = 1, bc_slip = 2, bc_inlet = 3, bc_outlet = 4, bc_wall = 5 };

struct Mesh {
  index_t n_cells = 0;
  index_t n_faces = 0;
  index_t n_interior = 0;
  std::vector<double> x, y, z;  // points
  std::vector<offset_t> face_node_offsets;  // n_faces + 1
  std::vector<index_t> face_nodes;
  std::vector<index_t> owner;      // n_faces
  std::vector<index_t> neighbour;  // n_interior
  std::vector<index_t> face_zone;  // n_faces - n_interior: zone of each boundary face
  std::vector<double> shift;       // 3 x n_interior (empty: no periodic face)
  // Derived by finish(): each cell's faces in increasing order (mesh contract, invariant 4).
  std::vector<offset_t> cell_face_offsets;  // n_cells + 1
  std::vector<index_t> cell_faces;

  [[nodiscard]] index_t n_boundary() const { return n_faces - n_interior; }
  [[nodiscard]] bool periodic() const { return !shift.empty(); }
  [[nodiscard]] std::array<double, 3> face_shift(index_t f) const {
    if (shift.empty() || f >= n_interior) return {0.0, 0.0, 0.0};
    const auto ni = static_cast<std::size_t>(n_interior), fi = static_cast<std::size_t>(f);
    return {shift[fi], shift[ni + fi], shift[2 * ni + fi]};
  }
};

/// Faces before assembly, flat: face i has nodes[off[i] .. off[i+1]) in order (right-hand normal
/// from cell0 towards cell1), cell1 = -1 on a boundary face with its zone in `zone`, and on a
/// periodic face the translation of cell1 as seen from cell0 in `shift` (3 per face, or empty).
struct RawFaces {
  std::vector<offset_t> off{0};
  std::vector<index_t> nodes;
  std::vector<index_t> c0, c1, zone;
  std::vector<double> shift;  // 3 x faces, face-major (x, y, z of face i at 3 i .. 3 i + 2)
  void add(std::initializer_list<index_t> n, index_t a, index_t b, index_t z = -1) {
    nodes.insert(nodes.end(), n);
    off.push_back(static_cast<offset_t>(nodes.size()));
    c0.push_back(a);
    c1.push_back(b);
    zone.push_back(z);
  }
};

/// Canonical order (core/include/autan/core/mesh.hpp:277-282): owner < neighbour on interior faces
/// (a face is reversed to get there: nodes reversed and, on a periodic face, moved to the new
/// owner's frame with the shift negated), interior faces sorted by (owner, neighbour), then
/// boundary faces by zone then owner. Cell order is kept.
Mesh assemble(index_t n_cells, std::vector<double> x, std::vector<double> y, std::vector<double> z,
              RawFaces faces);

/// Recomputes cell_face_offsets and cell_faces from owner / neighbour.
void finish(Mesh& m);

/// FP64 geometry of core/src/geometry.cpp: fan triangles about the node mean.
struct Geometry {
  std::vector<double> face_area;      // 3 x nf, area vectors (owner -> neighbour)
  std::vector<double> face_area_mag;  // nf
  std::vector<double> face_centroid;  // 3 x nf, owner frame
  std::vector<double> cell_volume;    // nc
  std::vector<double> cell_centroid;  // 3 x nc
};
Geometry compute_geometry(const Mesh& m);

/// The device-facing geometry of pack_geometry (codegen/cpp/geometry.cpp:166-307): unit normals,
/// per-face offsets r = x_f - x_c to each side (the FP64 difference `rel` rounds), the nine face
/// moments (signed zeros on faces proven planar) and whether every face is proven planar.
struct Packed {
  std::vector<double> face_normal;  // 3 x nf
  std::vector<double> face_area;    // nf (magnitude)
  std::vector<double> r_left;       // 3 x nf
  std::vector<double> r_right;      // 3 x nf (zero on boundary faces)
  std::vector<double> face_moment;  // 9 x nf
  std::vector<double> cell_volume;  // nc
  std::vector<double> cell_centroid;  // 3 x nc
  std::vector<double> face_centroid;  // 3 x nf
  bool all_planar = true;
};
Packed pack(const Mesh& m, const Geometry& g);

/// The face2 gradient stencil of cell_stencil (codegen/cpp/geometry.cpp:432-617): per cell its
/// face neighbours and their face neighbours in increasing index, then the boundary faces of both
/// as -(f + 1) in increasing index. On a periodic mesh `translation` holds, per entry, the
/// periodic shift of the image the cell sees (3 x entries); it is empty otherwise.
struct Stencil {
  std::vector<offset_t> offsets;  // nc + 1
  std::vector<index_t> entries;
  std::vector<double> translation;  // 3 x entries, periodic meshes only
};
Stencil face2_stencil(const Mesh& m, const Packed& p);

/// Euler::Impl::compute_gradient_weights (physics/src/euler.cpp:340-416): weights 3 x entries.
/// Boundary faces take part only where their condition is `state`, unless the remaining entries
/// leave the fit singular (det <= 1e-10 |a00 a11 a22|), then the cell keeps every boundary face.
/// `kind` has one code per boundary face. Returns the number of cells that kept every boundary face.
std::size_t gradient_weights(const Mesh& m, const Packed& p, const Stencil& s,
                             const std::vector<int>& kind, std::vector<double>& w);

/// Reverse Cuthill-McKee order of the cell graph (core/src/renumber.cpp:109-148): new_to_old.
std::vector<index_t> rcm_order(const Mesh& m);
/// Morton (Z-curve) order of the cell centres (mean of the cells' face nodes) quantised to 2^21
/// per axis over the bounding box: new_to_old. A run of consecutive cells is then a compact 3D
/// region, so the cells a GPU works on together share their stencil neighbours in cache.
std::vector<index_t> morton_order(const Mesh& m);
/// The mesh with cells renumbered (new_to_old), re-assembled in canonical order.
Mesh renumber(const Mesh& m, const std::vector<index_t>& new_to_old);

// Generators ------------------------------------------------------------------------------------

/// n^3 hexahedra filling [lo, hi]^3, periodic in x, y and z (n >= 3).
Mesh periodic_hex(int n, double lo, double hi);
/// n^3 cubes of [0,1]^3 each split into 6 Kuhn tetrahedra; interior vertices jittered uniformly by
/// +-jitter*h per axis from a seeded generator. Boundary zones 0..5 = xmin, xmax, ymin, ymax,
/// zmin, zmax.
Mesh kuhn_tets(int n, double jitter, std::uint64_t seed);
/// nx x ny x nz hexahedra of the box [0,lx] x [0,ly] x [0,lz] with the lower wall raised by a
/// Gaussian bump h exp(-((x - x0)/w)^2) blended linearly to zero at the top. Zones 0..5 = inflow
/// (xmin), outflow (xmax), lower wall (ymin), upper wall (ymax), zmin, zmax.
Mesh bump_hex(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0);
/// The same channel, hybrid: rows j < prism_layers above the lower wall split into two
/// triangular prisms each (the x-y quad cut along its diagonal, extruded in z), rows band_lo <= j < band_hi merged 2 x 2 x 2
/// into polyhedra (nx, nz and the band height even). A face between two merged cells is one
/// 8-node polygon (its quads' outer ring); a merged cell shows a fine neighbour its 4 quads. Zones
/// as bump_hex.
Mesh bump_hybrid(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                 int prism_layers, int band_lo, int band_hi);
/// OpenFOAM polyMesh directory, ASCII or binary: points, faces, owner, neighbour, boundary.
Mesh read_polymesh(const std::string& dir);
/// Gmsh MSH 4.1 file, ASCII or little-endian binary (io/src/gmsh.cpp): linear tetrahedra,
/// hexahedra, prisms and pyramids; boundary zones the 2-D physical groups in increasing tag order,
/// then one zone for boundary faces in no group; cells in file order; surface links of $Periodic
/// by a translation become periodic faces. Other versions and links that rotate are refused.
Mesh read_gmsh(const std::string& path);

}  // namespace mx
