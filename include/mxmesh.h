/* Host set-up of the explicit Euler residual: meshes, FP64 geometry, cell renumbering and the
 * face2 least-squares gradient stencil, behind a plain C ABI.
 *
 * These replace the host work the reference does before its device stepper exists: mesh
 * assembly and RCM renumbering (core/include/autan/core/mesh.hpp:277-282,
 * core/include/autan/core/renumber.hpp:16-33), compute_geometry
 * (core/include/autan/core/geometry.hpp:41), pack_geometry and cell_stencil
 * (codegen/include/autan/kernels/geometry.hpp:16-105) and Euler::set_boundary's gradient weights
 * (physics/include/autan/physics/euler.hpp:129, physics/src/euler.cpp:340-416). Their output is
 * exactly the array set that mxeuler_create (mxeuler.h) takes.
 *
 * Errors: every call returning int returns 0 on success and 1 on failure, with a message in
 * err[0 .. err_len) (the reference's capi convention, physics/mms/capi.cpp:126-347).
 */
#ifndef MXMESH_H
#define MXMESH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxmesh mxmesh;

/* n^3 hexahedra of [lo, hi]^3, periodic in every direction (n >= 3). */
int mxmesh_periodic_hex(int n, double lo, double hi, mxmesh** out, char* err, size_t err_len);
/* n^3 cubes of the unit cube, 6 Kuhn tetrahedra each, interior nodes jittered by +-jitter h. */
int mxmesh_kuhn_tets(int n, double jitter, uint64_t seed, mxmesh** out, char* err, size_t err_len);
/* nx x ny x nz hexahedra of a channel with a Gaussian bump h exp(-((x - x0) / w)^2) on y = 0. */
int mxmesh_bump_hex(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                    mxmesh** out, char* err, size_t err_len);
/* The bump channel, hybrid: prism_layers rows of prisms on the lower wall, rows [band_lo, band_hi)
 * merged 2 x 2 x 2 into polyhedra, hexahedra elsewhere. */
int mxmesh_bump_hybrid(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                       int prism_layers, int band_lo, int band_hi, mxmesh** out, char* err, size_t err_len);
/* An OpenFOAM polyMesh directory, ASCII or binary (io/src/polymesh.cpp): points, faces, owner,
 * neighbour, boundary; every count, label and patch range checked (E0130 a missing file, E0132
 * malformed). */
int mxmesh_read_polymesh(const char* dir, mxmesh** out, char* err, size_t err_len);
/* A Gmsh MSH 4.1 file, ASCII or binary (io/src/gmsh.cpp; docs/contracts/mesh.md, Readers): linear volume
 * elements; boundary zone k is the k-th 2-D physical group by increasing tag, then one zone for
 * boundary faces in no group; surface links of $Periodic by a translation become periodic faces.
 * Errors carry the reference's reader codes (E0130 missing file, E0131 not a mesh file, E0132
 * malformed, E0133 unsupported: MSH 2.2, big-endian, higher-order elements, periodic links that are
 * not translations; E0135 an entity in two physical groups). */
int mxmesh_read_gmsh(const char* path, mxmesh** out, char* err, size_t err_len);
void mxmesh_destroy(mxmesh* m);

/* Renumbers the cells by reverse Cuthill-McKee (before mxmesh_prepare). */
int mxmesh_renumber_rcm(mxmesh* m, char* err, size_t err_len);
/* Renumbers the cells along a Morton curve of their centres (before mxmesh_prepare): runs of
 * consecutive cells are compact regions, which keeps stencil gathers in cache. */
int mxmesh_renumber_morton(mxmesh* m, char* err, size_t err_len);
/* FP64 geometry, packed device geometry and the face2 stencil. */
int mxmesh_prepare(mxmesh* m, char* err, size_t err_len);
/* One condition code per boundary face (BoundaryKind values 0..5): recomputes the gradient
 * weights; *kept receives the number of cells whose fit keeps every boundary face. */
int mxmesh_set_boundary_kinds(mxmesh* m, const int32_t* kind, int64_t* kept, char* err, size_t err_len);

/* Sizes: n_cells, n_faces, n_interior, n_points, n_incidences, n_stencil, all_planar, n_face_nodes. */
void mxmesh_sizes(const mxmesh* m, int64_t out[8]);

/* A named array of the mesh: *data points into the mesh (valid until it changes), *count is the
 * number of elements, *dtype is 0 for int32, 1 for int64, 2 for double. Names: owner, neighbour,
 * cell_face_offsets, cell_faces, face_zone, shift, face_node_offsets, face_nodes, x, y, z,
 * face_normal, face_area, r_left, r_right, face_moment, cell_volume, cell_centroid, face_centroid,
 * stencil_offsets, stencil, stencil_weights. */
int mxmesh_array(const mxmesh* m, const char* name, const void** data, int64_t* count, int* dtype, char* err,
                 size_t err_len);

#ifdef __cplusplus
}
#endif

#endif
