// C ABI of the host set-up library (include/mxmesh.h).
#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "../../../include/mxmesh.h"
#include "mesh.hpp"

struct mxmesh {
  mx::Mesh mesh;
  bool prepared = false;
  mx::Packed packed;
  mx::Stencil stencil;
  std::vector<double> weights;
  std::vector<double> area_vector;  // 3 x nf, FP64 area vectors (owner -> neighbour)
};

namespace {

int fail(const std::exception& e, char* err, size_t err_len) {
  if (err != nullptr && err_len > 0) {
    std::strncpy(err, e.what(), err_len - 1);
    err[err_len - 1] = '\0';
  }
  return 1;
}

template <class F>
int guarded(char* err, size_t err_len, F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, err_len);
  }
}

int make(mx::Mesh&& m, mxmesh** out) {
  auto h = std::make_unique<mxmesh>();
  h->mesh = std::move(m);
  *out = h.release();
  return 0;
}

}  // namespace

extern "C" {

int mxmesh_periodic_hex(int n, double lo, double hi, mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { make(mx::periodic_hex(n, lo, hi), out); });
}

int mxmesh_kuhn_tets(int n, double jitter, uint64_t seed, mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { make(mx::kuhn_tets(n, jitter, seed), out); });
}

int mxmesh_bump_hex(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                    mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { make(mx::bump_hex(nx, ny, nz, lx, ly, lz, h, w, x0), out); });
}

int mxmesh_bump_hybrid(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                       int prism_layers, int band_lo, int band_hi, mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    make(mx::bump_hybrid(nx, ny, nz, lx, ly, lz, h, w, x0, prism_layers, band_lo, band_hi), out);
  });
}

int mxmesh_read_polymesh(const char* dir, mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { make(mx::read_polymesh(dir), out); });
}

int mxmesh_read_gmsh(const char* path, mxmesh** out, char* err, size_t err_len) {
  return guarded(err, err_len, [&] { make(mx::read_gmsh(path), out); });
}

void mxmesh_destroy(mxmesh* m) { delete m; }

int mxmesh_renumber_rcm(mxmesh* m, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (m->prepared) throw std::logic_error("renumber before prepare");
    m->mesh = mx::renumber(m->mesh, mx::rcm_order(m->mesh));
  });
}

int mxmesh_renumber_morton(mxmesh* m, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (m->prepared) throw std::logic_error("renumber before prepare");
    m->mesh = mx::renumber(m->mesh, mx::morton_order(m->mesh));
  });
}

int mxmesh_prepare(mxmesh* m, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    const mx::Geometry g = mx::compute_geometry(m->mesh);
    m->packed = mx::pack(m->mesh, g);
    m->area_vector = g.face_area;
    m->stencil = mx::face2_stencil(m->mesh, m->packed);
    m->prepared = true;
    // Every boundary face a `state` face until the caller says otherwise.
    std::vector<int> kind(static_cast<std::size_t>(m->mesh.n_boundary()), mx::bc_state);
    mx::gradient_weights(m->mesh, m->packed, m->stencil, kind, m->weights);
  });
}

int mxmesh_set_boundary_kinds(mxmesh* m, const int32_t* kind, int64_t* kept, char* err, size_t err_len) {
  return guarded(err, err_len, [&] {
    if (!m->prepared) throw std::logic_error("prepare before setting boundary kinds");
    std::vector<int> k(kind, kind + m->mesh.n_boundary());
    const std::size_t n = mx::gradient_weights(m->mesh, m->packed, m->stencil, k, m->weights);
    if (kept != nullptr) *kept = static_cast<int64_t>(n);
  });
}

void mxmesh_sizes(const mxmesh* m, int64_t out[8]) {
  const auto& me = m->mesh;
  out[0] = me.n_cells;
  out[1] = me.n_faces;
  out[2] = me.n_interior;
  out[3] = static_cast<int64_t>(me.x.size());
  out[4] = me.cell_face_offsets.empty() ? 0 : me.cell_face_offsets.back();
  out[5] = m->prepared ? static_cast<int64_t>(m->stencil.entries.size()) : 0;
  out[6] = m->prepared ? (m->packed.all_planar ? 1 : 0) : -1;
  out[7] = static_cast<int64_t>(me.face_nodes.size());
}

int mxmesh_array(const mxmesh* m, const char* name, const void** data, int64_t* count, int* dtype, char* err,
                 size_t err_len) {
  return guarded(err, err_len, [&] {
    const std::string n(name);
    auto i32 = [&](const std::vector<int32_t>& v) {
      *data = v.data();
      *count = static_cast<int64_t>(v.size());
      *dtype = 0;
    };
    auto i64 = [&](const std::vector<int64_t>& v) {
      *data = v.data();
      *count = static_cast<int64_t>(v.size());
      *dtype = 1;
    };
    auto f64 = [&](const std::vector<double>& v) {
      *data = v.data();
      *count = static_cast<int64_t>(v.size());
      *dtype = 2;
    };
    const auto& me = m->mesh;
    if (n == "owner") return i32(me.owner);
    if (n == "neighbour") return i32(me.neighbour);
    if (n == "cell_face_offsets") return i64(me.cell_face_offsets);
    if (n == "cell_faces") return i32(me.cell_faces);
    if (n == "face_zone") return i32(me.face_zone);
    if (n == "shift") return f64(me.shift);
    if (n == "face_node_offsets") return i64(me.face_node_offsets);
    if (n == "face_nodes") return i32(me.face_nodes);
    if (n == "x") return f64(me.x);
    if (n == "y") return f64(me.y);
    if (n == "z") return f64(me.z);
    if (!m->prepared) throw std::logic_error(n + ": unknown before mxmesh_prepare");
    const auto& p = m->packed;
    if (n == "face_normal") return f64(p.face_normal);
    if (n == "face_area") return f64(p.face_area);
    if (n == "r_left") return f64(p.r_left);
    if (n == "r_right") return f64(p.r_right);
    if (n == "face_moment") return f64(p.face_moment);
    if (n == "cell_volume") return f64(p.cell_volume);
    if (n == "cell_centroid") return f64(p.cell_centroid);
    if (n == "face_centroid") return f64(p.face_centroid);
    if (n == "face_area_vector") return f64(m->area_vector);
    if (n == "stencil_offsets") return i64(m->stencil.offsets);
    if (n == "stencil") return i32(m->stencil.entries);
    if (n == "stencil_weights") return f64(m->weights);
    throw std::invalid_argument("unknown mesh array '" + n + "'");
  });
}

}  // extern "C"
