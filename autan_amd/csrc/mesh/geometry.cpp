// FP64 geometry, device packing, the face2 stencil and its least-squares weights.
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

#include "mesh.hpp"

namespace mx {

namespace {

using Vec = std::array<double, 3>;

std::size_t at(index_t i) { return static_cast<std::size_t>(i); }
Vec add(Vec a, Vec b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vec sub(Vec a, Vec b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec scale(double s, Vec a) { return {s * a[0], s * a[1], s * a[2]}; }
double dot(Vec a, Vec b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec cross(Vec a, Vec b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}

struct FacePoints {
  const Mesh& m;
  index_t f;
  Vec offset;  // added to every node: a periodic face seen from its neighbour
  [[nodiscard]] std::size_t size() const {
    return static_cast<std::size_t>(m.face_node_offsets[at(f) + 1] - m.face_node_offsets[at(f)]);
  }
  [[nodiscard]] Vec operator[](std::size_t i) const {
    const auto p = at(m.face_nodes[static_cast<std::size_t>(m.face_node_offsets[at(f)]) + i]);
    return {m.x[p] + offset[0], m.y[p] + offset[1], m.z[p] + offset[2]};
  }
};

// core/src/geometry.cpp:36-42: node mean accumulated relative to the first node.
Vec node_mean(const FacePoints& p) {
  const Vec origin = p[0];
  Vec sum{0.0, 0.0, 0.0};
  for (std::size_t i = 0; i < p.size(); ++i) sum = add(sum, sub(p[i], origin));
  return add(origin, scale(1.0 / static_cast<double>(p.size()), sum));
}

struct FaceGeom {
  Vec area;
  double magnitude;
  Vec centroid;
};

// core/src/geometry.cpp:50-71: fan triangles about the node mean; the centroid weights each
// triangle by its area projected on the face normal.
FaceGeom face_geometry(const FacePoints& p) {
  const std::size_t n = p.size();
  const Vec m = node_mean(p);
  Vec area{0.0, 0.0, 0.0};
  for (std::size_t i = 0; i < n; ++i) area = add(area, scale(0.5, cross(sub(p[i], m), sub(p[(i + 1) % n], m))));
  const double magnitude = std::hypot(area[0], area[1], area[2]);
  if (!(magnitude > 0.0) || !std::isfinite(magnitude)) return {area, magnitude, m};
  const Vec normal = scale(1.0 / magnitude, area);
  Vec moment{0.0, 0.0, 0.0};
  double weight = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    const Vec a = sub(p[i], m);
    const Vec b = sub(p[(i + 1) % n], m);
    const double w = dot(scale(0.5, cross(a, b)), normal);
    moment = add(moment, scale(w / 3.0, add(a, b)));
    weight += w;
  }
  return {area, magnitude, add(m, scale(1.0 / weight, moment))};
}

}  // namespace

Geometry compute_geometry(const Mesh& m) {
  const std::size_t nf = at(m.n_faces), nc = at(m.n_cells), ni = at(m.n_interior);
  Geometry g;
  g.face_area.resize(3 * nf);
  g.face_area_mag.resize(nf);
  g.face_centroid.resize(3 * nf);
  for (std::size_t f = 0; f < nf; ++f) {
    const FaceGeom fg = face_geometry(FacePoints{m, static_cast<index_t>(f), {0.0, 0.0, 0.0}});
    if (!(fg.magnitude > 0.0) || !std::isfinite(fg.magnitude)) {
      throw std::runtime_error("face " + std::to_string(f) + " has zero or non-finite area");
    }
    for (std::size_t u = 0; u < 3; ++u) {
      g.face_area[u * nf + f] = fg.area[u];
      g.face_centroid[u * nf + f] = fg.centroid[u];
    }
    g.face_area_mag[f] = fg.magnitude;
  }
  g.cell_volume.resize(nc);
  g.cell_centroid.resize(3 * nc);
  // core/src/geometry.cpp:82-113: tetrahedra from the mean of the face centroids to every fan
  // triangle, in units of the cell size. A periodic face enters its neighbour moved by -shift.
  for (std::size_t c = 0; c < nc; ++c) {
    const auto k0 = static_cast<std::size_t>(m.cell_face_offsets[c]);
    const auto k1 = static_cast<std::size_t>(m.cell_face_offsets[c + 1]);
    auto frame = [&](index_t f) -> Vec {
      if (at(f) < ni && at(m.neighbour[at(f)]) == c && m.periodic()) {
        const Vec s = m.face_shift(f);
        return {-s[0], -s[1], -s[2]};
      }
      return {0.0, 0.0, 0.0};
    };
    auto centroid = [&](index_t f) -> Vec {
      const Vec o = frame(f);
      const auto fi = at(f);
      return {g.face_centroid[fi] + o[0], g.face_centroid[nf + fi] + o[1], g.face_centroid[2 * nf + fi] + o[2]};
    };
    Vec apex{0.0, 0.0, 0.0};
    for (std::size_t k = k0; k < k1; ++k) apex = add(apex, centroid(m.cell_faces[k]));
    apex = scale(1.0 / static_cast<double>(k1 - k0), apex);
    double length = 0.0;
    for (std::size_t k = k0; k < k1; ++k) length = std::max(length, std::sqrt(g.face_area_mag[at(m.cell_faces[k])]));
    const double inv = 1.0 / length;
    double scaled_volume = 0.0;
    Vec moment{0.0, 0.0, 0.0};
    for (std::size_t k = k0; k < k1; ++k) {
      const index_t f = m.cell_faces[k];
      const double sign = at(m.owner[at(f)]) == c ? 1.0 : -1.0;
      const FacePoints p{m, f, frame(f)};
      const std::size_t n = p.size();
      const Vec mid = scale(inv, sub(node_mean(p), apex));
      for (std::size_t i = 0; i < n; ++i) {
        const Vec a = scale(inv, sub(p[i], apex));
        const Vec b = scale(inv, sub(p[(i + 1) % n], apex));
        const double v = sign * dot(scale(0.5, cross(sub(a, mid), sub(b, mid))), mid) / 3.0;
        scaled_volume += v;
        moment = add(moment, scale(v / 4.0, add(add(mid, a), b)));
      }
    }
    const double volume = scaled_volume * length * length * length;
    if (!(scaled_volume > 0.0 && std::isfinite(volume) && volume > 0.0)) {
      throw std::runtime_error("cell " + std::to_string(c) + " has non-positive or non-finite volume");
    }
    const Vec xc = add(apex, scale(length / scaled_volume, moment));
    g.cell_volume[c] = volume;
    for (std::size_t u = 0; u < 3; ++u) g.cell_centroid[u * nc + c] = xc[u];
  }
  return g;
}

namespace {

// codegen/cpp/geometry.cpp:22-52: planarity proven with == on node coordinates: a triangle, nodes
// sharing one coordinate, or nodes with at most two distinct projections along an axis.
bool exactly_planar(const FacePoints& p) {
  const std::size_t nn = p.size();
  if (nn == 3) return true;
  const Vec p0 = p[0];
  for (std::size_t a = 0; a < 3; ++a) {
    const std::size_t b = (a + 1) % 3, c = (a + 2) % 3;
    bool shared = true, two_lines = true;
    std::size_t second = 0;
    for (std::size_t i = 1; i < nn && (shared || two_lines); ++i) {
      const Vec q = p[i];
      shared = shared && q[a] == p0[a];
      if (q[b] == p0[b] && q[c] == p0[c]) continue;
      if (second == 0) {
        second = i;
      } else {
        const Vec s = p[second];
        two_lines = two_lines && q[b] == s[b] && q[c] == s[c];
      }
    }
    if (shared || two_lines) return true;
  }
  return false;
}

}  // namespace

Packed pack(const Mesh& m, const Geometry& g) {
  const std::size_t nf = at(m.n_faces), nc = at(m.n_cells), ni = at(m.n_interior);
  Packed p;
  p.face_area = g.face_area_mag;
  p.cell_volume = g.cell_volume;
  p.cell_centroid = g.cell_centroid;
  p.face_centroid = g.face_centroid;
  p.face_normal.resize(3 * nf);
  p.r_left.resize(3 * nf);
  p.r_right.assign(3 * nf, 0.0);
  p.face_moment.assign(9 * nf, 0.0);
  for (std::size_t f = 0; f < nf; ++f) {
    const auto l = at(m.owner[f]);
    const Vec s = m.face_shift(static_cast<index_t>(f));
    for (std::size_t u = 0; u < 3; ++u) {
      p.face_normal[u * nf + f] = g.face_area[u * nf + f] / g.face_area_mag[f];
      p.r_left[u * nf + f] = g.face_centroid[u * nf + f] - g.cell_centroid[u * nc + l];
      if (f < ni) {
        const auto r = at(m.neighbour[f]);
        // The neighbour's image as the owner sees it: its centroid plus the face's shift (zero
        // off periodic faces, where this is the plain difference).
        const double xr = s[u] != 0.0 ? g.cell_centroid[u * nc + r] + s[u] : g.cell_centroid[u * nc + r];
        p.r_right[u * nf + f] = g.face_centroid[u * nf + f] - xr;
      }
    }
  }
  // codegen/cpp/geometry.cpp:212-266: moments over the fan (mean, p_i, p_i+1) from the node of
  // least (x, y, z); signed zeros where planarity is proven.
  for (std::size_t f = 0; f < nf; ++f) {
    const FacePoints pts{m, static_cast<index_t>(f), {0.0, 0.0, 0.0}};
    if (exactly_planar(pts)) continue;
    p.all_planar = false;
    const std::size_t nn = pts.size();
    std::size_t start = 0;
    for (std::size_t i = 1; i < nn; ++i) {
      if (pts[i] < pts[start]) start = i;
    }
    auto point = [&](std::size_t i) { return pts[(start + i) % nn]; };
    Vec mean{0.0, 0.0, 0.0};
    for (std::size_t i = 0; i < nn; ++i) mean = add(mean, point(i));
    const double k = static_cast<double>(nn);
    mean = {mean[0] / k, mean[1] / k, mean[2] / k};
    const Vec xf{g.face_centroid[f], g.face_centroid[nf + f], g.face_centroid[2 * nf + f]};
    std::array<double, 9> moment{};
    for (std::size_t i = 0; i < nn; ++i) {
      const Vec a = point(i), b = point(i + 1);
      const Vec s = cross(sub(a, mean), sub(b, mean));
      const Vec c{(mean[0] + a[0] + b[0]) / 3.0, (mean[1] + a[1] + b[1]) / 3.0, (mean[2] + a[2] + b[2]) / 3.0};
      const Vec r = sub(c, xf);
      for (std::size_t u = 0; u < 3; ++u) {
        for (std::size_t v = 0; v < 3; ++v) moment[3 * u + v] += 0.5 * r[u] * s[v];
      }
    }
    for (std::size_t j = 0; j < 9; ++j) p.face_moment[j * nf + f] = moment[j];
  }
  return p;
}

Stencil face2_stencil(const Mesh& m, const Packed& p) {
  const std::size_t nc = at(m.n_cells), nf = at(m.n_faces), ni = at(m.n_interior);
  Stencil st;
  st.offsets.assign(nc + 1, 0);
  st.entries.reserve(nc * 16);
  std::vector<double> shift;  // entry-major while building
  // Per cell under construction: the translation each listed cell or boundary face was reached
  // with (a periodic path), to check that one index never stands for two images.
  std::vector<index_t> listed(nc, -1), bound_listed(nf - ni, -1);
  struct Item {
    index_t id;
    Vec t;
  };
  std::vector<Item> cells, bfaces;
  auto across = [&](std::size_t c, std::size_t f) {
    return at(m.owner[f]) == c ? at(m.neighbour[f]) : at(m.owner[f]);
  };
  // Translation of the cell across f as seen from c.
  auto hop = [&](std::size_t c, std::size_t f) -> Vec {
    if (!m.periodic()) return {0.0, 0.0, 0.0};
    const Vec s = m.face_shift(static_cast<index_t>(f));
    return at(m.owner[f]) == c ? s : Vec{-s[0], -s[1], -s[2]};
  };
  auto check = [&](const std::vector<Item>& items, index_t id, const Vec& t) {
    for (const Item& it : items) {
      if (it.id == id && (it.t[0] != t[0] || it.t[1] != t[1] || it.t[2] != t[2])) {
        throw std::runtime_error("face2 stencil: a cell reached as two periodic images (mesh too small)");
      }
    }
  };
  for (std::size_t c = 0; c < nc; ++c) {
    cells.clear();
    bfaces.clear();
    listed[c] = static_cast<index_t>(c);
    auto add_cell = [&](std::size_t n, const Vec& t) {
      if (n == c) {
        if (t[0] != 0.0 || t[1] != 0.0 || t[2] != 0.0) {
          throw std::runtime_error("face2 stencil: a cell reaches its own periodic image (mesh too small)");
        }
        return;
      }
      if (listed[n] != static_cast<index_t>(c)) {
        listed[n] = static_cast<index_t>(c);
        cells.push_back({static_cast<index_t>(n), t});
      } else if (m.periodic()) {
        check(cells, static_cast<index_t>(n), t);
      }
    };
    auto add_bface = [&](std::size_t f, const Vec& t) {
      auto& mark = bound_listed[f - ni];
      if (mark != static_cast<index_t>(c)) {
        mark = static_cast<index_t>(c);
        bfaces.push_back({static_cast<index_t>(f), t});
      } else if (m.periodic()) {
        check(bfaces, static_cast<index_t>(f), t);
      }
    };
    for (auto k = m.cell_face_offsets[c]; k < m.cell_face_offsets[c + 1]; ++k) {
      const auto f = at(m.cell_faces[static_cast<std::size_t>(k)]);
      if (f >= ni) {
        add_bface(f, {0.0, 0.0, 0.0});
        continue;
      }
      const std::size_t n = across(c, f);
      const Vec t1 = hop(c, f);
      add_cell(n, t1);
      for (auto k2 = m.cell_face_offsets[n]; k2 < m.cell_face_offsets[n + 1]; ++k2) {
        const auto f2 = at(m.cell_faces[static_cast<std::size_t>(k2)]);
        if (f2 >= ni) {
          add_bface(f2, t1);
        } else {
          add_cell(across(n, f2), add(t1, hop(n, f2)));
        }
      }
    }
    std::sort(cells.begin(), cells.end(), [](const Item& a, const Item& b) { return a.id < b.id; });
    std::sort(bfaces.begin(), bfaces.end(), [](const Item& a, const Item& b) { return a.id < b.id; });
    for (const Item& it : cells) {
      st.entries.push_back(it.id);
      if (m.periodic()) shift.insert(shift.end(), it.t.begin(), it.t.end());
    }
    for (const Item& it : bfaces) {
      st.entries.push_back(-it.id - 1);
      if (m.periodic()) shift.insert(shift.end(), it.t.begin(), it.t.end());
    }
    st.offsets[c + 1] = static_cast<offset_t>(st.entries.size());
  }
  (void)p;
  if (m.periodic()) {
    const std::size_t ns = st.entries.size();
    st.translation.resize(3 * ns);
    for (std::size_t e = 0; e < ns; ++e) {
      for (std::size_t u = 0; u < 3; ++u) st.translation[u * ns + e] = shift[3 * e + u];
    }
  }
  return st;
}

std::size_t gradient_weights(const Mesh& m, const Packed& p, const Stencil& s, const std::vector<int>& kind,
                             std::vector<double>& w) {
  const std::size_t nc = at(m.n_cells), ni = at(m.n_interior), nf = at(m.n_faces);
  const std::size_t ns = s.entries.size();
  if (kind.size() != at(m.n_boundary())) throw std::invalid_argument("gradient_weights: one kind per boundary face");
  w.assign(3 * ns, 0.0);
  constexpr double coordinate_eps = std::numeric_limits<double>::epsilon();
  std::size_t kept = 0;
  std::vector<std::array<double, 4>> offs;
  for (std::size_t c = 0; c < nc; ++c) {
    const auto begin = static_cast<std::size_t>(s.offsets[c]);
    const auto end = static_cast<std::size_t>(s.offsets[c + 1]);
    auto weights = [&](bool all_boundaries) {
      std::array<double, 9> a{};
      offs.clear();
      for (std::size_t k = begin; k < end; ++k) {
        const index_t e = s.entries[k];
        bool used = true;
        if (e < 0) {
          const auto f = static_cast<std::size_t>(-e - 1);
          used = all_boundaries || kind[f - ni] == bc_state;
        }
        // Offsets within 64 ulps of the coordinates are round-off (physics/src/euler.cpp:371-380).
        // The entry's centroid: a cell's or a boundary face's, moved to the image the cell sees.
        std::array<double, 4> dk{};
        for (std::size_t j = 0; j < 3; ++j) {
          double x = e >= 0 ? p.cell_centroid[j * nc + at(e)] : p.face_centroid[j * nf + at(-e - 1)];
          if (!s.translation.empty() && s.translation[j * ns + k] != 0.0) x += s.translation[j * ns + k];
          const double xc = p.cell_centroid[j * nc + c];
          const double round_off = 64 * coordinate_eps * std::max(std::abs(x), std::abs(xc));
          dk[j] = x - xc;
          if (std::abs(dk[j]) <= round_off) dk[j] = 0.0;
        }
        dk[3] = used ? 1.0 / (dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2]) : 0.0;
        for (std::size_t u = 0; u < 3; ++u) {
          for (std::size_t v = 0; v < 3; ++v) a[3 * u + v] += dk[3] * dk[u] * dk[v];
        }
        offs.push_back(dk);
      }
      const double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                         a[2] * (a[3] * a[7] - a[4] * a[6]);
      const double sc = a[0] * a[4] * a[8];
      if (!(std::abs(det) > 1e-10 * std::abs(sc))) return false;
      const std::array<double, 9> inv = {
          (a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det, (a[1] * a[5] - a[2] * a[4]) / det,
          (a[5] * a[6] - a[3] * a[8]) / det, (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
          (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det, (a[0] * a[4] - a[1] * a[3]) / det};
      for (std::size_t k = begin; k < end; ++k) {
        const auto& dk = offs[k - begin];
        for (std::size_t u = 0; u < 3; ++u) {
          w[u * ns + k] = dk[3] * (inv[3 * u] * dk[0] + inv[3 * u + 1] * dk[1] + inv[3 * u + 2] * dk[2]);
        }
      }
      return true;
    };
    if (!weights(false)) {
      weights(true);
      ++kept;
    }
  }
  return kept;
}

}  // namespace mx
