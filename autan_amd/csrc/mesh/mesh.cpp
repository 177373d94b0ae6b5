// Mesh assembly, generators, polyMesh reading and RCM renumbering.
#include "mesh.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace mx {

namespace {

std::size_t at(index_t i) { return static_cast<std::size_t>(i); }

}  // namespace

void finish(Mesh& m) {
  const std::size_t nc = at(m.n_cells), nf = at(m.n_faces), ni = at(m.n_interior);
  m.cell_face_offsets.assign(nc + 1, 0);
  for (std::size_t f = 0; f < nf; ++f) {
    ++m.cell_face_offsets[at(m.owner[f]) + 1];
    if (f < ni) ++m.cell_face_offsets[at(m.neighbour[f]) + 1];
  }
  for (std::size_t c = 0; c < nc; ++c) m.cell_face_offsets[c + 1] += m.cell_face_offsets[c];
  m.cell_faces.assign(static_cast<std::size_t>(m.cell_face_offsets[nc]), 0);
  std::vector<offset_t> next(m.cell_face_offsets.begin(), m.cell_face_offsets.end() - 1);
  // Faces in increasing order, so every cell's list comes out ascending.
  for (std::size_t f = 0; f < nf; ++f) {
    m.cell_faces[static_cast<std::size_t>(next[at(m.owner[f])]++)] = static_cast<index_t>(f);
    if (f < ni) m.cell_faces[static_cast<std::size_t>(next[at(m.neighbour[f])]++)] = static_cast<index_t>(f);
  }
}

Mesh assemble(index_t n_cells, std::vector<double> x, std::vector<double> y, std::vector<double> z,
              RawFaces in) {
  const std::size_t n = in.c0.size();
  if (in.off.size() != n + 1 || in.c1.size() != n || in.zone.size() != n ||
      (!in.shift.empty() && in.shift.size() != 3 * n)) {
    throw std::invalid_argument("assemble: face arrays of different lengths");
  }
  const bool periodic = !in.shift.empty();
  // Points moved into another cell's frame by a reversed periodic face: (point, shift) -> index.
  std::map<std::tuple<index_t, double, double, double>, index_t> moved;
  for (std::size_t i = 0; i < n; ++i) {
    const index_t a = in.c0[i], b = in.c1[i];
    if (a < 0 || a >= n_cells || b >= n_cells || a == b) {
      throw std::invalid_argument("assemble: face " + std::to_string(i) + " has invalid cells");
    }
    if (b < 0 || a < b) continue;
    std::swap(in.c0[i], in.c1[i]);
    auto first = in.nodes.begin() + in.off[i], last = in.nodes.begin() + in.off[i + 1];
    std::reverse(first, last);
    if (periodic) {
      double* s = &in.shift[3 * i];
      if (s[0] != 0.0 || s[1] != 0.0 || s[2] != 0.0) {
        // The nodes were in the old owner's frame; the new owner sees them moved by -shift.
        for (auto it = first; it != last; ++it) {
          const auto key = std::make_tuple(*it, s[0], s[1], s[2]);
          auto found = moved.find(key);
          if (found == moved.end()) {
            const std::size_t p = at(*it);
            x.push_back(x[p] - s[0]);
            y.push_back(y[p] - s[1]);
            z.push_back(z[p] - s[2]);
            found = moved.emplace(key, static_cast<index_t>(x.size() - 1)).first;
          }
          *it = found->second;
        }
      }
      for (int k = 0; k < 3; ++k) s[k] = -s[k];
    }
  }
  std::vector<std::size_t> order(n);
  std::iota(order.begin(), order.end(), std::size_t{0});
  std::stable_sort(order.begin(), order.end(), [&](std::size_t p, std::size_t q) {
    const bool bp = in.c1[p] < 0, bq = in.c1[q] < 0;
    if (bp != bq) return bq;  // interior first
    if (!bp) return in.c0[p] != in.c0[q] ? in.c0[p] < in.c0[q] : in.c1[p] < in.c1[q];
    return in.zone[p] != in.zone[q] ? in.zone[p] < in.zone[q] : in.c0[p] < in.c0[q];
  });
  Mesh m;
  m.n_cells = n_cells;
  m.n_faces = static_cast<index_t>(n);
  m.n_interior = static_cast<index_t>(std::count_if(in.c1.begin(), in.c1.end(), [](index_t b) { return b >= 0; }));
  m.x = std::move(x);
  m.y = std::move(y);
  m.z = std::move(z);
  m.owner.resize(n);
  m.neighbour.resize(at(m.n_interior));
  m.face_zone.resize(n - at(m.n_interior));
  m.face_node_offsets.assign(n + 1, 0);
  m.face_nodes.reserve(in.nodes.size());
  bool any_shift = false;
  if (periodic) {
    for (double s : in.shift) any_shift = any_shift || s != 0.0;
  }
  if (any_shift) m.shift.assign(3 * at(m.n_interior), 0.0);
  const std::size_t ni = at(m.n_interior);
  for (std::size_t k = 0; k < n; ++k) {
    const std::size_t i = order[k];
    m.owner[k] = in.c0[i];
    if (k < ni) {
      m.neighbour[k] = in.c1[i];
      if (any_shift) {
        for (std::size_t u = 0; u < 3; ++u) m.shift[u * ni + k] = in.shift[3 * i + u];
      }
    } else {
      m.face_zone[k - ni] = in.zone[i];
    }
    m.face_nodes.insert(m.face_nodes.end(), in.nodes.begin() + in.off[i], in.nodes.begin() + in.off[i + 1]);
    m.face_node_offsets[k + 1] = static_cast<offset_t>(m.face_nodes.size());
  }
  finish(m);
  return m;
}

// Generators ------------------------------------------------------------------------------------

Mesh periodic_hex(int n, double lo, double hi) {
  if (n < 3) throw std::invalid_argument("periodic_hex needs n >= 3");
  const double len = hi - lo;
  const int p = n + 1;
  std::vector<double> x, y, z;
  x.reserve(static_cast<std::size_t>(p) * p * p);
  for (int k = 0; k < p; ++k) {
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i < p; ++i) {
        x.push_back(lo + len * i / n);
        y.push_back(lo + len * j / n);
        z.push_back(lo + len * k / n);
      }
    }
  }
  y.reserve(x.size());
  z.reserve(x.size());
  auto node = [&](int i, int j, int k) { return static_cast<index_t>(i + p * (j + p * k)); };
  auto cell = [&](int i, int j, int k) { return static_cast<index_t>((i % n) + n * ((j % n) + n * (k % n))); };
  RawFaces f;
  const std::size_t nf = 3 * static_cast<std::size_t>(n) * n * n;
  f.off.reserve(nf + 1);
  f.nodes.reserve(4 * nf);
  f.c0.reserve(nf);
  f.c1.reserve(nf);
  f.zone.reserve(nf);
  f.shift.reserve(3 * nf);
  auto push_shift = [&](double sx, double sy, double sz) {
    f.shift.push_back(sx);
    f.shift.push_back(sy);
    f.shift.push_back(sz);
  };
  for (int k = 0; k < n; ++k) {
    for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
        const index_t c = cell(i, j, k);
        // +x face, normal +x: (j,k) (j+1,k) (j+1,k+1) (j,k+1) in the y-z plane.
        f.add({node(i + 1, j, k), node(i + 1, j + 1, k), node(i + 1, j + 1, k + 1), node(i + 1, j, k + 1)}, c,
              cell(i + 1, j, k));
        push_shift(i == n - 1 ? len : 0.0, 0.0, 0.0);
        // +y face, normal +y: (k,i) order: (i,k) (i,k+1) (i+1,k+1) (i+1,k).
        f.add({node(i, j + 1, k), node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1), node(i + 1, j + 1, k)}, c,
              cell(i, j + 1, k));
        push_shift(0.0, j == n - 1 ? len : 0.0, 0.0);
        // +z face, normal +z: (i,j) (i+1,j) (i+1,j+1) (i,j+1).
        f.add({node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)}, c,
              cell(i, j, k + 1));
        push_shift(0.0, 0.0, k == n - 1 ? len : 0.0);
      }
    }
  }
  return assemble(static_cast<index_t>(n) * n * n, std::move(x), std::move(y), std::move(z), std::move(f));
}

namespace {

// splitmix64: a small seeded generator, fixed so that meshes are reproducible everywhere.
struct SplitMix {
  std::uint64_t s;
  std::uint64_t next() {
    std::uint64_t v = (s += 0x9e3779b97f4a7c15ULL);
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
  }
  // Uniform in [-1, 1).
  double symmetric() { return static_cast<double>(next() >> 11) * 0x1.0p-52 - 1.0; }
};

// Faces of cells given as outward-oriented polygons, matched by their three smallest node ids.
struct CellFaces {
  std::vector<offset_t> off{0};
  std::vector<index_t> nodes;
  std::vector<index_t> cell;
  void add(index_t c, std::initializer_list<index_t> n) {
    nodes.insert(nodes.end(), n);
    off.push_back(static_cast<offset_t>(nodes.size()));
    cell.push_back(c);
  }
};

// Boundary zone of an unmatched face, from its nodes.
template <class ZoneOf>
RawFaces match_faces(const CellFaces& cf, const ZoneOf& zone_of) {
  const std::size_t n = cf.cell.size();
  // A face is known by its three smallest nodes, its largest node and its node count; two sides
  // with equal keys must also hold equal node sets, which the pairing below checks.
  struct Key {
    index_t a, b, c, z, count;
    std::uint32_t i;
  };
  std::vector<Key> keys(n);
  for (std::size_t i = 0; i < n; ++i) {
    std::array<index_t, 3> s{INT32_MAX, INT32_MAX, INT32_MAX};
    index_t z = -1;
    for (auto k = cf.off[i]; k < cf.off[i + 1]; ++k) {
      index_t v = cf.nodes[static_cast<std::size_t>(k)];
      z = std::max(z, v);
      for (int q = 0; q < 3; ++q) {
        if (v < s[q]) std::swap(v, s[q]);
      }
    }
    keys[i] = {s[0], s[1], s[2], z, static_cast<index_t>(cf.off[i + 1] - cf.off[i]), static_cast<std::uint32_t>(i)};
  }
  auto same = [](const Key& p, const Key& q) {
    return p.a == q.a && p.b == q.b && p.c == q.c && p.z == q.z && p.count == q.count;
  };
  std::sort(keys.begin(), keys.end(), [](const Key& p, const Key& q) {
    if (p.a != q.a) return p.a < q.a;
    if (p.b != q.b) return p.b < q.b;
    if (p.c != q.c) return p.c < q.c;
    if (p.z != q.z) return p.z < q.z;
    if (p.count != q.count) return p.count < q.count;
    return p.i < q.i;
  });
  auto sorted_nodes = [&](std::size_t i) {
    std::vector<index_t> v(cf.nodes.begin() + cf.off[i], cf.nodes.begin() + cf.off[i + 1]);
    std::sort(v.begin(), v.end());
    return v;
  };
  RawFaces out;
  out.off.reserve(n / 2 + 16);
  for (std::size_t k = 0; k < n;) {
    const Key& p = keys[k];
    const std::size_t i = p.i;
    const bool pair = k + 1 < n && same(keys[k + 1], p);
    if (pair && k + 2 < n && same(keys[k + 2], p)) {
      throw std::runtime_error("match_faces: a face shared by more than two cells");
    }
    if (pair && p.count > 3 && sorted_nodes(i) != sorted_nodes(keys[k + 1].i)) {
      throw std::runtime_error("match_faces: two distinct faces share a key");
    }
    out.nodes.insert(out.nodes.end(), cf.nodes.begin() + cf.off[i], cf.nodes.begin() + cf.off[i + 1]);
    out.off.push_back(static_cast<offset_t>(out.nodes.size()));
    out.c0.push_back(cf.cell[i]);
    if (pair) {
      out.c1.push_back(cf.cell[keys[k + 1].i]);
      out.zone.push_back(-1);
      k += 2;
    } else {
      out.c1.push_back(-1);
      out.zone.push_back(zone_of(&cf.nodes[static_cast<std::size_t>(cf.off[i])],
                                 static_cast<std::size_t>(cf.off[i + 1] - cf.off[i])));
      k += 1;
    }
  }
  return out;
}

}  // namespace

Mesh kuhn_tets(int n, double jitter, std::uint64_t seed) {
  if (n < 1) throw std::invalid_argument("kuhn_tets needs n >= 1");
  const int p = n + 1;
  const double h = 1.0 / n;
  const std::size_t np = static_cast<std::size_t>(p) * p * p;
  std::vector<double> x(np), y(np), z(np);
  SplitMix rng{seed};
  for (int k = 0; k < p; ++k) {
    for (int j = 0; j < p; ++j) {
      for (int i = 0; i < p; ++i) {
        const std::size_t v = static_cast<std::size_t>(i + p * (j + p * k));
        // Three draws per node in node order, used only on interior nodes, so a boundary node
        // stays on its plane whatever the seed.
        const double dx = rng.symmetric(), dy = rng.symmetric(), dz = rng.symmetric();
        const bool inside = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
        x[v] = h * i + (inside ? jitter * h * dx : 0.0);
        y[v] = h * j + (inside ? jitter * h * dy : 0.0);
        z[v] = h * k + (inside ? jitter * h * dz : 0.0);
      }
    }
  }
  auto node = [&](int i, int j, int k) { return static_cast<index_t>(i + p * (j + p * k)); };
  CellFaces cf;
  const std::size_t nt = 6 * static_cast<std::size_t>(n) * n * n;
  cf.off.reserve(4 * nt + 1);
  cf.nodes.reserve(12 * nt);
  cf.cell.reserve(4 * nt);
  // The six tetrahedra about the diagonal v000 - v111: for each order (a, b, c) of the axes,
  // {v000, v000 + e_a, v000 + e_a + e_b, v111}.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  index_t c = 0;
  for (int k = 0; k < n; ++k) {
    for (int j = 0; j < n; ++j) {
      for (int i = 0; i < n; ++i) {
        for (const auto& pm : perms) {
          int q1[3] = {i, j, k};
          q1[pm[0]] += 1;
          int q2[3] = {q1[0], q1[1], q1[2]};
          q2[pm[1]] += 1;
          index_t t[4] = {node(i, j, k), node(q1[0], q1[1], q1[2]), node(q2[0], q2[1], q2[2]),
                          node(i + 1, j + 1, k + 1)};
          // Orient so that (t1 - t0) x (t2 - t0) . (t3 - t0) > 0 on the unjittered cube: the
          // parity of the axis permutation decides it.
          const bool even = (pm[0] == 0 && pm[1] == 1) || (pm[0] == 1 && pm[1] == 2) || (pm[0] == 2 && pm[1] == 0);
          if (!even) std::swap(t[1], t[2]);
          // Outward faces of a positive tetrahedron (0,1,2,3): (0,2,1), (0,1,3), (1,2,3), (0,3,2).
          cf.add(c, {t[0], t[2], t[1]});
          cf.add(c, {t[0], t[1], t[3]});
          cf.add(c, {t[1], t[2], t[3]});
          cf.add(c, {t[0], t[3], t[2]});
          ++c;
        }
      }
    }
  }
  const double eps = 0.25 * h;
  auto zone_of = [&](const index_t* v, std::size_t count) {
    auto all = [&](auto pred) {
      for (std::size_t q = 0; q < count; ++q) {
        if (!pred(at(v[q]))) return false;
      }
      return true;
    };
    if (all([&](std::size_t w) { return x[w] < eps; })) return 0;
    if (all([&](std::size_t w) { return x[w] > 1.0 - eps; })) return 1;
    if (all([&](std::size_t w) { return y[w] < eps; })) return 2;
    if (all([&](std::size_t w) { return y[w] > 1.0 - eps; })) return 3;
    if (all([&](std::size_t w) { return z[w] < eps; })) return 4;
    return 5;
  };
  RawFaces faces = match_faces(cf, zone_of);
  cf = CellFaces();
  return assemble(c, std::move(x), std::move(y), std::move(z), std::move(faces));
}

Mesh bump_hex(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0) {
  if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("bump_hex needs positive counts");
  const int px = nx + 1, py = ny + 1, pz = nz + 1;
  const std::size_t np = static_cast<std::size_t>(px) * py * pz;
  std::vector<double> x(np), y(np), z(np);
  for (int k = 0; k < pz; ++k) {
    for (int j = 0; j < py; ++j) {
      for (int i = 0; i < px; ++i) {
        const std::size_t v = static_cast<std::size_t>(i + px * (j + py * k));
        const double xi = lx * i / nx;
        const double t = (xi - x0) / w;
        const double wall = h * std::exp(-t * t);
        x[v] = xi;
        y[v] = wall + (ly - wall) * j / ny;
        z[v] = lz * k / nz;
      }
    }
  }
  auto node = [&](int i, int j, int k) { return static_cast<index_t>(i + px * (j + py * k)); };
  auto cell = [&](int i, int j, int k) { return static_cast<index_t>(i + nx * (j + ny * k)); };
  RawFaces f;
  for (int k = 0; k < nz; ++k) {
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        const index_t c = cell(i, j, k);
        const index_t xp = i + 1 < nx ? cell(i + 1, j, k) : -1;
        const index_t yp = j + 1 < ny ? cell(i, j + 1, k) : -1;
        const index_t zp = k + 1 < nz ? cell(i, j, k + 1) : -1;
        f.add({node(i + 1, j, k), node(i + 1, j + 1, k), node(i + 1, j + 1, k + 1), node(i + 1, j, k + 1)}, c, xp,
              xp < 0 ? 1 : -1);
        f.add({node(i, j + 1, k), node(i, j + 1, k + 1), node(i + 1, j + 1, k + 1), node(i + 1, j + 1, k)}, c, yp,
              yp < 0 ? 3 : -1);
        f.add({node(i, j, k + 1), node(i + 1, j, k + 1), node(i + 1, j + 1, k + 1), node(i, j + 1, k + 1)}, c, zp,
              zp < 0 ? 5 : -1);
        if (i == 0) {
          f.add({node(0, j, k), node(0, j, k + 1), node(0, j + 1, k + 1), node(0, j + 1, k)}, c, -1, 0);
        }
        if (j == 0) {
          f.add({node(i, 0, k), node(i + 1, 0, k), node(i + 1, 0, k + 1), node(i, 0, k + 1)}, c, -1, 2);
        }
        if (k == 0) {
          f.add({node(i, j, 0), node(i, j + 1, 0), node(i + 1, j + 1, 0), node(i + 1, j, 0)}, c, -1, 4);
        }
      }
    }
  }
  return assemble(static_cast<index_t>(nx) * ny * nz, std::move(x), std::move(y), std::move(z), std::move(f));
}

Mesh bump_hybrid(int nx, int ny, int nz, double lx, double ly, double lz, double h, double w, double x0,
                 int prism_layers, int band_lo, int band_hi) {
  if (nx < 2 || ny < 2 || nz < 2 || nx % 2 || nz % 2 || (band_hi - band_lo) % 2 || band_lo < prism_layers ||
      band_hi > ny || prism_layers < 0 || band_lo > band_hi) {
    throw std::invalid_argument("bump_hybrid: nx, nz and the band height even, prisms below the band");
  }
  const int px = nx + 1, py = ny + 1, pz = nz + 1;
  const std::size_t np = static_cast<std::size_t>(px) * py * pz;
  std::vector<double> x(np), y(np), z(np);
  for (int k = 0; k < pz; ++k) {
    for (int j = 0; j < py; ++j) {
      for (int i = 0; i < px; ++i) {
        const std::size_t v = static_cast<std::size_t>(i + px * (j + py * k));
        const double xi = lx * i / nx;
        const double t = (xi - x0) / w;
        const double wall = h * std::exp(-t * t);
        x[v] = xi;
        y[v] = wall + (ly - wall) * j / ny;
        z[v] = lz * k / nz;
      }
    }
  }
  auto node = [&](int i, int j, int k) { return static_cast<index_t>(i + px * (j + py * k)); };
  auto in_band = [&](int j) { return j >= band_lo && j < band_hi; };
  // Cell ids: prisms (two per hex of the prism rows), merged cells, then hexes.
  CellFaces cf;
  index_t nc = 0;
  // Outward orientation by Newell's normal against the vector from the cell centre.
  std::vector<index_t> poly;
  double ccx = 0, ccy = 0, ccz = 0;
  auto emit = [&](index_t c) {
    double nx_ = 0, ny_ = 0, nz_ = 0, fx = 0, fy = 0, fz = 0;
    const std::size_t n = poly.size();
    for (std::size_t a = 0; a < n; ++a) {
      const auto p = static_cast<std::size_t>(poly[a]), q = static_cast<std::size_t>(poly[(a + 1) % n]);
      nx_ += (y[p] - y[q]) * (z[p] + z[q]);
      ny_ += (z[p] - z[q]) * (x[p] + x[q]);
      nz_ += (x[p] - x[q]) * (y[p] + y[q]);
      fx += x[p];
      fy += y[p];
      fz += z[p];
    }
    const double d = nx_ * (fx / n - ccx) + ny_ * (fy / n - ccy) + nz_ * (fz / n - ccz);
    if (d < 0) std::reverse(poly.begin(), poly.end());
    cf.nodes.insert(cf.nodes.end(), poly.begin(), poly.end());
    cf.off.push_back(static_cast<offset_t>(cf.nodes.size()));
    cf.cell.push_back(c);
  };
  auto centre = [&](std::initializer_list<index_t> nodes) {
    ccx = ccy = ccz = 0;
    for (const index_t v : nodes) {
      ccx += x[at(v)];
      ccy += y[at(v)];
      ccz += z[at(v)];
    }
    ccx /= static_cast<double>(nodes.size());
    ccy /= static_cast<double>(nodes.size());
    ccz /= static_cast<double>(nodes.size());
  };
  for (int k = 0; k < nz; ++k) {
    for (int j = 0; j < ny; ++j) {
      for (int i = 0; i < nx; ++i) {
        const index_t a000 = node(i, j, k), a100 = node(i + 1, j, k), a010 = node(i, j + 1, k),
                      a110 = node(i + 1, j + 1, k), a001 = node(i, j, k + 1), a101 = node(i + 1, j, k + 1),
                      a011 = node(i, j + 1, k + 1), a111 = node(i + 1, j + 1, k + 1);
        if (j < prism_layers) {
          // Two prisms: the x-y quad split along (i, j) - (i+1, j+1), extruded over [k, k+1].
          const index_t c0 = nc++, c1 = nc++;
          centre({a000, a100, a110, a001, a101, a111});
          poly = {a000, a100, a110}; emit(c0);
          poly = {a001, a101, a111}; emit(c0);
          poly = {a000, a100, a101, a001}; emit(c0);
          poly = {a100, a110, a111, a101}; emit(c0);
          poly = {a000, a110, a111, a001}; emit(c0);
          centre({a000, a110, a010, a001, a111, a011});
          poly = {a000, a110, a010}; emit(c1);
          poly = {a001, a111, a011}; emit(c1);
          poly = {a000, a110, a111, a001}; emit(c1);
          poly = {a110, a010, a011, a111}; emit(c1);
          poly = {a010, a000, a001, a011}; emit(c1);
          continue;
        }
        if (in_band(j)) {
          if (i % 2 || (j - band_lo) % 2 || k % 2) continue;
          // One merged cell over [i, i+2] x [j, j+2] x [k, k+2]; a side facing another merged cell
          // is the 8-node ring of its 2 x 2 quads, else its 4 quads.
          const index_t c = nc++;
          centre({node(i, j, k), node(i + 2, j, k), node(i, j + 2, k), node(i + 2, j + 2, k), node(i, j, k + 2),
                  node(i + 2, j, k + 2), node(i, j + 2, k + 2), node(i + 2, j + 2, k + 2)});
          // side: axis (0 x, 1 y, 2 z), at the low (0) or high (2) end.
          for (int axis = 0; axis < 3; ++axis) {
            for (int end = 0; end <= 2; end += 2) {
              auto at3 = [&](int u, int v) {
                int q[3];
                q[axis] = end;
                q[(axis + 1) % 3] = u;
                q[(axis + 2) % 3] = v;
                return node(i + q[0], j + q[1], k + q[2]);
              };
              bool merged_nb;
              if (axis == 0) {
                merged_nb = end == 0 ? i > 0 : i + 2 < nx;
              } else if (axis == 1) {
                merged_nb = end == 0 ? j > band_lo : j + 2 < band_hi;
              } else {
                merged_nb = end == 0 ? k > 0 : k + 2 < nz;
              }
              if (merged_nb) {
                poly = {at3(0, 0), at3(1, 0), at3(2, 0), at3(2, 1), at3(2, 2), at3(1, 2), at3(0, 2), at3(0, 1)};
                emit(c);
              } else {
                for (int u = 0; u < 2; ++u) {
                  for (int v = 0; v < 2; ++v) {
                    poly = {at3(u, v), at3(u + 1, v), at3(u + 1, v + 1), at3(u, v + 1)};
                    emit(c);
                  }
                }
              }
            }
          }
          continue;
        }
        const index_t c = nc++;
        centre({a000, a100, a010, a110, a001, a101, a011, a111});
        poly = {a000, a001, a011, a010}; emit(c);
        poly = {a100, a110, a111, a101}; emit(c);
        poly = {a000, a100, a101, a001}; emit(c);
        poly = {a010, a011, a111, a110}; emit(c);
        poly = {a000, a010, a110, a100}; emit(c);
        poly = {a001, a101, a111, a011}; emit(c);
      }
    }
  }
  auto zone_of = [&](const index_t* v, std::size_t count) {
    auto all = [&](auto pred) {
      for (std::size_t q = 0; q < count; ++q) {
        const int vi = v[q] % px, vj = (v[q] / px) % py, vk = v[q] / (px * py);
        if (!pred(vi, vj, vk)) return false;
      }
      return true;
    };
    if (all([](int i, int, int) { return i == 0; })) return 0;
    if (all([&](int i, int, int) { return i == nx; })) return 1;
    if (all([](int, int j, int) { return j == 0; })) return 2;
    if (all([&](int, int j, int) { return j == ny; })) return 3;
    if (all([](int, int, int k) { return k == 0; })) return 4;
    return 5;
  };
  RawFaces faces = match_faces(cf, zone_of);
  cf = CellFaces();
  return assemble(nc, std::move(x), std::move(y), std::move(z), std::move(faces));
}

// polyMesh ---------------------------------------------------------------------------------------

namespace {

std::string slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("E0130: cannot open " + path);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

// A cursor over one polyMesh file's text, every read checked: a damaged file is an error
// (E0132), never a read past its end.
class FoamText {
 public:
  FoamText(std::string text, std::string name) : s_(std::move(text)), name_(std::move(name)) {
    // The list starts after the FoamFile header and comments, at its count.
    std::size_t p = s_.find("FoamFile");
    p = p == std::string::npos ? 0 : s_.find('}', p);
    if (p == std::string::npos) fail("an unterminated FoamFile header");
    if (p > 0) {
      // A binary file's lists hold raw little-endian labels (32 or 64 bits) and 64-bit scalars.
      const std::string header = s_.substr(0, p);
      const std::size_t f = header.find("format");
      const std::size_t v = f == std::string::npos ? std::string::npos : header.find_first_not_of(" \t", f + 6);
      binary_ = v != std::string::npos && header.compare(v, 6, "binary") == 0;
      if (binary_) {
        if (header.find("label=64") != std::string::npos) label_bytes_ = 8;
        if (header.find("scalar=32") != std::string::npos) scalar_bytes_ = 4;  // read, widened to double
        if (header.find("MSB") != std::string::npos) unsupported("a big-endian file");
      }
    }
    p_ = p == 0 ? 0 : p + 1;
    while (p_ < s_.size()) {
      if (s_.compare(p_, 2, "//") == 0) {
        p_ = s_.find('\n', p_);
        if (p_ == std::string::npos) p_ = s_.size();
      } else if (s_.compare(p_, 2, "/*") == 0) {
        const std::size_t q = s_.find("*/", p_);
        if (q == std::string::npos) fail("an unterminated comment");
        p_ = q + 2;
      } else if (std::isdigit(static_cast<unsigned char>(s_[p_]))) {
        return;
      } else {
        ++p_;
      }
    }
    fail("no list");
  }
  [[noreturn]] void fail(const std::string& why) const {
    throw std::runtime_error("E0132: polyMesh " + name_ + ": " + why);
  }
  [[noreturn]] void unsupported(const std::string& what) const {
    throw std::runtime_error("E0133: polyMesh " + name_ + ": unsupported " + what);
  }
  long integer() {
    const char* b = s_.c_str() + p_;
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(b, &e, 10);
    if (e == b || errno != 0) fail("expected an integer");
    p_ += static_cast<std::size_t>(e - b);
    return v;
  }
  double real() {
    const char* b = s_.c_str() + p_;
    char* e = nullptr;
    const double v = std::strtod(b, &e);
    if (e == b) fail("expected a number");
    p_ += static_cast<std::size_t>(e - b);
    return v;
  }
  void open(char c) {
    const std::size_t q = s_.find(c, p_);
    if (q == std::string::npos) fail(std::string("expected '") + c + "'");
    p_ = q + 1;
  }
  // A list count, no larger than the items of at least `each` characters (or bytes, in a binary
  // file) that the rest holds.
  std::size_t count(std::size_t each) {
    const long n = integer();
    if (n < 0 || static_cast<std::size_t>(n) > (s_.size() - p_) / each) fail("a count larger than the file can hold");
    return static_cast<std::size_t>(n);
  }
  // After a list's count: whether the list is written uniform, N{value}; consumes the '{'.
  bool uniform() {
    std::size_t q = p_;
    while (q < s_.size() && std::isspace(static_cast<unsigned char>(s_[q]))) ++q;
    if (q >= s_.size() || s_[q] != '{') return false;
    p_ = q + 1;
    return true;
  }
  [[nodiscard]] bool binary() const { return binary_; }
  [[nodiscard]] std::size_t label_bytes() const { return label_bytes_; }
  [[nodiscard]] std::size_t scalar_bytes() const { return scalar_bytes_; }
  // A binary list's items: right after its '(', raw, then its ')'.
  long raw_label() {
    if (label_bytes_ == 8) return static_cast<long>(raw<std::int64_t>());
    return static_cast<long>(raw<std::int32_t>());
  }
  double raw_scalar() { return scalar_bytes_ == 4 ? static_cast<double>(raw<float>()) : raw<double>(); }
  void close() {
    if (p_ >= s_.size() || s_[p_] != ')') fail("expected ')' after a binary list");
    ++p_;
  }

 private:
  template <class T>
  T raw() {
    if (p_ + sizeof(T) > s_.size()) fail("the file ends inside a list");
    T v;
    std::memcpy(&v, s_.data() + p_, sizeof(T));
    p_ += sizeof(T);
    return v;
  }
  bool binary_ = false;
  std::size_t label_bytes_ = 4;
  std::size_t scalar_bytes_ = 8;
  std::string s_;
  std::string name_;
  std::size_t p_ = 0;
};

std::vector<long> label_list(FoamText& t) {
  const std::size_t n = t.count(t.binary() ? t.label_bytes() : 2);
  t.open('(');
  std::vector<long> out(n);
  for (auto& v : out) v = t.binary() ? t.raw_label() : t.integer();
  if (t.binary()) t.close();
  return out;
}

// A labelList that may also be written uniform, N{v} (as OpenFOAM writes a list of one repeated
// value, in either format), with N at most `most`.
std::vector<long> labels(const std::string& dir, const char* file, std::size_t most) {
  FoamText t(slurp(dir + "/" + file), file);
  const long n = t.integer();
  if (t.uniform()) {
    if (n < 0 || static_cast<std::size_t>(n) > most) t.fail("a uniform list of more labels than the mesh has faces");
    const long v = t.integer();
    t.open('}');
    return std::vector<long>(static_cast<std::size_t>(n), v);
  }
  FoamText again(slurp(dir + "/" + file), file);
  return label_list(again);
}

// The boundary file's patch dictionaries, each read as a unit: name { key value; ... }.
struct Patch {
  std::string name, type;
  long start = -1, count = -1;
};

std::vector<Patch> boundary_patches(const std::string& text) {
  const auto fail = [](const std::string& why) -> void {
    throw std::runtime_error("E0132: polyMesh boundary: " + why);
  };
  std::size_t p = text.find("FoamFile");
  p = p == std::string::npos ? 0 : text.find('}', p);
  if (p == std::string::npos) fail("an unterminated FoamFile header");
  // Tokens: words, '{', '}', ';', '(' ... ')' groups kept whole; comments skipped.
  std::vector<std::string> tok;
  for (std::size_t i = p == 0 ? 0 : p + 1; i < text.size();) {
    const char c = text[i];
    if (std::isspace(static_cast<unsigned char>(c))) {
      ++i;
    } else if (text.compare(i, 2, "//") == 0) {
      const std::size_t q = text.find('\n', i);
      i = q == std::string::npos ? text.size() : q;
    } else if (text.compare(i, 2, "/*") == 0) {
      const std::size_t q = text.find("*/", i);
      if (q == std::string::npos) fail("an unterminated comment");
      i = q + 2;
    } else if (c == '{' || c == '}' || c == ';' || c == '(' || c == ')') {
      tok.emplace_back(1, c);
      ++i;
    } else if (c == '"') {
      const std::size_t q = text.find('"', i + 1);
      if (q == std::string::npos) fail("an unterminated string");
      tok.push_back(text.substr(i, q + 1 - i));
      i = q + 1;
    } else {
      const std::size_t b = i;
      while (i < text.size() && !std::isspace(static_cast<unsigned char>(text[i])) &&
             std::strchr("{};()\"", text[i]) == nullptr)
        ++i;
      tok.push_back(text.substr(b, i - b));
    }
  }
  std::size_t k = 0;
  const auto next = [&]() -> const std::string& {
    if (k >= tok.size()) fail("the file ends inside the patch list");
    return tok[k++];
  };
  const auto integer = [&](const std::string& w, const std::string& what) {
    char* e = nullptr;
    errno = 0;
    const long v = std::strtol(w.c_str(), &e, 10);
    if (w.empty() || *e != '\0' || errno != 0) fail(what + " is not an integer");
    return v;
  };
  const long n = integer(next(), "the patch count");
  if (n < 0 || static_cast<std::size_t>(n) > tok.size()) fail("a patch count larger than the file can hold");
  if (next() != "(") fail("expected '(' after the patch count");
  std::vector<Patch> out;
  for (long i = 0; i < n; ++i) {
    Patch patch;
    patch.name = next();
    if (patch.name == "(" || patch.name == ")" || patch.name == "{" || patch.name == "}" || patch.name == ";")
      fail("expected a patch name");
    if (next() != "{") fail("patch '" + patch.name + "' has no dictionary");
    // key value ... ; entries until the closing brace (nested dictionaries and lists skipped whole)
    for (;;) {
      const std::string key = next();
      if (key == "}") break;
      std::vector<std::string> value;
      int depth = 0;
      for (;;) {
        const std::string w = next();
        if (w == "{" || w == "(") ++depth;
        if (w == "}" || w == ")") {
          if (depth == 0) fail("patch '" + patch.name + "': unbalanced '" + w + "'");
          --depth;
          if (depth == 0 && w == "}") break;  // a sub-dictionary ends its entry without ';'
          continue;
        }
        if (depth == 0 && w == ";") break;
        if (depth == 0) value.push_back(w);
      }
      if (key == "type" || key == "nFaces" || key == "startFace") {
        if (value.size() != 1) fail("patch '" + patch.name + "': '" + key + "' takes one value");
        if (key == "type") patch.type = value[0];
        if (key == "nFaces") patch.count = integer(value[0], "patch '" + patch.name + "' nFaces");
        if (key == "startFace") patch.start = integer(value[0], "patch '" + patch.name + "' startFace");
      }
    }
    if (patch.type.empty()) fail("patch '" + patch.name + "' has no 'type'");
    if (patch.count < 0) fail("patch '" + patch.name + "' has no non-negative 'nFaces'");
    if (patch.start < 0) fail("patch '" + patch.name + "' has no non-negative 'startFace'");
    out.push_back(std::move(patch));
  }
  if (next() != ")") fail("more patches than the count says");
  return out;
}

}  // namespace

Mesh read_polymesh(const std::string& dir) {
  std::vector<double> x, y, z;
  {
    FoamText t(slurp(dir + "/points"), "points");
    const std::size_t n = t.count(t.binary() ? 3 * t.scalar_bytes() : 7);  // three scalars, or "(a b c)"
    t.open('(');
    x.reserve(n);
    y.reserve(n);
    z.reserve(n);
    for (std::size_t i = 0; i < n; ++i) {
      if (t.binary()) {
        x.push_back(t.raw_scalar());
        y.push_back(t.raw_scalar());
        z.push_back(t.raw_scalar());
        continue;
      }
      t.open('(');
      x.push_back(t.real());
      y.push_back(t.real());
      z.push_back(t.real());
      t.open(')');
    }
    if (t.binary()) t.close();
  }
  const long np = static_cast<long>(x.size());
  RawFaces f;
  {
    FoamText t(slurp(dir + "/faces"), "faces");
    if (t.binary()) {
      // faceCompactList: the faces' offsets (one more than the faces), then their nodes.
      const std::vector<long> off = label_list(t);
      const std::vector<long> nodes = label_list(t);
      if (off.empty() || off[0] != 0 || off.back() != static_cast<long>(nodes.size())) t.fail("face offsets");
      for (std::size_t i = 0; i + 1 < off.size(); ++i) {
        const long k = off[i + 1] - off[i];
        if (k < 3 || k > 1024) t.fail("a face of " + std::to_string(k) + " nodes");
        for (long j = off[i]; j < off[i + 1]; ++j) {
          const long v = nodes[static_cast<std::size_t>(j)];
          if (v < 0 || v >= np) t.fail("a face node out of range");
          f.nodes.push_back(static_cast<index_t>(v));
        }
        f.off.push_back(static_cast<offset_t>(f.nodes.size()));
      }
    }
    const std::size_t n = t.binary() ? 0 : t.count(8);  // "3(a b c)"
    if (!t.binary()) t.open('(');
    for (std::size_t i = 0; i < n; ++i) {
      const long k = t.integer();
      if (k < 3 || k > 1024) t.fail("a face of " + std::to_string(k) + " nodes");
      t.open('(');
      for (long j = 0; j < k; ++j) {
        const long v = t.integer();
        if (v < 0 || v >= np) t.fail("a face node out of range");
        f.nodes.push_back(static_cast<index_t>(v));
      }
      t.open(')');
      f.off.push_back(static_cast<offset_t>(f.nodes.size()));
    }
  }
  const std::size_t nf = f.off.size() - 1;
  const auto own = labels(dir, "owner", nf);
  const auto nei = labels(dir, "neighbour", nf);
  if (own.size() != nf) throw std::runtime_error("E0132: polyMesh: owner and faces differ in length");
  if (nei.size() > nf) throw std::runtime_error("E0132: polyMesh: more neighbours than faces");
  long n_cells = 0;
  for (long c : own) n_cells = std::max(n_cells, c + 1);
  for (long c : nei) n_cells = std::max(n_cells, c + 1);
  for (long c : own) {
    if (c < 0) throw std::runtime_error("E0132: polyMesh: a negative owner");
  }
  for (long c : nei) {
    if (c < 0) throw std::runtime_error("E0132: polyMesh: a negative neighbour");
  }
  if (n_cells > INT32_MAX) throw std::runtime_error("E0132: polyMesh: too many cells");
  // A cell has 4 faces at least and a face 2 cells at most: a label past (faces + interior faces) / 4
  // cannot be a cell of this mesh, and is refused before anything is sized by it.
  if (n_cells > static_cast<long>((nf + nei.size()) / 4)) {
    throw std::runtime_error("E0132: polyMesh: a cell label beyond the cells the faces can close");
  }
  {
    std::vector<int> faces_of(static_cast<std::size_t>(n_cells), 0);
    for (long c : own) ++faces_of[static_cast<std::size_t>(c)];
    for (long c : nei) ++faces_of[static_cast<std::size_t>(c)];
    for (int k : faces_of) {
      if (k < 4) throw std::runtime_error("E0132: polyMesh: a cell with fewer than 4 faces");
    }
  }
  // boundary: the patches in order, each continuing the boundary face range where the previous
  // one ends (the first at the first boundary face), together covering every boundary face, as
  // io/src/polymesh.cpp:486-509 requires; processor patches (a decomposed mesh) are refused.
  std::vector<index_t> zone(nf, -1);
  {
    long expected = static_cast<long>(nei.size());
    index_t patch = 0;
    for (const Patch& pt : boundary_patches(slurp(dir + "/boundary"))) {
      if (pt.type == "processor" || pt.type == "processorCyclic") {
        throw std::runtime_error("E0133: polyMesh boundary: unsupported decomposed mesh (processor patch '" +
                                 pt.name + "')");
      }
      if (pt.start != expected || pt.count > static_cast<long>(nf) - pt.start) {
        throw std::runtime_error("E0132: polyMesh boundary: patch '" + pt.name +
                                 "' does not continue the boundary face range");
      }
      for (long i = pt.start; i < pt.start + pt.count; ++i) zone[static_cast<std::size_t>(i)] = patch;
      expected = pt.start + pt.count;
      ++patch;
    }
    if (expected != static_cast<long>(nf)) {
      throw std::runtime_error("E0132: polyMesh boundary: the patches do not cover every boundary face");
    }
  }
  for (std::size_t i = 0; i < nf; ++i) {
    f.c0.push_back(static_cast<index_t>(own[i]));
    f.c1.push_back(i < nei.size() ? static_cast<index_t>(nei[i]) : -1);
    if (i >= nei.size() && zone[i] < 0) throw std::runtime_error("E0132: polyMesh boundary: a face in no patch");
    f.zone.push_back(i < nei.size() ? -1 : zone[i]);
  }
  return assemble(static_cast<index_t>(n_cells), std::move(x), std::move(y), std::move(z), std::move(f));
}

// Gmsh MSH 4.1 (ASCII) ------------------------------------------------------------------------------

namespace {

// Whitespace-separated tokens of a text; a "quoted string" is one token.
class Tokens {
 public:
  explicit Tokens(std::string text) : s_(std::move(text)) {}
  bool done() {
    skip();
    return p_ >= s_.size();
  }
  std::string word() {
    skip();
    if (p_ >= s_.size()) throw std::runtime_error("E0132: gmsh: the file ends inside a section");
    if (s_[p_] == '"') {
      const std::size_t q = s_.find('"', p_ + 1);
      if (q == std::string::npos) throw std::runtime_error("E0132: gmsh: an unterminated string");
      std::string w = s_.substr(p_ + 1, q - p_ - 1);
      p_ = q + 1;
      return w;
    }
    const std::size_t b = p_;
    while (p_ < s_.size() && !std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
    return s_.substr(b, p_ - b);
  }
  long long integer() {
    const std::string w = word();
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(w.c_str(), &end, 10);
    if (end == w.c_str() || *end != '\0' || errno != 0) {
      throw std::runtime_error("E0132: gmsh: expected an integer, found '" + w + "'");
    }
    return v;
  }
  // A count or tag that must be >= lo and small enough to index with.
  std::size_t size(long long lo = 0) {
    const long long v = integer();
    if (v < lo || v > INT32_MAX) throw std::runtime_error("E0132: gmsh: a count or tag out of range");
    return static_cast<std::size_t>(v);
  }
  double real() {
    const std::string w = word();
    char* end = nullptr;
    const double v = std::strtod(w.c_str(), &end);
    if (end == w.c_str() || *end != '\0') throw std::runtime_error("E0132: gmsh: expected a number, found '" + w + "'");
    return v;
  }
  void expect(const char* w) {
    if (word() != w) throw std::runtime_error(std::string("E0132: gmsh: expected ") + w);
  }

  // The typed fields of $Entities, $Nodes and $Elements: text tokens, or in a binary file's
  // sections the raw little-endian values (int for tags, dimensions and types, size_t for counts
  // and node and element tags, double for coordinates).
  void set_binary(bool on) { binary_ = on; }
  long long i32() { return binary_ ? static_cast<long long>(raw<std::int32_t>()) : integer(); }
  long long u64() {
    if (!binary_) return integer();
    const std::uint64_t v = raw<std::uint64_t>();
    if (v > static_cast<std::uint64_t>(INT64_MAX)) throw std::runtime_error("E0132: gmsh: a count or tag out of range");
    return static_cast<long long>(v);
  }
  std::size_t count(long long lo = 0) {
    const long long v = u64();
    if (v < lo || v > INT32_MAX) throw std::runtime_error("E0132: gmsh: a count or tag out of range");
    return static_cast<std::size_t>(v);
  }
  double f64() { return binary_ ? raw<double>() : real(); }
  // A declared count of items that take at least `text_bytes` (ASCII) or `binary_bytes` each must
  // fit in what is left of the file: checked before anything is allocated for them.
  void check_fits(std::size_t n, std::size_t text_bytes, std::size_t binary_bytes) const {
    const std::size_t each = binary_ ? binary_bytes : text_bytes;
    if (n > (s_.size() - std::min(p_, s_.size())) / each) {
      throw std::runtime_error("E0132: gmsh: a count larger than the file can hold");
    }
  }
  // After a binary block: the newline before its end marker.
  void skip_newline() {
    if (p_ < s_.size() && s_[p_] == '\n') ++p_;
  }
  std::int32_t probe_int() {
    std::int32_t v = raw<std::int32_t>();
    return v;
  }

 private:
  template <class T>
  T raw() {
    if (p_ + sizeof(T) > s_.size()) throw std::runtime_error("E0132: gmsh: the file ends inside a section");
    T v;
    std::memcpy(&v, s_.data() + p_, sizeof(T));
    p_ += sizeof(T);
    return v;
  }
  void skip() {
    while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
  }
  std::string s_;
  std::size_t p_ = 0;
  bool binary_ = false;
};

// Linear element types of the format: node count, dimension, and for volume elements the faces of
// an element in the format's node ordering, outward for a positively oriented element
// (count, then that many local nodes).
struct Shape {
  int nodes, dim, n_faces;
  int face[6][5];
};

bool shape_of(long long type, Shape& out) {
  static const Shape tet{4, 3, 4, {{3, 0, 2, 1}, {3, 0, 1, 3}, {3, 1, 2, 3}, {3, 0, 3, 2}}};
  static const Shape hex{
      8, 3, 6, {{4, 0, 3, 2, 1}, {4, 4, 5, 6, 7}, {4, 0, 1, 5, 4}, {4, 2, 3, 7, 6}, {4, 0, 4, 7, 3}, {4, 1, 2, 6, 5}}};
  static const Shape prism{6, 3, 5, {{3, 0, 2, 1}, {3, 3, 4, 5}, {4, 0, 1, 4, 3}, {4, 0, 3, 5, 2}, {4, 1, 2, 5, 4}}};
  static const Shape pyramid{5, 3, 5, {{4, 0, 3, 2, 1}, {3, 0, 1, 4}, {3, 1, 2, 4}, {3, 2, 3, 4}, {3, 3, 0, 4}}};
  switch (type) {
    case 15: out = Shape{1, 0, 0, {}}; return true;  // point
    case 1: out = Shape{2, 1, 0, {}}; return true;   // line
    case 2: out = Shape{3, 2, 0, {}}; return true;   // triangle
    case 3: out = Shape{4, 2, 0, {}}; return true;   // quadrangle
    case 4: out = tet; return true;
    case 5: out = hex; return true;
    case 6: out = prism; return true;
    case 7: out = pyramid; return true;
    default: return false;
  }
}

// A boundary face known by its sorted nodes (-1 padded).
using FaceKey = std::array<index_t, 4>;

FaceKey face_key(const index_t* v, std::size_t count) {
  FaceKey k{-1, -1, -1, -1};
  for (std::size_t i = 0; i < count && i < 4; ++i) k[i] = v[i];
  std::sort(k.begin(), k.begin() + static_cast<std::ptrdiff_t>(std::min<std::size_t>(count, 4)));
  return k;
}

// The surface links of $Periodic as periodic faces (the mesh's model: an interior face whose
// neighbour sits, as seen from the owner, at its own position plus `shift`): each boundary face of
// a slave surface is paired with the master surface's face at its position minus the link's
// translation t, node by node within 1e-9 |t|; the master face becomes the interior face (owner the
// master side's cell, neighbour the slave side's, shift -t) and the slave face goes. Node pairs
// that the translation does not carry onto each other, a slave face without its image, and two
// links of one surface pair with different translations are refused.
template <class Link>
RawFaces periodic_faces(RawFaces in, const std::vector<Link>& links, const std::map<FaceKey, long long>& surface_of,
                        const std::unordered_map<long long, index_t>& node_of, const std::vector<double>& x,
                        const std::vector<double>& y, const std::vector<double>& z) {
  const std::size_t n = in.c0.size();
  std::vector<long long> surface(n, -1);
  for (std::size_t i = 0; i < n; ++i) {
    if (in.c1[i] >= 0) continue;
    const auto it = surface_of.find(face_key(&in.nodes[static_cast<std::size_t>(in.off[i])],
                                             static_cast<std::size_t>(in.off[i + 1] - in.off[i])));
    if (it != surface_of.end()) surface[i] = it->second;
  }
  auto point = [&](index_t v) { return std::array<double, 3>{x[at(v)], y[at(v)], z[at(v)]}; };
  std::vector<double> shift(3 * n, 0.0);
  std::vector<char> gone(n, 0);
  std::map<std::pair<long long, long long>, std::array<double, 3>> seen;
  for (const Link& l : links) {
    const double len = std::sqrt(l.t[0] * l.t[0] + l.t[1] * l.t[1] + l.t[2] * l.t[2]);
    const double tol = 1e-9 * len;
    const auto key = std::make_pair(std::min(l.slave, l.master), std::max(l.slave, l.master));
    const std::array<double, 3> tt{l.t[0], l.t[1], l.t[2]};
    const auto prior = seen.find(key);
    if (prior != seen.end() && prior->second != tt) {
      throw std::runtime_error("E0132: gmsh: two periodic links of one surface pair with different translations");
    }
    seen[key] = tt;
    auto carried = [&](index_t s, index_t m) {
      const auto ps = point(s), pm = point(m);
      const double d[3] = {ps[0] - pm[0] - l.t[0], ps[1] - pm[1] - l.t[1], ps[2] - pm[2] - l.t[2]};
      return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= tol;
    };
    for (const auto& [ts, tm] : l.pairs) {
      const auto s = node_of.find(ts), m = node_of.find(tm);
      if (s == node_of.end() || m == node_of.end()) throw std::runtime_error("E0132: gmsh: a periodic node pair of undefined nodes");
      if (!carried(s->second, m->second)) throw std::runtime_error("E0132: gmsh: a periodic node pair the translation does not carry onto each other");
    }
    if (len == 0.0) throw std::runtime_error("E0132: gmsh: a periodic link without a translation");
    // Master faces by the cell of space their centroid falls in (side 4 tol): an image lies in
    // the same cell or a neighbouring one.
    const double cell = 4.0 * tol;
    auto centroid = [&](std::size_t i) {
      std::array<double, 3> c{0.0, 0.0, 0.0};
      for (auto k = in.off[i]; k < in.off[i + 1]; ++k) {
        const auto p = point(in.nodes[static_cast<std::size_t>(k)]);
        for (int q = 0; q < 3; ++q) c[q] += p[q];
      }
      const double m = static_cast<double>(in.off[i + 1] - in.off[i]);
      for (int q = 0; q < 3; ++q) c[q] /= m;
      return c;
    };
    auto bin = [&](const std::array<double, 3>& c) {
      return std::array<long long, 3>{static_cast<long long>(std::floor(c[0] / cell)),
                                      static_cast<long long>(std::floor(c[1] / cell)),
                                      static_cast<long long>(std::floor(c[2] / cell))};
    };
    std::map<std::array<long long, 3>, std::vector<std::size_t>> masters;
    for (std::size_t i = 0; i < n; ++i) {
      if (surface[i] == l.master && !gone[i] && in.c1[i] < 0) masters[bin(centroid(i))].push_back(i);
    }
    for (std::size_t i = 0; i < n; ++i) {
      if (surface[i] != l.slave || gone[i] || in.c1[i] >= 0) continue;
      auto c = centroid(i);
      for (int q = 0; q < 3; ++q) c[q] -= l.t[q];
      const auto b = bin(c);
      std::size_t found = n;
      for (long long dx = -1; dx <= 1 && found == n; ++dx) {
        for (long long dy = -1; dy <= 1 && found == n; ++dy) {
          for (long long dz = -1; dz <= 1 && found == n; ++dz) {
            const auto it = masters.find({b[0] + dx, b[1] + dy, b[2] + dz});
            if (it == masters.end()) continue;
            for (const std::size_t j : it->second) {
              if (in.c1[j] >= 0 || in.off[j + 1] - in.off[j] != in.off[i + 1] - in.off[i]) continue;
              bool all = true;  // every slave node carried onto some node of the master face
              for (auto k = in.off[i]; k < in.off[i + 1] && all; ++k) {
                bool any = false;
                for (auto h = in.off[j]; h < in.off[j + 1] && !any; ++h) {
                  any = carried(in.nodes[static_cast<std::size_t>(k)], in.nodes[static_cast<std::size_t>(h)]);
                }
                all = any;
              }
              if (all) {
                found = j;
                break;
              }
            }
          }
        }
      }
      if (found == n) throw std::runtime_error("E0132: gmsh: a periodic face without its image on the master surface");
      in.c1[found] = in.c0[i];
      in.zone[found] = -1;
      for (int q = 0; q < 3; ++q) shift[3 * found + static_cast<std::size_t>(q)] = -l.t[q];
      gone[i] = 1;
    }
  }
  RawFaces out;
  for (std::size_t i = 0; i < n; ++i) {
    if (gone[i]) continue;
    out.nodes.insert(out.nodes.end(), in.nodes.begin() + in.off[i], in.nodes.begin() + in.off[i + 1]);
    out.off.push_back(static_cast<offset_t>(out.nodes.size()));
    out.c0.push_back(in.c0[i]);
    out.c1.push_back(in.c1[i]);
    out.zone.push_back(in.zone[i]);
    out.shift.insert(out.shift.end(), shift.begin() + static_cast<std::ptrdiff_t>(3 * i),
                     shift.begin() + static_cast<std::ptrdiff_t>(3 * i + 3));
  }
  return out;
}

}  // namespace

Mesh read_gmsh(const std::string& path) {
  std::ifstream probe(path, std::ios::binary);
  if (!probe) throw std::runtime_error("E0130: cannot open " + path);
  probe.close();
  Tokens t(slurp(path));
  if (t.done() || t.word() != "$MeshFormat") throw std::runtime_error("E0131: gmsh: no $MeshFormat at the start");
  const std::string version = t.word();
  if (version != "4.1") throw std::runtime_error("E0133: gmsh: MSH " + version + " is not read (MSH 4.1 only)");
  const long long file_type = t.integer();
  if (file_type != 0 && file_type != 1) throw std::runtime_error("E0132: gmsh: file type 0 (ASCII) or 1 (binary)");
  const bool binary = file_type == 1;
  if (t.integer() != 8) throw std::runtime_error("E0133: gmsh: a data size other than 8");
  if (binary) {
    // The integer 1 written in binary after the header line tells the byte order.
    t.skip_newline();
    if (t.probe_int() != 1) throw std::runtime_error("E0133: gmsh: a big-endian binary file");
  }
  t.expect("$EndMeshFormat");

  std::vector<int> surface_groups;                // 2-D physical tags
  std::map<std::pair<int, long long>, int> physical;  // (dim, entity) -> its physical tag, 0: none
  std::vector<double> x, y, z;
  std::unordered_map<long long, index_t> node_of;  // node tag -> index
  std::vector<index_t> cell_nodes;
  std::vector<const Shape*> cell_shape;
  std::vector<std::pair<FaceKey, int>> labelled;  // boundary elements and their physical tags
  std::map<FaceKey, long long> surface_of;         // every 2-D element's surface entity
  struct Link {
    long long slave, master;
    double t[3];
    std::vector<std::pair<long long, long long>> pairs;  // node tags: slave, master
  };
  std::vector<Link> links;  // the surface links of $Periodic
  std::array<Shape, 8> shapes{};                  // the volume shapes met, by element type
  bool have_nodes = false, have_elements = false;

  while (!t.done()) {
    const std::string section = t.word();
    if (section == "$PhysicalNames") {
      const std::size_t n = t.size();
      for (std::size_t i = 0; i < n; ++i) {
        const long long dim = t.integer();
        const long long tag = t.integer();
        (void)t.word();  // the name
        if (tag <= 0) throw std::runtime_error("E0132: gmsh: a physical tag must be positive");
        if (dim == 2) surface_groups.push_back(static_cast<int>(tag));
      }
      t.expect("$EndPhysicalNames");
    } else if (section == "$Entities") {
      t.skip_newline();
      t.set_binary(binary);
      std::size_t count[4];
      for (auto& c : count) {
        c = t.count();
        t.check_fits(c, 10, 36);  // a tag, coordinates and a physical count at least
      }
      for (int dim = 0; dim < 4; ++dim) {
        for (std::size_t i = 0; i < count[dim]; ++i) {
          const long long tag = t.i32();
          for (int k = 0; k < (dim == 0 ? 3 : 6); ++k) (void)t.f64();
          const std::size_t np = t.count();
          if (np > 1) throw std::runtime_error("E0135: gmsh: an entity in two physical groups");
          const int ph = np == 1 ? static_cast<int>(t.i32()) : 0;
          if (np == 1 && ph <= 0) throw std::runtime_error("E0132: gmsh: a physical tag must be positive");
          physical[{dim, tag}] = ph;
          if (dim == 2 && ph != 0) surface_groups.push_back(ph);
          if (dim > 0) {
            const std::size_t nb = t.count();
            for (std::size_t k = 0; k < nb; ++k) (void)t.i32();
          }
        }
      }
      t.set_binary(false);
      t.expect("$EndEntities");
    } else if (section == "$Nodes") {
      t.skip_newline();
      t.set_binary(binary);
      const std::size_t blocks = t.count();
      const std::size_t n = t.count();
      (void)t.u64();
      (void)t.u64();
      t.check_fits(blocks, 8, 20);
      t.check_fits(n, 8, 32);  // a tag and three coordinates
      x.reserve(n);
      y.reserve(n);
      z.reserve(n);
      node_of.reserve(n);
      for (std::size_t b = 0; b < blocks; ++b) {
        const long long dim = t.i32();
        (void)t.i32();  // entity
        const long long parametric = t.i32();
        const std::size_t m = t.count();
        if (x.size() + m > n) throw std::runtime_error("E0132: gmsh: $Nodes blocks hold more nodes than declared");
        t.check_fits(m, 8, 32);
        std::vector<long long> tags(m);
        for (auto& g : tags) g = t.u64();
        for (std::size_t i = 0; i < m; ++i) {
          if (!node_of.emplace(tags[i], static_cast<index_t>(x.size())).second) {
            throw std::runtime_error("E0132: gmsh: node tag " + std::to_string(tags[i]) + " is defined twice");
          }
          x.push_back(t.f64());
          y.push_back(t.f64());
          z.push_back(t.f64());
          for (long long k = 0; parametric != 0 && k < dim; ++k) (void)t.f64();
        }
      }
      if (x.size() != n) throw std::runtime_error("E0132: gmsh: $Nodes count does not match its blocks");
      t.set_binary(false);
      t.expect("$EndNodes");
      have_nodes = true;
    } else if (section == "$Elements") {
      if (!have_nodes) throw std::runtime_error("E0132: gmsh: $Elements before $Nodes");
      t.skip_newline();
      t.set_binary(binary);
      const std::size_t blocks = t.count();
      const std::size_t n = t.count();
      (void)t.u64();
      (void)t.u64();
      t.check_fits(blocks, 8, 20);
      t.check_fits(n, 4, 16);  // a tag and a node at least
      std::size_t read = 0;
      std::vector<index_t> v;
      for (std::size_t b = 0; b < blocks; ++b) {
        const long long dim = t.i32();
        const long long entity = t.i32();
        const long long type = t.i32();
        const std::size_t m = t.count();
        read += m;
        if (read > n) throw std::runtime_error("E0132: gmsh: $Elements blocks hold more elements than declared");
        Shape sh;
        if (!shape_of(type, sh)) {
          throw std::runtime_error("E0133: gmsh: element type " + std::to_string(type) + " (linear elements only)");
        }
        if (sh.dim != dim) throw std::runtime_error("E0132: gmsh: an element type does not match its entity's dimension");
        const Shape* kept = nullptr;
        if (dim == 3) {
          shapes[static_cast<std::size_t>(type)] = sh;
          kept = &shapes[static_cast<std::size_t>(type)];
        }
        const auto it = physical.find({static_cast<int>(dim), entity});
        const int ph = it == physical.end() ? 0 : it->second;
        v.resize(static_cast<std::size_t>(sh.nodes));
        for (std::size_t e = 0; e < m; ++e) {
          (void)t.u64();  // element tag
          for (auto& q : v) {
            const long long g = t.u64();
            const auto f = node_of.find(g);
            if (f == node_of.end()) throw std::runtime_error("E0132: gmsh: an element refers to undefined node " + std::to_string(g));
            q = f->second;
          }
          if (dim == 3) {
            cell_nodes.insert(cell_nodes.end(), v.begin(), v.end());
            cell_shape.push_back(kept);
          } else if (dim == 2) {
            surface_of[face_key(v.data(), v.size())] = entity;
            if (ph != 0) labelled.emplace_back(face_key(v.data(), v.size()), ph);
          }
        }
      }
      if (read != n) throw std::runtime_error("E0132: gmsh: $Elements count does not match its blocks");
      t.set_binary(false);
      t.expect("$EndElements");
      have_elements = true;
    } else if (section == "$Periodic") {
      // Links of a slave entity to its master by an affine map (16 values, row-major 4 x 4) and
      // their node pairs; the surface links become periodic faces below.
      t.skip_newline();
      t.set_binary(binary);
      const std::size_t nl = t.count();
      t.check_fits(nl, 8, 20);
      for (std::size_t l = 0; l < nl; ++l) {
        const long long dim = t.i32();
        const long long slave = t.i32();
        const long long master = t.i32();
        if (dim < 0 || dim > 2) throw std::runtime_error("E0132: gmsh: a periodic link of dimension " + std::to_string(dim));
        if (slave <= 0 || master <= 0) throw std::runtime_error("E0132: gmsh: a periodic link to entity tag 0 or less");
        const std::size_t na = t.count();
        if (na != 0 && na != 16) throw std::runtime_error("E0132: gmsh: an affine map of " + std::to_string(na) + " values");
        double a16[16] = {};
        for (std::size_t k = 0; k < na; ++k) a16[k] = t.f64();
        const std::size_t np = t.count();
        t.check_fits(np, 4, 16);
        Link link{slave, master, {a16[3], a16[7], a16[11]}, {}};
        link.pairs.reserve(np);
        for (std::size_t k = 0; k < np; ++k) {
          const long long ns = t.u64();
          const long long nm = t.u64();
          link.pairs.emplace_back(ns, nm);
        }
        if (dim != 2) continue;
        if (na != 16) throw std::runtime_error("E0133: gmsh: a periodic surface link without its affine map");
        const double id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
        for (int k : {0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14, 15}) {
          if (a16[k] != id[k]) throw std::runtime_error("E0133: gmsh: a periodic link that is not a translation");
        }
        links.push_back(std::move(link));
      }
      t.set_binary(false);
      t.expect("$EndPeriodic");
    } else if (section.size() > 1 && section[0] == '$') {
      // A section this reader does not use ($Comments, $NodeData, ...): skipped to its end.
      const std::string end = "$End" + section.substr(1);
      while (t.word() != end) {
      }
    } else {
      throw std::runtime_error("E0131: gmsh: '" + section + "' where a section should start");
    }
  }
  if (!have_elements || cell_shape.empty()) throw std::runtime_error("E0132: gmsh: no volume elements");
  // Periodic surfaces pair zones: both surfaces of a link must be in a physical group
  // (io/src/gmsh.cpp:178-182).
  for (const Link& l : links) {
    for (const long long e : {l.master, l.slave}) {
      const auto it = physical.find({2, e});
      if (it == physical.end() || it->second == 0) {
        throw std::runtime_error("E0132: gmsh: periodic surface " + std::to_string(e) + " is in no physical group");
      }
    }
  }

  // Zones: the 2-D physical groups in increasing tag order, then one zone for boundary faces of no group.
  std::sort(surface_groups.begin(), surface_groups.end());
  surface_groups.erase(std::unique(surface_groups.begin(), surface_groups.end()), surface_groups.end());
  std::map<FaceKey, int> zone_of_face;
  for (const auto& [key, ph] : labelled) {
    const int zone = static_cast<int>(std::lower_bound(surface_groups.begin(), surface_groups.end(), ph) -
                                      surface_groups.begin());
    zone_of_face[key] = zone;
  }
  const int unlabelled = static_cast<int>(surface_groups.size());

  CellFaces cf;
  const index_t n_cells = static_cast<index_t>(cell_shape.size());
  std::size_t base = 0;
  for (index_t c = 0; c < n_cells; ++c) {
    const Shape& sh = *cell_shape[at(c)];
    const index_t* v = &cell_nodes[base];
    // Orientation: the outward-flux sum of the faces' Newell normals against the node average is
    // positive for a positively oriented element; an inverted element has its faces reversed.
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < sh.nodes; ++i) {
      cx += x[at(v[i])];
      cy += y[at(v[i])];
      cz += z[at(v[i])];
    }
    cx /= sh.nodes;
    cy /= sh.nodes;
    cz /= sh.nodes;
    double flux = 0.0;
    for (int f = 0; f < sh.n_faces; ++f) {
      const int k = sh.face[f][0];
      double nx = 0, ny = 0, nz = 0, fx = 0, fy = 0, fz = 0;
      for (int i = 0; i < k; ++i) {
        const std::size_t a = at(v[sh.face[f][1 + i]]), b = at(v[sh.face[f][1 + (i + 1) % k]]);
        nx += (y[a] - y[b]) * (z[a] + z[b]);
        ny += (z[a] - z[b]) * (x[a] + x[b]);
        nz += (x[a] - x[b]) * (y[a] + y[b]);
        fx += x[a];
        fy += y[a];
        fz += z[a];
      }
      flux += nx * (fx / k - cx) + ny * (fy / k - cy) + nz * (fz / k - cz);
    }
    const bool inverted = flux < 0.0;
    for (int f = 0; f < sh.n_faces; ++f) {
      const int k = sh.face[f][0];
      for (int i = 0; i < k; ++i) cf.nodes.push_back(v[sh.face[f][1 + (inverted ? k - 1 - i : i)]]);
      cf.off.push_back(static_cast<offset_t>(cf.nodes.size()));
      cf.cell.push_back(c);
    }
    base += static_cast<std::size_t>(sh.nodes);
  }
  auto zone_of = [&](const index_t* v, std::size_t count) {
    const auto it = zone_of_face.find(face_key(v, count));
    return it == zone_of_face.end() ? unlabelled : it->second;
  };
  RawFaces faces = match_faces(cf, zone_of);
  cf = CellFaces();
  if (!links.empty()) faces = periodic_faces(std::move(faces), links, surface_of, node_of, x, y, z);
  return assemble(n_cells, std::move(x), std::move(y), std::move(z), std::move(faces));
}

// Renumbering ------------------------------------------------------------------------------------

namespace {

struct Graph {
  std::vector<offset_t> off;
  std::vector<index_t> adj;
  [[nodiscard]] int degree(index_t c) const { return static_cast<int>(off[at(c) + 1] - off[at(c)]); }
};

Graph cell_graph(const Mesh& m) {
  const std::size_t nc = at(m.n_cells);
  Graph g;
  g.off.assign(nc + 1, 0);
  for (index_t f = 0; f < m.n_interior; ++f) {
    ++g.off[at(m.owner[at(f)]) + 1];
    ++g.off[at(m.neighbour[at(f)]) + 1];
  }
  for (std::size_t c = 0; c < nc; ++c) g.off[c + 1] += g.off[c];
  g.adj.resize(static_cast<std::size_t>(g.off[nc]));
  std::vector<offset_t> next(g.off.begin(), g.off.end() - 1);
  for (index_t f = 0; f < m.n_interior; ++f) {
    const index_t a = m.owner[at(f)], b = m.neighbour[at(f)];
    g.adj[static_cast<std::size_t>(next[at(a)]++)] = b;
    g.adj[static_cast<std::size_t>(next[at(b)]++)] = a;
  }
  // Two faces between the same cells count once.
  std::vector<offset_t> off2(nc + 1, 0);
  std::size_t w = 0;
  for (std::size_t c = 0; c < nc; ++c) {
    auto first = g.adj.begin() + g.off[c], last = g.adj.begin() + g.off[c + 1];
    std::sort(first, last);
    const auto uend = std::unique(first, last);
    for (auto it = first; it != uend; ++it) g.adj[w++] = *it;
    off2[c + 1] = static_cast<offset_t>(w);
  }
  g.adj.resize(w);
  g.off.swap(off2);
  return g;
}

// Breadth-first levels from root: the last level's cells and the depth.
int bfs(const Graph& g, index_t root, std::vector<int>& level, std::vector<index_t>& last) {
  std::vector<index_t> queue{root};
  level[at(root)] = 0;
  int depth = 0;
  for (std::size_t h = 0; h < queue.size(); ++h) {
    const index_t c = queue[h];
    for (auto k = g.off[at(c)]; k < g.off[at(c) + 1]; ++k) {
      const index_t d = g.adj[static_cast<std::size_t>(k)];
      if (level[at(d)] < 0) {
        level[at(d)] = level[at(c)] + 1;
        depth = std::max(depth, level[at(d)]);
        queue.push_back(d);
      }
    }
  }
  last.clear();
  for (const index_t c : queue) {
    if (level[at(c)] == depth) last.push_back(c);
  }
  for (const index_t c : queue) level[at(c)] = -1;
  return depth;
}

// George and Liu (1979): move the root to a least-degree cell of the last level while the depth grows.
index_t pseudo_peripheral(const Graph& g, index_t seed, std::vector<int>& level) {
  std::vector<index_t> last;
  index_t root = seed;
  int depth = bfs(g, root, level, last);
  for (;;) {
    index_t best = last.front();
    for (const index_t c : last) {
      if (g.degree(c) < g.degree(best) || (g.degree(c) == g.degree(best) && c < best)) best = c;
    }
    std::vector<index_t> cand_last;
    const int d = bfs(g, best, level, cand_last);
    if (d <= depth) return root;
    root = best;
    depth = d;
    last.swap(cand_last);
  }
}

}  // namespace

std::vector<index_t> rcm_order(const Mesh& m) {
  const Graph g = cell_graph(m);
  const std::size_t nc = at(m.n_cells);
  std::vector<int> level(nc, -1);
  std::vector<char> placed(nc, 0);
  std::vector<index_t> order;
  order.reserve(nc);
  std::vector<index_t> by_degree(nc);
  std::iota(by_degree.begin(), by_degree.end(), 0);
  std::stable_sort(by_degree.begin(), by_degree.end(),
                   [&](index_t a, index_t b) { return g.degree(a) < g.degree(b); });
  std::vector<index_t> nb;
  for (const index_t seed : by_degree) {
    if (placed[at(seed)] != 0) continue;
    const index_t root = pseudo_peripheral(g, seed, level);
    std::size_t head = order.size();
    order.push_back(root);
    placed[at(root)] = 1;
    while (head < order.size()) {
      const index_t c = order[head++];
      nb.clear();
      for (auto k = g.off[at(c)]; k < g.off[at(c) + 1]; ++k) {
        const index_t d = g.adj[static_cast<std::size_t>(k)];
        if (placed[at(d)] == 0) nb.push_back(d);
      }
      std::sort(nb.begin(), nb.end(), [&](index_t a, index_t b) {
        return g.degree(a) != g.degree(b) ? g.degree(a) < g.degree(b) : a < b;
      });
      for (const index_t d : nb) {
        placed[at(d)] = 1;
        order.push_back(d);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

namespace {

// The 21 low bits of v spread to every third bit.
std::uint64_t spread3(std::uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

}  // namespace

std::vector<index_t> morton_order(const Mesh& m) {
  const std::size_t nc = at(m.n_cells), ni = at(m.n_interior);
  std::vector<double> cx(nc, 0.0), cy(nc, 0.0), cz(nc, 0.0), cnt(nc, 0.0);
  for (std::size_t f = 0; f < at(m.n_faces); ++f) {
    const auto a = static_cast<std::size_t>(m.face_node_offsets[f]), b = static_cast<std::size_t>(m.face_node_offsets[f + 1]);
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (std::size_t k = a; k < b; ++k) {
      const auto p = at(m.face_nodes[k]);
      sx += m.x[p];
      sy += m.y[p];
      sz += m.z[p];
    }
    const double nn = static_cast<double>(b - a);
    const std::size_t o = at(m.owner[f]);
    cx[o] += sx;
    cy[o] += sy;
    cz[o] += sz;
    cnt[o] += nn;
    if (f < ni) {
      const std::size_t r = at(m.neighbour[f]);
      const auto s = m.face_shift(static_cast<index_t>(f));
      cx[r] += sx - nn * s[0];
      cy[r] += sy - nn * s[1];
      cz[r] += sz - nn * s[2];
      cnt[r] += nn;
    }
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (std::size_t c = 0; c < nc; ++c) {
    cx[c] /= cnt[c];
    cy[c] /= cnt[c];
    cz[c] /= cnt[c];
    const double v[3] = {cx[c], cy[c], cz[c]};
    for (int u = 0; u < 3; ++u) {
      lo[u] = std::min(lo[u], v[u]);
      hi[u] = std::max(hi[u], v[u]);
    }
  }
  const double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  const double q = static_cast<double>((1u << 21) - 1) / span;
  std::vector<std::pair<std::uint64_t, index_t>> key(nc);
  for (std::size_t c = 0; c < nc; ++c) {
    const auto ix = static_cast<std::uint64_t>((cx[c] - lo[0]) * q);
    const auto iy = static_cast<std::uint64_t>((cy[c] - lo[1]) * q);
    const auto iz = static_cast<std::uint64_t>((cz[c] - lo[2]) * q);
    key[c] = {spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2), static_cast<index_t>(c)};
  }
  std::sort(key.begin(), key.end());
  std::vector<index_t> order(nc);
  for (std::size_t c = 0; c < nc; ++c) order[c] = key[c].second;
  return order;
}

Mesh renumber(const Mesh& m, const std::vector<index_t>& new_to_old) {
  const std::size_t nc = at(m.n_cells), nf = at(m.n_faces), ni = at(m.n_interior);
  if (new_to_old.size() != nc) throw std::invalid_argument("renumber: permutation size");
  std::vector<index_t> old_to_new(nc, -1);
  for (std::size_t i = 0; i < nc; ++i) {
    const index_t o = new_to_old[i];
    if (o < 0 || at(o) >= nc || old_to_new[at(o)] != -1) throw std::invalid_argument("renumber: not a permutation");
    old_to_new[at(o)] = static_cast<index_t>(i);
  }
  RawFaces f;
  f.off = m.face_node_offsets;
  f.nodes = m.face_nodes;
  f.c0.resize(nf);
  f.c1.resize(nf);
  f.zone.resize(nf);
  if (m.periodic()) f.shift.resize(3 * nf, 0.0);
  for (std::size_t i = 0; i < nf; ++i) {
    f.c0[i] = old_to_new[at(m.owner[i])];
    f.c1[i] = i < ni ? old_to_new[at(m.neighbour[i])] : -1;
    f.zone[i] = i < ni ? -1 : m.face_zone[i - ni];
    if (m.periodic() && i < ni) {
      for (std::size_t u = 0; u < 3; ++u) f.shift[3 * i + u] = m.shift[u * ni + i];
    }
  }
  return assemble(m.n_cells, m.x, m.y, m.z, std::move(f));
}

}  // namespace mx
