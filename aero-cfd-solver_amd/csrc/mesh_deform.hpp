// mesh_deform.hpp -- Hicks-Henne surface bumps and inverse-distance-weighted (IDW) volume deformation of
// an O-grid, with the exact derivatives of the deformed geometry with respect to the bump amplitudes.
//
// The reference has no shape parameterisation.  Definitions used here:
//   * bump i of a surface: b_i(x) = sin^3(pi x^{m_i}), m_i = ln 0.5 / ln h_i, with x = (x - x_le) / (x_te - x_le)
//     in [0, 1] along the chord and h_i = linspace(h_min, h_max, n) the position of the bump's peak;
//   * parameter i (upper surface, i < n) or n + i (lower surface) moves the wall nodes of that surface by
//     a * b_i(x) in +y; a wall node is on the upper surface when its outward normal (minus the mean of
//     its wall faces' normals, which point into the body) has y >= 0;
//   * every other node moves by the IDW average of the boundary-node displacements,
//     d(q) = sum_j w_j d_j / sum_j w_j with w_j = |q - p_j|^-k over the wall nodes (displaced) and the
//     farfield nodes (fixed).
// IDW with fixed weights is linear in the amplitudes, so the node derivative for parameter i is the IDW
// average of bump i alone; the geometry derivatives follow by differentiating the geometry formulas of
// the generator (face normal, length and midpoint, triangle / polygon area and vertex-mean centre).
#pragma once

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <vector>

namespace aero_mesh {

// Host dual number for the geometry derivatives.
struct GDual {
  double v = 0.0, d = 0.0;
  GDual() = default;
  GDual(double x) : v(x), d(0.0) {}  // NOLINT
  GDual(double x, double dx) : v(x), d(dx) {}
};
inline GDual operator+(const GDual& a, const GDual& b) { return {a.v + b.v, a.d + b.d}; }
inline GDual operator-(const GDual& a, const GDual& b) { return {a.v - b.v, a.d - b.d}; }
inline GDual operator-(const GDual& a) { return {-a.v, -a.d}; }
inline GDual operator*(const GDual& a, const GDual& b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline GDual operator/(const GDual& a, const GDual& b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
inline bool operator<(const GDual& a, const GDual& b) { return a.v < b.v; }
inline GDual gsqrt(const GDual& a) {
  const double s = std::sqrt(a.v);
  return {s, s > 0.0 ? a.d / (2.0 * s) : 0.0};
}
inline double gsqrt(double a) { return std::sqrt(a); }
inline double value(double a) { return a; }
inline double value(const GDual& a) { return a.v; }

// Topology of a generated mesh (the arrays of aero_mesh_arrays).
struct Topology {
  int num_cells = 0, num_faces = 0, num_points = 0;
  const int* face_v = nullptr;        // [2F], in the owner's counter-clockwise order
  const int* cell_offsets = nullptr;  // [N] running end offsets
  const int* cell_conn = nullptr;     // counter-clockwise vertices
};

template <class T> struct Geometry {
  std::vector<T> normal, face_center, area, volume, cell_center;  // reference layouts (xyz, z = 0)
};

// The generator's geometry formulas (mesh_gen.hpp build_mesh) on given point coordinates.
template <class T>
void compute_geometry(const Topology& t, const std::vector<T>& px, const std::vector<T>& py, Geometry<T>* g) {
  const int N = t.num_cells, F = t.num_faces;
  g->volume.assign(N, T(0.0));
  g->cell_center.assign(3 * static_cast<size_t>(N), T(0.0));
  for (int c = 0; c < N; ++c) {
    const int b = c == 0 ? 0 : t.cell_offsets[c - 1];
    const int k = t.cell_offsets[c] - b;
    const int* cv = t.cell_conn + b;
    if (k == 3) {
      const T ax = px[cv[0]], ay = py[cv[0]], bx = px[cv[1]], by = py[cv[1]], dx = px[cv[2]], dy = py[cv[2]];
      const T b2 = (bx - ax) * (dy - ay) - (by - ay) * (dx - ax);
      g->volume[c] = 0.5 * (value(b2) < 0.0 ? -b2 : b2);
      g->cell_center[3 * c + 0] = (ax + bx + dx) / T(3.0);
      g->cell_center[3 * c + 1] = (ay + by + dy) / T(3.0);
    } else {
      T s(0.0), cx(0.0), cy(0.0);
      for (int i = 0; i < k; ++i) {
        const int p = cv[i], q = cv[(i + 1) % k];
        s = s + (px[p] * py[q] - px[q] * py[p]);
      }
      for (int i = 0; i < k; ++i) {
        cx = cx + px[cv[i]];
        cy = cy + py[cv[i]];
      }
      g->volume[c] = 0.5 * s;
      g->cell_center[3 * c + 0] = cx / T(static_cast<double>(k));
      g->cell_center[3 * c + 1] = cy / T(static_cast<double>(k));
    }
  }
  g->normal.assign(3 * static_cast<size_t>(F), T(0.0));
  g->face_center.assign(3 * static_cast<size_t>(F), T(0.0));
  g->area.assign(F, T(0.0));
  for (int f = 0; f < F; ++f) {
    const int v0 = t.face_v[2 * f], v1 = t.face_v[2 * f + 1];
    const T dx = px[v1] - px[v0];
    const T dy = py[v1] - py[v0];
    const T len = gsqrt(dx * dx + dy * dy);
    g->normal[3 * f + 0] = dy / len;
    g->normal[3 * f + 1] = -dx / len;
    g->area[f] = len;
    g->face_center[3 * f + 0] = 0.5 * (px[v0] + px[v1]);
    g->face_center[3 * f + 1] = 0.5 * (py[v0] + py[v1]);
  }
}

struct ShapeSetup {
  int n_params = 0;
  std::vector<int> sources;           // boundary nodes: wall nodes, then farfield nodes
  std::vector<double> src_dy;         // [n_params][n_sources] unit-amplitude y displacement of each source
  std::vector<int> src_index;         // node -> source index or -1
};

// Wall / farfield nodes and the unit bump displacements of the wall nodes.
inline ShapeSetup shape_setup(const Topology& t, const double* px, const double* py, const double* normal, int wall_start,
                              int wall_count, int far_start, int far_count, int n_per_surface, double h_min,
                              double h_max) {
  if (n_per_surface < 1) throw std::invalid_argument("bumps_per_surface must be >= 1");
  if (!(h_min > 0.0 && h_max < 1.0 && h_min <= h_max)) throw std::invalid_argument("need 0 < h_min <= h_max < 1");
  ShapeSetup s;
  s.n_params = 2 * n_per_surface;
  s.src_index.assign(t.num_points, -1);
  std::vector<double> ny_sum(t.num_points, 0.0);
  auto add = [&](int v) {
    if (s.src_index[v] < 0) {
      s.src_index[v] = static_cast<int>(s.sources.size());
      s.sources.push_back(v);
    }
  };
  for (int i = 0; i < wall_count; ++i) {
    const int f = wall_start + i;
    for (int e = 0; e < 2; ++e) {
      add(t.face_v[2 * f + e]);
      ny_sum[t.face_v[2 * f + e]] += normal[3 * f + 1];
    }
  }
  const int n_wall = static_cast<int>(s.sources.size());
  for (int i = 0; i < far_count; ++i) {
    const int f = far_start + i;
    add(t.face_v[2 * f]);
    add(t.face_v[2 * f + 1]);
  }
  if (n_wall == 0) throw std::invalid_argument("mesh has no wall patch");
  double xle = px[s.sources[0]], xte = xle;
  for (int j = 0; j < n_wall; ++j) {
    xle = std::min(xle, px[s.sources[j]]);
    xte = std::max(xte, px[s.sources[j]]);
  }
  const double chord = xte - xle;
  if (!(chord > 0.0)) throw std::invalid_argument("degenerate wall");
  const int ns = static_cast<int>(s.sources.size());
  s.src_dy.assign(static_cast<size_t>(s.n_params) * ns, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < n_per_surface; ++i) {
    const double h = n_per_surface == 1 ? h_min : h_min + (h_max - h_min) * i / (n_per_surface - 1);
    const double m = std::log(0.5) / std::log(h);
    for (int j = 0; j < n_wall; ++j) {
      const int v = s.sources[j];
      double x = (px[v] - xle) / chord;
      x = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
      const double sn = std::sin(pi * std::pow(x, m));
      const double b = sn * sn * sn;
      const bool upper = -ny_sum[v] >= 0.0;
      s.src_dy[static_cast<size_t>(upper ? i : n_per_surface + i) * ns + j] = b;
    }
  }
  (void)py;
  return s;
}

// Node derivatives: ty[p][node] = IDW average of parameter p's unit source displacements (y only).
inline std::vector<double> idw_tangents(const Topology& t, const double* px, const double* py, const ShapeSetup& s,
                                        double k) {
  const int P = s.n_params, np = t.num_points;
  const int ns = static_cast<int>(s.sources.size());
  std::vector<double> ty(static_cast<size_t>(P) * np, 0.0);
  std::vector<double> sx(ns), sy(ns);
  for (int j = 0; j < ns; ++j) {
    sx[j] = px[s.sources[j]];
    sy[j] = py[s.sources[j]];
  }
  const double half_k = 0.5 * k;
  auto work = [&](int lo, int hi) {
    std::vector<double> acc(P);
    for (int q = lo; q < hi; ++q) {
      const int si = s.src_index[q];
      if (si >= 0) {
        for (int p = 0; p < P; ++p) ty[static_cast<size_t>(p) * np + q] = s.src_dy[static_cast<size_t>(p) * ns + si];
        continue;
      }
      std::fill(acc.begin(), acc.end(), 0.0);
      double wsum = 0.0;
      for (int j = 0; j < ns; ++j) {
        const double dx = px[q] - sx[j], dy = py[q] - sy[j];
        const double d2 = dx * dx + dy * dy;
        const double w = 1.0 / std::pow(d2, half_k);
        wsum += w;
        for (int p = 0; p < P; ++p) acc[p] += w * s.src_dy[static_cast<size_t>(p) * ns + j];
      }
      for (int p = 0; p < P; ++p) ty[static_cast<size_t>(p) * np + q] = acc[p] / wsum;
    }
  };
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int chunk = (np + static_cast<int>(hw) - 1) / static_cast<int>(hw);
  std::vector<std::thread> th;
  for (unsigned w = 0; w < hw; ++w) {
    const int lo = static_cast<int>(w) * chunk, hi = std::min(np, lo + chunk);
    if (lo < hi) th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
  return ty;
}

}  // namespace aero_mesh
