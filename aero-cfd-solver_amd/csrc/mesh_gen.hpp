// mesh_gen.hpp -- host-side NACA-4 O-grid generator, templated on the scalar type.
//
// Follows make_airfoil_ogrid_mesh (src/cfd_core/mesh/airfoil_mesh.cpp:459-544) and its helpers
// (:56-135 curve resampling, :184-392 face build and ordering, :395-457 NACA-4 contour).  With
// R = float and triangle cells it reproduces the reference generator's arrays exactly (checked in
// tests/test_mesh_gen.py against golden mesh arrays).  Beyond the reference it builds quad and
// mixed (quad wall block + triangles) O-grids, applies seeded interior-node jitter and renumbers
// cells by reverse Cuthill-McKee before faces are discovered, for the large FP64 workloads.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace aero_mesh {

template <class R> constexpr R kc(float f, double d);
template <> constexpr float kc<float>(float f, double) { return f; }
template <> constexpr double kc<double>(float, double d) { return d; }
#define AM_K(x) ::aero_mesh::kc<R>(x##f, x)

enum CellKind : int { kTri = 0, kQuad = 1, kMixed = 2 };

struct GenConfig {
  std::string naca = "0012";
  double chord = 1.0;
  int n_theta = 160;
  int n_radial = 48;
  double farfield_radius = 15.0;
  double radial_stretch = 1.4;
  int cells = kTri;
  int quad_layers = 30;
  double jitter = 0.0;
  std::uint64_t seed = 12345;
  bool rcm = false;
};

template <class R> struct Mesh {
  int num_cells = 0, num_faces = 0;
  std::vector<int> owner, neighbor, bc, face_v;
  std::vector<R> normal, face_center, area, volume, cell_center, points;
  std::vector<int> cell_offsets, cell_conn;
  int wall_start = 0, wall_count = 0, far_start = 0, far_count = 0;
};

template <class R> using P2 = std::array<R, 2>;

template <class R> R signed_area(const std::vector<P2<R>>& p) {  // airfoil_mesh.cpp:56-69
  if (p.size() < 3) return R(0);
  double a2 = 0.0;
  const std::size_t n = p.size();
  for (std::size_t i = 0; i < n; ++i) {
    const std::size_t j = (i + 1) % n;
    a2 += static_cast<double>(p[i][0]) * static_cast<double>(p[j][1]) -
          static_cast<double>(p[j][0]) * static_cast<double>(p[i][1]);
  }
  return static_cast<R>(0.5 * a2);
}

template <class R> R dist2(const P2<R>& a, const P2<R>& b) {
  const R dx = a[0] - b[0];
  const R dy = a[1] - b[1];
  return dx * dx + dy * dy;
}

template <class R> std::vector<P2<R>> dedup(const std::vector<P2<R>>& in) {  // airfoil_mesh.cpp:77-91
  std::vector<P2<R>> u;
  u.reserve(in.size());
  for (const auto& p : in) {
    if (!u.empty() && dist2(u.back(), p) < AM_K(1.0e-12)) continue;
    u.push_back(p);
  }
  if (u.size() >= 2 && dist2(u.front(), u.back()) < AM_K(1.0e-12)) u.pop_back();
  return u;
}

template <class R> std::vector<P2<R>> resample(const std::vector<P2<R>>& in, int count) {  // :93-135
  if (count < 8) throw std::invalid_argument("Target curve resolution must be >= 8.");
  const auto p = dedup(in);
  if (p.size() < 8) throw std::invalid_argument("Airfoil contour has too few points.");
  const std::size_t n = p.size();
  std::vector<R> cum(n + 1, R(0));
  for (std::size_t i = 0; i < n; ++i) cum[i + 1] = cum[i] + std::sqrt(dist2(p[i], p[(i + 1) % n]));
  const R total = cum.back();
  if (total < AM_K(1.0e-7)) throw std::invalid_argument("Airfoil contour has near-zero length.");
  std::vector<P2<R>> out;
  out.reserve(static_cast<std::size_t>(count));
  std::size_t seg = 0;
  for (int k = 0; k < count; ++k) {
    const R ts = total * static_cast<R>(k) / static_cast<R>(count);
    while (seg + 1 < cum.size() && cum[seg + 1] < ts) ++seg;
    const std::size_t i0 = seg % n, i1 = (seg + 1) % n;
    const R s0 = cum[seg], s1 = cum[seg + 1];
    const R den = std::max(s1 - s0, AM_K(1.0e-7));
    const R t = (ts - s0) / den;
    out.push_back({p[i0][0] + t * (p[i1][0] - p[i0][0]), p[i0][1] + t * (p[i1][1] - p[i0][1])});
  }
  return out;
}

template <class R> std::vector<P2<R>> naca4(const std::string& code, int n, R chord) {  // :395-457
  if (chord <= R(0)) throw std::invalid_argument("Airfoil chord must be positive.");
  std::string d;
  for (char ch : code)
    if (ch >= '0' && ch <= '9') d.push_back(ch);
  if (d.size() < 4) throw std::invalid_argument("NACA code must contain four digits.");
  const R pi = kc<R>(3.14159265358979323846f, 3.14159265358979323846);
  const R m = static_cast<R>(d[0] - '0') / AM_K(100.0);
  const R p = static_cast<R>(d[1] - '0') / AM_K(10.0);
  const R t = static_cast<R>(10 * (d[2] - '0') + (d[3] - '0')) / AM_K(100.0);
  const int pps = std::max(16, n / 2 + 1);
  std::vector<P2<R>> up(static_cast<std::size_t>(pps)), lo(static_cast<std::size_t>(pps));
  for (int i = 0; i < pps; ++i) {
    const R beta = pi * static_cast<R>(i) / static_cast<R>(pps - 1);
    const R x = AM_K(0.5) * (AM_K(1.0) - std::cos(beta));
    const R sx = std::sqrt(std::max(x, R(0)));
    const R yt = AM_K(5.0) * t *
                 (AM_K(0.2969) * sx - AM_K(0.1260) * x - AM_K(0.3516) * x * x + AM_K(0.2843) * x * x * x -
                  AM_K(0.1036) * x * x * x * x);
    R yc = R(0), dyc = R(0);
    if (m > R(0) && p > R(0) && p < AM_K(1.0)) {
      if (x < p) {
        yc = m / (p * p) * (AM_K(2.0) * p * x - x * x);
        dyc = AM_K(2.0) * m / (p * p) * (p - x);
      } else {
        const R q = AM_K(1.0) - p;
        yc = m / (q * q) * (AM_K(1.0) - AM_K(2.0) * p + AM_K(2.0) * p * x - x * x);
        dyc = AM_K(2.0) * m / (q * q) * (p - x);
      }
    }
    const R th = std::atan(dyc);
    const R xu = x - yt * std::sin(th);
    const R yu = yc + yt * std::cos(th);
    const R xl = x + yt * std::sin(th);
    const R yl = yc - yt * std::cos(th);
    up[i] = {xu * chord, yu * chord};
    lo[i] = {xl * chord, yl * chord};
  }
  std::vector<P2<R>> c;
  c.reserve(static_cast<std::size_t>(2 * pps - 1));
  for (int i = pps - 1; i >= 0; --i) c.push_back(up[i]);
  for (int i = 1; i < pps; ++i) c.push_back(lo[i]);
  if (signed_area(c) < R(0)) std::reverse(c.begin(), c.end());
  return resample(c, std::max(32, n));
}

// splitmix64 -> uniform in [-1, 1)
struct Rng {
  std::uint64_t s;
  std::uint64_t next() {
    std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double sym() { return (static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0; }
};

// Reverse Cuthill-McKee order of a graph given as CSR adjacency; returns new->old.
inline std::vector<int> rcm_order(int n, const std::vector<int>& off, const std::vector<int>& adj) {
  std::vector<int> deg(static_cast<std::size_t>(n));
  for (int i = 0; i < n; ++i) deg[i] = off[i + 1] - off[i];
  std::vector<int> order;
  order.reserve(static_cast<std::size_t>(n));
  std::vector<char> seen(static_cast<std::size_t>(n), 0);
  std::vector<int> level(static_cast<std::size_t>(n), -1);
  auto bfs_last_level = [&](int root, int* depth) {
    std::vector<int> q{root};
    std::fill(level.begin(), level.end(), -1);
    level[root] = 0;
    for (std::size_t h = 0; h < q.size(); ++h) {
      const int v = q[h];
      for (int e = off[v]; e < off[v + 1]; ++e)
        if (level[adj[e]] < 0) {
          level[adj[e]] = level[v] + 1;
          q.push_back(adj[e]);
        }
    }
    *depth = level[q.back()];
    int best = q.back();
    for (int v : q)
      if (level[v] == *depth && deg[v] < deg[best]) best = v;
    return best;
  };
  for (int start = 0; start < n; ++start) {
    if (seen[start]) continue;
    int root = start;
    for (int i = 0; i < n; ++i)  // lowest-degree unvisited vertex of this component is found by BFS below
      if (!seen[i] && deg[i] < deg[root] && i >= start) root = i;
    int d0 = -1, d1 = 0;
    root = bfs_last_level(root, &d1);  // pseudo-peripheral root (George-Liu)
    for (int guard = 0; guard < 8 && d1 > d0; ++guard) {
      d0 = d1;
      const int cand = bfs_last_level(root, &d1);
      if (d1 > d0) root = cand;
    }
    const std::size_t base = order.size();
    order.push_back(root);
    seen[root] = 1;
    std::vector<int> nb;
    for (std::size_t h = base; h < order.size(); ++h) {
      const int v = order[h];
      nb.clear();
      for (int e = off[v]; e < off[v + 1]; ++e)
        if (!seen[adj[e]]) nb.push_back(adj[e]);
      std::sort(nb.begin(), nb.end(), [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      for (int w : nb) {
        seen[w] = 1;
        order.push_back(w);
      }
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Face build and ordering for polygon cells (airfoil_mesh.cpp:184-392, generalised from triangles).
template <class R>
Mesh<R> build_mesh(const std::vector<P2<R>>& pts, std::vector<std::vector<int>> cells,
                   const std::map<std::pair<int, int>, int>& patch_of_edge /* 1 wall, 2 farfield */) {
  if (pts.empty() || cells.empty()) throw std::invalid_argument("Cannot build mesh from empty point/cell arrays.");
  Mesh<R> m;
  m.num_cells = static_cast<int>(cells.size());
  m.volume.assign(static_cast<std::size_t>(m.num_cells), R(0));
  m.cell_center.assign(static_cast<std::size_t>(m.num_cells) * 3, R(0));
  for (int c = 0; c < m.num_cells; ++c) {
    auto& cv = cells[c];
    if (cv.size() == 3) {  // the reference's triangle path, :191-224
      const auto& a = pts[cv[0]];
      const auto& b = pts[cv[1]];
      const auto& d = pts[cv[2]];
      const R a2 = (b[0] - a[0]) * (d[1] - a[1]) - (b[1] - a[1]) * (d[0] - a[0]);
      if (a2 < R(0)) std::swap(cv[1], cv[2]);
      const auto& A = pts[cv[0]];
      const auto& B = pts[cv[1]];
      const auto& D = pts[cv[2]];
      const R b2 = (B[0] - A[0]) * (D[1] - A[1]) - (B[1] - A[1]) * (D[0] - A[0]);
      const R area = AM_K(0.5) * std::abs(b2);
      if (area < (sizeof(R) == 4 ? AM_K(1.0e-7) : AM_K(1.0e-20))) throw std::runtime_error("Degenerate triangle detected while building airfoil mesh.");
      m.volume[c] = area;
      m.cell_center[3 * c + 0] = (A[0] + B[0] + D[0]) / AM_K(3.0);
      m.cell_center[3 * c + 1] = (A[1] + B[1] + D[1]) / AM_K(3.0);
    } else {  // general polygon: shoelace area, vertex-mean centre
      R s = R(0);
      const std::size_t k = cv.size();
      for (std::size_t i = 0; i < k; ++i) {
        const auto& p = pts[cv[i]];
        const auto& q = pts[cv[(i + 1) % k]];
        s += p[0] * q[1] - q[0] * p[1];
      }
      if (s < R(0)) {
        std::reverse(cv.begin(), cv.end());
        s = -s;
      }
      const R area = AM_K(0.5) * s;
      if (!(area > R(0))) throw std::runtime_error("Degenerate polygon detected while building airfoil mesh.");
      R cx = R(0), cy = R(0);
      for (int v : cv) {
        cx += pts[v][0];
        cy += pts[v][1];
      }
      m.volume[c] = area;
      m.cell_center[3 * c + 0] = cx / static_cast<R>(k);
      m.cell_center[3 * c + 1] = cy / static_cast<R>(k);
    }
  }
  struct FB {
    int o = -1, n = -1, v0 = -1, v1 = -1;
    R nx = 0, ny = 0, a = 0;
  };
  std::map<std::pair<int, int>, int> e2f;
  std::vector<FB> faces;
  faces.reserve(static_cast<std::size_t>(m.num_cells) * 2);
  for (int c = 0; c < m.num_cells; ++c) {
    const auto& cv = cells[c];
    const std::size_t k = cv.size();
    for (std::size_t i = 0; i < k; ++i) {
      const int v0 = cv[i], v1 = cv[(i + 1) % k];
      const std::pair<int, int> key = v0 < v1 ? std::make_pair(v0, v1) : std::make_pair(v1, v0);
      const auto it = e2f.find(key);
      if (it == e2f.end()) {
        FB f;
        f.o = c;
        f.v0 = v0;
        f.v1 = v1;
        const R dx = pts[v1][0] - pts[v0][0];
        const R dy = pts[v1][1] - pts[v0][1];
        const R len = std::sqrt(dx * dx + dy * dy);
        if (len < (sizeof(R) == 4 ? AM_K(1.0e-7) : AM_K(1.0e-12)))
          throw std::runtime_error("Degenerate face detected while building airfoil mesh.");
        f.nx = dy / len;
        f.ny = -dx / len;
        f.a = len;
        e2f.emplace(key, static_cast<int>(faces.size()));
        faces.push_back(f);
      } else {
        faces[it->second].n = c;
      }
    }
  }
  std::vector<int> interior, wall, far, other;
  interior.reserve(faces.size());
  for (int f = 0; f < static_cast<int>(faces.size()); ++f) {
    if (faces[f].n >= 0) {
      interior.push_back(f);
      continue;
    }
    const std::pair<int, int> key =
      faces[f].v0 < faces[f].v1 ? std::make_pair(faces[f].v0, faces[f].v1) : std::make_pair(faces[f].v1, faces[f].v0);
    const auto it = patch_of_edge.find(key);
    if (it == patch_of_edge.end())
      other.push_back(f);  // "boundary" patch: treated as farfield (mesh_geometry.cpp:27-30)
    else if (it->second == 1)
      wall.push_back(f);
    else
      far.push_back(f);
  }
  std::vector<int> order;
  order.reserve(faces.size());
  order.insert(order.end(), interior.begin(), interior.end());
  order.insert(order.end(), wall.begin(), wall.end());
  order.insert(order.end(), far.begin(), far.end());
  order.insert(order.end(), other.begin(), other.end());
  m.num_faces = static_cast<int>(order.size());
  m.owner.resize(m.num_faces);
  m.neighbor.resize(m.num_faces);
  m.bc.resize(m.num_faces);
  m.face_v.resize(static_cast<std::size_t>(m.num_faces) * 2);
  m.normal.assign(static_cast<std::size_t>(m.num_faces) * 3, R(0));
  m.face_center.assign(static_cast<std::size_t>(m.num_faces) * 3, R(0));
  m.area.resize(m.num_faces);
  for (int nf = 0; nf < m.num_faces; ++nf) {
    const FB& s = faces[order[nf]];
    m.owner[nf] = s.o;
    m.neighbor[nf] = s.n;
    m.face_v[2 * nf + 0] = s.v0;
    m.face_v[2 * nf + 1] = s.v1;
    m.normal[3 * nf + 0] = s.nx;
    m.normal[3 * nf + 1] = s.ny;
    m.area[nf] = s.a;
    m.face_center[3 * nf + 0] = AM_K(0.5) * (pts[s.v0][0] + pts[s.v1][0]);
    m.face_center[3 * nf + 1] = AM_K(0.5) * (pts[s.v0][1] + pts[s.v1][1]);
  }
  const int ni = static_cast<int>(interior.size());
  m.wall_start = ni;
  m.wall_count = static_cast<int>(wall.size());
  m.far_start = ni + m.wall_count;
  m.far_count = static_cast<int>(far.size());
  for (int f = 0; f < m.num_faces; ++f) m.bc[f] = f < ni ? 0 : (f < m.far_start ? 2 : 1);
  m.points.reserve(pts.size() * 3);
  for (const auto& p : pts) {
    m.points.push_back(p[0]);
    m.points.push_back(p[1]);
    m.points.push_back(R(0));
  }
  int run = 0;
  for (const auto& cv : cells) {
    for (int v : cv) m.cell_conn.push_back(v);
    run += static_cast<int>(cv.size());
    m.cell_offsets.push_back(run);
  }
  return m;
}

template <class R> Mesh<R> generate(const GenConfig& g) {  // make_airfoil_ogrid_mesh, :459-544
  if (g.n_theta < 32) throw std::invalid_argument("num_circumferential must be >= 32.");
  if (g.n_radial < 4) throw std::invalid_argument("num_radial must be >= 4.");
  const R chord = static_cast<R>(g.chord);
  std::vector<P2<R>> contour = naca4<R>(g.naca, g.n_theta, chord);
  if (signed_area(contour) < R(0)) std::reverse(contour.begin(), contour.end());
  P2<R> ctr{R(0), R(0)};
  for (const auto& p : contour) {
    ctr[0] += p[0];
    ctr[1] += p[1];
  }
  ctr[0] /= static_cast<R>(contour.size());
  ctr[1] /= static_cast<R>(contour.size());
  const R ffr = std::max(static_cast<R>(g.farfield_radius), AM_K(2.5)) * chord;
  const R stretch = std::max(static_cast<R>(g.radial_stretch), AM_K(1.0));
  const int nt = g.n_theta, nr = g.n_radial;
  auto pid = [nt](int i, int j) { return j * nt + i; };
  std::vector<P2<R>> pts;
  pts.reserve(static_cast<std::size_t>((nr + 1) * nt));
  for (int j = 0; j <= nr; ++j) {
    const R eta = static_cast<R>(j) / static_cast<R>(nr);
    R sigma = eta;
    if (stretch > AM_K(1.0) + AM_K(1.0e-4)) sigma = (std::pow(stretch, eta) - AM_K(1.0)) / (stretch - AM_K(1.0));
    for (int i = 0; i < nt; ++i) {
      const auto& in = contour[static_cast<std::size_t>(i)];
      R dx = in[0] - ctr[0];
      R dy = in[1] - ctr[1];
      R rad = std::sqrt(dx * dx + dy * dy);
      if (rad < AM_K(1.0e-7)) {
        dx = AM_K(1.0);
        dy = R(0);
        rad = AM_K(1.0);
      }
      const R ir = AM_K(1.0) / rad;
      const R ox = ctr[0] + dx * ir * ffr;
      const R oy = ctr[1] + dy * ir * ffr;
      pts.push_back({in[0] + sigma * (ox - in[0]), in[1] + sigma * (oy - in[1])});
    }
  }
  std::vector<std::vector<int>> cells;
  cells.reserve(static_cast<std::size_t>(2 * nt * nr));
  for (int j = 0; j < nr; ++j)
    for (int i = 0; i < nt; ++i) {
      const int i1 = (i + 1) % nt;
      const int a = pid(i, j), b = pid(i1, j), c = pid(i1, j + 1), d = pid(i, j + 1);
      const bool quad = g.cells == kQuad || (g.cells == kMixed && j < g.quad_layers);
      if (quad) {
        cells.push_back({a, b, c, d});
      } else {
        cells.push_back({a, b, c});
        cells.push_back({a, c, d});
      }
    }
  if (g.jitter > 0.0) {  // interior nodes only; displacements from the unjittered grid
    Rng rng{g.seed};
    std::vector<P2<R>> moved = pts;
    for (int j = 1; j < nr; ++j)
      for (int i = 0; i < nt; ++i) {
        const auto& pm = pts[pid((i + nt - 1) % nt, j)];
        const auto& pp = pts[pid((i + 1) % nt, j)];
        const auto& qm = pts[pid(i, j - 1)];
        const auto& qp = pts[pid(i, j + 1)];
        const R ht = std::min(std::sqrt(dist2(pts[pid(i, j)], pm)), std::sqrt(dist2(pts[pid(i, j)], pp)));
        const R hr = std::min(std::sqrt(dist2(pts[pid(i, j)], qm)), std::sqrt(dist2(pts[pid(i, j)], qp)));
        const R tx = (pp[0] - pm[0]), ty = (pp[1] - pm[1]);
        const R tl = std::sqrt(tx * tx + ty * ty);
        const R rx = (qp[0] - qm[0]), ry = (qp[1] - qm[1]);
        const R rl = std::sqrt(rx * rx + ry * ry);
        const R a = static_cast<R>(g.jitter * rng.sym()) * ht;
        const R b = static_cast<R>(g.jitter * rng.sym()) * hr;
        moved[pid(i, j)][0] += a * tx / tl + b * rx / rl;
        moved[pid(i, j)][1] += a * ty / tl + b * ry / rl;
      }
    // Thin trailing-edge wake cells can be narrower than 10% of their edge lengths: a node keeps its
    // displacement only if every cell around it keeps at least half its unjittered signed area.
    auto sarea = [](const std::vector<P2<R>>& p, const std::vector<int>& cv) {
      R s = R(0);
      for (std::size_t q = 0; q < cv.size(); ++q) {
        const auto& a = p[cv[q]];
        const auto& b = p[cv[(q + 1) % cv.size()]];
        s += a[0] * b[1] - b[0] * a[1];
      }
      return s;
    };
    for (int pass = 0; pass < 16; ++pass) {
      bool changed = false;
      for (const auto& cv : cells) {
        const R a0 = sarea(pts, cv), a1 = sarea(moved, cv);
        if (!(a1 / a0 >= AM_K(0.5))) {
          for (int v : cv)
            if (moved[v][0] != pts[v][0] || moved[v][1] != pts[v][1]) {
              moved[v] = pts[v];
              changed = true;
            }
        }
      }
      if (!changed) break;
    }
    pts.swap(moved);
  }
  if (g.rcm) {
    std::map<std::pair<int, int>, std::vector<int>> ec;
    for (int c = 0; c < static_cast<int>(cells.size()); ++c) {
      const auto& cv = cells[c];
      for (std::size_t i = 0; i < cv.size(); ++i) {
        const int v0 = cv[i], v1 = cv[(i + 1) % cv.size()];
        ec[v0 < v1 ? std::make_pair(v0, v1) : std::make_pair(v1, v0)].push_back(c);
      }
    }
    const int n = static_cast<int>(cells.size());
    std::vector<std::vector<int>> nbr(static_cast<std::size_t>(n));
    for (const auto& kv : ec)
      if (kv.second.size() == 2) {
        nbr[kv.second[0]].push_back(kv.second[1]);
        nbr[kv.second[1]].push_back(kv.second[0]);
      }
    std::vector<int> off(static_cast<std::size_t>(n) + 1, 0), adj;
    for (int i = 0; i < n; ++i) {
      off[i + 1] = off[i] + static_cast<int>(nbr[i].size());
      adj.insert(adj.end(), nbr[i].begin(), nbr[i].end());
    }
    const std::vector<int> order = rcm_order(n, off, adj);
    std::vector<std::vector<int>> re(static_cast<std::size_t>(n));
    for (int i = 0; i < n; ++i) re[i] = cells[order[i]];
    cells.swap(re);
  }
  std::map<std::pair<int, int>, int> patch;
  for (int i = 0; i < nt; ++i) {
    const int i1 = (i + 1) % nt;
    const int a = pid(i, 0), b = pid(i1, 0);
    patch.emplace(a < b ? std::make_pair(a, b) : std::make_pair(b, a), 1);
    const int c = pid(i, nr), d = pid(i1, nr);
    patch.emplace(c < d ? std::make_pair(c, d) : std::make_pair(d, c), 2);
  }
  return build_mesh<R>(pts, std::move(cells), patch);
}

}  // namespace aero_mesh
