// mesh_capi.cpp -- extern "C" export of the host-side O-grid generator (aero_mesh_generate).
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <string>

#include "aero_hip.h"
#include "mesh_gen.hpp"

namespace {

template <class T> T* dup(const std::vector<T>& v) {
  T* p = static_cast<T*>(std::malloc(std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (p && !v.empty()) std::memcpy(p, v.data(), v.size() * sizeof(T));
  return p;
}

template <class R> void fill(const aero_mesh::Mesh<R>& m, aero_mesh_arrays* o) {
  o->precision = static_cast<int32_t>(sizeof(R));
  o->num_cells = m.num_cells;
  o->num_faces = m.num_faces;
  o->num_points = static_cast<int32_t>(m.points.size() / 3);
  o->face_owner = dup(m.owner);
  o->face_neighbor = dup(m.neighbor);
  o->face_bc = dup(m.bc);
  o->face_vertices = dup(m.face_v);
  o->face_normal = dup(m.normal);
  o->face_center = dup(m.face_center);
  o->face_area = dup(m.area);
  o->cell_volume = dup(m.volume);
  o->cell_center = dup(m.cell_center);
  o->points = dup(m.points);
  o->cell_offsets = dup(m.cell_offsets);
  o->cell_connectivity = dup(m.cell_conn);
  o->wall_start = m.wall_start;
  o->wall_count = m.wall_count;
  o->farfield_start = m.far_start;
  o->farfield_count = m.far_count;
}

}  // namespace

extern "C" int aero_mesh_generate(const aero_mesh_gen_config* c, int32_t precision, aero_mesh_arrays* out, char* err,
                                  size_t errlen) {
  if (!c || !out || (precision != 4 && precision != 8)) {
    if (err && errlen) std::snprintf(err, errlen, "aero_mesh_generate: bad arguments");
    return AERO_EINVAL;
  }
  std::memset(out, 0, sizeof(*out));
  aero_mesh::GenConfig g;
  g.naca = std::string(c->naca_code, strnlen(c->naca_code, sizeof(c->naca_code)));
  g.chord = c->chord;
  g.n_theta = c->num_circumferential;
  g.n_radial = c->num_radial;
  g.farfield_radius = c->farfield_radius;
  g.radial_stretch = c->radial_stretch;
  g.cells = c->cell_type;
  g.quad_layers = c->quad_layers;
  g.jitter = c->jitter;
  g.seed = c->seed;
  g.rcm = c->renumber_rcm != 0;
  try {
    if (precision == 4)
      fill(aero_mesh::generate<float>(g), out);
    else
      fill(aero_mesh::generate<double>(g), out);
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return AERO_EINVAL;
  }
  return AERO_OK;
}

extern "C" void aero_mesh_free(aero_mesh_arrays* a) {
  if (!a) return;
  void* ptrs[] = {a->face_owner, a->face_neighbor, a->face_bc, a->face_vertices, a->face_normal, a->face_center,
                  a->face_area,  a->cell_volume,   a->cell_center, a->points,      a->cell_offsets, a->cell_connectivity};
  for (void* p : ptrs) std::free(p);
  std::memset(a, 0, sizeof(*a));
}

// ---- binary mesh file ------------------------------------------------------------------------------
// Layout: the 8-byte magic, int32 precision, num_cells, num_faces, num_points, connectivity length,
// wall_start, wall_count, farfield_start, farfield_count, then the arrays of aero_mesh_arrays in
// declaration order, the floating ones at the file's precision, all in the host's byte order (little
// endian on x86-64 and the GPU hosts).
namespace {

constexpr char kMeshMagic[8] = {'A', 'E', 'R', 'O', 'M', 'S', 'H', '1'};

struct MeshFileHead {
  int32_t precision, num_cells, num_faces, num_points, conn_len, wall_start, wall_count, far_start, far_count;
};

// Element counts of the twelve arrays in declaration order.
void mesh_counts(const MeshFileHead& h, size_t n[12]) {
  const size_t N = h.num_cells, F = h.num_faces, P = h.num_points;
  const size_t c[12] = {F, F, F, 2 * F, 3 * F, 3 * F, F, N, 3 * N, 3 * P, N, static_cast<size_t>(h.conn_len)};
  for (int i = 0; i < 12; ++i) n[i] = c[i];
}

bool mesh_is_float(int i) { return i >= 4 && i <= 9; }

void set_msg(char* err, size_t errlen, const std::string& m) {
  if (err && errlen) std::snprintf(err, errlen, "%s", m.c_str());
}

// Every index in range and every count consistent, so that a mesh read from a file can go to the
// device kernels without further checks.
std::string mesh_invalid(const MeshFileHead& h, const aero_mesh_arrays& a) {
  const int N = h.num_cells, F = h.num_faces, P = h.num_points;
  for (int f = 0; f < F; ++f) {
    if (a.face_owner[f] < 0 || a.face_owner[f] >= N) return "face owner out of range";
    if (a.face_neighbor[f] < -1 || a.face_neighbor[f] >= N) return "face neighbour out of range";
    if (a.face_bc[f] < 0 || a.face_bc[f] > 2) return "face boundary code out of range";
    if ((a.face_neighbor[f] < 0) != (a.face_bc[f] != AERO_BC_INTERIOR)) return "face boundary code mismatch";
    for (int k = 0; k < 2; ++k)
      if (a.face_vertices[2 * f + k] < 0 || a.face_vertices[2 * f + k] >= P) return "face vertex out of range";
  }
  int prev = 0;
  for (int c = 0; c < N; ++c) {
    if (a.cell_offsets[c] < prev) return "cell offsets decrease";
    prev = a.cell_offsets[c];
  }
  if (prev != h.conn_len) return "cell offsets do not end at the connectivity length";
  for (int i = 0; i < h.conn_len; ++i)
    if (a.cell_connectivity[i] < 0 || a.cell_connectivity[i] >= P) return "cell connectivity out of range";
  auto range_ok = [F](int s, int n) { return s >= 0 && n >= 0 && s <= F && n <= F - s; };
  if (!range_ok(h.wall_start, h.wall_count) || !range_ok(h.far_start, h.far_count)) return "patch range invalid";
  return "";
}

}  // namespace

extern "C" int aero_mesh_write(const aero_mesh_arrays* a, const char* path, char* err, size_t errlen) {
  if (!a || !path || (a->precision != 4 && a->precision != 8) || a->num_cells < 0 || a->num_faces < 0 ||
      a->num_points < 0) {
    set_msg(err, errlen, "aero_mesh_write: bad arguments");
    return AERO_EINVAL;
  }
  MeshFileHead h{a->precision, a->num_cells, a->num_faces, a->num_points,
                 a->num_cells > 0 ? a->cell_offsets[a->num_cells - 1] : 0, a->wall_start, a->wall_count,
                 a->farfield_start, a->farfield_count};
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_msg(err, errlen, std::string("aero_mesh_write: cannot open ") + path);
    return AERO_EINVAL;
  }
  size_t n[12];
  mesh_counts(h, n);
  const void* arrays[12] = {a->face_owner, a->face_neighbor, a->face_bc, a->face_vertices, a->face_normal,
                            a->face_center, a->face_area, a->cell_volume, a->cell_center, a->points,
                            a->cell_offsets, a->cell_connectivity};
  bool ok = std::fwrite(kMeshMagic, 1, 8, fp) == 8 && std::fwrite(&h, sizeof(h), 1, fp) == 1;
  for (int i = 0; ok && i < 12; ++i) {
    const size_t es = mesh_is_float(i) ? static_cast<size_t>(h.precision) : sizeof(int32_t);
    ok = n[i] == 0 || std::fwrite(arrays[i], es, n[i], fp) == n[i];
  }
  ok = (std::fclose(fp) == 0) && ok;
  if (!ok) {
    set_msg(err, errlen, std::string("aero_mesh_write: write failed: ") + path);
    return AERO_EINVAL;
  }
  return AERO_OK;
}

extern "C" int aero_mesh_read(const char* path, aero_mesh_arrays* out, char* err, size_t errlen) {
  if (!path || !out) {
    set_msg(err, errlen, "aero_mesh_read: null argument");
    return AERO_EINVAL;
  }
  std::memset(out, 0, sizeof(*out));
  std::FILE* fp = std::fopen(path, "rb");
  if (!fp) {
    set_msg(err, errlen, std::string("aero_mesh_read: cannot open ") + path);
    return AERO_EINVAL;
  }
  char magic[8];
  MeshFileHead h{};
  const bool head_ok = std::fread(magic, 1, 8, fp) == 8 && std::memcmp(magic, kMeshMagic, 8) == 0 &&
                       std::fread(&h, sizeof(h), 1, fp) == 1 && (h.precision == 4 || h.precision == 8) &&
                       h.num_cells >= 0 && h.num_faces >= 0 && h.num_points >= 0 && h.conn_len >= 0;
  if (!head_ok) {
    std::fclose(fp);
    set_msg(err, errlen, std::string("aero_mesh_read: not a mesh file (or an unsupported version): ") + path);
    return AERO_EINVAL;
  }
  size_t n[12];
  mesh_counts(h, n);
  void* arrays[12] = {};
  bool ok = true;
  for (int i = 0; ok && i < 12; ++i) {
    const size_t es = mesh_is_float(i) ? static_cast<size_t>(h.precision) : sizeof(int32_t);
    arrays[i] = std::malloc(std::max<size_t>(n[i], 1) * es);
    ok = arrays[i] && (n[i] == 0 || std::fread(arrays[i], es, n[i], fp) == n[i]);
  }
  char extra;
  ok = ok && std::fread(&extra, 1, 1, fp) == 0;  // nothing after the last array
  std::fclose(fp);
  out->precision = h.precision;
  out->num_cells = h.num_cells;
  out->num_faces = h.num_faces;
  out->num_points = h.num_points;
  out->face_owner = static_cast<int32_t*>(arrays[0]);
  out->face_neighbor = static_cast<int32_t*>(arrays[1]);
  out->face_bc = static_cast<int32_t*>(arrays[2]);
  out->face_vertices = static_cast<int32_t*>(arrays[3]);
  out->face_normal = arrays[4];
  out->face_center = arrays[5];
  out->face_area = arrays[6];
  out->cell_volume = arrays[7];
  out->cell_center = arrays[8];
  out->points = arrays[9];
  out->cell_offsets = static_cast<int32_t*>(arrays[10]);
  out->cell_connectivity = static_cast<int32_t*>(arrays[11]);
  out->wall_start = h.wall_start;
  out->wall_count = h.wall_count;
  out->farfield_start = h.far_start;
  out->farfield_count = h.far_count;
  std::string why = ok ? mesh_invalid(h, *out) : std::string("truncated or oversized file");
  if (!why.empty()) {
    aero_mesh_free(out);
    set_msg(err, errlen, std::string("aero_mesh_read: ") + why + ": " + path);
    return AERO_EINVAL;
  }
  return AERO_OK;
}

// ---- shape deformation ------------------------------------------------------------------------------
#include "mesh_deform.hpp"

extern "C" int aero_mesh_deform(const aero_mesh_arrays* base, const aero_shape_config* shape, const double* amp,
                                aero_mesh_arrays* out, aero_geometry_tangents* tan, char* err, size_t errlen) {
  if (!base || !shape || !amp || !out || base->precision != 8) {
    if (err && errlen) std::snprintf(err, errlen, "aero_mesh_deform: bad arguments (base must be precision 8)");
    return AERO_EINVAL;
  }
  std::memset(out, 0, sizeof(*out));
  try {
    using namespace aero_mesh;
    Topology t;
    t.num_cells = base->num_cells;
    t.num_faces = base->num_faces;
    t.num_points = base->num_points;
    t.face_v = base->face_vertices;
    t.cell_offsets = base->cell_offsets;
    t.cell_conn = base->cell_connectivity;
    const int N = t.num_cells, F = t.num_faces, NP = t.num_points;
    const double* pts = static_cast<const double*>(base->points);
    std::vector<double> px(NP), py(NP);
    for (int i = 0; i < NP; ++i) {
      px[i] = pts[3 * i];
      py[i] = pts[3 * i + 1];
    }
    const ShapeSetup s = shape_setup(t, px.data(), py.data(), static_cast<const double*>(base->face_normal),
                                     base->wall_start, base->wall_count, base->farfield_start, base->farfield_count,
                                     shape->bumps_per_surface, shape->h_min, shape->h_max);
    const int P = s.n_params;
    if (tan && tan->n_params != P) throw std::invalid_argument("tangents.n_params must be 2 * bumps_per_surface");
    const double k = shape->idw_exponent > 0.0 ? shape->idw_exponent : 3.0;
    const std::vector<double> ty = idw_tangents(t, px.data(), py.data(), s, k);
    std::vector<double> qx = px, qy = py;
    for (int i = 0; i < NP; ++i) {
      double d = 0.0;
      for (int p = 0; p < P; ++p) d += amp[p] * ty[static_cast<size_t>(p) * NP + i];
      qy[i] = py[i] + d;
    }
    Geometry<double> g;
    compute_geometry(t, qx, qy, &g);
    for (int c = 0; c < N; ++c)
      if (!(g.volume[c] > 0.0)) throw std::runtime_error("the deformation inverts cell " + std::to_string(c));
    Mesh<double> m;
    m.num_cells = N;
    m.num_faces = F;
    m.owner.assign(base->face_owner, base->face_owner + F);
    m.neighbor.assign(base->face_neighbor, base->face_neighbor + F);
    m.bc.assign(base->face_bc, base->face_bc + F);
    m.face_v.assign(base->face_vertices, base->face_vertices + 2 * static_cast<size_t>(F));
    m.normal = g.normal;
    m.face_center = g.face_center;
    m.area = g.area;
    m.volume = g.volume;
    m.cell_center = g.cell_center;
    m.points.resize(3 * static_cast<size_t>(NP));
    for (int i = 0; i < NP; ++i) {
      m.points[3 * i] = qx[i];
      m.points[3 * i + 1] = qy[i];
      m.points[3 * i + 2] = 0.0;
    }
    m.cell_offsets.assign(base->cell_offsets, base->cell_offsets + N);
    m.cell_conn.assign(base->cell_connectivity, base->cell_connectivity + (N ? base->cell_offsets[N - 1] : 0));
    m.wall_start = base->wall_start;
    m.wall_count = base->wall_count;
    m.far_start = base->farfield_start;
    m.far_count = base->farfield_count;
    if (tan) {
      std::vector<GDual> dx(NP), dy(NP);
      for (int p = 0; p < P; ++p) {
        for (int i = 0; i < NP; ++i) {
          dx[i] = GDual(qx[i], 0.0);
          dy[i] = GDual(qy[i], ty[static_cast<size_t>(p) * NP + i]);
        }
        Geometry<GDual> gd;
        compute_geometry(t, dx, dy, &gd);
        auto put = [&](double* dst, const std::vector<GDual>& src) {
          if (!dst) return;
          double* o = dst + static_cast<size_t>(p) * src.size();
          for (size_t i = 0; i < src.size(); ++i) o[i] = src[i].d;
        };
        put(tan->d_face_normal, gd.normal);
        put(tan->d_face_center, gd.face_center);
        put(tan->d_face_area, gd.area);
        put(tan->d_cell_center, gd.cell_center);
        put(tan->d_cell_volume, gd.volume);
        if (tan->d_points) {
          double* o = tan->d_points + static_cast<size_t>(p) * 3 * NP;
          for (int i = 0; i < NP; ++i) {
            o[3 * i] = 0.0;
            o[3 * i + 1] = ty[static_cast<size_t>(p) * NP + i];
            o[3 * i + 2] = 0.0;
          }
        }
      }
    }
    fill(m, out);
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return AERO_EINVAL;
  }
  return AERO_OK;
}
