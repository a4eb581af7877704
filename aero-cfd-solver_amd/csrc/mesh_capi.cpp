// mesh_capi.cpp -- extern "C" export of the host-side O-grid generator (aero_mesh_generate).
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
