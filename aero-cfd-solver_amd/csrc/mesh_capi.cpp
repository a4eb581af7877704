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
