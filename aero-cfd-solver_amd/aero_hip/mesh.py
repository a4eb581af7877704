"""NACA-4 O-grid meshes from the native generator (aero_mesh_generate), as numpy arrays.

The arrays follow the reference UnstructuredMesh layout (src/cfd_core/include/cfd_core/mesh.hpp:14-36):
xyz-interleaved normals and centres, interior faces first, then the "wall" and "farfield" patches.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Mesh:
    num_cells: int
    num_faces: int
    owner: np.ndarray
    neighbor: np.ndarray
    bc: np.ndarray
    face_vertices: np.ndarray
    normal: np.ndarray        # (F, 3)
    face_center: np.ndarray   # (F, 3)
    area: np.ndarray          # (F,)
    volume: np.ndarray        # (N,)
    cell_center: np.ndarray   # (N, 3)
    points: np.ndarray        # (P, 3)
    cell_offsets: np.ndarray
    cell_connectivity: np.ndarray
    wall_start: int
    wall_count: int
    farfield_start: int
    farfield_count: int
    meta: dict = field(default_factory=dict)

    def astype(self, dtype) -> "Mesh":
        f = lambda a: np.ascontiguousarray(a, dtype=dtype)
        return Mesh(self.num_cells, self.num_faces, self.owner, self.neighbor, self.bc, self.face_vertices,
                    f(self.normal), f(self.face_center), f(self.area), f(self.volume), f(self.cell_center),
                    f(self.points), self.cell_offsets, self.cell_connectivity, self.wall_start, self.wall_count,
                    self.farfield_start, self.farfield_count, dict(self.meta))

    def desc(self) -> "_lib.MeshDesc":
        """aero_mesh_desc over this mesh's FP64 arrays (keep the Mesh alive while the desc is used)."""
        m = self if self.area.dtype == np.float64 else self.astype(np.float64)
        self._keep = m
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        return _lib.MeshDesc(m.num_cells, m.num_faces, P(m.owner, C.c_int32), P(m.neighbor, C.c_int32),
                             P(m.bc, C.c_int32), P(m.normal, C.c_double), P(m.face_center, C.c_double),
                             P(m.area, C.c_double), P(m.volume, C.c_double), P(m.cell_center, C.c_double),
                             m.wall_start, m.wall_count)


def generate_mesh(naca: str = "0012", num_circumferential: int = 160, num_radial: int = 48,
                  farfield_radius: float = 15.0, radial_stretch: float = 1.4, chord: float = 1.0,
                  cells: str = "tri", quad_layers: int = 30, jitter: float = 0.0, seed: int = 12345,
                  rcm: bool = False, precision: int = 8) -> Mesh:
    """Generate an O-grid.  Defaults are AirfoilMeshConfig's (mesh/airfoil_mesh.hpp:15-24).

    cells="tri" with precision=4 reproduces make_airfoil_ogrid_mesh bit for bit."""
    lib = _lib.lib()
    cfg = _lib.MeshGenConfig(naca.encode()[:7], chord, num_circumferential, num_radial, farfield_radius,
                             radial_stretch, _lib.CELLS[cells], quad_layers, jitter, seed, 1 if rcm else 0)
    out = _lib.MeshArrays()
    err = _lib.errbuf()
    _lib.check(lib.aero_mesh_generate(C.byref(cfg), precision, C.byref(out), err, len(err)), err, "aero_mesh_generate")
    try:
        m = _from_arrays(out, dict(naca=naca, num_circumferential=num_circumferential, num_radial=num_radial,
                                   farfield_radius=farfield_radius, radial_stretch=radial_stretch, chord=chord,
                                   cells=cells, quad_layers=quad_layers, jitter=jitter, seed=seed, rcm=rcm))
    finally:
        lib.aero_mesh_free(C.byref(out))
    return m


def _from_arrays(out: "_lib.MeshArrays", meta: dict) -> Mesh:
    ft = np.float32 if out.precision == 4 else np.float64
    ct = C.c_float if out.precision == 4 else C.c_double
    N, F, P = out.num_cells, out.num_faces, out.num_points

    def ia(p, n):
        return np.ctypeslib.as_array(p, shape=(n,)).copy()

    def fa(p, n):
        return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(n,)).astype(ft, copy=True)

    offs = ia(out.cell_offsets, N)
    return Mesh(N, F, ia(out.face_owner, F), ia(out.face_neighbor, F), ia(out.face_bc, F),
                ia(out.face_vertices, 2 * F).reshape(F, 2), fa(out.face_normal, 3 * F).reshape(F, 3),
                fa(out.face_center, 3 * F).reshape(F, 3), fa(out.face_area, F), fa(out.cell_volume, N),
                fa(out.cell_center, 3 * N).reshape(N, 3), fa(out.points, 3 * P).reshape(P, 3), offs,
                ia(out.cell_connectivity, int(offs[-1]) if N else 0), out.wall_start, out.wall_count,
                out.farfield_start, out.farfield_count, meta)


def _to_arrays(m: Mesh, precision: int = 8):
    """aero_mesh_arrays over a mesh's arrays at `precision` (4 or 8) (returns the struct and the arrays to
    keep alive)."""
    ft = np.float32 if precision == 4 else np.float64
    keep = [np.ascontiguousarray(a, dtype=np.int32) for a in (m.owner, m.neighbor, m.bc, m.face_vertices)]
    keep += [np.ascontiguousarray(a, dtype=ft) for a in (m.normal, m.face_center, m.area, m.volume,
                                                          m.cell_center, m.points)]
    keep += [np.ascontiguousarray(a, dtype=np.int32) for a in (m.cell_offsets, m.cell_connectivity)]
    PI = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    arr = _lib.MeshArrays(precision, m.num_cells, m.num_faces, len(m.points), PI(keep[0]), PI(keep[1]), PI(keep[2]),
                          PI(keep[3]), vp(keep[4]), vp(keep[5]), vp(keep[6]), vp(keep[7]), vp(keep[8]), vp(keep[9]),
                          PI(keep[10]), PI(keep[11]), m.wall_start, m.wall_count, m.farfield_start, m.farfield_count)
    return arr, keep


def deform_mesh(mesh: Mesh, amplitudes, bumps_per_surface: int = 10, h_min: float = 0.05, h_max: float = 0.95,
                idw_exponent: float = 3.0, tangents: bool = True):
    """Hicks-Henne bumps + IDW deformation (aero_mesh_deform): returns (deformed Mesh, tangents) where
    tangents holds d(points, normal, face_center, area, cell_center, volume)/d(amplitude_i), shaped
    [2 * bumps_per_surface, ...] like the mesh arrays, or None."""
    if mesh.area.dtype != np.float64:
        mesh = mesh.astype(np.float64)
    P = 2 * bumps_per_surface
    amp = np.ascontiguousarray(amplitudes, dtype=np.float64)
    if amp.shape != (P,):
        raise ValueError(f"need {P} amplitudes")
    base, keep = _to_arrays(mesh)
    shape = _lib.ShapeConfig(bumps_per_surface, h_min, h_max, idw_exponent)
    tan = None
    tc = None
    if tangents:
        N, F, NP = mesh.num_cells, mesh.num_faces, len(mesh.points)
        tan = dict(points=np.zeros((P, NP, 3)), normal=np.zeros((P, F, 3)), face_center=np.zeros((P, F, 3)),
                   area=np.zeros((P, F)), cell_center=np.zeros((P, N, 3)), volume=np.zeros((P, N)))
        pd = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        tc = _lib.GeometryTangents(P, pd(tan["points"]), pd(tan["normal"]), pd(tan["face_center"]), pd(tan["area"]),
                                   pd(tan["cell_center"]), pd(tan["volume"]))
    out = _lib.MeshArrays()
    lib = _lib.lib()
    err = _lib.errbuf()
    _lib.check(lib.aero_mesh_deform(C.byref(base), C.byref(shape), amp.ctypes.data_as(C.POINTER(C.c_double)),
                                    C.byref(out), C.byref(tc) if tc is not None else None, err, len(err)),
               err, "aero_mesh_deform")
    try:
        m = _from_arrays(out, dict(mesh.meta, shape=dict(amplitudes=amp.tolist(), bumps_per_surface=bumps_per_surface,
                                                         h_min=h_min, h_max=h_max, idw_exponent=idw_exponent)))
    finally:
        lib.aero_mesh_free(C.byref(out))
    del keep
    return m, tan


def save_mesh(mesh: Mesh, path) -> None:
    """Write `mesh` to a binary mesh file (aero_mesh_write), at the mesh's own precision."""
    precision = 4 if np.asarray(mesh.area).dtype == np.float32 else 8
    arr, _keep = _to_arrays(mesh, precision)
    err = _lib.errbuf()
    _lib.check(_lib.lib().aero_mesh_write(C.byref(arr), str(path).encode(), err, len(err)), err, "aero_mesh_write")


def load_mesh(path) -> Mesh:
    """Read a binary mesh file (aero_mesh_read); the file's indices and ranges are validated."""
    out = _lib.MeshArrays()
    lib = _lib.lib()
    err = _lib.errbuf()
    _lib.check(lib.aero_mesh_read(str(path).encode(), C.byref(out), err, len(err)), err, "aero_mesh_read")
    try:
        return _from_arrays(out, dict(file=str(path)))
    finally:
        lib.aero_mesh_free(C.byref(out))
